"""The AXT autotuner: model ranking (C host library) + device timing.

For one matrix and k right-hand sides, ``tune_axt`` ranks the candidate tile
heights and right-hand-side blockings kc by the bytes they move
(``intspmv_tune_axt``: exact padded lengths, no arrays built).  The best
``top`` of them are then crossed with the wave mapping -- how many AXT units
one wave's block takes (``umaxes``; rows above half of it get waves of their
own) -- and every combination is timed on the device with the batched kernel
itself.  So that the timing reflects throughput inside a batch and not the
latency of one small launch, a small matrix is replicated until the timed
launch holds at least ``min_items`` work items.  The fastest combination is
returned; the batched launch takes one such choice per matrix.
"""
from __future__ import annotations

import numpy as np

from . import device as D
from . import host as H

_BYTES = {D.F64: 8, D.F32: 4, D.BF16: 2}
UMAXES = (32, 64, 128)


def tune_axt(eng: D.Engine, m: H.CSR, k: int, val_type: int = D.F32, x_type: int = D.F32,
             top: int = 2, reps: int = 5, umaxes=UMAXES, min_items: int = 2048,
             max_copies: int = 64) -> dict:
    import torch
    sx = 8 if x_type == D.F64 else 4
    ranked = H.tune_model(m, k, _BYTES[val_type], sx)
    n = m.nrows
    xdt = torch.float64 if x_type == D.F64 else torch.float32
    ydt = torch.float64 if val_type == D.F64 else torch.float32
    X = torch.rand(n * k, dtype=xdt, device=eng.torch_device)
    best, built = None, {}
    for c in ranked[:top]:
        th = c["tileH"]
        if th not in built:
            a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
            built[th] = (a, D.upload_batch_arrays(eng, a, val_type))
        a, dev = built[th]
        Y = torch.empty(n * k, dtype=ydt, device=eng.torch_device)
        for umax in umaxes:
            ent = dict(a=a, dev=dev, X=X, Y=Y, k=k, kc=c["kc"], umax=umax)
            one = D.DeviceBatch(eng, [ent], val_type, x_type)
            copies = int(min(max_copies, max(1, -(-min_items // max(1, one.info()["nitems"])))))
            if copies > 1:
                # copies read the same arrays and write their own y
                del one
                ents = [dict(ent, Y=torch.empty(n * k, dtype=ydt, device=eng.torch_device))
                        for _ in range(copies)]
                b = D.DeviceBatch(eng, ents, val_type, x_type)
            else:
                b = one
            for _ in range(2):
                b.launch()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                b.launch()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / reps / copies
            del b
            if best is None or us < best["us"]:
                best = dict(tileH=th, kc=c["kc"], umax=umax, us=us, model_bytes=c["bytes"])
    best["ranked"] = ranked[:top]
    return best
