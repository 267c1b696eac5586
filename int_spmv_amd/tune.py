"""The AXT autotuner: model ranking (C host library) + device timing.

For one matrix and k right-hand sides, ``tune_axt`` ranks the candidate tile
heights and right-hand-side blockings by the bytes they move
(``intspmv_tune_axt``: exact padded lengths, no arrays built), then times the
best ``top`` candidates with the column-map kernel on the device and returns
the fastest.  The batched launch takes one such choice per matrix.
"""
from __future__ import annotations

import numpy as np

from . import device as D
from . import host as H

_BYTES = {D.F64: 8, D.F32: 4, D.BF16: 2}


def tune_axt(eng: D.Engine, m: H.CSR, k: int, val_type: int = D.F32, x_type: int = D.F32,
             top: int = 3, reps: int = 5) -> dict:
    import torch
    sx = 8 if x_type == D.F64 else 4
    ranked = H.tune_model(m, k, _BYTES[val_type], sx)
    n = m.nrows
    dt = torch.float64 if x_type == D.F64 else torch.float32
    X = torch.rand(n * k, dtype=dt, device=eng.torch_device)
    Y = torch.empty(n * k, dtype=dt if val_type == D.F64 else torch.float32,
                    device=eng.torch_device)
    best, built = None, {}
    for c in ranked[:top]:
        th = c["tileH"]
        if th not in built:
            a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
            built[th] = (a, D.DeviceAXTGather(eng, a, val_type, x_type))
        g = built[th][1]
        for _ in range(2):
            g.matvec(X, Y, k=k, kc=c["kc"])
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            g.matvec(X, Y, k=k, kc=c["kc"])
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        if best is None or us < best["us"]:
            best = dict(tileH=th, kc=c["kc"], us=us, model_bytes=c["bytes"])
    best["ranked"] = ranked[:top]
    return best
