#!/usr/bin/env python3
"""SpMV HBM GB/s (% of 8 TB/s) + GFLOP/s on one MI355X: AXT vs CSR, FP64 vs FP32/BF16.

One "step" is one SpMV y = A x over the whole matrix with A and x resident in
HBM.  The headline `value` is AXT NOC tileHeight 1 in FP64 (the drop-in for
the reference's int_axt_noc_h1, intSpmv_main.c:1385) on the power-law
workload: n = 2^20, ~30M entries, Zipf(2) row lengths up to 8192.  GB/s are
algorithmic bytes (int_spmv_amd/metrics.py) per step time.  The other formats,
tile heights and storage precisions are timed beside it and reported under
"formats"; every timed product is then checked against the FP64 oracle
(per-row scaled error, as the reference's getErrors after each kernel), and
the 256-matrix batch against cf_CSR of its FP32-rounded inputs.  The CPU
baseline is the AVX-512 + OpenMP restatement of the reference's kernels
(oracle/cpu_avx512.c) on the host's cores; the reference itself needs
<mkl.h> to build.

Usage: python bench.py [--steps K] [--warmup W] [--workload NAME] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="powerlaw_1m")
    p.add_argument("--quick", action="store_true", help="headline kernel only (profiling)")
    p.add_argument("--batch-only", action="store_true", help="the batched launch only (profiling)")
    p.add_argument("--gather-only", type=int, default=0, metavar="K",
                   help="only the FP64 column-map product with K right-hand sides on the workload's "
                        "matrix, tile height 1 (profiling)")
    p.add_argument("--placement", choices=("tuned", "plain", "xcd"), default="tuned",
                   help="--gather-only: the column-map plan's placement (tuned: the faster)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-precision", action="store_true",
                   help="skip the stencil / R-MAT precision runs (FP64, FP32, BF16 storage)")
    p.add_argument("--batch-count", type=int, default=256,
                   help="matrices in the batched-launch measurement (0 = skip)")
    p.add_argument("--batch-tune", choices=("device", "model"), default="model",
                   help="batch autotuner: byte model only (default; 128 units per wave), or "
                        "byte model + device timing of each matrix alone (tile height, RHS "
                        "blocking and units per wave)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only run a plain device copy of this many MiB (counter calibration)")
    return p.parse_args()


def cores() -> int:
    """Threads the CPU baseline runs on: OMP_NUM_THREADS, else the cores this
    process may use (its affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    return len(os.sched_getaffinity(0))


def physical_cores():
    """Physical cores of the host (distinct core/socket pairs in lscpu -p),
    reported beside the threads the baseline uses."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        pairs = {line.strip() for line in out.splitlines() if line and not line.startswith("#")}
        return len(pairs) or None
    except (OSError, subprocess.SubprocessError):
        return None


def timed(fn, steps, warmup, torch):
    """Returns (wall seconds per step, event seconds per step).  The wall
    clock brackets exactly `steps` launches between two synchronisations.
    The HIP events bracket launches 2..steps on the launch stream: the start
    event is queued behind launch 1, so it fires when the GPU is already busy
    and the events see the kernels back to back, not the host's latency to
    issue the first launch onto an idle GPU."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    if steps < 2:
        e0.record()
    fn()
    if steps >= 2:
        e0.record()
    for _ in range(steps - 1):
        fn()
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, e0.elapsed_time(e1) * 1e-3 / max(steps - 1, 1)


def timed_each(fn, steps, warmup, torch, between=None):
    """Per-launch HIP-event times (seconds), each launch bracketed by its own
    events on the current stream; `between` (untimed) runs before each launch."""
    for _ in range(warmup):
        fn()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    for a, b in evs:
        if between is not None:
            between()
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return sorted(a.elapsed_time(b) * 1e-3 for a, b in evs)


def bench_batch(eng, count, torch, log, tune="model"):
    """256 matrices (n log-uniform in [1e3, 2e5], 4..64 entries per row, 70 %
    banded), FP32 values and x, k in {1,4,8,16,32} each; per-matrix tile
    height, right-hand-side blocking and units per wave from the autotuner
    (tune="device": byte model + device timing; "model": byte model only, 128
    units per wave); one launch (two concurrent persistent kernels) captured in a hipGraph and
    replayed."""
    import collections

    from int_spmv_amd.tune import tune_axt
    import numpy as np

    from int_spmv_amd import device as D
    from int_spmv_amd import host as H
    from int_spmv_amd import metrics as M
    from int_spmv_amd.workloads import batch_256
    t0 = time.perf_counter()
    mats = batch_256(seed=0, count=count)
    gen_s = time.perf_counter() - t0
    ents, flops, bytes_, rows, nnz = [], 0.0, 0, 0, 0
    chosen = collections.Counter()
    t1 = time.perf_counter()
    for i, (m, k) in enumerate(mats):
        if tune == "device":
            c = tune_axt(eng, m, k, D.F32, D.F32, top=2, reps=3)
        else:
            c = dict(H.tune_model(m, k, 4, 4)[0], umax=128)
        chosen[(c["tileH"], c["kc"], c["umax"])] += 1
        a = H.AXT(m, np.zeros(m.nrows), th=c["tileH"], want_acol=True)
        X = torch.rand(m.nrows * k, dtype=torch.float32, device=eng.torch_device) * 2 - 1
        Y = torch.full((m.nrows * k,), float("nan"), dtype=torch.float32,
                       device=eng.torch_device)
        ents.append(dict(a=a, X=X, Y=Y, k=k, kc=c["kc"], umax=c["umax"]))
        flops += M.flops(m.nnz, k)
        bytes_ += M.bytes_axt_gather(m.nrows, a.lenAX, a.lenSEC, k=k, sv=4, sx=4, sy=4)
        rows += m.nrows
        nnz += m.nnz
    b = D.DeviceBatch(eng, ents, D.F32, D.F32)
    for e in ents:
        e.pop("a")
    info = b.info()
    tune_s = time.perf_counter() - t1
    b.capture()
    for _ in range(3):
        b.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        b.replay()
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    r = {"matrices": count, "rows": rows, "nnz": nnz, "us": round(t * 1e6, 1),
         "GFLOPs": round(flops / t / 1e9, 1), "GBs": round(bytes_ / t / 1e9, 1),
         "pct_hbm": round(100 * bytes_ / t / 1e9 / M.HBM_PEAK_GBS, 1), "items": info["nitems"],
         "grid": info["grid"], "generate_s": round(gen_s, 1), "tune_build_s": round(tune_s, 1),
         "tuner": tune,
         "choices": {f"th{t}_kc{kc}_u{u}": c for (t, kc, u), c in sorted(chosen.items())}}
    log(f"[bench] BATCH: {r}")
    r["_check"] = [(m, e["X"], e["Y"], e["k"]) for (m, _), e in zip(mats, ents)]
    return r


# The precision workloads: a 27-point stencil and an R-MAT graph, values
# stored in FP64, FP32 or BF16 (x FP32 beside the last two).
PRECISION_WORKLOADS = ("stencil27_128", "rmat_22")


def bench_precision(eng, torch, log, names=PRECISION_WORKLOADS):
    """The embedded AXT product at tile height 1 on each workload with values
    stored in FP64 (the drop-in), FP32 (x FP32) and BF16 (x FP32), FP32 sums
    (compensated on long rows) for the last two.  Returns the timings and the
    outputs to check against the FP64 oracle of the unrounded matrix."""
    import numpy as np

    from int_spmv_amd import device as D
    from int_spmv_amd import host as H
    from int_spmv_amd import metrics as M
    from int_spmv_amd.workloads import WORKLOADS
    out, pending = {}, []
    for name in names:
        t0 = time.perf_counter()
        m, x = WORKLOADS[name]()
        a = H.AXT(m, x, th=1)
        n = m.nrows
        fb = M.bytes_axt(n, a.lenAX, a.lenSEC)
        r, ys, fe = {}, {}, None
        for vt, xt, prec, sv, sx, sy, dt, tol in (
                (D.F64, D.F64, "F64", 8, 8, 8, torch.float64, 1e-12),
                (D.F32, D.F32, "F32", 4, 4, 4, torch.float32, 1e-6),
                (D.BF16, D.F32, "BF16", 2, 4, 4, torch.float32, 1e-2)):
            d = D.DeviceAXT(eng, a, val_type=vt, x_type=xt)
            y = torch.full((max(n, 1),), float("nan"), dtype=dt, device=eng.torch_device)
            w, e = timed(lambda: d.int_axt(y), 20, 3, torch)
            fe = e if prec == "F64" else fe
            b = M.bytes_axt(n, a.lenAX, a.lenSEC, sv=sv, sx=sx, sy=sy)
            r[prec] = {"us": round(e * 1e6, 2), "GBs": round(b / e / 1e9, 1),
                       "pct_hbm": round(100 * b / e / 1e9 / M.HBM_PEAK_GBS, 1),
                       "GBs_eq_fp64": round(fb / e / 1e9, 1), "vs_fp64": round(fe / e, 2),
                       "tol": tol}
            torch.cuda.synchronize()
            ys[prec] = y[:n].double().cpu().numpy()
            del d, y
        out[name] = {"n": n, "nnz": m.nnz, "tileHeight": 1, "x": "FP64 for F64, FP32 else",
                     "formats": r, "setup_s": round(time.perf_counter() - t0, 1)}
        log(f"[bench] PRECISION {name}: {r}")
        pending.append((name, m, np.asarray(x), ys))
        del a
    return out, pending


def bench_reference_case(eng, torch, log):
    """The reference-sized case: the 2-D Laplacian on 256 x 256 (n = 65,536),
    x ~ U[-1, 1] seed 0, through the ordered CSR product and the embedded AXT
    product at tile heights 1 and 4 (FP64).  Returns the timings and the
    outputs for the checker: CSR bit for bit against cf_CSR, AXT bit for bit
    against the CPU restatement of the reference kernel and within 1e-12 of
    cf_CSR.  Correctness only: the matrix (~9 MB) stays in cache."""
    from int_spmv_amd import device as D
    from int_spmv_amd import host as H
    from int_spmv_amd.workloads import laplace_256
    m, x = laplace_256()
    n = m.nrows
    xd = eng.upload(x)
    out, ys, hosts = {"n": n, "nnz": m.nnz}, {}, {}
    dc = D.DeviceCSR(eng, m)
    y = torch.full((n,), float("nan"), dtype=torch.float64, device=eng.torch_device)
    _, e = timed(lambda: dc.cf_csr(xd, y), 10, 2, torch)
    out["CSR_ORDERED_F64"] = {"us": round(e * 1e6, 2)}
    torch.cuda.synchronize()
    ys["CSR_ORDERED_F64"] = y.cpu().numpy()
    for th in (1, 4):
        a = H.AXT(m, x, th=th)
        d = D.DeviceAXT(eng, a)
        y = torch.full((n,), float("nan"), dtype=torch.float64, device=eng.torch_device)
        _, e = timed(lambda: d.int_axt(y), 10, 2, torch)
        out[f"AXT_NOC_H{th}_F64"] = {"us": round(e * 1e6, 2)}
        torch.cuda.synchronize()
        ys[f"AXT_NOC_H{th}_F64"] = y.cpu().numpy()
        hosts[f"AXT_NOC_H{th}_F64"] = a
        del d
    log(f"[bench] REFERENCE CASE laplace_256: {out}")
    return out, (m, x, ys, hosts)


def check_reference_case(out, pending):
    """cf_CSR's bits for the ordered CSR product; the reference AXT kernel's
    bits (its CPU restatement) and 1e-12 of cf_CSR for the AXT products.
    Returns the names that fail."""
    import numpy as np

    import oracle as O
    m, x, ys, hosts = pending
    yref = O.cf_csr(m.row, m.col, m.val, x)
    bad = []
    for name, y in ys.items():
        f = out[name]
        f["max_scaled_err_vs_fp64_oracle"] = O.scaled_err(m.row, m.col, m.val, x, yref, y)[0]
        want = yref if name.startswith("CSR") else O.int_axt(hosts[name])
        f["bitwise_vs_" + ("cf_csr" if name.startswith("CSR") else "reference_kernel")] = bool(
            np.array_equal(y, want))
        if not (np.array_equal(y, want) and f["max_scaled_err_vs_fp64_oracle"] <= 1e-12):
            bad.append(f"laplace_256:{name}")
    return bad


def bench_differentiation(eng, m, torch, log):
    """L(v, x) = y_bar^T A(v) x on the benchmark matrix through the autograd
    operator (column map, tile height 1), FP64 and FP32: the largest relative
    gap between central differences along 64 random directions (h = 1e-6 /
    1e-3) and the directional derivative from the backward's (v_bar, x_bar);
    and the times of the forward, of x_bar = A^T y_bar and of the whole
    backward (x_bar reading the current values through the slot map, plus
    v_bar), each against the forward.  L is bilinear, so the differences are
    exact up to rounding; they are taken along the perturbations the inputs
    actually carry after rounding to the storage type.  The times are also
    given for the pair operator, which keeps A's values in A^T's slot order
    too (both copies updated alike), so x_bar reads its own copy: x_bar
    alone, and x_bar plus v_bar in both orders (its backward).  Each plan's
    placement is tuned first."""
    from int_spmv_amd import autograd as AG
    from int_spmv_amd import device as D
    out = {}
    n = m.nrows
    for vt, h, tol, name in ((D.F64, 1e-6, 1e-6, "F64"), (D.F32, 1e-3, 1e-4, "F32")):
        op = AG.AxtOperator(eng, m, th=1, val_type=vt)
        placement = op.tune_placement()
        dev, dt = eng.torch_device, op.dtype
        gen = torch.Generator(device=dev)
        gen.manual_seed(21)
        v = op.initial_values().requires_grad_(True)
        x = (torch.rand(n, device=dev, dtype=torch.float64, generator=gen) * 2 - 1).to(dt)
        x.requires_grad_(True)
        ybar = torch.randn(n, device=dev, dtype=torch.float64, generator=gen).to(dt)
        (AG.spmv(v, x, op) * ybar).sum().backward()
        gv, gx = v.grad.detach().double(), x.grad.detach().double()
        occupied = (op.fwd.acol.view(torch.int32) != -1).to(torch.float64)
        worst = 0.0
        with torch.no_grad():
            yb64 = ybar.double()
            for _ in range(64):
                dv = torch.randn(len(occupied), device=dev, dtype=torch.float64,
                                 generator=gen) * occupied
                dx = torch.randn(n, device=dev, dtype=torch.float64, generator=gen)
                vp, vm = (v.double() + h * dv).to(dt), (v.double() - h * dv).to(dt)
                xp, xm = (x.double() + h * dx).to(dt), (x.double() - h * dx).to(dt)
                dv_e = (vp.double() - vm.double()) / (2 * h)
                dx_e = (xp.double() - xm.double()) / (2 * h)
                lp = (AG.spmv(vp, xp, op).double() * yb64).sum()
                lm = (AG.spmv(vm, xm, op).double() * yb64).sum()
                fd = ((lp - lm) / (2 * h)).item()
                an = ((gv * dv_e).sum() + (gx * dx_e).sum()).item()
                scale = ((gv * dv_e).abs().sum() + (gx * dx_e).abs().sum()).item()
                worst = max(worst, abs(fd - an) / scale)
            vd, xd = v.detach(), x.detach()
            y = torch.empty(n, dtype=dt, device=dev)
            gxb, gvb = torch.empty_like(xd), torch.empty_like(vd)
            _, fwd = timed(lambda: op.fwd.matvec(xd, y, vals=vd), 20, 3, torch)
            _, adj = timed(lambda: op.adj.adjoint(ybar, gxb, vals=vd), 20, 3, torch)

            def backward():
                op.adj.adjoint(ybar, gxb, vals=vd)
                op.fwd.grad_vals(xd, ybar, gvb)
            _, bwd = timed(backward, 20, 3, torch)
            # the pair operator: A's values also in A^T's slot order
            _, vt_pair = op.initial_values_pair()
            gvt = torch.empty_like(vt_pair)
            _, adj_pair = timed(lambda: op.adj.gT.matvec(ybar, gxb, vals=vt_pair), 20, 3, torch)

            def backward_pair():
                op.fwd.grad_vals(xd, ybar, gvb)
                op.adj.adjoint_grad(ybar, gxb, xd, gvt, vals_t=vt_pair)
            _, bwd_pair = timed(backward_pair, 20, 3, torch)
        out[name] = {"fd_directions": 64, "h": h, "fd_worst_rel_err": worst, "tol": tol,
                     "placement": placement, "forward_us": round(fwd * 1e6, 2),
                     "adjoint_x_us": round(adj_pair * 1e6, 2),
                     "adjoint_vs_forward": round(adj_pair / fwd, 3),
                     "backward_pair_us": round(bwd_pair * 1e6, 2),
                     "backward_pair_vs_forward": round(bwd_pair / fwd, 3),
                     "adjoint_x_slot_map_us": round(adj * 1e6, 2),
                     "adjoint_slot_map_vs_forward": round(adj / fwd, 3),
                     "backward_slot_map_us": round(bwd * 1e6, 2),
                     "backward_slot_map_vs_forward": round(bwd / fwd, 3)}
        log(f"[bench] DIFFERENTIATION {name}: {out[name]}")
        del op, v, x, ybar, gv, gx
        torch.cuda.empty_cache()
    return out


def check_precision(out, pending):
    """Max per-row scaled error of every precision result against the FP64
    oracle (cf_CSR) of the unrounded matrix; returns the names over their
    tolerance."""
    import oracle as O
    over = []
    for name, m, x, ys in pending:
        yref = O.cf_csr(m.row, m.col, m.val, x)
        for prec, y in ys.items():
            e = O.scaled_err(m.row, m.col, m.val, x, yref, y)[0]
            f = out[name]["formats"][prec]
            f["max_scaled_err_vs_fp64_oracle"] = e
            if not e <= f["tol"]:
                over.append(f"{name}:{prec}")
    return over


def check_batch(items):
    """Every result column of every batch matrix against cf_CSR of its
    FP32-rounded values and x (the accumulation error of FP32 sums)."""
    import numpy as np

    import oracle as O
    worst, where, cols = 0.0, None, 0
    for i, (m, X, Y, k) in enumerate(items):
        v32 = m.val.astype(np.float32).astype(np.float64)
        Xh = X.double().cpu().numpy().reshape(m.nrows, k)
        Yr = O.csr_mm(m.row, m.col, v32, Xh)
        e, pos = O.scaled_err_mm(m.row, m.col, v32, Xh, Yr,
                                 Y.double().cpu().numpy().reshape(m.nrows, k))
        cols += k
        if not e <= worst:
            worst, where = e, [i, int(pos[0]), int(pos[1])]
    log(f"[bench] batch check: {len(items)} matrices, {cols} columns, max scaled err {worst:.2e}")
    return {"max_scaled_err_vs_fp32_oracle": worst, "worst_at_matrix_row_col": where,
            "checked_matrices": len(items), "checked_columns": cols}


def rocsparse_leg(dc, xd, y_csr, n, nnz):
    """rocSPARSE csrmv (vendor comparator, run and timed only) on the same
    device CSR arrays as spmv_csr; each algorithm's result is compared with
    the ordered product already in y_csr (max |diff| / max |y|, as getErrors)."""
    import torch

    from int_spmv_amd import comparators as CMP
    if not CMP.rocsparse_available():
        return None
    ref = y_csr[:n].clone()
    yr = torch.empty_like(ref)
    out = {}
    for name, alg in CMP.ROCSPARSE_CSR_ALGS.items():
        yr.zero_()  # rocSPARSE forms beta*y even for beta = 0 in some algorithms
        r = CMP.rocsparse_csr(dc.row, dc.col, dc.val, xd[:n], yr, n, nnz, torch,
                              algs={name: alg})[name]
        if isinstance(r, float):
            rel = float((yr - ref).abs().max() / ref.abs().max().clamp_min(1e-300))
            out[name] = (r, rel)
        else:
            out[name] = (None, r)
    return out


def check_results(m, x, checks, formats):
    """Every timed product against the FP64 oracle, as the reference's
    getErrors after each kernel (intSpmv_main.c:1507-1510): per-row scaled
    error max |y - y_ref| / sum |a x| (all right-hand sides; the adjoint
    against the sequential transpose product; the value gradient bit for
    bit), held to 1e-12 (FP64), 1e-6 (FP32 storage) or 1e-2 (BF16).
    Returns the names over their tolerance."""
    import numpy as np

    import oracle as O
    n = m.nrows
    yref = O.cf_csr(m.row, m.col, m.val, x)
    at_csr = None
    over = []
    for name, c in checks.items():
        tol, kind, y = c.get("tol", 1e-12), c["kind"], c["y"]
        f = formats[name]
        if kind == "y":
            e = O.scaled_err(m.row, m.col, m.val, x, yref, y)[0]
            if c.get("bitwise"):
                f["bitwise_vs_cf_csr"] = bool(np.array_equal(y, yref))
                if not f["bitwise_vs_cf_csr"]:
                    over.append(name)
        elif kind == "mm":
            X = c["X"]
            Yr = O.csr_mm(m.row, m.col, m.val, X)
            e = O.scaled_err_mm(m.row, m.col, m.val, X, Yr, y.reshape(n, X.shape[1]))[0]
        elif kind == "adj":
            if at_csr is None:
                at_csr, _ = m.transpose()
            ref = O.csr_transpose_product(m.row, m.col, m.val, c["ybar"])
            e = O.scaled_err(at_csr.row, at_csr.col, at_csr.val, c["ybar"], ref, y)[0]
        else:  # value gradient: ybar[row] * x[col] on the slots, bit for bit
            rows = np.repeat(np.arange(n), np.diff(m.row.astype(np.int64)))
            want = np.zeros(len(y))
            want[c["nzslot"]] = c["ybar"][rows] * x[m.col]
            e = 0.0 if np.array_equal(y, want) else float("inf")
        f["max_scaled_err_vs_fp64_oracle"] = e
        f["tol"] = tol
        if not e <= tol:
            over.append(name)
    if "AXT_NOC_H1_F64" in checks:  # the reference's own figures for the headline
        a_err, r_err, pos = O.get_errors(yref, checks["AXT_NOC_H1_F64"]["y"])
        formats["AXT_NOC_H1_F64"]["getErrors"] = {"aErr": a_err, "rErr": r_err, "rowInd": pos}
    return over


def cpu_leg(m, x, a1, hosts, head_bytes, workload, quick):
    """The CPU baseline: the oracle's OpenMP restatements of the reference
    kernels (and MKL's CSR kernels when libmkl_rt is present) timed on this
    host's cores."""
    import ctypes as C

    import numpy as np

    import oracle as O
    from int_spmv_amd import comparators as CMP
    from int_spmv_amd import metrics as M
    n, nnz = m.nrows, m.nnz
    yref = O.cf_csr(m.row, m.col, m.val, x)

    def serr(y):
        return O.scaled_err(m.row, m.col, m.val, x, yref, y)[0]
    threads = cores()
    L = O.lib_avx512()
    if L is None:  # no scalar stand-in for the reference's AVX-512 kernels
        return {"value": None, "unit": "GB/s", "cores": threads, "kind": "port",
                "sample": "n/a: the host CPU has no AVX-512 (the reference's kernels are "
                          "AVX-512 intrinsics)", "host_physical_cores": physical_cores()}

    def run(kind, thr, budget, len_, th, arrs):
        keep = [np.ascontiguousarray(a) for a in arrs]
        ptrs = [C.c_void_p(a.ctypes.data) for a in keep] + [None] * (6 - len(keep))
        y = np.zeros(n)
        t1 = L.orc_bench_avx512(kind, thr, 1, n, len_, th, *ptrs, C.c_void_p(y.ctypes.data))
        if t1 < 0:
            raise RuntimeError(f"orc_bench_avx512 kind {kind}")
        iters = int(max(1, min(250, budget / max(t1, 1e-6))))
        t = L.orc_bench_avx512(kind, thr, iters, n, len_, th, *ptrs, C.c_void_p(y.ctypes.data))
        return t, iters, y

    t, iters, ycpu = run(4, threads, 10.0, a1.tileN, 1, [a1.ax, a1.sec])
    cpu = {"value": round(head_bytes / t / 1e9, 2), "unit": "GB/s", "cores": threads,
           "kind": "port",
           "sample": f"int_axt_noc_h1 AVX-512 + OpenMP restatement (oracle/cpu_avx512.c: "
                     f"dynamic schedule, atomic y), full {workload} matrix, {iters} calls, "
                     f"y zeroed untimed; {t * 1e3:.2f} ms/call",
           "host_physical_cores": physical_cores(),
           "max_scaled_err_vs_csr": serr(ycpu)}
    log(f"[bench] cpu baseline: {cpu}")
    if quick:
        return cpu
    others = {}

    def rec(name, t, calls, thr, bytes_, y, kind="port"):
        others[name] = {"ms": round(t * 1e3, 3), "GBs": round(bytes_ / t / 1e9, 2),
                        "GFLOPs": round(M.flops(nnz) / t / 1e9, 2), "calls": calls,
                        "threads": thr, "kind": kind, "max_scaled_err_vs_csr": serr(y)}
        log(f"[bench] cpu {name}: {others[name]}")
    t, it, y = run(0, 1, 2.0, 0, 0, [m.row, m.col, m.val, x])
    rec("cf_CSR", t, it, 1, M.bytes_csr(n, nnz), y)
    if "AXC" in hosts:
        c = hosts["AXC"]
        t, it, y = run(1, threads, 2.0, 0, 0, [c.ax, c.brp])
        rec("int_axc", t, it, threads, M.bytes_axc(n, c.lenAX), y)
    if "K1" in hosts:
        k = hosts["K1"]
        arrs = [k.val, k.col, k.chp, k.nmc, x, k.permi]
        t, it, y = run(2, threads, 2.0, k.chunkNum, 0, arrs)
        rec("int_k1 (+orderArrayK1 untimed)", t, it, threads, M.bytes_k1(n, k.lenVC, k.chunkNum), y)
        t, it, y = run(3, threads, 2.0, k.chunkNum, 0, arrs)
        rec("int_k1p", t, it, threads, M.bytes_k1(n, k.lenVC, k.chunkNum), y)
    if "AXT_H4" in hosts:
        a4 = hosts["AXT_H4"]
        t, it, y = run(5, threads, 2.0, a4.tileN, 4, [a4.ax, a4.sec])
        rec("int_axt_noc th4", t, it, threads, M.bytes_axt(n, a4.lenAX, a4.lenSEC), y)
    # the rest of the reference's tile-height sweep, and its COM kernel at
    # three block sizes (the reference's COM defects included: its errors
    # against CSR are reported, not hidden)
    from int_spmv_amd import host as H
    for th in (2, 8, 16):
        a = H.AXT(m, x, th=th)
        t, it, y = run(5, threads, 1.5, a.tileN, th, [a.ax, a.sec])
        rec(f"int_axt_noc th{th}", t, it, threads, M.bytes_axt(n, a.lenAX, a.lenSEC), y)
        del a
    for bs in (128, 256, 512):
        a = H.AXT(m, x, th=1, mode="COM", bs=bs)
        bn = (a.tileN * a.tileHW + bs - 1) // bs
        prm = np.array([bs, a.log, a.tileHW, a.lenAX, a.lenSEC], np.uint32)
        t, it, y = run(6, threads, 1.5, bn, 1, [a.ax, a.sec, prm])
        rec(f"int_axt_com_h1 bs{bs} (reference defects kept)", t, it, threads,
            M.bytes_axt(n, a.lenAX, a.lenSEC), y)
        del a
    if CMP.mkl_available():
        for thr in sorted({1, threads}):
            res = CMP.mkl_csr(m.row, m.col, m.val, x, thr, budget_s=1.5)
            for nm, r in (res or {}).items():
                rec(f"{nm} ({thr} thr)", r["ms"] * 1e-3, r["calls"], thr, M.bytes_csr(n, nnz),
                    r["y"], kind="vendor")
    else:
        others["mkl"] = "libmkl_rt not found"
    cpu["formats"] = others
    return cpu


def formats_checked(formats):
    return [k for k, v in formats.items()
            if isinstance(v, dict) and "max_scaled_err_vs_fp64_oracle" in v]


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures the SpMV path on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is designed but not built).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("no HIP GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        w, e = timed(lambda: dst.copy_(src), args.steps, args.warmup, torch)
        # a read-only stream (torch's sum) for the read-dominated kernels' ceiling
        w2, e2 = timed(lambda: torch.sum(src), args.steps, args.warmup, torch)
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                          "GBs_rw": 2 * n * 8 / e / 1e9, "sum_GBs_read": n * 8 / e2 / 1e9}))
        return

    from int_spmv_amd import device as D
    from int_spmv_amd import host as H
    from int_spmv_amd import metrics as M
    from int_spmv_amd.workloads import WORKLOADS
    if args.batch_only:  # the batched launch alone (profiling)
        eng = D.Engine(0)
        r = bench_batch(eng, args.batch_count, torch, log, args.batch_tune)
        r.pop("_check")
        print(json.dumps({"batch": r}), flush=True)
        return

    t = time.perf_counter()
    m, x = WORKLOADS[args.workload]()
    n, nnz = m.nrows, m.nnz
    log(f"[bench] workload {args.workload}: n={n} nnz={nnz} rmax={m.stats()['rmax']} "
        f"({time.perf_counter() - t:.1f}s)")
    eng = D.Engine(0)
    if args.gather_only:  # the column-map product alone (profiling)
        k = args.gather_only
        ag = H.AXT(m, x, th=1, want_acol=True)
        g = D.DeviceAXTGather(eng, ag)
        if args.placement == "tuned":
            g.tune_placement(k=k)
        else:
            g.set_placement(args.placement)
        X = torch.rand(n * k, dtype=torch.float64, device=eng.torch_device)
        Y = eng.empty(n * k)
        w, e = timed(lambda: g.matvec(X, Y, k=k), args.steps, args.warmup, torch)
        b = M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC, k=k)
        print(json.dumps({"gather": {"workload": args.workload, "k": k, "us": round(e * 1e6, 2),
                                     "placement": g.placement, "bytes": b,
                                     "GBs": round(b / e / 1e9, 1)}}), flush=True)
        return
    xd = eng.upload(x)
    formats = {}
    # results kept for the oracle checks of the CPU leg: name -> (kind, data)
    checks = {}

    def nanv(count, dt=np.float64):
        """An output vector filled with NaN, so an unwritten entry fails its check."""
        t = eng.empty(count, dt)
        t.fill_(float("nan"))
        return t

    def keep(name, kind, t, count, **kw):
        """Host copy of a product's output (taken after its timing)."""
        torch.cuda.synchronize()
        checks[name] = dict(kind=kind, y=t.reshape(-1)[:count].double().cpu().numpy(), **kw)

    def record(name, bytes_, wall, ev, extra=None):
        r = {"us": round(ev * 1e6, 2), "GBs": round(bytes_ / ev / 1e9, 1),
             "pct_hbm": round(100 * bytes_ / ev / 1e9 / M.HBM_PEAK_GBS, 1),
             "GFLOPs": round(M.flops(nnz) / ev / 1e9, 1), "bytes": int(bytes_)}
        if extra:
            r.update(extra)
        formats[name] = r
        log(f"[bench] {name}: {r}")
        return r

    # ---- headline: AXT NOC th=1, FP64, reference layout ----------------------
    a1 = H.AXT(m, x, th=1)
    d1 = D.DeviceAXT(eng, a1)
    y1 = nanv(n)
    head_bytes = M.bytes_axt(n, a1.lenAX, a1.lenSEC)
    wall, ev = timed(lambda: d1.int_axt(y1), args.steps, args.warmup, torch)
    record("AXT_NOC_H1_F64", head_bytes, wall, ev,
           {"beta": round(a1.fd.beta, 4), "plan": d1.plan.info()})
    keep("AXT_NOC_H1_F64", "y", y1, n)
    head_wall, head_ev = wall, ev
    # what a plain read of as many bytes reaches on this device, back to back
    # like the timed launches: the streaming ceiling beside the 8 TB/s peak
    stream_bps = eng.stream_rate(int(head_bytes), 20)
    hosts = {}
    if args.quick:  # headline launches only (profiling)
        warm = cold = None
    else:
        # per-launch events: median warm, and cold (a 1 GiB buffer written between
        # launches evicts the matrix from the 256 MB MALL and the L2s)
        warm = timed_each(lambda: d1.int_axt(y1), args.steps, 1, torch)
        scratch = torch.empty(1 << 27, dtype=torch.float64, device=eng.torch_device)
        cold = timed_each(lambda: d1.int_axt(y1), args.steps, 1, torch,
                          between=lambda: scratch.fill_(1.0))
        del scratch
        med, cmed = warm[len(warm) // 2], cold[len(cold) // 2]
        formats["AXT_NOC_H1_F64"].update({
            "us_median": round(med * 1e6, 2), "GBs_median": round(head_bytes / med / 1e9, 1),
            "us_cold_median": round(cmed * 1e6, 2), "GBs_cold": round(head_bytes / cmed / 1e9, 1)})
        log(f"[bench] AXT_NOC_H1_F64 per-launch median {med * 1e6:.2f} us, cold {cmed * 1e6:.2f} us")

    if not args.quick:
        # the reference's sweep (tile heights 1, 4, 8, 16, 32, 48 and COM
        # with blocks of 128/256/512, intSpmv_main.c:1668-1676) plus th = 2
        for th in (2, 4, 8, 16, 32, 48):
            a = H.AXT(m, x, th=th)
            d = D.DeviceAXT(eng, a)
            y = nanv(n)
            w, e = timed(lambda: d.int_axt(y), 20, 3, torch)
            record(f"AXT_NOC_H{th}_F64", M.bytes_axt(n, a.lenAX, a.lenSEC), w, e,
                   {"beta": round(a.fd.beta, 4)})
            keep(f"AXT_NOC_H{th}_F64", "y", y, n)
            if th == 4:
                hosts["AXT_H4"] = a
                h4_ev = e
            del d, a, y
        for bs in (128, 256, 512):
            a = H.AXT(m, x, th=1, mode="COM", bs=bs)
            d = D.DeviceAXT(eng, a)
            y = nanv(n)
            w, e = timed(lambda: d.int_axt(y), 20, 3, torch)
            record(f"AXT_COM_BS{bs}_F64", M.bytes_axt(n, a.lenAX, a.lenSEC), w, e,
                   {"beta": round(a.fd.beta, 4)})
            keep(f"AXT_COM_BS{bs}_F64", "y", y, n)
            del d, a, y
        a4 = hosts["AXT_H4"]
        h4_bytes = M.bytes_axt(n, a4.lenAX, a4.lenSEC)
        for ah, th, fb, fe in ((a1, 1, head_bytes, head_ev), (a4, 4, h4_bytes, h4_ev)):
            for vt, xt, prec, sv, sx in ((D.F32, D.F32, "F32", 4, 4), (D.BF16, D.F32, "BF16", 2, 4)):
                nm = f"AXT_NOC_H{th}_{prec}"
                d = D.DeviceAXT(eng, ah, val_type=vt, x_type=xt)
                y = nanv(n, np.float32)
                w, e = timed(lambda: d.int_axt(y), 20, 3, torch)
                own = M.bytes_axt(n, ah.lenAX, ah.lenSEC, sv=sv, sx=sx, sy=4)
                record(nm, own, w, e, {"GBs_eq_fp64": round(fb / e / 1e9, 1),
                                       "vs_fp64": round(fe / e, 2)})
                keep(nm, "y", y, n, tol=1e-6 if prec == "F32" else 1e-2)
                del d
        # column-map form: 1 and 8 right-hand sides, adjoint on the transposed map
        # each column-map product first picks its plan's placement on this
        # matrix (long rows dealt over all XCDs, or row groups pinned to XCDs;
        # same bits) by timing both, as a caller would at set-up
        ag = H.AXT(m, x, th=1, want_acol=True)
        g = D.DeviceAXTGather(eng, ag)
        pl1 = g.tune_placement(k=1)
        y = nanv(n)
        w, e = timed(lambda: g.matvec(xd, y), 20, 3, torch)
        fwd_g = e
        # the ceilings of a product that gathers x once per entry: random
        # 8-byte gathers within one XCD's L2 and over all of x
        ceil = {"l2_1MiB_G_per_s": round(eng.gather_rate(1 << 20, 8) / 1e9, 1),
                "x_span_G_per_s": round(eng.gather_rate(n * 8, 8) / 1e9, 1),
                "x_span_bytes": n * 8}
        record("AXT_GATHER_H1_F64_K1", M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC), w, e,
               {"vs_embedded": round(e / head_ev, 3), "placement": pl1,
                "G_gathers_per_s": round(nnz / e / 1e9, 1), "gather_ceilings": ceil})
        keep("AXT_GATHER_H1_F64_K1", "y", y, n)
        # conversion: a new x into the embedded AXT arrays, on the device
        # (spmv_axt_refresh_x) versus the host builder (getFormatDataAXT)
        t0 = time.perf_counter()
        H.AXT(m, x, th=1)
        host_build = time.perf_counter() - t0
        dacol = eng.upload(ag.acol)  # same slots as the headline's arrays
        w, e = timed(lambda: D.refresh_x(d1, dacol, xd), 20, 3, torch)
        formats["AXT_X_REFRESH_H1"] = {"us": round(e * 1e6, 2),
                                       "host_build_ms": round(host_build * 1e3, 1)}
        log(f"[bench] AXT_X_REFRESH_H1: {formats['AXT_X_REFRESH_H1']}")
        del dacol
        k8 = 8
        X8 = torch.rand(n * k8, dtype=torch.float64, device=eng.torch_device) * 2 - 1
        Y8 = nanv(n * k8)
        pl8 = g.tune_placement(k=k8)
        w, e = timed(lambda: g.matvec(X8, Y8, k=k8), 10, 2, torch)
        record("AXT_GATHER_H1_F64_K8", M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC, k=k8), w, e,
               {"GFLOPs": round(M.flops(nnz, k8) / e / 1e9, 1), "placement": pl8})
        g.set_placement(pl1["placement"])
        keep("AXT_GATHER_H1_F64_K8", "mm", Y8, n * k8, X=X8.cpu().numpy().reshape(n, k8))
        adj = D.AxtAdjoint(eng, m, g)
        pla = adj.gT.tune_placement(k=1)
        yb = torch.randn(n, dtype=torch.float64, device=eng.torch_device)
        xb = nanv(n)
        w, e = timed(lambda: adj.adjoint(yb, xb), 20, 3, torch)
        at = adj.at
        record("AXT_ADJOINT_H1_F64", M.bytes_axt_gather(n, at.lenAX, at.lenSEC), w, e,
               {"vs_forward_gather": round(e / fwd_g, 3), "vs_forward_embedded": round(e / head_ev, 3),
                "placement": pla})
        ybh = yb.cpu().numpy()
        keep("AXT_ADJOINT_H1_F64", "adj", xb, n, ybar=ybh)
        # right after a value update (a training step): A's values read through
        # the slot map, or first copied into Aᵀ's order (a permutation)
        xb.fill_(float("nan"))
        w, e = timed(lambda: adj.adjoint(yb, xb, vals=g.aval), 20, 3, torch)
        perm_t = e
        record("AXT_ADJOINT_PERMUTED_H1_F64",
               M.bytes_axt_gather(n, at.lenAX, at.lenSEC) + (at.lenAX // 2) * 4, w, e,
               {"vs_forward_gather": round(e / fwd_g, 3),
                "vs_forward_embedded": round(e / head_ev, 3)})
        keep("AXT_ADJOINT_PERMUTED_H1_F64", "adj", xb, n, ybar=ybh)
        w, e = timed(lambda: adj.refresh(g.aval), 20, 3, torch)
        record("AXT_ADJOINT_VALUES_H1_F64", (at.lenAX // 2) * (4 + 8 + 8), w, e,
               {"vs_forward_gather": round(e / fwd_g, 3)})
        # which of the two the autograd backward takes on this matrix
        formats["AXT_ADJOINT_VALUES_PATH_H1_F64"] = adj.tune_values_path()
        log(f"[bench] AXT_ADJOINT_VALUES_PATH_H1_F64: {formats['AXT_ADJOINT_VALUES_PATH_H1_F64']}")
        vb = nanv(ag.lenAX // 2)
        w, e = timed(lambda: g.grad_vals(xd, yb, vb), 20, 3, torch)
        gv_t = e
        record("AXT_GRAD_VALS_H1_F64", (ag.lenAX // 2) * (4 + 8) + ag.lenSEC * 4 + 2 * n * 8, w, e,
               {"vs_forward_gather": round(e / fwd_g, 3)})
        keep("AXT_GRAD_VALS_H1_F64", "gv", vb, ag.lenAX // 2, ybar=ybh,
             nzslot=ag.nzslot.copy())
        # values held in both slot orders (the autograd pair operator): x_bar
        # on A^T's own copy with A_bar in A^T's order written in the same pass
        # (spmv_axt_gather_grad); backward = that + A_bar in A's order
        vbt = nanv(at.lenAX // 2)
        xb.fill_(float("nan"))
        w, e = timed(lambda: adj.adjoint_grad(yb, xb, xd, vbt), 20, 3, torch)
        record("AXT_ADJOINT_GRAD_H1_F64",
               M.bytes_axt_gather(n, at.lenAX, at.lenSEC) + (at.lenAX // 2) * 8 + n * 8, w, e,
               {"vs_forward_gather": round(e / fwd_g, 3),
                "vs_forward_embedded": round(e / head_ev, 3),
                "backward_pair_us": round((e + gv_t) * 1e6, 2),
                "backward_permuted_us": round((perm_t + gv_t) * 1e6, 2)})
        keep("AXT_ADJOINT_GRAD_H1_F64", "adj", xb, n, ybar=ybh)
        del vbt
        del g, adj, X8, Y8, vb
        # the same in FP32 storage (values and x FP32, FP32 sums)
        g32 = D.DeviceAXTGather(eng, ag, D.F32, D.F32)
        x32 = xd.float()
        y32 = nanv(n, np.float32)
        pl32 = g32.tune_placement(k=1)
        w, e = timed(lambda: g32.matvec(x32, y32), 20, 3, torch)
        fwd32 = e
        record("AXT_GATHER_H1_F32_K1", M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC, sv=4, sx=4, sy=4),
               w, e, {"placement": pl32})
        keep("AXT_GATHER_H1_F32_K1", "y", y32, n, tol=1e-6)
        k32 = 32  # 16 right-hand sides per pass
        X32 = torch.rand(n * k32, dtype=torch.float32, device=eng.torch_device) * 2 - 1
        Y32 = nanv(n * k32, np.float32)
        pl32k = g32.tune_placement(k=k32)
        w, e = timed(lambda: g32.matvec(X32, Y32, k=k32), 10, 2, torch)
        record("AXT_GATHER_H1_F32_K32",
               M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC, k=k32, sv=4, sx=4, sy=4), w, e,
               {"GFLOPs": round(M.flops(nnz, k32) / e / 1e9, 1), "placement": pl32k})
        g32.set_placement(pl32["placement"])
        keep("AXT_GATHER_H1_F32_K32", "mm", Y32, n * k32,
             X=X32.double().cpu().numpy().reshape(n, k32), tol=1e-6)
        del X32, Y32
        adj32 = D.AxtAdjoint(eng, m, g32)
        pla32 = adj32.gT.tune_placement(k=1)
        yb32 = yb.float()
        xb32 = nanv(n, np.float32)
        w, e = timed(lambda: adj32.adjoint(yb32, xb32), 20, 3, torch)
        at = adj32.at
        record("AXT_ADJOINT_H1_F32", M.bytes_axt_gather(n, at.lenAX, at.lenSEC, sv=4, sx=4, sy=4),
               w, e, {"vs_forward_gather": round(e / fwd32, 3), "placement": pla32})
        keep("AXT_ADJOINT_H1_F32", "adj", xb32, n, ybar=yb32.double().cpu().numpy(), tol=1e-6)
        del g32, adj32, ag, x32, y32, yb32, xb32
        dc = D.DeviceCSR(eng, m)
        y = nanv(n)
        plc = dc.tune_placement(xd, y)
        y.fill_(float("nan"))
        w, e = timed(lambda: dc.spmv(xd, y), 20, 3, torch)
        record("CSR_F64", M.bytes_csr(n, nnz), w, e, {"placement": plc})
        keep("CSR_F64", "y", y, n)
        y.fill_(float("nan"))
        w, e = timed(lambda: dc.cf_csr(xd, y), 3, 1, torch)
        record("CSR_ORDERED_F64", M.bytes_csr(n, nnz), w, e)
        keep("CSR_ORDERED_F64", "y", y, n, bitwise=True)
        # cf_CSR's bits from the block kernel: an exact plan (no split rows,
        # one fma chain per long row through LDS)
        y.fill_(float("nan"))
        w, e = timed(lambda: dc.spmv_exact(xd, y), 20, 3, torch)
        record("CSR_EXACT_F64", M.bytes_csr(n, nnz), w, e)
        keep("CSR_EXACT_F64", "y", y, n, bitwise=True)
        rs = rocsparse_leg(dc, xd, y, n, nnz)
        if rs is None:
            formats["ROCSPARSE_CSR_F64"] = "librocsparse not found"
        for alg, (t, rel) in (rs or {}).items():
            if t is None:
                formats[f"ROCSPARSE_CSR_F64_{alg}"] = rel
            else:
                record(f"ROCSPARSE_CSR_F64_{alg}", M.bytes_csr(n, nnz), t, t,
                       {"rel_err_vs_csr": rel})
        c = H.AXC(m, x)
        dcc = D.DeviceAXC(eng, c)
        y.fill_(float("nan"))
        w, e = timed(lambda: dcc.int_axc(y), 20, 3, torch)
        record("AXC_F64", M.bytes_axc(n, c.lenAX), w, e)
        keep("AXC_F64", "y", y, n)
        del dcc
        hosts["AXC"] = c
        t0 = time.perf_counter()
        k = H.K1(m)
        host_build = time.perf_counter() - t0
        dk = D.DeviceK1(eng, k)
        y.fill_(float("nan"))
        w, e = timed(lambda: dk.int_k1p(xd, y), 20, 3, torch)
        keep("K1P_F64", "y", y, n)
        # the K1 conversion on the device from the device CSR (intspmv_k1_build_device)
        D.DeviceK1.from_device_csr(eng, dc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dk2 = D.DeviceK1.from_device_csr(eng, dc)
        torch.cuda.synchronize()
        dev_build = time.perf_counter() - t0
        record("K1P_F64", M.bytes_k1(n, k.lenVC, k.chunkNum), w, e,
               {"build_host_ms": round(host_build * 1e3, 1),
                "build_device_ms": round(dev_build * 1e3, 2)})
        del dk, dk2, dc
        hosts["K1"] = k

    batch = None
    if not args.quick and args.batch_count > 0:
        batch = bench_batch(eng, args.batch_count, torch, log, args.batch_tune)
    precision, prec_pending = None, []
    if not args.quick and not args.no_precision:
        precision, prec_pending = bench_precision(eng, torch, log)
    differentiation, reference_case, ref_pending = None, None, None
    if not args.quick:
        differentiation = bench_differentiation(eng, m, torch, log)
        reference_case, ref_pending = bench_reference_case(eng, torch, log)

    # ---- CPU baseline + correctness (oracle as the checker) -------------------
    cpu, over = None, None
    if not args.no_cpu:
        over = check_results(m, x, checks, formats)
        del checks
        if batch is not None:
            batch.update(check_batch(batch.pop("_check")))
            if not batch["max_scaled_err_vs_fp32_oracle"] <= 1e-6:
                over.append("batch")
        if precision is not None:
            over += check_precision(precision, prec_pending)
        if reference_case is not None:
            over += check_reference_case(reference_case, ref_pending)
        for name, r in (differentiation or {}).items():
            if not r["fd_worst_rel_err"] < r["tol"]:
                over.append(f"differentiation:{name}")
        if over:
            log(f"[bench] OVER TOLERANCE: {over}")
        cpu = cpu_leg(m, x, a1, hosts, head_bytes, args.workload, args.quick)
    elif batch is not None:
        batch.pop("_check")

    achieved = head_bytes / head_ev / 1e9
    traffic_file = f"profiles/axt_h1_f64_{args.workload}_traffic.txt"
    traffic = read_traffic(os.path.basename(traffic_file))
    cold_us = formats["AXT_NOC_H1_F64"].get("us_cold_median")
    out = {
        "metric": "SpMV HBM GB/s (% of 8 TB/s) + GFLOP/s: AXT vs CSR, FP64 vs FP32/BF16, 1 GPU",
        "value": round(head_bytes / head_wall / 1e9, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head_wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: AXT NOC tileHeight 1, FP64, x embedded",
                   "n": n, "nnz": nnz, "lenAX": a1.lenAX, "tileN": a1.tileN,
                   "parallelism": "single GPU"},
        "gflops": round(M.flops(nnz) / head_wall / 1e9, 1),
        "pct_of_hbm_peak": round(100 * head_bytes / head_wall / 1e9 / M.HBM_PEAK_GBS, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": M.HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / M.HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": (f"{traffic_file}: rocprofv3 --pmc FETCH_SIZE, then "
                                        f"WRITE_SIZE, per dispatch of the same kernel on the same "
                                        f"workload, scaled by a 512 MiB device copy's counts; "
                                        f"not measured in this run") if traffic else None,
                     "traffic_vs_algorithmic": round(traffic / head_bytes, 4) if traffic else None,
                     "bytes_per_launch": int(head_bytes),
                     "us_per_launch": round(head_ev * 1e6, 2),
                     "cold_us_median": cold_us,
                     "cold_frac": (round(head_bytes / (cold_us * 1e-6) / 1e9 / M.HBM_PEAK_GBS, 4)
                                   if cold_us else None),
                     "kernel": "axt_h1_kernel<double,double,double,false> (+ split fix-up)",
                     "stream_read_GBs": round(stream_bps / 1e9, 1),
                     "frac_of_stream_read": round(achieved * 1e9 / stream_bps, 4)},
        "errors": ({"checked": len(formats_checked(formats)) + sum(
                        "max_scaled_err_vs_fp64_oracle" in f for w in (precision or {}).values()
                        for f in w["formats"].values()) + sum(
                        isinstance(f, dict) and "max_scaled_err_vs_fp64_oracle" in f
                        for f in (reference_case or {}).values()), "over_tolerance": over}
                   if over is not None else None),
        "cpu_baseline": cpu,
        "formats": formats,
        "batch": batch,
        "precision": precision,
        "differentiation": differentiation,
        "reference_case": reference_case,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
