#!/usr/bin/env python3
"""SpMV HBM GB/s (% of 8 TB/s) + GFLOP/s on one MI355X: AXT vs CSR, FP64 vs FP32/BF16.

One "step" is one SpMV y = A x over the whole matrix with A and x resident in
HBM.  The headline `value` is AXT NOC tileHeight 1 in FP64 (the drop-in for
the reference's int_axt_noc_h1, intSpmv_main.c:1385) on the power-law
workload: n = 2^20, ~30M entries, Zipf(2) row lengths up to 8192.  GB/s are
algorithmic bytes (int_spmv_amd/metrics.py) per step time.  The other formats,
tile heights and storage precisions are timed beside it and reported under
"formats".  The CPU baseline is the OpenMP restatement of the reference's
kernel on the host's cores (the reference itself needs <mkl.h> to build).

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
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="powerlaw_1m")
    p.add_argument("--quick", action="store_true", help="headline kernel only (profiling)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--batch-count", type=int, default=256,
                   help="matrices in the batched-launch measurement (0 = skip)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only run a plain device copy of this many MiB (counter calibration)")
    return p.parse_args()


def cores() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    return len(os.sched_getaffinity(0))


def timed(fn, steps, warmup, torch):
    """Returns (wall seconds per step, event seconds per step)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, e0.elapsed_time(e1) * 1e-3 / steps


def bench_batch(eng, count, torch, log):
    """256 matrices (n log-uniform in [1e3, 2e5], 4..64 entries per row, 70 %
    banded), FP32 values and x, k in {1,4,8,16,32} each; per-matrix tile height
    and right-hand-side blocking from the autotuner's byte model; one
    persistent launch captured in a hipGraph and replayed."""
    import numpy as np

    from int_spmv_amd import device as D
    from int_spmv_amd import host as H
    from int_spmv_amd import metrics as M
    from int_spmv_amd.workloads import batch_256
    t0 = time.perf_counter()
    mats = batch_256(seed=0, count=count)
    ents, flops, bytes_, rows, nnz = [], 0.0, 0, 0, 0
    for i, (m, k) in enumerate(mats):
        c = H.tune_model(m, k, 4, 4)[0]
        a = H.AXT(m, np.zeros(m.nrows), th=c["tileH"], want_acol=True)
        X = torch.rand(m.nrows * k, dtype=torch.float32, device=eng.torch_device)
        Y = torch.empty(m.nrows * k, dtype=torch.float32, device=eng.torch_device)
        ents.append(dict(a=a, X=X, Y=Y, k=k, kc=c["kc"]))
        flops += M.flops(m.nnz, k)
        bytes_ += M.bytes_axt_gather(m.nrows, a.lenAX, a.lenSEC, k=k, sv=4, sx=4, sy=4)
        rows += m.nrows
        nnz += m.nnz
    b = D.DeviceBatch(eng, ents, D.F32, D.F32)
    for e in ents:
        e.pop("a")
    info = b.info()
    setup = time.perf_counter() - t0
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
         "grid": info["grid"], "setup_s": round(setup, 1)}
    log(f"[bench] BATCH: {r}")
    return r


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
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (every box here "
              f"has a single MI355X).", file=sys.stderr)
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
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                          "GBs_rw": 2 * n * 8 / e / 1e9}))
        return

    from int_spmv_amd import device as D
    from int_spmv_amd import host as H
    from int_spmv_amd import metrics as M
    from int_spmv_amd.workloads import WORKLOADS

    t = time.perf_counter()
    m, x = WORKLOADS[args.workload]()
    n, nnz = m.nrows, m.nnz
    log(f"[bench] workload {args.workload}: n={n} nnz={nnz} rmax={m.stats()['rmax']} "
        f"({time.perf_counter() - t:.1f}s)")
    eng = D.Engine(0)
    xd = eng.upload(x)
    formats = {}

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
    y1 = eng.empty(n)
    head_bytes = M.bytes_axt(n, a1.lenAX, a1.lenSEC)
    wall, ev = timed(lambda: d1.int_axt(y1), args.steps, args.warmup, torch)
    record("AXT_NOC_H1_F64", head_bytes, wall, ev,
           {"beta": round(a1.fd.beta, 4), "plan": d1.plan.info()})
    head_wall, head_ev = wall, ev

    if not args.quick:
        for th in (2, 4, 8, 16):
            a = H.AXT(m, x, th=th)
            d = D.DeviceAXT(eng, a)
            y = eng.empty(n)
            w, e = timed(lambda: d.int_axt(y), 20, 3, torch)
            record(f"AXT_NOC_H{th}_F64", M.bytes_axt(n, a.lenAX, a.lenSEC), w, e,
                   {"beta": round(a.fd.beta, 4)})
            del d, a, y
        for vt, xt, nm, sv, sx in ((D.F32, D.F32, "AXT_NOC_H1_F32", 4, 4),
                                  (D.BF16, D.F32, "AXT_NOC_H1_BF16", 2, 4)):
            d = D.DeviceAXT(eng, a1, val_type=vt, x_type=xt)
            y = eng.empty(n, np.float32)
            w, e = timed(lambda: d.int_axt(y), 20, 3, torch)
            own = M.bytes_axt(n, a1.lenAX, a1.lenSEC, sv=sv, sx=sx, sy=4)
            record(nm, own, w, e, {"GBs_eq_fp64": round(head_bytes / e / 1e9, 1),
                                   "vs_fp64": round(head_ev / e, 2)})
            formats[nm]["_y"] = y
            del d
        # column-map form: 1 and 8 right-hand sides, adjoint on the transposed map
        ag = H.AXT(m, x, th=1, want_acol=True)
        g = D.DeviceAXTGather(eng, ag)
        y = eng.empty(n)
        w, e = timed(lambda: g.matvec(xd, y), 20, 3, torch)
        fwd_g = e
        record("AXT_GATHER_H1_F64_K1", M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC), w, e)
        k8 = 8
        X8 = torch.rand(n * k8, dtype=torch.float64, device=eng.torch_device)
        Y8 = eng.empty(n * k8)
        w, e = timed(lambda: g.matvec(X8, Y8, k=k8), 10, 2, torch)
        record("AXT_GATHER_H1_F64_K8", M.bytes_axt_gather(n, ag.lenAX, ag.lenSEC, k=k8), w, e,
               {"GFLOPs": round(M.flops(nnz, k8) / e / 1e9, 1)})
        adj = D.AxtAdjoint(eng, m, g)
        yb = torch.randn(n, dtype=torch.float64, device=eng.torch_device)
        xb = eng.empty(n)
        w, e = timed(lambda: adj.adjoint(yb, xb), 20, 3, torch)
        at = adj.at
        record("AXT_ADJOINT_H1_F64", M.bytes_axt_gather(n, at.lenAX, at.lenSEC), w, e,
               {"vs_forward_gather": round(e / fwd_g, 3), "vs_forward_embedded": round(e / head_ev, 3)})
        vb = eng.empty(ag.lenAX // 2)
        w, e = timed(lambda: g.grad_vals(xd, yb, vb), 20, 3, torch)
        record("AXT_GRAD_VALS_H1_F64", (ag.lenAX // 2) * (4 + 8) + ag.lenSEC * 4 + 2 * n * 8, w, e,
               {"vs_forward_gather": round(e / fwd_g, 3)})
        del g, adj, ag, X8, Y8, vb
        dc = D.DeviceCSR(eng, m)
        y = eng.empty(n)
        w, e = timed(lambda: dc.spmv(xd, y), 20, 3, torch)
        record("CSR_F64", M.bytes_csr(n, nnz), w, e)
        w, e = timed(lambda: dc.cf_csr(xd, y), 3, 1, torch)
        record("CSR_ORDERED_F64", M.bytes_csr(n, nnz), w, e)
        del dc
        c = H.AXC(m, x)
        dcc = D.DeviceAXC(eng, c)
        w, e = timed(lambda: dcc.int_axc(y), 20, 3, torch)
        record("AXC_F64", M.bytes_axc(n, c.lenAX), w, e)
        del dcc, c
        k = H.K1(m)
        dk = D.DeviceK1(eng, k)
        w, e = timed(lambda: dk.int_k1p(xd, y), 20, 3, torch)
        record("K1P_F64", M.bytes_k1(n, k.lenVC, k.chunkNum), w, e)
        del dk, k

    batch = None
    if not args.quick and args.batch_count > 0:
        batch = bench_batch(eng, args.batch_count, torch, log)

    # ---- CPU baseline + correctness (oracle as the checker) -------------------
    cpu = None
    if not args.no_cpu:
        import numpy as np

        import oracle as O
        torch.cuda.synchronize()
        yg = y1[:n].cpu().numpy()
        yref = O.cf_csr(m.row, m.col, m.val, x)
        err, _ = O.scaled_err(m.row, m.col, m.val, x, yref, yg)
        formats["AXT_NOC_H1_F64"]["max_scaled_err_vs_csr"] = err
        for nm in ("AXT_NOC_H1_F32", "AXT_NOC_H1_BF16"):
            if nm in formats:
                yy = formats[nm].pop("_y")[:n].double().cpu().numpy()
                formats[nm]["max_scaled_err_vs_fp64"] = O.scaled_err(m.row, m.col, m.val, x,
                                                                     yref, yy)[0]
        threads = cores()
        ycpu = np.zeros(n)
        import ctypes as C
        L = O.lib()

        def p(a):
            return C.c_void_p(a.ctypes.data)
        ax = np.ascontiguousarray(a1.ax)
        sec = np.ascontiguousarray(a1.sec)
        t1 = L.orc_bench(4, threads, 1, n, a1.tileN, 1, p(ax), p(sec), None, None, None, None,
                         p(ycpu))
        iters = int(max(1, min(250, 10.0 / max(t1, 1e-6))))
        tc = L.orc_bench(4, threads, iters, n, a1.tileN, 1, p(ax), p(sec), None, None, None,
                         None, p(ycpu))
        cpu = {"value": round(head_bytes / tc / 1e9, 2), "unit": "GB/s", "cores": threads,
               "kind": "port",
               "sample": f"int_axt_noc_h1 OpenMP restatement (dynamic schedule, atomic y), "
                         f"full {args.workload} matrix, {iters} calls, y zeroed untimed; "
                         f"{tc * 1e3:.2f} ms/call"}
        cerr, _ = O.scaled_err(m.row, m.col, m.val, x, yref, ycpu)
        cpu["max_scaled_err_vs_csr"] = cerr
        log(f"[bench] cpu baseline: {cpu}")
    for f in formats.values():
        f.pop("_y", None)

    achieved = head_bytes / head_ev / 1e9
    traffic = read_traffic(f"axt_h1_f64_{args.workload}_traffic.txt")
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
                     "kernel": "axt_h1_kernel<double,double,double,false> (+ split fix-up)"},
        "cpu_baseline": cpu,
        "formats": formats,
        "batch": batch,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
