"""GPU: the batched persistent launch (many matrices, uneven sizes, own tile
heights, 1..32 right-hand sides), its hipGraph capture and replay, and the
autotuner."""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H
from int_spmv_amd.workloads import batch_256

from .util import bitwise_equal, csr_long_rows, csr_sparse_gaps, scaled

pytestmark = pytest.mark.gpu


def _entries(engine, mats, ths, dtype):
    import torch
    ents = []
    for i, (m, k) in enumerate(mats):
        n = m.nrows
        X = np.stack([H.uniform(n, seed=50 + i * 40 + j) for j in range(k)], axis=1)
        if dtype == torch.float32:
            X = X.astype(np.float32).astype(np.float64)
        Xd = torch.tensor(X.reshape(-1), dtype=dtype, device=engine.torch_device)
        Yd = torch.full((n * k,), float("nan"), dtype=dtype, device=engine.torch_device)
        a = H.AXT(m, np.zeros(n), th=ths[i % len(ths)], want_acol=True)
        ents.append(dict(a=a, X=Xd, Y=Yd, k=k, Xh=X, m=m))
    return ents


def _y(e):
    import torch
    torch.cuda.synchronize()
    return e["Y"].double().cpu().numpy().reshape(e["m"].nrows, e["k"])


def test_batch_f64_equals_single(engine):
    import torch
    mats = [(m, k) for (m, k) in batch_256(seed=1, count=10, nmax=20000)]
    mats.append((csr_long_rows(), 3))  # split rows inside a batch
    mats.append((csr_sparse_gaps(), 5))  # long runs of empty rows
    ents = _entries(engine, mats, (1, 2, 4, 8), torch.float64)
    b = D.DeviceBatch(engine, ents, D.F64, D.F64)
    b.launch()
    assert b.info()["nsplit"] > 0
    for e in ents:
        g = D.DeviceAXTGather(engine, e["a"], D.F64, D.F64)
        Y1 = torch.empty_like(e["Y"])
        g.matvec(e["X"], Y1, k=e["k"])
        torch.cuda.synchronize()
        Yb = _y(e)
        assert bitwise_equal(Yb.reshape(-1), Y1.cpu().numpy()), e["m"].nrows
        m = e["m"]
        for j in range(e["k"]):
            yr = O.cf_csr(m.row, m.col, m.val, e["Xh"][:, j])
            assert scaled(m, e["Xh"][:, j], yr, Yb[:, j]) <= 1e-12


@pytest.mark.parametrize("vt,ks,tol", [(D.F64, (8, 5, 4), 1e-12), (D.F32, (32, 13, 8), 1e-6)])
def test_batch_wide_long_rows_equal_single(engine, vt, ks, tol):
    """Wide passes over long rows, split rows among them (tile height 1, one
    pass of 4..32 right-hand sides, lanes past k idle): the batch holds the
    single-matrix launch's bits, FP32 included (the batch keeps one tile per
    group of lanes in flight, the single-matrix kernel two for FP64: the same
    order), and both are within tolerance of cf_CSR in every column."""
    import torch
    dt = torch.float64 if vt == D.F64 else torch.float32
    mats = [(csr_long_rows(), ks[0]), (csr_long_rows(n=9000, seed=6), ks[1]),
            (H.CSR.powerlaw(n=1 << 13, target_nnz=200_000, maxlen=4096, seed=2), ks[2]),
            (csr_sparse_gaps(), ks[0])]
    ents = _entries(engine, mats, (1,), dt)
    b = D.DeviceBatch(engine, ents, vt, vt)
    b.launch()
    assert b.info()["nsplit"] > 0
    for i, e in enumerate(ents):
        print(f"[wide long rows] matrix {i} n={e['m'].nrows} k={e['k']}", flush=True)
        g = D.DeviceAXTGather(engine, e["a"], vt, vt)
        Y1 = torch.full_like(e["Y"], float("nan"))
        g.matvec(e["X"], Y1, k=e["k"])
        torch.cuda.synchronize()
        Yb = _y(e)
        assert bitwise_equal(Yb.reshape(-1), Y1.double().cpu().numpy()), i
        m = e["m"]
        for j in range(e["k"]):
            yr = O.cf_csr(m.row, m.col, m.val, e["Xh"][:, j])
            assert scaled(m, e["Xh"][:, j], yr, Yb[:, j]) <= tol, (i, j)


def test_batch_f32_and_graph_replay(engine):
    import torch
    mats = batch_256(seed=2, count=24, nmax=30000)
    ents = _entries(engine, mats, (1, 4), torch.float32)
    b = D.DeviceBatch(engine, ents, D.F32, D.F32)
    b.launch()
    eager = [_y(e).copy() for e in ents]
    for e in ents:
        m = e["m"]
        v32 = m.val.astype(np.float32).astype(np.float64)
        for j in range(e["k"]):
            yr = O.cf_csr(m.row, m.col, v32, e["Xh"][:, j])
            assert scaled(m, e["Xh"][:, j], yr, eager[ents.index(e)][:, j]) <= 1e-6
    b.capture()
    for rep in range(3):
        for e in ents:
            e["Y"].fill_(float("nan"))       # eager work between replays
        torch.cuda.synchronize()
        b.replay()
        for i, e in enumerate(ents):
            assert bitwise_equal(_y(e).reshape(-1), eager[i].reshape(-1)), (rep, i)


def test_autotuner_picks_a_measured_candidate(engine):
    from int_spmv_amd.tune import tune_axt
    m = H.CSR.banded(30000, 4, 64, band_frac=0.7, band_w=1500, seed=3)
    best = tune_axt(engine, m, 8, D.F32, D.F32, top=3, reps=3)
    assert best["tileH"] in (1, 2, 4, 8, 16) and best["kc"] in (1, 2, 4, 8)
    assert best["umax"] in (32, 64, 128)
    assert best["us"] > 0 and len(best["ranked"]) == 3
    assert (best["tileH"], best["kc"]) in [(c["tileH"], c["kc"]) for c in best["ranked"]]


def test_batch_wave_mapping_is_exact(engine):
    """Every block size the tuner may pick gives the single-matrix bits."""
    import torch
    m = H.CSR.powerlaw(n=6000, target_nnz=90_000, maxlen=1500, seed=4)
    k = 3
    X = np.random.default_rng(4).uniform(-1, 1, (m.nrows, k))
    Xd = torch.tensor(X.reshape(-1), dtype=torch.float64, device=engine.torch_device)
    a = H.AXT(m, np.zeros(m.nrows), th=1, want_acol=True)
    g = D.DeviceAXTGather(engine, a, D.F64, D.F64)
    Yref = torch.full((m.nrows * k,), float("nan"), dtype=torch.float64, device=engine.torch_device)
    g.matvec(Xd, Yref, k=k, kc=4)
    for umax in (32, 64, 128):
        Y = torch.full_like(Yref, float("nan"))
        b = D.DeviceBatch(engine, [dict(a=a, X=Xd, Y=Y, k=k, kc=4, umax=umax)], D.F64, D.F64)
        b.launch()
        torch.cuda.synchronize()
        yb, yr = Y.cpu().numpy(), Yref.cpu().numpy()
        yref = np.stack([O.cf_csr(m.row, m.col, m.val, X[:, j].copy()) for j in range(k)], 1)
        for j in range(k):
            assert scaled(m, X[:, j].copy(), yref[:, j], yb.reshape(-1, k)[:, j]) <= 1e-12, umax
        if umax == 128:
            assert bitwise_equal(yb, yr)


def test_batch_wide_x_plain_loads(engine):
    """A batch holding an x of 2 GiB or more gathers every matrix with plain
    loads; the bits equal the buffer-resource launch."""
    import torch
    mats = [(m, k) for (m, k) in batch_256(seed=5, count=6, nmax=8000)]
    ents = _entries(engine, mats, (1, 4), torch.float64)
    b = D.DeviceBatch(engine, ents, D.F64, D.F64)
    b.launch()
    ref = [_y(e).copy() for e in ents]
    e0 = ents[0]
    n, k = e0["m"].nrows, e0["k"]
    wide = -(-(1 << 31) // (n * 8))
    Xw = torch.zeros((n, wide), dtype=torch.float64, device=engine.torch_device)
    Xw[:, :k] = e0["X"].view(n, k)
    ents2 = [dict(e, Y=torch.full_like(e["Y"], float("nan"))) for e in ents]
    ents2[0] = dict(ents2[0], X=Xw.view(-1), ldx=wide)
    b2 = D.DeviceBatch(engine, ents2, D.F64, D.F64)
    b2.launch()
    for i, e in enumerate(ents2):
        assert bitwise_equal(_y(e).reshape(-1), ref[i].reshape(-1)), i
    del b2, ents2, Xw
    torch.cuda.empty_cache()
