"""GPU: seeded random ragged matrices through every product entry point,
and random mixes of them in one batched launch; the precision kernels, the
adjoint and the value gradient on them; the device-side conversions (K1
build, x refresh) against the host builders.

Each seed draws a matrix of 1..3000 rows with runs of empty rows, duplicate
columns and, for most seeds, one long row (up to 6000 entries, so plans split
it across waves), a tile height and a COM block size.  Bar, against the
oracle's CSR-naive (cf_CSR): the ordered CSR product bit-exact; K1's chunk
results bit-exact for chunks of at most 64 steps; every other product within
1e-12 * sum_j |a_ij x_j| per row (FP64).  y is pre-filled with NaN so a row
a kernel forgets to write fails.
"""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal, scaled

pytestmark = pytest.mark.gpu

SEEDS = range(96)


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([1, 2, 7, 64, 257, 1000, 3000]))
    lens = rng.integers(0, 30, size=n)
    lens[rng.random(n) < 0.2] = 0
    if n > 8:
        a = int(rng.integers(0, n - 8))
        lens[a:a + int(rng.integers(1, 8))] = 0  # a run of empty rows
    if seed % 4 != 3:
        lens[int(rng.integers(0, n))] = int(rng.integers(100, 6000))
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, n, size=k)) for k in lens] +
                         [np.zeros(0, np.int64)]).astype(np.uint32)
    m = H.CSR.from_arrays(row, col, rng.standard_normal(len(col)))
    x = rng.uniform(-1, 1, n)
    th = int(rng.choice([1, 2, 3, 4, 8, 16, 32, 48]))
    bs = int(rng.choice([2, 8, 128, 256, 512]))
    return m, x, th, bs


def _y(eng, n, dtype=np.float64):
    t = eng.empty(n, dtype)
    t.fill_(float("nan"))
    return t


def _host(t, n):
    import torch
    torch.cuda.synchronize()
    return t.reshape(-1)[:n].double().cpu().numpy()


@pytest.mark.parametrize("seed", SEEDS)
def test_random_matrix_all_products(engine, seed):
    m, x, th, bs = _case(seed)
    n = m.nrows
    print(f"[random] seed={seed} n={n} nnz={m.nnz} th={th} bs={bs}", flush=True)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    xd = engine.upload(x)
    # CSR: ordered (the oracle's order) and blocked
    d = D.DeviceCSR(engine, m)
    y1, y2 = _y(engine, n), _y(engine, n)
    d.cf_csr(xd, y1)
    d.spmv(xd, y2)
    assert bitwise_equal(_host(y1, n), yr), seed
    assert scaled(m, x, yr, _host(y2, n)) <= 1e-12, seed
    y3 = _y(engine, n)
    d.spmv_exact(xd, y3)  # the exact plan: cf_CSR's bits, long rows included
    assert bitwise_equal(_host(y3, n), yr), seed
    # AXC
    a = H.AXC(m, x)
    y = _y(engine, n)
    D.DeviceAXC(engine, a).int_axc(y)
    assert scaled(m, x, yr, _host(y, n)) <= 1e-12, seed
    # K1: chunk results, reordered y, and the fused product
    k = H.K1(m)
    dk = D.DeviceK1(engine, k)
    res, yk, yp = _y(engine, k.chunkNum * 8), _y(engine, n), _y(engine, n)
    dk.int_k1(xd, res)
    dk.order(res, yk)
    dk.int_k1p(xd, yp)
    ref_res = O.int_k1(k.chunkNum, k.val, k.col, k.chp, k.nmc, x)
    narrow = ~np.repeat(k.nmc > 64, 8)
    assert bitwise_equal(_host(res, k.chunkNum * 8)[narrow], ref_res[narrow]), seed
    assert scaled(m, x, yr, _host(yk, n)) <= 1e-12, seed
    assert scaled(m, x, yr, _host(yp, n)) <= 1e-12, seed
    # AXT NOC at this tile height (x embedded), and COM at this block size
    for mode, t in (("NOC", th), ("COM", 1)):
        ax = H.AXT(m, x, th=t, mode=mode, bs=bs)
        y = _y(engine, n)
        D.DeviceAXT(engine, ax).int_axt(y)
        assert scaled(m, x, yr, _host(y, n)) <= 1e-12, (seed, mode)
    # column map, 1 and 3 right-hand sides
    ag = H.AXT(m, np.zeros(n), th=th, want_acol=True)
    g = D.DeviceAXTGather(engine, ag)
    y = _y(engine, n)
    g.matvec(xd, y)
    assert scaled(m, x, yr, _host(y, n)) <= 1e-12, seed
    X = np.stack([x, -x, np.linspace(-1, 1, n)], axis=1)
    Y = _y(engine, 3 * n)
    g.matvec(engine.upload(X.reshape(-1)), Y, k=3)
    Yg = _host(Y, 3 * n).reshape(n, 3)
    for j in range(3):
        assert scaled(m, X[:, j], O.cf_csr(m.row, m.col, m.val, X[:, j]), Yg[:, j]) <= 1e-12, (seed, j)


@pytest.mark.parametrize("seed", range(24))
def test_random_many_empty_rows(engine, seed):
    """Matrices with 30-70 % empty rows take the plans of 128-unit blocks,
    whose load rounds the th = 1 FP64 and AXC kernels keep in flight
    together, and whose rows run up to 256 per block.  Rows in normal blocks
    keep the reference kernels' bits (AXT NOC at th 1 and at a drawn th, AXC),
    COM rows of at most 128 entries cf_CSR's; everything within 1e-12, FP32
    storage within 1e-6 and bf16 values within 1e-2 of the FP64 oracle."""
    from .test_gpu_fullsize import _many_empty, _normal_rows
    rng = np.random.default_rng(5000 + seed)
    n = int(rng.choice([500, 3000, 20000]))
    lens = rng.integers(1, 60, size=n)
    lens[rng.random(n) < rng.uniform(0.3, 0.7)] = 0
    for r in rng.integers(0, n, size=3):  # long rows: some past 64 tiles
        lens[r] = int(rng.integers(200, 3000))
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, n, size=k)) for k in lens]).astype(np.uint32)
    m = H.CSR.from_arrays(row, col, rng.standard_normal(len(col)))
    x = rng.uniform(-1, 1, n)
    th = int(rng.choice([2, 4, 8, 16, 32, 48]))
    bs = int(rng.choice([2, 128, 512]))
    print(f"[random-empty] seed={seed} n={n} nnz={m.nnz} th={th} bs={bs}", flush=True)
    assert _many_empty(m)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    for t in (1, th):
        a = H.AXT(m, x, th=t)
        y = _y(engine, n)
        D.DeviceAXT(engine, a).int_axt(y)
        g = _host(y, n)
        ok = _normal_rows(m, t)
        assert bitwise_equal(g[ok], O.int_axt(a)[ok]), (seed, t)
        assert scaled(m, x, yr, g) <= 1e-12, (seed, t)
        if t == 1:
            for vt, xt, tol in ((D.F32, D.F32, 1e-6), (D.BF16, D.F32, 1e-2)):
                y = _y(engine, n, np.float32)
                D.DeviceAXT(engine, a, val_type=vt, x_type=xt).int_axt(y)
                assert scaled(m, x, yr, _host(y, n)) <= tol, (seed, vt)
    c = H.AXC(m, x)
    y = _y(engine, n)
    D.DeviceAXC(engine, c).int_axc(y)
    g = _host(y, n)
    ok = np.diff(c.brp.astype(np.int64)) // 16 <= 64  # normal blocks of the sparse-row plan
    assert bitwise_equal(g[ok], O.int_axc(n, c.ax, c.brp)[ok]), seed
    assert scaled(m, x, yr, g) <= 1e-12, seed
    a = H.AXT(m, x, th=1, mode="COM", bs=bs)
    y = _y(engine, n)
    D.DeviceAXT(engine, a).int_axt(y)
    g = _host(y, n)
    short = np.diff(m.row.astype(np.int64)) <= 128
    assert bitwise_equal(g[short], yr[short]), seed
    assert scaled(m, x, yr, g) <= 1e-12, seed


@pytest.mark.parametrize("seed", range(18))
def test_random_batch(engine, seed):
    """One batched launch over 1..12 of the random matrices, each with its own
    tile height, 1..32 right-hand sides, a leading dimension of x up to two
    past k, a drawn RHS blocking kc and units per wave block; by seed: FP64
    (1e-12), FP32 storage and sums (1e-6 against CSR-naive of the
    FP32-rounded values) or BF16 values with FP32 x and sums (1e-2)."""
    import torch
    rng = np.random.default_rng(5000 + seed)
    mode = seed % 3
    f64 = mode == 0
    vt, xt, dt, tol = [(D.F64, D.F64, torch.float64, 1e-12), (D.F32, D.F32, torch.float32, 1e-6),
                       (D.BF16, D.F32, torch.float32, 1e-2)][mode]
    dev = engine.torch_device
    ents = []
    for i in range(int(rng.integers(1, 13))):
        m, _, th, _ = _case(int(rng.integers(0, 1 << 20)))
        n = m.nrows
        k = int(rng.choice([1, 2, 3, 5, 8, 13, 32]))
        ldx = k + int(rng.integers(0, 3))
        X = rng.uniform(-1, 1, (n, ldx))
        if not f64:
            X = X.astype(np.float32).astype(np.float64)
        a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
        ents.append(dict(a=a, X=torch.tensor(X.reshape(-1), dtype=dt, device=dev),
                         Y=torch.full((n * k,), float("nan"), dtype=dt, device=dev), k=k, ldx=ldx,
                         kc=int(rng.choice([0, 1, 2, 4, 8])), umax=int(rng.choice([0, 16, 32, 64, 128])),
                         m=m, Xh=X[:, :k]))
    print(f"[random batch] seed={seed} " + " ".join(
        f"(n={e['m'].nrows} th={e['a'].tileH} k={e['k']} ldx={e['ldx']} kc={e['kc']} umax={e['umax']})"
        for e in ents), flush=True)
    b = D.DeviceBatch(engine, ents, vt, xt)
    b.launch()
    torch.cuda.synchronize()
    for i, e in enumerate(ents):
        m, k = e["m"], e["k"]
        Y = e["Y"].double().cpu().numpy().reshape(m.nrows, k)
        vals = m.val.astype(np.float32).astype(np.float64) if mode == 1 else m.val
        for j in range(k):
            yr = O.cf_csr(m.row, m.col, vals, e["Xh"][:, j])
            assert scaled(m, e["Xh"][:, j], yr, Y[:, j]) <= tol, (seed, i, j)


@pytest.mark.parametrize("seed", range(32))
def test_random_precision_and_adjoint(engine, seed):
    """The random matrices through the FP32/BF16 embedded kernels (1e-6 /
    1e-2), the planar layout with FP64 and FP32 values (1e-12 / 1e-6), COM in
    FP32 (1e-6), the BF16 column map (1e-2), the FP32 column
    map with 20 right-hand sides (a pass of 16 and a pass of 4, 1e-6), the adjoint on the transposed tile map (1e-12 against
    the sequential CSR transpose product; the slot-map read gives the same
    bits) and the value gradient (bit-exact: one product per slot)."""
    import torch
    m, x, th, bs = _case(seed)
    n = m.nrows
    print(f"[random precision/adjoint] seed={seed} n={n} nnz={m.nnz} th={th} bs={bs}", flush=True)
    dev = engine.torch_device
    yr = O.cf_csr(m.row, m.col, m.val, x)
    a = H.AXT(m, x, th=th)
    for vt, tol in ((D.F32, 1e-6), (D.BF16, 1e-2)):
        y = _y(engine, n, np.float32)
        D.DeviceAXT(engine, a, val_type=vt, x_type=D.F32).int_axt(y)
        assert scaled(m, x, yr, _host(y, n)) <= tol, (seed, vt)
    # planar value / x arrays with FP64 and FP32 values
    for vt, dt, tol in ((D.F64, np.float64, 1e-12), (D.F32, np.float32, 1e-6)):
        y = _y(engine, n, dt)
        D.DeviceAXT(engine, a, val_type=vt, x_type=vt, layout=D.PLANAR).int_axt(y)
        assert scaled(m, x, yr, _host(y, n)) <= tol, (seed, vt, "planar")
    y = _y(engine, n, np.float32)
    D.DeviceAXT(engine, H.AXT(m, x, th=1, mode="COM", bs=bs), val_type=D.F32,
                x_type=D.F32).int_axt(y)
    assert scaled(m, x, yr, _host(y, n)) <= 1e-6, (seed, "COM F32")
    ag = H.AXT(m, np.zeros(n), th=th, want_acol=True)
    y = _y(engine, n, np.float32)
    D.DeviceAXTGather(engine, ag, D.BF16, D.F32).matvec(
        torch.tensor(x, dtype=torch.float32, device=dev), y)
    x32 = x.astype(np.float32).astype(np.float64)
    assert scaled(m, x32, O.cf_csr(m.row, m.col, m.val, x32), _host(y, n)) <= 1e-2, (seed, "BF16")
    g32 = D.DeviceAXTGather(engine, ag, D.F32, D.F32)
    k = 20
    X = np.random.default_rng(seed).uniform(-1, 1, (n, k)).astype(np.float32).astype(np.float64)
    Y = torch.full((n * k,), float("nan"), dtype=torch.float32, device=dev)
    g32.matvec(torch.tensor(X.reshape(-1), dtype=torch.float32, device=dev), Y, k=k)
    Yg = _host(Y, n * k).reshape(n, k)
    v32 = m.val.astype(np.float32).astype(np.float64)
    for j in range(k):
        assert scaled(m, X[:, j], O.cf_csr(m.row, m.col, v32, X[:, j]), Yg[:, j]) <= 1e-6, (seed, j)
    g = D.DeviceAXTGather(engine, ag)
    adj = D.AxtAdjoint(engine, m, g)
    ybar = np.random.default_rng(seed + 1).standard_normal(n)
    ybd = engine.upload(ybar)
    xb1, xb2 = _y(engine, n), _y(engine, n)
    adj.adjoint(ybd, xb1)
    adj.adjoint(ybd, xb2, vals=g.aval)
    at, _ = m.transpose()
    got = _host(xb1, n)
    assert scaled(at, ybar, O.csr_transpose_product(m.row, m.col, m.val, ybar), got) <= 1e-12, seed
    assert bitwise_equal(_host(xb2, n), got), seed
    vb = _y(engine, ag.lenAX // 2)
    g.grad_vals(engine.upload(x), ybd, vb)
    rows = np.repeat(np.arange(n), np.diff(m.row.astype(np.int64)))
    want = np.zeros(ag.lenAX // 2)
    want[ag.nzslot] = ybar[rows] * x[m.col]
    assert bitwise_equal(_host(vb, ag.lenAX // 2), want), seed


@pytest.mark.parametrize("seed", range(32))
def test_random_device_conversion(engine, seed):
    """Device-side conversions on the random matrices, bit for bit against the
    host builders: the K1 build from device CSR (ties of equal row length
    are many here; the device sort keeps the host's stable order) and the x
    refresh of an embedded AXT (every x half first set to NaN)."""
    import torch
    m, x, th, _ = _case(seed)
    n = m.nrows
    print(f"[random conversion] seed={seed} n={n} nnz={m.nnz} th={th}", flush=True)
    k = H.K1(m)
    dk = D.DeviceK1.from_device_csr(engine, D.DeviceCSR(engine, m))
    torch.cuda.synchronize()

    def u32(t, c):
        return t[:c].cpu().numpy().view(np.uint32)

    assert (dk.chunkNum, dk.lenVC) == (k.chunkNum, k.lenVC), seed
    assert np.array_equal(u32(dk.permi, (k.chunkNum + 1) * 8), k.permi), seed
    for f, c in (("nmc", k.chunkNum), ("chp", k.chunkNum), ("col", k.lenVC)):
        assert np.array_equal(u32(getattr(dk, f), c), getattr(k, f)), (seed, f)
    assert bitwise_equal(dk.val[: k.lenVC].cpu().numpy(), k.val), seed
    x1 = np.random.default_rng(seed + 7).uniform(-1, 1, n)
    a = H.AXT(m, x, th=th, want_acol=True)
    e = D.DeviceAXT(engine, a)
    stale = a.ax.copy()
    stale.reshape(-1, 16)[:, 8:] = np.nan
    e.ax.copy_(engine.upload(stale))
    D.refresh_x(e, engine.upload(a.acol), engine.upload(x1))
    torch.cuda.synchronize()
    assert bitwise_equal(e.ax.cpu().numpy(), H.AXT(m, x1, th=th).ax), seed
