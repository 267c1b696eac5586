"""GPU parity: every kernel behind the C ABI against the oracle.

Bar: bit-exact where the oracle fixes the rounding order (CSR ordered and
block rows, AXC, K1, AXT rows that stay inside one wave block); otherwise
|y_gpu - y_ref| <= 1e-12 * sum_j |a_ij x_j| per row against CSR-naive (FP64),
1e-6 (FP32 storage) and 1e-2 (BF16 storage).  y is pre-filled with NaN so a
row the kernel forgets to write fails.
"""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal, csr_long_rows, csr_sparse_gaps, csr_with_empty_rows, scaled

pytestmark = pytest.mark.gpu

TOL64, TOL32, TOL16 = 1e-12, 1e-6, 1e-2


def _y(eng, n, dtype=np.float64):
    torch = pytest.importorskip("torch")
    t = eng.empty(n, dtype)
    t.fill_(float("nan"))
    return t


def _host(t, n):
    t.device  # noqa: B018
    import torch
    torch.cuda.synchronize()
    return t[:n].double().cpu().numpy()


def _matrices():
    yield "powerlaw16k", H.CSR.powerlaw(n=1 << 14, target_nnz=460_000, maxlen=8192, seed=1)
    yield "empty_rows", csr_with_empty_rows()
    yield "banded", H.CSR.banded(5000, 4, 64, seed=2)
    yield "long_rows", csr_long_rows()
    yield "sparse_gaps", csr_sparse_gaps()


@pytest.fixture(scope="module")
def mats():
    out = []
    for name, m in _matrices():
        x = H.uniform(m.nrows, -1.0, 1.0, seed=7)
        out.append((name, m, x, O.cf_csr(m.row, m.col, m.val, x)))
    return out


# ---------------------------------------------------------------- CSR


def test_csr_ordered_bitwise_lap256(engine, lap256):
    m, x = lap256
    d = D.DeviceCSR(engine, m)
    xd, y = engine.upload(x), _y(engine, m.nrows)
    d.cf_csr(xd, y)
    assert bitwise_equal(_host(y, m.nrows), O.cf_csr(m.row, m.col, m.val, x))


def test_csr_block_lap256(engine, lap256):
    m, x = lap256
    d = D.DeviceCSR(engine, m)
    xd, y = engine.upload(x), _y(engine, m.nrows)
    d.spmv(xd, y)
    assert bitwise_equal(_host(y, m.nrows), O.cf_csr(m.row, m.col, m.val, x))


def test_csr_exact_plan_bitwise(engine, mats):
    """spmv_csr under intspmv_csr_plan_exact: cf_CSR's bits for every row,
    rows of 9,000-30,000 entries (split under the plain plan) included."""
    for name, m, x, yr in mats:
        d = D.DeviceCSR(engine, m)
        y = _y(engine, m.nrows)
        d.spmv_exact(engine.upload(x), y)
        assert bitwise_equal(_host(y, m.nrows), yr), name


def test_csr_mats(engine, mats):
    for name, m, x, yr in mats:
        d = D.DeviceCSR(engine, m)
        xd = engine.upload(x)
        y1, y2 = _y(engine, m.nrows), _y(engine, m.nrows)
        d.cf_csr(xd, y1)
        d.spmv(xd, y2)
        g1, g2 = _host(y1, m.nrows), _host(y2, m.nrows)
        assert bitwise_equal(g1, yr), name
        assert scaled(m, x, yr, g2) <= TOL64, name
        short = np.diff(m.row.astype(np.int64)) <= 256
        assert bitwise_equal(g2[short], yr[short]), name


def _small_csr(kind):
    rng = np.random.default_rng(11)
    if kind == "one":
        return H.CSR.from_arrays(np.array([0, 1], np.uint32), np.array([0], np.uint32),
                                 np.array([2.5]))
    if kind == "three":
        return H.CSR.from_arrays(np.array([0, 1, 2, 3], np.uint32),
                                 np.array([0, 2, 1], np.uint32), np.array([1.5, -2.0, 0.25]))
    # odd entry count, rows of 0..40 entries, a long row in the last block
    n = 1001
    lens = rng.integers(0, 40, size=n)
    lens[-1] = 301
    if lens.sum() % 2 == 0:
        lens[3] += 1
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, n, size=l)) for l in lens]).astype(np.uint32)
    return H.CSR.from_arrays(row, col, rng.standard_normal(len(col)))


@pytest.mark.parametrize("kind", ["one", "three", "odd"])
def test_csr_block_odd_entry_counts(engine, kind):
    """Pairwise loads never read past the arrays; the lone last entry of an
    odd count is substituted; unaligned user arrays take the one-entry path."""
    import ctypes as C

    import torch
    m = _small_csr(kind)
    assert m.nnz % 2 == 1
    x = H.uniform(m.nrows, seed=5)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    d = D.DeviceCSR(engine, m)
    xd, y = engine.upload(x), _y(engine, m.nrows)
    d.spmv(xd, y)
    g = _host(y, m.nrows)
    short = np.diff(m.row.astype(np.int64)) <= 256
    assert bitwise_equal(g[short], yr[short]) and scaled(m, x, yr, g) <= TOL64
    # the same arrays one element into a larger buffer: 8-byte aligned values
    vbuf = torch.zeros(m.nnz + 1, dtype=torch.float64, device=engine.torch_device)
    cbuf = torch.zeros(m.nnz + 1, dtype=torch.int32, device=engine.torch_device)
    vbuf[1:] = d.val
    cbuf[1:] = d.col
    y2 = _y(engine, m.nrows)
    st = D.lib().spmv_csr(engine.h, d.plan.h, C.c_void_p(vbuf.data_ptr() + 8),
                          C.c_void_p(cbuf.data_ptr() + 4), D._dptr(d.row), D._dptr(xd),
                          D._dptr(y2))
    assert st == 0
    assert bitwise_equal(_host(y2, m.nrows), g)


# ---------------------------------------------------------------- AXC / K1


def test_axc(engine, lap256, mats, tiny10):
    fx, tm, tx = tiny10
    cases = [("lap256",) + lap256, ("tiny10", tm, tx)] + [(n, m, x) for n, m, x, _ in mats]
    for name, m, x in cases:
        a = H.AXC(m, x)
        d = D.DeviceAXC(engine, a)
        y = _y(engine, m.nrows)
        d.int_axc(y)
        g = _host(y, m.nrows)
        ref = O.int_axc(m.nrows, a.ax, a.brp)
        long_rows = np.diff(a.brp.astype(np.int64)) // 16 > 48  # rows on waves of their own
        assert bitwise_equal(g[~long_rows], ref[~long_rows]), name
        assert scaled(m, x, O.cf_csr(m.row, m.col, m.val, x), g) <= TOL64, name


def test_k1(engine, lap256, mats, tiny10):
    fx, tm, tx = tiny10
    cases = [("lap256",) + lap256, ("tiny10", tm, tx)] + [(n, m, x) for n, m, x, _ in mats]
    for name, m, x in cases:
        k = H.K1(m)
        d = D.DeviceK1(engine, k)
        xd = engine.upload(x)
        res = _y(engine, k.chunkNum * 8)
        d.int_k1(xd, res)
        y = _y(engine, m.nrows)
        d.order(res, y)
        y2 = _y(engine, m.nrows)
        d.int_k1p(xd, y2)
        ref_res = O.int_k1(k.chunkNum, k.val, k.col, k.chp, k.nmc, x)
        g_res = _host(res, k.chunkNum * 8)
        split = np.repeat(k.nmc > 64, 8)  # wide chunks: tree of eight step slices
        assert bitwise_equal(g_res[~split], ref_res[~split]), name
        yr = O.cf_csr(m.row, m.col, m.val, x)
        assert scaled(m, x, yr, _host(y, m.nrows)) <= TOL64, name
        assert scaled(m, x, yr, _host(y2, m.nrows)) <= TOL64, name


# ---------------------------------------------------------------- AXT


@pytest.mark.parametrize("th", [1, 2, 4, 8, 16, 32, 48])
def test_axt_lap256(engine, lap256, th):
    m, x = lap256
    a = H.AXT(m, x, th=th)
    d = D.DeviceAXT(engine, a)
    y = _y(engine, m.nrows)
    d.int_axt(y)
    g = _host(y, m.nrows)
    assert bitwise_equal(g, O.int_axt(a))
    assert scaled(m, x, O.cf_csr(m.row, m.col, m.val, x), g) <= TOL64


@pytest.mark.parametrize("th", [1, 2, 4, 8, 16, 32, 48])
def test_axt_mats(engine, mats, tiny10, th):
    fx, tm, tx = tiny10
    cases = [("tiny10", tm, tx, O.cf_csr(tm.row, tm.col, tm.val, tx))] + list(mats)
    for name, m, x, yr in cases:
        a = H.AXT(m, x, th=th)
        d = D.DeviceAXT(engine, a)
        y = _y(engine, m.nrows)
        d.int_axt(y)
        g = _host(y, m.nrows)
        assert scaled(m, x, yr, g) <= TOL64, (name, th)
        # rows that fit one wave block have the oracle's bits
        per = 8 if th == 1 else th
        units = (np.diff(m.row.astype(np.int64)) + per - 1) // per
        # long-row thresholds of the plans (th = 8: 64-column blocks on dense
        # rows; th >= 16: 32-column blocks on dense rows)
        lim = 48 if th == 1 else max(4, min(32 if th == 8 else 16 if th >= 16 else 64, 512 // th))
        ok = units <= lim
        assert bitwise_equal(g[ok], O.int_axt(a)[ok]), (name, th)


def test_axt_tiny10_exact(engine, tiny10):
    fx, m, x = tiny10
    for th in (1, 2):
        a = H.AXT(m, x, th=th)
        d = D.DeviceAXT(engine, a)
        y = _y(engine, m.nrows)
        d.int_axt(y)
        assert np.array_equal(_host(y, m.nrows), np.array(fx["y_exact"], np.float64))


@pytest.mark.parametrize("th", [1, 2, 3, 4, 8, 16])
@pytest.mark.parametrize("vt,xt,tol", [(D.F32, D.F32, TOL32), (D.BF16, D.F32, TOL16)])
def test_axt_precision(engine, mats, th, vt, xt, tol):
    for name, m, x, yr in mats:
        a = H.AXT(m, x, th=th)
        d = D.DeviceAXT(engine, a, val_type=vt, x_type=xt)
        y = _y(engine, m.nrows, np.float32)
        d.int_axt(y)
        g = _host(y, m.nrows)
        e = scaled(m, x, yr, g)
        assert e <= tol, (name, th, vt, e)


def test_axt_repeat_is_deterministic(engine, mats):
    name, m, x, yr = mats[3]
    a = H.AXT(m, x, th=1)
    d = D.DeviceAXT(engine, a)
    outs = []
    for _ in range(3):
        y = _y(engine, m.nrows)
        d.int_axt(y)
        outs.append(_host(y, m.nrows))
    assert bitwise_equal(outs[0], outs[1]) and bitwise_equal(outs[1], outs[2])
    assert d.plan.info()["nsplit_rows"] > 0


@pytest.mark.parametrize("bs", [2, 8, 128, 256, 512])
def test_axt_com(engine, lap256, mats, tiny10, bs):
    """COM is checked against CSR-naive (the reference's COM kernel drops row 0
    segments whose head word is 0 and sums by prefix differences); rows that
    fit one block are bit-identical to cf_CSR."""
    fx, tm, tx = tiny10
    cases = [("tiny10", tm, tx, O.cf_csr(tm.row, tm.col, tm.val, tx)),
             ("lap256",) + lap256 + (O.cf_csr(lap256[0].row, lap256[0].col, lap256[0].val, lap256[1]),)]
    cases += list(mats)
    for name, m, x, yr in cases:
        a = H.AXT(m, x, th=1, mode="COM", bs=bs)
        d = D.DeviceAXT(engine, a)
        y = _y(engine, m.nrows)
        d.int_axt(y)
        g = _host(y, m.nrows)
        assert scaled(m, x, yr, g) <= TOL64, (name, bs)
        short = np.diff(m.row.astype(np.int64)) <= 128
        assert bitwise_equal(g[short], yr[short]), (name, bs)


def test_axt_com_f32(engine, mats):
    for name, m, x, yr in mats:
        a = H.AXT(m, x, th=1, mode="COM", bs=128)
        d = D.DeviceAXT(engine, a, val_type=D.F32, x_type=D.F32)
        y = _y(engine, m.nrows, np.float32)
        d.int_axt(y)
        assert scaled(m, x, yr, _host(y, m.nrows)) <= TOL32, name
