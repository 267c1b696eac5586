"""Pins the oracle (CPU restatement of the reference kernels) without running
the reference: closed forms, dense products on exact data, the hand-derived
10x10 fixture, and cases whose result depends on the exact rounding order
the oracle claims to follow."""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import host as H
from oracle import builders_np as B

from .util import csr_with_empty_rows


def _dense(m):
    A = np.zeros((m.nrows, m.nrows))
    for i in range(m.nrows):
        for k in range(m.row[i], m.row[i + 1]):
            A[i, m.col[k]] += m.val[k]
    return A


def _int_matrix(n, seed, maxlen=40):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen, size=n)
    row = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    col = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens]).astype(np.uint32)
    val = rng.integers(-9, 10, size=len(col)).astype(np.float64)
    x = rng.integers(-9, 10, size=n).astype(np.float64)
    return H.CSR.from_arrays(row, col, val), x


def test_cf_csr_laplacian_closed_forms():
    m = H.CSR.laplace2d(16)
    y = O.cf_csr(m.row, m.col, m.val, np.ones(m.nrows))
    g = 16
    for i in range(g):
        for j in range(g):
            edges = (i == 0) + (i == g - 1) + (j == 0) + (j == g - 1)
            assert y[i * g + j] == edges
    y = O.cf_csr(m.row, m.col, m.val, np.arange(m.nrows, dtype=np.float64))
    interior = [i * g + j for i in range(1, g - 1) for j in range(1, g - 1)]
    assert np.all(y[interior] == 0)


def test_cf_csr_dense_exact():
    for seed in range(3):
        m, x = _int_matrix(60, seed)
        assert np.array_equal(O.cf_csr(m.row, m.col, m.val, x), _dense(m) @ x)


def test_cf_csr_is_fused():
    # acc = -1 - 2^-29, then (1 + 2^-30)^2 = 1 + 2^-29 + 2^-60: only a fused
    # multiply-add keeps the 2^-60 term.
    row = np.array([0, 2], np.uint32)
    col = np.array([0, 1], np.uint32)
    a = 1 + 2.0 ** -30
    val = np.array([-1 - 2.0 ** -29, a])
    x = np.array([1.0, a])
    assert O.cf_csr(row, col, val, x)[0] == 2.0 ** -60


def test_reduce8_is_halving_tree():
    p = np.array([2.0 ** 53, -2.0 ** 53, 0, 0, 1, 1, 0, 0])
    assert O.reduce8(p) == 1.0            # ((p0+p4)+(p2+p6)) + ((p1+p5)+(p3+p7))
    assert sum(p.tolist()) == 2.0         # any left-to-right order gives 2


def test_kernels_exact_on_integers():
    for seed in range(3):
        m, x = _int_matrix(70, seed)
        yd = _dense(m) @ x
        a = H.AXC(m, x)
        assert np.array_equal(O.int_axc(m.nrows, a.ax, a.brp), yd)
        k = H.K1(m)
        res = O.int_k1(k.chunkNum, k.val, k.col, k.chp, k.nmc, x)
        assert np.array_equal(O.order_k1(m.nrows, k.permi, res), yd)
        assert np.array_equal(O.int_k1p(k.chunkNum, m.nrows, k.val, k.col, k.chp, k.nmc,
                                        k.permi, x), yd)
        for th in (1, 2, 3, 4, 8, 16):
            assert np.array_equal(O.int_axt(H.AXT(m, x, th=th)), yd), th


def test_tiny10_fixture(tiny10):
    fx, m, x = tiny10
    y = np.array(fx["y_exact"], np.float64)
    assert np.array_equal(O.cf_csr(m.row, m.col, m.val, x), y)
    for th in (1, 2, 4):
        assert np.array_equal(O.int_axt(H.AXT(m, x, th=th)), y)
    a = H.AXC(m, x)
    assert np.array_equal(O.int_axc(m.nrows, a.ax, a.brp), y)
    # the reference COM kernel drops row 0 (its sec word is 0); all else exact
    com = H.AXT(m, x, th=1, mode="COM", bs=8)
    assert np.array_equal(O.int_axt_com_h1(com), np.array(fx["com_bs8"]["y_reference_kernel"], np.float64))


def test_axt_th_orders_match_csr_on_laplacian(lap256):
    # products are exact (+-x, 4x) and a 5-entry row fits one column for
    # th >= 5, so the vertical fma chain equals CSR-naive bit for bit
    m, x = lap256
    yr = O.cf_csr(m.row, m.col, m.val, x)
    for th in (8, 16):
        assert np.array_equal(O.int_axt(H.AXT(m, x, th=th)), yr)
    k = H.K1(m)
    res = O.int_k1(k.chunkNum, k.val, k.col, k.chp, k.nmc, x)
    assert np.array_equal(O.order_k1(m.nrows, k.permi, res), yr)


def test_get_errors_compares_magnitudes():
    a, r, p = O.get_errors(np.array([1.0, -2.0, 4.0]), np.array([-1.0, 2.0, 3.0]))
    assert (a, r, p) == (1.0, 0.25, 2)


def test_scaled_err():
    row = np.array([0, 2, 2], np.uint32)
    col = np.array([0, 1], np.uint32)
    val = np.array([2.0, -1.0])
    x = np.array([1.0, 4.0])
    yref = np.array([-2.0, 0.0])
    e, arg = O.scaled_err(row, col, val, x, yref, np.array([-2.0 + 6e-3, 0.0]))
    assert arg == 0 and abs(e - 1e-3) < 1e-15
    e, arg = O.scaled_err(row, col, val, x, yref, np.array([-2.0, 1e-300]))
    assert e == np.inf and arg == 1  # empty row must be exactly zero


def test_transpose_product():
    m, x = _int_matrix(50, 9)
    yb = np.arange(50, dtype=np.float64) - 20
    assert np.array_equal(O.csr_transpose_product(m.row, m.col, m.val, yb), _dense(m).T @ yb)


def test_numpy_builders_agree_on_tiny10(tiny10):
    fx, m, x = tiny10
    b = B.axt(m.row, m.col, m.val, x, th=2)
    assert b["con"].tolist() == fx["noc_th2"]["con"]
    assert b["sec"].tolist() == fx["noc_th2"]["sec"]
    c = B.axt(m.row, m.col, m.val, x, th=1, mode="COM", bs=8)
    assert c["sec"].tolist() == fx["com_bs8"]["sec"]


def test_empty_rows_matrix_oracles_agree():
    m = csr_with_empty_rows()
    x = H.uniform(m.nrows, seed=4)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    for th in (1, 4):
        e, _ = O.scaled_err(m.row, m.col, m.val, x, yr, O.int_axt(H.AXT(m, x, th=th)))
        assert e <= 1e-12
    assert np.all(yr[np.diff(m.row.astype(np.int64)) == 0] == 0)


@pytest.mark.parametrize("kind", [1, 4, 5])
def test_cpu_baseline_port_matches_oracle(kind, lap256):
    import ctypes as C
    m, x = lap256
    L = O.lib()
    y = np.zeros(m.nrows)

    def p(a):
        return C.c_void_p(np.ascontiguousarray(a).ctypes.data)
    if kind == 1:
        a = H.AXC(m, x)
        ax, brp = np.ascontiguousarray(a.ax), np.ascontiguousarray(a.brp)
        L.orc_bench(1, 2, 2, m.nrows, 0, 0, p(ax), p(brp), None, None, None, None, p(y))
        ref = O.int_axc(m.nrows, a.ax, a.brp)
    else:
        th = 1 if kind == 4 else 4
        a = H.AXT(m, x, th=th)
        ax, sec = np.ascontiguousarray(a.ax), np.ascontiguousarray(a.sec)
        L.orc_bench(kind, 2, 2, m.nrows, a.tileN, th, p(ax), p(sec), None, None, None, None, p(y))
        ref = O.int_axt(a)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    assert O.scaled_err(m.row, m.col, m.val, x, yr, y)[0] <= 1e-12
    assert O.scaled_err(m.row, m.col, m.val, x, ref, y)[0] <= 1e-12


def _randn_matrix(n, seed, maxlen):
    """Random rows of 1..maxlen entries, values N(0,1), x ~ U[-1,1]: products
    are inexact, so every rounding step of a chain shows in its bits."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, maxlen + 1, size=n)
    row = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    col = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens]).astype(np.uint32)
    return H.CSR.from_arrays(row, col, rng.standard_normal(len(col))), rng.uniform(-1, 1, n)


def _mul_then_add_csr(m, x):
    """cf_CSR with the product rounded before the add (Python floats)."""
    y = np.zeros(m.nrows)
    for i in range(m.nrows):
        acc = 0.0
        for k in range(m.row[i], m.row[i + 1]):
            acc = acc + float(m.val[k]) * float(x[m.col[k]])
        y[i] = acc
    return y


def test_lanewise_chains_are_fused():
    # int_k1 / int_k1p keep each row's entries in CSR order in one lane, and
    # int_axt_noc with a tile height at least the longest row puts each row in
    # one column: their vtSum chains are then the same sequence of operations
    # as cf_CSR (padding adds 0*x, which is exact).  cf_CSR is pinned as fused
    # (test_cf_csr_is_fused), so these chains must be fused too; a
    # multiply-then-add chain gives other bits on this data (checked below),
    # so an oracle that rounded the products would fail here.
    m, x = _randn_matrix(400, 21, 48)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    ma = _mul_then_add_csr(m, x)
    assert np.count_nonzero(ma != yr) > 50, "data not rounding-sensitive"
    k = H.K1(m)
    res = O.int_k1(k.chunkNum, k.val, k.col, k.chp, k.nmc, x)
    assert np.array_equal(O.order_k1(m.nrows, k.permi, res), yr)
    assert np.array_equal(O.int_k1p(k.chunkNum, m.nrows, k.val, k.col, k.chp, k.nmc, k.permi, x), yr)
    for th in (48, 64):
        assert np.array_equal(O.int_axt(H.AXT(m, x, th=th)), yr), th


@pytest.mark.parametrize("kind", [0, 2, 3, 6])
def test_cpu_baseline_port_matches_oracle_more(kind):
    """The CPU-baseline kernels of cf_CSR, int_k1 (+orderArrayK1), int_k1p and
    int_axt_com_h1 give the oracle's results (K1: bit for bit; COM: its
    blocks' atomic order differs from the serial restatement's)."""
    import ctypes as C
    m, x = _randn_matrix(3000, 5, 40)
    L = O.lib()
    y = np.zeros(m.nrows)
    keep = []

    def p(a):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return C.c_void_p(a.ctypes.data)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    if kind == 0:
        L.orc_bench(0, 1, 2, m.nrows, 0, 0, p(m.row), p(m.col), p(m.val), p(x), None, None, p(y))
        assert np.array_equal(y, yr)
    elif kind in (2, 3):
        k = H.K1(m)
        L.orc_bench(kind, 3, 2, m.nrows, k.chunkNum, 0, p(k.val), p(k.col), p(k.chp), p(k.nmc),
                    p(x), p(k.permi), p(y))
        assert np.array_equal(y, yr)
    else:
        a = H.AXT(m, x, th=1, mode="COM", bs=128)
        bn = (a.tileN * a.tileHW + 127) // 128
        prm = np.array([128, a.log, a.tileHW, a.lenAX, a.lenSEC], np.uint32)
        L.orc_bench(6, 3, 2, m.nrows, bn, 1, p(a.ax), p(a.sec), p(prm), None, None, None, p(y))
        ref = O.int_axt_com_h1(a)
        assert O.scaled_err(m.row, m.col, m.val, x, ref, y)[0] <= 1e-12
        # the reference's COM drops row 0 (its head word is 0): everything else
        # is CSR-naive to rounding
        keep_rows = np.arange(1, m.nrows)
        assert np.max(np.abs(y - yr)[keep_rows]) <= 1e-12 * np.max(np.abs(yr))


def test_multi_rhs_checker_equals_single_column():
    """orc_csr_mm gives every column cf_CSR's bits; orc_scaled_err_mm is the
    largest of orc_scaled_err over the columns, at the same position."""
    m, _ = _randn_matrix(700, 31, 60)
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (m.nrows, 5))
    Y = O.csr_mm(m.row, m.col, m.val, X)
    for j in range(5):
        assert np.array_equal(Y[:, j], O.cf_csr(m.row, m.col, m.val, X[:, j]))
    Yp = Y.copy()
    Yp[123, 3] += 1e-9
    Yp[7, 1] += 1e-12
    e, (i, j) = O.scaled_err_mm(m.row, m.col, m.val, X, Y, Yp)
    per = [O.scaled_err(m.row, m.col, m.val, X[:, c], Y[:, c], Yp[:, c]) for c in range(5)]
    assert (i, j) == (123, 3) and e == per[3][0] == max(p[0] for p in per) and per[3][1] == 123
    Yp[0, 0] = np.nan
    assert O.scaled_err_mm(m.row, m.col, m.val, X, Y, Yp)[0] == np.inf


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5, 6])
def test_avx512_cpu_baseline_matches_oracle(kind):
    """The AVX-512 CPU baseline gives the scalar oracle's results: bit for bit
    where neither has atomics (cf_CSR, AXC, K1 and K1p: the same fused
    chains and halving trees), to 1e-12 where y is accumulated atomically in
    thread order (AXT), and COM with the reference's defects kept."""
    import ctypes as C
    L = O.lib_avx512()
    if L is None:
        pytest.skip("host CPU has no AVX-512")
    m, x = _randn_matrix(3001, 9, 70)
    y = np.zeros(m.nrows)
    keep = []

    def p(a):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return C.c_void_p(a.ctypes.data)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    if kind == 0:
        t = L.orc_bench_avx512(0, 1, 2, m.nrows, 0, 0, p(m.row), p(m.col), p(m.val), p(x), None,
                               None, p(y))
        ref, exact = yr, True
    elif kind == 1:
        a = H.AXC(m, x)
        t = L.orc_bench_avx512(1, 3, 2, m.nrows, 0, 0, p(a.ax), p(a.brp), None, None, None, None,
                               p(y))
        ref, exact = O.int_axc(m.nrows, a.ax, a.brp), True
    elif kind in (2, 3):
        k = H.K1(m)
        t = L.orc_bench_avx512(kind, 3, 2, m.nrows, k.chunkNum, 0, p(k.val), p(k.col), p(k.chp),
                               p(k.nmc), p(x), p(k.permi), p(y))
        ref, exact = yr, True  # K1 chains equal cf_CSR (test_lanewise_chains_are_fused)
    elif kind in (4, 5):
        th = 1 if kind == 4 else 4
        a = H.AXT(m, x, th=th)
        t = L.orc_bench_avx512(kind, 3, 2, m.nrows, a.tileN, th, p(a.ax), p(a.sec), None, None,
                               None, None, p(y))
        ref, exact = O.int_axt(a), False
    else:
        a = H.AXT(m, x, th=1, mode="COM", bs=128)
        bn = (a.tileN * a.tileHW + 127) // 128
        prm = np.array([128, a.log, a.tileHW, a.lenAX, a.lenSEC], np.uint32)
        t = L.orc_bench_avx512(6, 3, 2, m.nrows, bn, 1, p(a.ax), p(a.sec), p(prm), None, None,
                               None, p(y))
        ref, exact = O.int_axt_com_h1(a), False
    assert t > 0
    if exact:
        assert np.array_equal(y, ref), kind
    else:
        assert O.scaled_err(m.row, m.col, m.val, x, ref, y)[0] <= 1e-12, kind
