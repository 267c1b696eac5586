"""The oracle -- TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference's kernels (oracle.c, plain C) and of its
CSR -> AXC / K1 / AXT builders (builders_np.py, numpy).  It is the checker
of the GPU library and the CPU baseline timed by bench.py; the product
(int_spmv_amd/) never imports it.

Pinning (what makes the restatement trustworthy without running the
reference, which cannot be compiled here: it needs <mkl.h>):
  * cf_CSR       -- closed forms on the Laplacian, dense products on small
                    integer matrices, and an FMA-sensitive case that only a
                    fused multiply-add per term reproduces;
  * reduce8      -- a case whose result differs for every association other
                    than the halving tree of _mm512_reduce_add_pd;
  * int_axc / int_k1(p) / int_axt_noc(_h1) -- exact integer data against
                    dense products, the hand-derived 10x10 fixture
                    (tests/golden/tiny10.json), and bitwise agreement with
                    cf_CSR where the orders coincide;
  * int_axt_com_h1 -- the 10x10 fixture, whose y[0] the reference drops
                    (its sec word for row 0 is 0); every other row exact;
  * builders     -- the closed-form lengths of the 256x256 Laplacian and the
                    10x10 fixture, plus slot-by-slot invariants.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

_vp = C.c_void_p
_u32 = C.c_uint32


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build()
        L = C.CDLL(path)
        sig = {
            "orc_cf_csr": [_u32, _vp, _vp, _vp, _vp, _vp],
            "orc_get_errors": [_u32, _vp, _vp, _vp, _vp, _vp],
            "orc_scaled_err": [_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
            "orc_csr_transpose_product": [_u32, _vp, _vp, _vp, _vp, _vp],
            "orc_csr_mm": [_u32, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _u32],
            "orc_scaled_err_mm": [_u32, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _u32, _vp],
            "orc_reduce8": [_vp],
            "orc_int_axc": [_u32, _vp, _vp, _vp],
            "orc_int_k1": [_u32, _vp, _vp, _vp, _vp, _vp, _vp],
            "orc_order_k1": [_u32, _vp, _vp, _vp],
            "orc_int_k1p": [_u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
            "orc_int_axt_noc_h1": [_u32, _vp, _vp, _vp],
            "orc_int_axt_noc": [_u32, _u32, _vp, _vp, _vp],
            "orc_int_axt_com_h1": [_u32, _u32, _u32, _u32, _u32, _u32, _u32, _vp, _vp, _vp],
            "orc_bench": [C.c_int, C.c_int, C.c_int, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp,
                          _vp, _vp],
        }
        for k, a in sig.items():
            getattr(L, k).argtypes = a
            getattr(L, k).restype = None
        L.orc_scaled_err.restype = C.c_double
        L.orc_scaled_err_mm.restype = C.c_double
        L.orc_reduce8.restype = C.c_double
        L.orc_bench.restype = C.c_double
        _lib = L
    return _lib


_lib512 = None


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return any(line.startswith("flags") and " avx512f" in line for line in f)
    except OSError:
        return False


def lib_avx512():
    """The AVX-512 CPU baseline (cpu_avx512.c: orc_bench_avx512, the
    orc_bench protocol), or None when the host CPU has no AVX-512."""
    global _lib512
    if _lib512 is None:
        if not host_has_avx512():
            return None
        path = os.path.join(HERE, "liboracle_avx512.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build_avx512()
        L = C.CDLL(path)
        L.orc_bench_avx512.argtypes = [C.c_int, C.c_int, C.c_int, _u32, _u32, _u32, _vp, _vp,
                                       _vp, _vp, _vp, _vp, _vp]
        L.orc_bench_avx512.restype = C.c_double
        _lib512 = L
    return _lib512


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def cf_csr(row, col, val, x):
    row, col, val, x = _c(row, np.uint32), _c(col, np.uint32), _c(val, np.float64), _c(x, np.float64)
    n = len(row) - 1
    y = np.zeros(n, np.float64)
    lib().orc_cf_csr(n, _p(row), _p(col), _p(val), _p(x), _p(y))
    return y


def get_errors(ar, ac):
    ar, ac = _c(ar, np.float64), _c(ac, np.float64)
    a, r, p = C.c_double(), C.c_double(), C.c_uint32()
    lib().orc_get_errors(len(ar), _p(ar), _p(ac), C.byref(a), C.byref(r), C.byref(p))
    return a.value, r.value, p.value


def scaled_err(row, col, val, x, yref, y):
    row, col = _c(row, np.uint32), _c(col, np.uint32)
    val, x = _c(val, np.float64), _c(x, np.float64)
    yref, y = _c(yref, np.float64), _c(y, np.float64)
    am = C.c_uint32()
    e = lib().orc_scaled_err(len(row) - 1, _p(row), _p(col), _p(val), _p(x), _p(yref), _p(y),
                             C.byref(am))
    return e, am.value


def csr_mm(row, col, val, X):
    """cf_CSR of every column of X (n x k): Y (n x k), each column with the
    bits of cf_csr on that column."""
    row, col, val = _c(row, np.uint32), _c(col, np.uint32), _c(val, np.float64)
    X = _c(X, np.float64)
    if X.ndim == 1:
        X = X[:, None]
    n, k = len(row) - 1, X.shape[1]
    Y = np.zeros((n, k), np.float64)
    lib().orc_csr_mm(n, _p(row), _p(col), _p(val), _p(X), k, k, _p(Y), k)
    return Y


def scaled_err_mm(row, col, val, X, Yref, Y):
    """Largest per-entry scaled error over all columns: (err, (row, col))."""
    row, col, val = _c(row, np.uint32), _c(col, np.uint32), _c(val, np.float64)
    X, Yref, Y = _c(X, np.float64), _c(Yref, np.float64), _c(Y, np.float64)
    if X.ndim == 1:
        X, Yref, Y = X[:, None], Yref[:, None], Y[:, None]
    n, k = len(row) - 1, X.shape[1]
    am = C.c_uint64()
    e = lib().orc_scaled_err_mm(n, _p(row), _p(col), _p(val), _p(X), k, k, _p(Yref), _p(Y), k,
                                C.byref(am))
    return e, divmod(am.value, k)


def csr_transpose_product(row, col, val, ybar):
    row, col, val, ybar = (_c(row, np.uint32), _c(col, np.uint32), _c(val, np.float64),
                           _c(ybar, np.float64))
    n = len(row) - 1
    out = np.zeros(n, np.float64)
    lib().orc_csr_transpose_product(n, _p(row), _p(col), _p(val), _p(ybar), _p(out))
    return out


def reduce8(p):
    p = _c(p, np.float64)
    return lib().orc_reduce8(_p(p))


def int_axc(nrows, ax, brp):
    ax, brp = _c(ax, np.float64), _c(brp, np.uint32)
    y = np.zeros(nrows, np.float64)
    lib().orc_int_axc(nrows, _p(ax), _p(brp), _p(y))
    return y


def int_k1(chunkNum, val, col, chp, nmc, x):
    val, col, chp, nmc, x = (_c(val, np.float64), _c(col, np.uint32), _c(chp, np.uint32),
                             _c(nmc, np.uint32), _c(x, np.float64))
    res = np.zeros(chunkNum * 8, np.float64)
    lib().orc_int_k1(chunkNum, _p(val), _p(col), _p(chp), _p(nmc), _p(x), _p(res))
    return res


def order_k1(nrows, permi, res):
    permi, res = _c(permi, np.uint32), _c(res, np.float64)
    y = np.zeros(nrows, np.float64)
    lib().orc_order_k1(nrows, _p(permi), _p(res), _p(y))
    return y


def int_k1p(chunkNum, nrows, val, col, chp, nmc, permi, x):
    val, col, chp, nmc, permi, x = (_c(val, np.float64), _c(col, np.uint32), _c(chp, np.uint32),
                                    _c(nmc, np.uint32), _c(permi, np.uint32), _c(x, np.float64))
    y = np.zeros(nrows + 1, np.float64)
    lib().orc_int_k1p(chunkNum, nrows, _p(val), _p(col), _p(chp), _p(nmc), _p(permi), _p(x), _p(y))
    return y[:nrows]


def int_axt_noc_h1(nrows, tileN, ax, sec):
    ax, sec = _c(ax, np.float64), _c(sec, np.uint32)
    y = np.zeros(nrows, np.float64)
    lib().orc_int_axt_noc_h1(tileN, _p(ax), _p(sec), _p(y))
    return y


def int_axt_noc(nrows, tileN, th, ax, sec):
    ax, sec = _c(ax, np.float64), _c(sec, np.uint32)
    y = np.zeros(nrows, np.float64)
    lib().orc_int_axt_noc(tileN, th, _p(ax), _p(sec), _p(y))
    return y


def int_axt(a):
    """Dispatches on an AXT host object like the reference's main (NOC only)."""
    if a.mode != "NOC":
        raise ValueError("use int_axt_com_h1 for COM")
    if a.tileH == 1:
        return int_axt_noc_h1(a.nrows, a.tileN, a.ax, a.sec)
    return int_axt_noc(a.nrows, a.tileN, a.tileH, a.ax, a.sec)


def int_axt_com_h1(a):
    """The reference's COM kernel, defects included (see oracle.c)."""
    bs, thw, tn = a.bs, a.tileHW, a.tileN
    bn = (tn * 8 + bs - 1) // bs
    ax, sec = _c(a.ax, np.float64), _c(a.sec, np.uint32)
    y = np.zeros(a.nrows, np.float64)
    lib().orc_int_axt_com_h1(bn, bs, a.log, thw, tn, a.lenAX, a.lenSEC, _p(ax), _p(sec), _p(y))
    return y
