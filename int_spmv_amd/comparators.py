"""Vendor comparators, run and timed only; never part of the product path.

* MKL's CSR kernels on the host, as the reference's `test_mkl_CSR`,
  `test_sp_mkl_CSR` and `test_sp_insp_mkl_CSR` call them (intSpmv_main.c:461-569):
  `mkl_cspblas_dcsrgemv`, then `mkl_sparse_d_mv` without and with the
  inspector-executor hint + `mkl_sparse_optimize`.  MKL is loaded from
  `libmkl_rt.so` by name with its GNU threading layer (the process already
  holds libgomp through torch and the oracle).
* rocSPARSE's CSR SpMV on the GPU (`rocsparse_spmv`, generic API, 32-bit
  indices), every CSR algorithm the library offers, with its preprocessing
  done once outside the timed region.

Each function returns None when its library cannot be loaded.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
import time

import numpy as np

_MKL_PATHS = ("libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
              "/opt/conda/lib/libmkl_rt.so.1")
_mkl = None


class _MklDescr(C.Structure):
    # struct matrix_descr {sparse_matrix_type_t type; sparse_fill_mode_t mode; sparse_diag_type_t diag;}
    _fields_ = [("type", C.c_int), ("mode", C.c_int), ("diag", C.c_int)]


_SPARSE_OPERATION_NON_TRANSPOSE = 10
_SPARSE_MATRIX_TYPE_GENERAL = 20
_SPARSE_FILL_MODE_LOWER = 40
_SPARSE_DIAG_NON_UNIT = 50
_SPARSE_INDEX_BASE_ZERO = 0
_MKL_THREADING_GNU = 3  # MKL_THREADING_INTEL 0, SEQUENTIAL 1, PGI 2, GNU 3


def _load_mkl():
    global _mkl
    if _mkl is not None:
        return _mkl or None
    names = list(_MKL_PATHS)
    found = ctypes.util.find_library("mkl_rt")
    if found:
        names.insert(0, found)
    for nm in names:
        try:
            lib = C.CDLL(nm)
        except OSError:
            continue
        lib.MKL_Set_Threading_Layer.argtypes = [C.c_int]
        lib.MKL_Set_Threading_Layer(_MKL_THREADING_GNU)
        lib.MKL_Set_Num_Threads.argtypes = [C.c_int]
        lib.mkl_cspblas_dcsrgemv.restype = None
        lib.mkl_sparse_d_mv.argtypes = [C.c_int, C.c_double, C.c_void_p, _MklDescr, C.c_void_p,
                                        C.c_double, C.c_void_p]
        lib.mkl_sparse_set_mv_hint.argtypes = [C.c_void_p, C.c_int, _MklDescr, C.c_int]
        lib.mkl_sparse_d_create_csr.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.mkl_sparse_optimize.argtypes = [C.c_void_p]
        lib.mkl_sparse_destroy.argtypes = [C.c_void_p]
        _mkl = lib
        return lib
    _mkl = False
    return None


def mkl_available() -> bool:
    return _load_mkl() is not None


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def mkl_csr(row, col, val, x, threads: int, budget_s: float = 2.0):
    """Times the reference's three MKL CSR variants on `threads` host threads.

    Each variant runs until about `budget_s` seconds of calls have been timed
    (per-call wall time summed, as the reference's test loops do).  Returns
    {name: {"ms": per-call ms, "calls": count, "y": result}} or None.
    """
    L = _load_mkl()
    if L is None:
        return None
    n = len(row) - 1
    if len(val) >= 2**31 or n >= 2**31:
        return None
    row32 = np.ascontiguousarray(row, dtype=np.int32)
    col32 = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    L.MKL_Set_Num_Threads(int(threads))
    out = {}

    def run(name, call):
        y = np.zeros(n)
        t0 = time.perf_counter()
        if call(y) != 0:
            raise RuntimeError(f"MKL {name} failed")
        one = max(time.perf_counter() - t0, 1e-6)
        calls = int(max(1, min(250, budget_s / one)))
        tot = 0.0
        for _ in range(calls):
            t0 = time.perf_counter()
            call(y)
            tot += time.perf_counter() - t0
        out[name] = {"ms": tot / calls * 1e3, "calls": calls, "y": y}

    trans = C.c_char(b"N")
    nn = C.c_int(n)
    run("mkl_CSR", lambda y: L.mkl_cspblas_dcsrgemv(C.byref(trans), C.byref(nn), _p(val),
                                                     _p(row32), _p(col32), _p(x), _p(y)) or 0)
    descr = _MklDescr(_SPARSE_MATRIX_TYPE_GENERAL, _SPARSE_FILL_MODE_LOWER, _SPARSE_DIAG_NON_UNIT)
    row_end = row32[1:]
    for name, hint in (("sp_mkl_CSR", False), ("sp_insp_mkl_CSR", True)):
        h = C.c_void_p()
        st = L.mkl_sparse_d_create_csr(C.byref(h), _SPARSE_INDEX_BASE_ZERO, n, n, _p(row32),
                                       _p(row_end), _p(col32), _p(val))
        if st != 0:
            raise RuntimeError(f"mkl_sparse_d_create_csr: status {st}")
        try:
            if hint:
                L.mkl_sparse_set_mv_hint(h, _SPARSE_OPERATION_NON_TRANSPOSE, descr, 250)
                L.mkl_sparse_optimize(h)
            run(name, lambda y: L.mkl_sparse_d_mv(_SPARSE_OPERATION_NON_TRANSPOSE, 1.0, h, descr,
                                                  _p(x), 0.0, _p(y)))
        finally:
            L.mkl_sparse_destroy(h)
    return out


# ---- rocSPARSE -----------------------------------------------------------------

_ROCSPARSE_OPERATION_NONE = 111
_ROCSPARSE_INDEXTYPE_I32 = 2
_ROCSPARSE_INDEX_BASE_ZERO = 0
_ROCSPARSE_F64 = 152
_ROCSPARSE_F32 = 151
_ROCSPARSE_STAGE_BUFFER_SIZE = 1
_ROCSPARSE_STAGE_PREPROCESS = 2
_ROCSPARSE_STAGE_COMPUTE = 3
ROCSPARSE_CSR_ALGS = {"adaptive": 2, "rowsplit": 3, "lrb": 7, "nnzsplit": 8}
_rs = None


def _load_rocsparse():
    global _rs
    if _rs is not None:
        return _rs or None
    cands = ["librocsparse.so", "librocsparse.so.1", "/opt/rocm/lib/librocsparse.so"]
    rocm = os.environ.get("ROCM_PATH")
    if rocm:
        cands.insert(0, os.path.join(rocm, "lib", "librocsparse.so"))
    for nm in cands:
        try:
            lib = C.CDLL(nm)
        except OSError:
            continue
        vp = C.c_void_p
        lib.rocsparse_create_handle.argtypes = [C.POINTER(vp)]
        lib.rocsparse_destroy_handle.argtypes = [vp]
        lib.rocsparse_set_stream.argtypes = [vp, vp]
        lib.rocsparse_create_csr_descr.argtypes = [C.POINTER(vp), C.c_int64, C.c_int64, C.c_int64,
                                                   vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.rocsparse_destroy_spmat_descr.argtypes = [vp]
        lib.rocsparse_create_dnvec_descr.argtypes = [C.POINTER(vp), C.c_int64, vp, C.c_int]
        lib.rocsparse_destroy_dnvec_descr.argtypes = [vp]
        lib.rocsparse_spmv.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_size_t), vp]
        _rs = lib
        return lib
    _rs = False
    return None


def rocsparse_available() -> bool:
    return _load_rocsparse() is not None


def _chk(st, what):
    if st != 0:
        raise RuntimeError(f"rocSPARSE {what}: status {st}")


def rocsparse_csr(d_row, d_col, d_val, d_x, d_y, n: int, nnz: int, torch, steps=20, warmup=3,
                  algs=None):
    """Times rocsparse_spmv (y = A x, FP64 or FP32 from d_val's dtype) for each
    CSR algorithm on torch's current stream.  Device tensors: d_row (int32,
    n+1), d_col (int32, nnz), d_val, d_x, d_y.  Returns {alg: seconds per call}
    (None when the library is missing); an algorithm the library rejects is
    reported as its status code."""
    L = _load_rocsparse()
    if L is None:
        return None
    assert d_row.dtype == torch.int32 and d_col.dtype == torch.int32
    assert d_row.numel() == n + 1 and d_col.numel() == nnz and d_val.numel() == nnz
    dt = _ROCSPARSE_F64 if d_val.dtype == torch.float64 else _ROCSPARSE_F32
    one = C.c_double(1.0) if dt == _ROCSPARSE_F64 else C.c_float(1.0)
    zero = C.c_double(0.0) if dt == _ROCSPARSE_F64 else C.c_float(0.0)
    h, A, X, Y = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    _chk(L.rocsparse_create_handle(C.byref(h)), "create_handle")
    out = {}
    try:
        stream = torch.cuda.current_stream()
        _chk(L.rocsparse_set_stream(h, C.c_void_p(stream.cuda_stream)), "set_stream")
        _chk(L.rocsparse_create_csr_descr(C.byref(A), n, n, nnz, C.c_void_p(d_row.data_ptr()),
                                          C.c_void_p(d_col.data_ptr()),
                                          C.c_void_p(d_val.data_ptr()), _ROCSPARSE_INDEXTYPE_I32,
                                          _ROCSPARSE_INDEXTYPE_I32, _ROCSPARSE_INDEX_BASE_ZERO,
                                          dt), "create_csr_descr")
        _chk(L.rocsparse_create_dnvec_descr(C.byref(X), n, C.c_void_p(d_x.data_ptr()), dt), "x")
        _chk(L.rocsparse_create_dnvec_descr(C.byref(Y), n, C.c_void_p(d_y.data_ptr()), dt), "y")
        for name, alg in (algs or ROCSPARSE_CSR_ALGS).items():
            sz = C.c_size_t(0)

            def call(stage, buf):
                return L.rocsparse_spmv(h, _ROCSPARSE_OPERATION_NONE, C.byref(one), A, X,
                                        C.byref(zero), Y, dt, alg, stage, C.byref(sz), buf)
            st = call(_ROCSPARSE_STAGE_BUFFER_SIZE, None)
            if st != 0:
                out[name] = f"status {st}"
                continue
            buf = torch.empty(max(int(sz.value), 1), dtype=torch.uint8, device=d_val.device)
            bp = C.c_void_p(buf.data_ptr())
            st = call(_ROCSPARSE_STAGE_PREPROCESS, bp)
            if st != 0:
                out[name] = f"status {st}"
                continue
            for _ in range(warmup):
                _chk(call(_ROCSPARSE_STAGE_COMPUTE, bp), "spmv")
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(steps):
                call(_ROCSPARSE_STAGE_COMPUTE, bp)
            e1.record()
            torch.cuda.synchronize()
            out[name] = e0.elapsed_time(e1) * 1e-3 / steps
            del buf
    finally:
        for d, f in ((X, L.rocsparse_destroy_dnvec_descr), (Y, L.rocsparse_destroy_dnvec_descr),
                     (A, L.rocsparse_destroy_spmat_descr)):
            if d.value:
                f(d)
        L.rocsparse_destroy_handle(h)
    return out
