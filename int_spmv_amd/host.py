"""ctypes bindings of the C host library (include/intspmv_host.h).

Python objects here own the C structs and free them on deletion; array
fields are exposed as numpy views of the C memory.  Names follow the
reference's structs (str_matCSR, str_matAXC, str_matK1, str_matAXT in
intSpmv_main.c) and builder functions (getFormatData*).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._loader import load_library

u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)

HBRICK_SIZE = 8
CHUNK_SIZE = 8
ACOL_PAD = 0xFFFFFFFF


class HostError(RuntimeError):
    pass


class _CSR(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("nnz", C.c_uint32), ("rmin", C.c_uint32),
                ("rave", C.c_double), ("rmax", C.c_uint32), ("rsd", C.c_double),
                ("bw", C.c_uint32), ("val", f64p), ("row", u32p), ("col", u32p), ("rl", u32p)]


class FormatData(C.Structure):
    _fields_ = [("name_", C.c_char * 32), ("mfp", C.c_double), ("beta", C.c_double),
                ("ct", C.c_double)]

    @property
    def name(self) -> str:
        return self.name_.decode()


class _AXC(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("nnz", C.c_uint32), ("lenAX", C.c_uint32),
                ("lenBRP", C.c_uint32), ("ax", f64p), ("brp", u32p)]


class _K1(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("nnz", C.c_uint32), ("chunkNum", C.c_uint32),
                ("lenVC", C.c_uint32), ("permi", u32p), ("nmc", u32p), ("chp", u32p),
                ("val", f64p), ("col", u32p)]


class _AXT(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("nnz", C.c_uint32), ("mode_", C.c_char * 8),
                ("tileHW", C.c_uint32), ("tileH", C.c_uint32), ("logTH", C.c_uint32),
                ("tileN", C.c_uint32), ("lenAX", C.c_uint32), ("lenSEC", C.c_uint32),
                ("lenCON", C.c_uint32), ("log", C.c_uint32), ("ax", f64p), ("sec", u32p),
                ("con", u32p), ("bs", C.c_uint32), ("colN", C.c_uint32), ("acol", u32p),
                ("nzslot", u32p)]


class _TMAP(C.Structure):
    _fields_ = [("at", _AXT), ("lenPERM", C.c_uint32), ("permT", u32p), ("ct", C.c_double)]


class TuneChoice(C.Structure):
    _fields_ = [("tileH", C.c_uint32), ("kc", C.c_uint32), ("slots", C.c_uint64),
                ("bytes", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = load_library("libintspmv_host.so")
        L.intspmv_csr_read.argtypes = [C.c_char_p, C.POINTER(_CSR)]
        for f in ("intspmv_csr_write_bin", "intspmv_csr_write_csr", "intspmv_csr_write_mtx"):
            getattr(L, f).argtypes = [C.c_char_p, C.POINTER(_CSR)]
        L.intspmv_csr_alloc.argtypes = [C.POINTER(_CSR), C.c_uint32, C.c_uint32]
        L.intspmv_csr_free.argtypes = [C.POINTER(_CSR)]
        L.intspmv_csr_free.restype = None
        L.intspmv_csr_from_coo.argtypes = [C.c_uint32, C.c_size_t, u32p, u32p, f64p,
                                           C.POINTER(_CSR)]
        L.intspmv_csr_stats.argtypes = [C.POINTER(_CSR), C.c_char_p]
        L.intspmv_csr_transpose.argtypes = [C.POINTER(_CSR), C.POINTER(_CSR), u32p]
        L.intspmv_csr_format_data.argtypes = [C.POINTER(_CSR), C.POINTER(FormatData)]
        L.intspmv_csr_format_data.restype = None
        L.intspmv_axc_build.argtypes = [C.POINTER(_CSR), f64p, C.POINTER(_AXC),
                                        C.POINTER(FormatData)]
        L.intspmv_axc_free.argtypes = [C.POINTER(_AXC)]
        L.intspmv_axc_free.restype = None
        L.intspmv_k1_build.argtypes = [C.POINTER(_CSR), C.POINTER(_K1), C.POINTER(FormatData)]
        L.intspmv_k1_free.argtypes = [C.POINTER(_K1)]
        L.intspmv_k1_free.restype = None
        L.intspmv_axt_build.argtypes = [C.POINTER(_CSR), f64p, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_char_p, C.c_int, C.POINTER(_AXT),
                                        C.POINTER(FormatData)]
        L.intspmv_axt_free.argtypes = [C.POINTER(_AXT)]
        L.intspmv_axt_free.restype = None
        L.intspmv_axt_refresh_x.argtypes = [C.POINTER(_AXT), f64p]
        L.intspmv_axt_value_plane.argtypes = [C.POINTER(_AXT), f64p]
        L.intspmv_axt_transpose_map.argtypes = [C.POINTER(_CSR), C.POINTER(_AXT),
                                                C.POINTER(_TMAP)]
        L.intspmv_axt_tmap_free.argtypes = [C.POINTER(_TMAP)]
        L.intspmv_axt_tmap_free.restype = None
        L.intspmv_fill_uniform.argtypes = [f64p, C.c_size_t, C.c_double, C.c_double,
                                           C.c_uint64]
        L.intspmv_fill_uniform.restype = None
        L.intspmv_fill_normal.argtypes = [f64p, C.c_size_t, C.c_double, C.c_double,
                                          C.c_uint64]
        L.intspmv_fill_normal.restype = None
        L.intspmv_gen_laplace2d.argtypes = [C.c_uint32, C.POINTER(_CSR)]
        L.intspmv_gen_powerlaw.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_uint32,
                                           C.c_double, C.c_uint32, C.c_uint64,
                                           C.POINTER(_CSR)]
        L.intspmv_gen_stencil27.argtypes = [C.c_uint32, C.c_double, C.c_uint64,
                                            C.POINTER(_CSR)]
        L.intspmv_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                       C.c_double, C.c_uint64, C.POINTER(_CSR)]
        L.intspmv_gen_banded.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                         C.c_uint32, C.c_uint64, C.POINTER(_CSR)]
        L.intspmv_wtime.restype = C.c_double
        L.intspmv_tune_axt.argtypes = [C.POINTER(_CSR), C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(TuneChoice), C.c_uint32, u32p]
        _lib = L
    return _lib


def _check(st: int, what: str) -> None:
    if st != 0:
        raise HostError(f"{what} failed with status {st}")


def _view(ptr, n: int, dtype) -> np.ndarray:
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class CSR:
    """CSR matrix (str_matCSR, intSpmv_main.c:183) owned by the C library."""

    def __init__(self, s: _CSR):
        self._s = s

    def __del__(self):
        if getattr(self, "_s", None) is not None and _lib is not None:
            _lib.intspmv_csr_free(C.byref(self._s))
            self._s = None

    # --- constructors -------------------------------------------------------
    @classmethod
    def read(cls, path: str) -> "CSR":
        s = _CSR()
        _check(lib().intspmv_csr_read(os.fsencode(path), C.byref(s)), f"read {path}")
        m = cls(s)
        m.stats()
        return m

    @classmethod
    def from_arrays(cls, row, col, val) -> "CSR":
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        n = len(row) - 1
        if n < 0 or row[0] != 0 or row[-1] != len(col) or len(col) != len(val):
            raise ValueError("inconsistent CSR arrays")
        if np.any(np.diff(row.astype(np.int64)) < 0) or (len(col) and col.max() >= n):
            raise ValueError("invalid CSR row pointers or column indices")
        s = _CSR()
        _check(lib().intspmv_csr_alloc(C.byref(s), n, len(col)), "csr alloc")
        m = cls(s)
        m.row[:] = row
        m.col[:] = col
        m.val[:] = val
        m.stats()
        return m

    @classmethod
    def from_coo(cls, n, ri, ci, v) -> "CSR":
        ri = np.ascontiguousarray(ri, dtype=np.uint32)
        ci = np.ascontiguousarray(ci, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        s = _CSR()
        _check(lib().intspmv_csr_from_coo(n, len(ri), _ptr(ri, C.c_uint32),
                                          _ptr(ci, C.c_uint32), _ptr(v, C.c_double),
                                          C.byref(s)), "csr from coo")
        m = cls(s)
        m.stats()
        return m

    @classmethod
    def laplace2d(cls, g: int) -> "CSR":
        s = _CSR()
        _check(lib().intspmv_gen_laplace2d(g, C.byref(s)), "laplace2d")
        return cls(s)

    @classmethod
    def powerlaw(cls, n=1 << 20, target_nnz=30_000_000, s=2.0, maxlen=8192, band_frac=0.8,
                 band_half=None, seed=0) -> "CSR":
        if band_half is None:
            band_half = int(0.025 * n)
        st = _CSR()
        _check(lib().intspmv_gen_powerlaw(n, target_nnz, s, maxlen, band_frac, band_half,
                                          seed, C.byref(st)), "powerlaw")
        return cls(st)

    @classmethod
    def stencil27(cls, g=128, sigma=0.01, seed=0) -> "CSR":
        s = _CSR()
        _check(lib().intspmv_gen_stencil27(g, sigma, seed, C.byref(s)), "stencil27")
        return cls(s)

    @classmethod
    def rmat(cls, scale=22, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=0) -> "CSR":
        s = _CSR()
        _check(lib().intspmv_gen_rmat(scale, edgefactor, a, b, c, seed, C.byref(s)), "rmat")
        return cls(s)

    @classmethod
    def banded(cls, n, lmin=4, lmax=64, band_frac=0.7, band_w=None, seed=0) -> "CSR":
        if band_w is None:
            band_w = max(2, int(0.05 * n))
        s = _CSR()
        _check(lib().intspmv_gen_banded(n, lmin, lmax, band_frac, band_w, seed, C.byref(s)),
               "banded")
        return cls(s)

    # --- fields -------------------------------------------------------------
    @property
    def nrows(self) -> int:
        return self._s.nrows

    @property
    def nnz(self) -> int:
        return self._s.nnz

    @property
    def row(self) -> np.ndarray:
        return _view(self._s.row, self.nrows + 1, np.uint32)

    @property
    def col(self) -> np.ndarray:
        return _view(self._s.col, self.nnz, np.uint32)

    @property
    def val(self) -> np.ndarray:
        return _view(self._s.val, self.nnz, np.float64)

    @property
    def rl(self) -> np.ndarray:
        return _view(self._s.rl, self.nrows, np.uint32)

    def stats(self, mat_file_name: str | None = None) -> dict:
        """Row statistics (printMatrixStats, intSpmv_main.c:253-302); with a
        matrix file name also writes <name>.sta beside it."""
        _check(lib().intspmv_csr_stats(C.byref(self._s),
                                       os.fsencode(mat_file_name) if mat_file_name else None),
               "stats")
        s = self._s
        return dict(nrows=s.nrows, nnz=s.nnz, rmin=s.rmin, rave=s.rave, rmax=s.rmax,
                    rsd=s.rsd, bw=s.bw)

    def transpose(self):
        """(A^T, perm): perm[k'] = index in A of entry k' of A^T."""
        s = _CSR()
        perm = np.empty(max(self.nnz, 1), np.uint32)
        _check(lib().intspmv_csr_transpose(C.byref(self._s), C.byref(s), _ptr(perm, C.c_uint32)),
               "transpose")
        return CSR(s), perm[: self.nnz]

    def format_data(self) -> FormatData:
        fd = FormatData()
        lib().intspmv_csr_format_data(C.byref(self._s), C.byref(fd))
        return fd

    def write(self, path: str) -> None:
        f = {".bin": "intspmv_csr_write_bin", ".csr": "intspmv_csr_write_csr",
             ".mtx": "intspmv_csr_write_mtx"}[os.path.splitext(path)[1]]
        _check(getattr(lib(), f)(os.fsencode(path), C.byref(self._s)), f"write {path}")


def uniform(n: int, lo=-1.0, hi=1.0, seed=0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().intspmv_fill_uniform(_ptr(out, C.c_double), n, lo, hi, seed)
    return out


def normal(n: int, mean=0.0, sigma=1.0, seed=0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().intspmv_fill_normal(_ptr(out, C.c_double), n, mean, sigma, seed)
    return out


class AXC:
    """AXC format (str_matAXC, intSpmv_main.c:615), built by getFormatDataAXC."""

    def __init__(self, m: CSR, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if len(x) != m.nrows:
            raise ValueError("x must have nrows entries")
        self._s, self.fd = _AXC(), FormatData()
        _check(lib().intspmv_axc_build(C.byref(m._s), _ptr(x, C.c_double), C.byref(self._s),
                                       C.byref(self.fd)), "axc build")

    def __del__(self):
        if getattr(self, "_s", None) is not None and _lib is not None:
            _lib.intspmv_axc_free(C.byref(self._s))
            self._s = None

    nrows = property(lambda self: self._s.nrows)
    nnz = property(lambda self: self._s.nnz)
    lenAX = property(lambda self: self._s.lenAX)
    lenBRP = property(lambda self: self._s.lenBRP)
    ax = property(lambda self: _view(self._s.ax, self._s.lenAX, np.float64))
    brp = property(lambda self: _view(self._s.brp, self._s.lenBRP, np.uint32))


class K1:
    """K1 format (str_matK1, intSpmv_main.c:771), built by getFormatDataK1."""

    def __init__(self, m: CSR):
        self._s, self.fd = _K1(), FormatData()
        _check(lib().intspmv_k1_build(C.byref(m._s), C.byref(self._s), C.byref(self.fd)),
               "k1 build")

    def __del__(self):
        if getattr(self, "_s", None) is not None and _lib is not None:
            _lib.intspmv_k1_free(C.byref(self._s))
            self._s = None

    nrows = property(lambda self: self._s.nrows)
    nnz = property(lambda self: self._s.nnz)
    chunkNum = property(lambda self: self._s.chunkNum)
    lenVC = property(lambda self: self._s.lenVC)
    permi = property(lambda self: _view(self._s.permi, (self._s.chunkNum + 1) * CHUNK_SIZE,
                                        np.uint32))
    nmc = property(lambda self: _view(self._s.nmc, self._s.chunkNum, np.uint32))
    chp = property(lambda self: _view(self._s.chp, self._s.chunkNum, np.uint32))
    val = property(lambda self: _view(self._s.val, self._s.lenVC, np.float64))
    col = property(lambda self: _view(self._s.col, self._s.lenVC, np.uint32))


class AXT:
    """AXT format (str_matAXT, intSpmv_main.c:1114), built by getFormatDataAXT.

    ``mode`` is "NOC" (any tile height) or "COM" (tile height 1, block size bs).
    ``acol`` (column of every slot) is built when ``want_acol`` is true.
    """

    def __init__(self, m: CSR, x: np.ndarray, th: int = 1, mode: str = "NOC", bs: int = 128,
                 thw: int = HBRICK_SIZE, want_acol: bool = False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if len(x) != m.nrows:
            raise ValueError("x must have nrows entries")
        self._s, self.fd = _AXT(), FormatData()
        _check(lib().intspmv_axt_build(C.byref(m._s), _ptr(x, C.c_double), bs, thw, th,
                                       mode.encode(), int(want_acol), C.byref(self._s),
                                       C.byref(self.fd)), f"axt build {mode} th={th}")

    _owner = None  # set on views of a struct that another object frees

    @classmethod
    def _borrow(cls, s: _AXT, owner) -> "AXT":
        """A view of an AXT struct that ``owner`` owns (and frees)."""
        self = cls.__new__(cls)
        self._s, self._owner, self.fd = s, owner, FormatData()
        return self

    def __del__(self):
        if self._owner is not None:
            return
        if getattr(self, "_s", None) is not None and _lib is not None:
            _lib.intspmv_axt_free(C.byref(self._s))
            self._s = None

    def refresh_x(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().intspmv_axt_refresh_x(C.byref(self._s), _ptr(x, C.c_double)), "refresh")

    mode = property(lambda self: self._s.mode_.decode())
    nrows = property(lambda self: self._s.nrows)
    nnz = property(lambda self: self._s.nnz)
    tileHW = property(lambda self: self._s.tileHW)
    tileH = property(lambda self: self._s.tileH)
    logTH = property(lambda self: self._s.logTH)
    tileN = property(lambda self: self._s.tileN)
    lenAX = property(lambda self: self._s.lenAX)
    lenSEC = property(lambda self: self._s.lenSEC)
    lenCON = property(lambda self: self._s.lenCON)
    log = property(lambda self: self._s.log)
    bs = property(lambda self: self._s.bs)
    colN = property(lambda self: self._s.colN)
    ax = property(lambda self: _view(self._s.ax, self._s.lenAX, np.float64))
    sec = property(lambda self: _view(self._s.sec, self._s.lenSEC, np.uint32))
    con = property(lambda self: _view(self._s.con, self._s.lenCON, np.uint32))
    acol = property(lambda self: _view(self._s.acol, self._s.lenAX // 2, np.uint32))
    nzslot = property(lambda self: _view(self._s.nzslot, self._s.nnz, np.uint32))

    def value_plane(self) -> np.ndarray:
        """Stored values in slot order (slot = T*tileH*8 + R*8 + C),
        intspmv_axt_value_plane."""
        out = np.empty(max(self.lenAX // 2, 1), np.float64)
        _check(lib().intspmv_axt_value_plane(C.byref(self._s), _ptr(out, C.c_double)),
               "value plane")
        return out[: self.lenAX // 2]


class AxtTransposeMap:
    """The transposed tile map of an AXT NOC matrix (intspmv_axt_transpose_map):
    ``at`` is A^T's AXT NOC layout (A's tile height, with acol) and ``permT``
    maps each slot of A^T's layout to the slot of A's layout holding the same
    entry (ACOL_PAD on padding).  ``a`` must be built with want_acol."""

    def __init__(self, m: CSR, a: AXT):
        self._s = _TMAP()
        _check(lib().intspmv_axt_transpose_map(C.byref(m._s), C.byref(a._s), C.byref(self._s)),
               "axt transpose map")
        self.at = AXT._borrow(self._s.at, self)

    def __del__(self):
        if getattr(self, "_s", None) is not None and _lib is not None:
            _lib.intspmv_axt_tmap_free(C.byref(self._s))
            self._s = None

    permT = property(lambda self: _view(self._s.permT, self._s.lenPERM, np.uint32))
    ct = property(lambda self: self._s.ct)


def tune_model(m: CSR, k: int, sv: int = 8, sx: int = 8) -> list[dict]:
    """Candidate (tileH, kc) layouts ranked by modelled bytes (intspmv_tune_axt)."""
    out = (TuneChoice * 32)()
    n = C.c_uint32()
    _check(lib().intspmv_tune_axt(C.byref(m._s), k, sv, sx, out, 32, C.byref(n)), "tune")
    return [dict(tileH=o.tileH, kc=o.kc, slots=o.slots, bytes=o.bytes) for o in out[: n.value]]


def wtime() -> float:
    return lib().intspmv_wtime()
