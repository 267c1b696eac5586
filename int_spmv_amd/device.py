"""Device side: ctypes bindings of libintspmv.so (include/intspmv.h).

Device memory and streams come from PyTorch (plumbing only): arrays are
uploaded as torch tensors on ``cuda:N`` and their raw pointers are passed to
the C ABI, which launches the hand-written gfx950 kernels on the context's
stream (torch's current stream unless set otherwise).

The classes mirror the reference's formats and kernels (intSpmv_main.c):

  DeviceCSR.cf_csr / .spmv         cf_CSR              (382-396)
  DeviceAXC.int_axc                int_axc             (711-732)
  DeviceK1.int_k1 / .int_k1p       int_k1, int_k1p     (943-1066)
  DeviceAXT.int_axt                int_axt_noc_h1 / int_axt_noc (1385-1474)

Every method overwrites y (y = A x).  There is no CPU fallback: without the
built library or a GPU the constructors raise.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from ._loader import load_library

F64, F32, BF16 = 0, 1, 2
NOC, COM = 0, 1
INTERLEAVED, PLANAR, COLMAP = 0, 1, 2

_vp = C.c_void_p
_u32 = C.c_uint32


class IntSpmvError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        msg = f"{what}: {lib().intspmv_strerror(status).decode()} ({status})"
        if detail:
            msg += f" [{detail}]"
        super().__init__(msg)
        self.status = status


class AxtDesc(C.Structure):
    _fields_ = [("nrows", _u32), ("nnz", _u32), ("tileHW", _u32), ("tileH", _u32),
                ("tileN", _u32), ("lenAX", _u32), ("lenSEC", _u32), ("colN", _u32),
                ("bs", _u32), ("log", _u32), ("mode", C.c_int32), ("val_type", C.c_int32),
                ("x_type", C.c_int32), ("layout", C.c_int32)]


# (name, restype, argtypes) of every exported symbol of include/intspmv.h
SIGNATURES = [
    ("intspmv_create", C.c_int, [C.POINTER(_vp), C.c_int]),
    ("intspmv_destroy", C.c_int, [_vp]),
    ("intspmv_set_stream", C.c_int, [_vp, _vp]),
    ("intspmv_sync", C.c_int, [_vp]),
    ("intspmv_strerror", C.c_char_p, [C.c_int]),
    ("intspmv_last_hip_error", C.c_char_p, [_vp]),
    ("intspmv_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("intspmv_malloc", C.c_int, [_vp, C.POINTER(_vp), C.c_size_t]),
    ("intspmv_free", C.c_int, [_vp, _vp]),
    ("intspmv_memcpy_htod", C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    ("intspmv_memcpy_dtoh", C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    ("intspmv_csr_plan", C.c_int, [_vp, _u32, _vp, C.POINTER(_vp)]),
    ("intspmv_csr_plan_grouped", C.c_int, [_vp, _u32, _vp, _u32, _vp, C.POINTER(_vp)]),
    ("intspmv_csr_plan_exact", C.c_int, [_vp, _u32, _vp, C.POINTER(_vp)]),
    ("spmv_csr_ordered", C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    ("spmv_csr", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("intspmv_axc_plan", C.c_int, [_vp, _u32, _vp, C.POINTER(_vp)]),
    ("spmv_axc", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("intspmv_k1_plan", C.c_int, [_vp, _u32, _u32, _vp, C.POINTER(_vp)]),
    ("spmv_k1", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("spmv_k1_order", C.c_int, [_vp, _u32, _vp, _vp, _vp]),
    ("intspmv_k1_build_device", C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          _u32, C.POINTER(_u32)]),
    ("intspmv_axt_plan", C.c_int, [_vp, C.POINTER(AxtDesc), _vp, C.POINTER(_vp)]),
    ("intspmv_axt_plan_grouped", C.c_int,
     [_vp, C.POINTER(AxtDesc), _vp, _u32, _vp, C.POINTER(_vp)]),
    ("intspmv_axt_plan_tuned", C.c_int,
     [_vp, C.POINTER(AxtDesc), _vp, _vp, _vp, _vp, _u32, _u32, C.POINTER(_vp),
      C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("spmv_axt", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("spmv_axt_gather", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _u32, _u32]),
    ("spmv_axt_gather_permuted", C.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _u32, _u32]),
    ("spmv_axt_grad_vals", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("spmv_axt_gather_grad", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("spmv_axt_grad_vals_embedded", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("spmv_permute_vals", C.c_int, [_vp, _u32, C.c_int, _vp, _vp, _vp]),
    ("spmv_axt_refresh_x", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("intspmv_batch_create", C.c_int, [_vp, C.c_int, C.c_int, _u32, _vp, C.POINTER(_vp)]),
    ("intspmv_batch_launch", C.c_int, [_vp, _vp]),
    ("intspmv_batch_graph_capture", C.c_int, [_vp, _vp]),
    ("intspmv_batch_graph_launch", C.c_int, [_vp, _vp]),
    ("intspmv_batch_info", C.c_int, [_vp, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    ("intspmv_batch_destroy", C.c_int, [_vp]),
    ("spmv_axt_h1_noalloc", C.c_int, [_vp, _u32, _u32, _vp, _vp, _vp]),
    ("intspmv_gather_rate", C.c_int, [_vp, C.c_size_t, _u32, C.c_size_t, _u32,
                                      C.POINTER(C.c_double)]),
    ("intspmv_stream_rate", C.c_int, [_vp, C.c_size_t, _u32, C.POINTER(C.c_double)]),
    ("intspmv_plan_destroy", C.c_int, [_vp]),
    ("intspmv_plan_info", C.c_int, [_vp, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    ("intspmv_plan_layout", C.c_int, [_u32, _vp, _u32, _u32, _u32, _u32, _vp, C.c_size_t,
                                      C.POINTER(_u32), _vp, C.c_size_t, C.POINTER(_u32),
                                      C.POINTER(_u32)]),
    ("intspmv_plan_layout_grouped", C.c_int,
     [_u32, _vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp, C.c_size_t, C.POINTER(_u32), _vp,
      C.c_size_t, C.POINTER(_u32), C.POINTER(_u32)]),
]


def plan_layout(uptr: np.ndarray, umax: int, rmax: int, long_units: int, long_max: int):
    """Host-only: (blocks[nb, 8], splits[ns, 4], nlong_wg) of a block plan."""
    uptr = np.ascontiguousarray(uptr, dtype=np.uint32)
    nb, ns, nl = _u32(), _u32(), _u32()
    n = len(uptr) - 1
    st = lib().intspmv_plan_layout(n, _np_ptr(uptr), umax, rmax, long_units, long_max, None, 0,
                                   C.byref(nb), None, 0, C.byref(ns), C.byref(nl))
    if st:
        raise IntSpmvError(st, "plan_layout")
    b = np.zeros((max(nb.value, 1), 8), np.uint32)
    s = np.zeros((max(ns.value, 1), 4), np.uint32)
    lib().intspmv_plan_layout(n, _np_ptr(uptr), umax, rmax, long_units, long_max, _np_ptr(b),
                              nb.value, C.byref(nb), _np_ptr(s), ns.value, C.byref(ns),
                              C.byref(nl))
    return b[: nb.value], s[: ns.value], nl.value

def plan_layout_grouped(uptr: np.ndarray, umax: int, rmax: int, long_units: int, long_max: int,
                        group_rows: np.ndarray | None, ngroups: int = 8):
    """Host-only: (blocks[nb, 8], splits[ns, 4], nlong_wg) of a plan whose row
    groups run on XCD g % 8 (group_rows None: ngroups groups of about equal
    units)."""
    uptr = np.ascontiguousarray(uptr, dtype=np.uint32)
    nb, ns, nl = _u32(), _u32(), _u32()
    n = len(uptr) - 1
    if group_rows is None:
        args = (n, _np_ptr(uptr), umax, rmax, long_units, long_max, ngroups, None)
    else:
        g = np.ascontiguousarray(group_rows, dtype=np.uint32)
        args = (n, _np_ptr(uptr), umax, rmax, long_units, long_max, len(g) - 1, _np_ptr(g))
    st = lib().intspmv_plan_layout_grouped(*args, None, 0, C.byref(nb), None, 0, C.byref(ns),
                                           C.byref(nl))
    if st:
        raise IntSpmvError(st, "plan_layout_grouped")
    b = np.zeros((max(nb.value, 1), 8), np.uint32)
    s = np.zeros((max(ns.value, 1), 4), np.uint32)
    lib().intspmv_plan_layout_grouped(*args, _np_ptr(b), nb.value, C.byref(nb), _np_ptr(s),
                                      ns.value, C.byref(ns), C.byref(nl))
    return b[: nb.value], s[: ns.value], nl.value


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = load_library("libintspmv.so")
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    n = C.c_int(0)
    lib().intspmv_device_count(C.byref(n))
    return n.value


def _torch():
    import torch
    return torch


def _np_ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _dptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Engine:
    """Owns an intspmv context on one GPU.  Launches go to the stream last
    given to :meth:`set_stream`: torch's current stream of the device at the
    time the Engine is made, until :meth:`follow_current_stream` or
    :meth:`set_stream` points it elsewhere."""

    def __init__(self, device: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("int_spmv_amd needs a HIP GPU; none is visible")
        self.device = device
        self.torch_device = torch.device("cuda", device)
        h = _vp()
        self._check(lib().intspmv_create(C.byref(h), device), "intspmv_create", ctx=False)
        self.h = h
        self.set_stream(torch.cuda.current_stream(self.torch_device))

    def _check(self, st: int, what: str, ctx: bool = True) -> None:
        if st != 0:
            detail = lib().intspmv_last_hip_error(self.h).decode() if ctx else ""
            raise IntSpmvError(st, what, detail)

    def set_stream(self, stream) -> None:
        self.stream = stream
        self._check(lib().intspmv_set_stream(self.h, _vp(stream.cuda_stream)), "set_stream")

    def follow_current_stream(self) -> None:
        """Points the context at torch's current stream of the device (if it
        is not there already), so that launches are ordered with the torch
        work around them."""
        cur = _torch().cuda.current_stream(self.torch_device)
        if cur.cuda_stream != self.stream.cuda_stream:
            self.set_stream(cur)

    def gather_rate(self, span_bytes: int, elem_bytes: int = 8, count: int = 1 << 25,
                    reps: int = 10) -> float:
        """Random gathers per second over a buffer of span_bytes
        (intspmv_gather_rate): the ceiling of gather-bound products."""
        r = C.c_double()
        self._check(lib().intspmv_gather_rate(self.h, span_bytes, elem_bytes, count, reps,
                                              C.byref(r)), "gather_rate")
        return r.value

    def stream_rate(self, nbytes: int, reps: int = 20) -> float:
        """Bytes per second of a plain read of an nbytes buffer
        (intspmv_stream_rate): the ceiling of the streaming products."""
        r = C.c_double()
        self._check(lib().intspmv_stream_rate(self.h, nbytes, reps, C.byref(r)), "stream_rate")
        return r.value

    def sync(self) -> None:
        self._check(lib().intspmv_sync(self.h), "sync")

    def upload(self, a: np.ndarray):
        torch = _torch()
        a = np.ascontiguousarray(a)
        if a.dtype == np.uint32:  # torch has no uint32 arithmetic; keep the bits
            t = torch.from_numpy(a.view(np.int32)).to(self.torch_device)
        elif a.dtype == np.uint16:
            t = torch.from_numpy(a.view(np.int16)).to(self.torch_device)
        else:
            t = torch.from_numpy(a).to(self.torch_device)
        return t

    def empty(self, n: int, dtype=np.float64):
        torch = _torch()
        tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
               np.dtype(np.uint32): torch.int32}[np.dtype(dtype)]
        return torch.empty(max(n, 1), dtype=tdt, device=self.torch_device)

    def __del__(self):
        # at interpreter exit the HIP runtime and the other objects go in no
        # fixed order: leave the memory to the process teardown
        if getattr(self, "h", None) and _lib is not None and not sys.is_finalizing():
            _lib.intspmv_destroy(self.h)
            self.h = None


class _Plan:
    def __init__(self, eng: Engine, h: _vp):
        self.eng, self.h = eng, h

    def info(self) -> dict:
        a, b, c = _u32(), _u32(), _u32()
        lib().intspmv_plan_info(self.h, C.byref(a), C.byref(b), C.byref(c))
        return dict(nblocks=a.value, nsplit_rows=b.value, nlong_rows=c.value)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None and not sys.is_finalizing():
            _lib.intspmv_plan_destroy(self.h)
            self.h = None


class DeviceCSR:
    def __init__(self, eng: Engine, m, placement: str = "plain"):
        self.eng, self.nrows, self.nnz = eng, m.nrows, m.nnz
        self.val = eng.upload(m.val)
        self.col = eng.upload(m.col)
        self.row = eng.upload(m.row)
        self._row_host = np.ascontiguousarray(m.row)
        self.plans = {}
        self.set_placement(placement)

    def set_placement(self, placement: str) -> None:
        """"plain" (intspmv_csr_plan) or "xcd" (intspmv_csr_plan_grouped, eight
        row groups of equal entries, each on one XCD); same bits."""
        if placement not in self.plans:
            h = _vp()
            row = self._row_host
            if placement == "plain":
                self.eng._check(lib().intspmv_csr_plan(self.eng.h, self.nrows, _np_ptr(row),
                                                       C.byref(h)), "csr_plan")
            elif placement == "xcd":
                self.eng._check(lib().intspmv_csr_plan_grouped(self.eng.h, self.nrows,
                                                               _np_ptr(row), 8, None, C.byref(h)),
                                "csr_plan_grouped")
            else:
                raise ValueError(f"unknown placement {placement!r}")
            self.plans[placement] = _Plan(self.eng, h)
        self.plan = self.plans[placement]
        self.placement = placement

    def tune_placement(self, x, y, reps: int = 10) -> dict:
        """Times spmv under both plans and keeps the faster (microseconds)."""
        torch = _torch()
        times = {}
        for pl in ("plain", "xcd"):
            self.set_placement(pl)
            for _ in range(2):
                self.spmv(x, y)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(self.eng.stream)
            for _ in range(reps):
                self.spmv(x, y)
            e1.record(self.eng.stream)
            e1.synchronize()
            times[pl] = e0.elapsed_time(e1) * 1e3 / reps
        self.set_placement(min(times, key=times.get))
        return dict(placement=self.placement, plain_us=round(times["plain"], 2),
                    xcd_us=round(times["xcd"], 2))

    def cf_csr(self, x, y) -> None:
        """Ordered product, same bits as cf_CSR (intSpmv_main.c:382-396)."""
        self.eng._check(lib().spmv_csr_ordered(self.eng.h, self.nrows, _dptr(self.val),
                                               _dptr(self.col), _dptr(self.row), _dptr(x),
                                               _dptr(y)), "spmv_csr_ordered")

    def spmv(self, x, y) -> None:
        self.eng._check(lib().spmv_csr(self.eng.h, self.plan.h, _dptr(self.val), _dptr(self.col),
                                       _dptr(self.row), _dptr(x), _dptr(y)), "spmv_csr")

    def spmv_exact(self, x, y) -> None:
        """spmv_csr under an exact plan (intspmv_csr_plan_exact): the bits of
        cf_CSR (intSpmv_main.c:382-396) for every row, long rows included."""
        if getattr(self, "_exact_plan", None) is None:
            h = _vp()
            self.eng._check(lib().intspmv_csr_plan_exact(self.eng.h, self.nrows,
                                                         _np_ptr(self._row_host), C.byref(h)),
                            "csr_plan_exact")
            self._exact_plan = _Plan(self.eng, h)
        self.eng._check(lib().spmv_csr(self.eng.h, self._exact_plan.h, _dptr(self.val),
                                       _dptr(self.col), _dptr(self.row), _dptr(x), _dptr(y)),
                        "spmv_csr (exact plan)")


class DeviceAXC:
    def __init__(self, eng: Engine, a):
        self.eng, self.nrows = eng, a.nrows
        self.ax = eng.upload(a.ax)
        self.brp = eng.upload(a.brp)
        h = _vp()
        brp = np.ascontiguousarray(a.brp)
        eng._check(lib().intspmv_axc_plan(eng.h, a.nrows, _np_ptr(brp), C.byref(h)), "axc_plan")
        self.plan = _Plan(eng, h)

    def int_axc(self, y) -> None:
        self.eng._check(lib().spmv_axc(self.eng.h, self.plan.h, _dptr(self.ax), _dptr(self.brp),
                                       _dptr(y)), "spmv_axc")


class DeviceK1:
    def __init__(self, eng: Engine, k):
        self.eng, self.nrows, self.chunkNum = eng, k.nrows, k.chunkNum
        self.val = eng.upload(k.val)
        self.col = eng.upload(k.col)
        self.chp = eng.upload(k.chp)
        self.nmc = eng.upload(k.nmc)
        self.permi = eng.upload(k.permi)
        h = _vp()
        nmc = np.ascontiguousarray(k.nmc)
        eng._check(lib().intspmv_k1_plan(eng.h, k.nrows, k.chunkNum, _np_ptr(nmc), C.byref(h)),
                   "k1_plan")
        self.plan = _Plan(eng, h)

    @classmethod
    def from_device_csr(cls, eng: Engine, dc: "DeviceCSR") -> "DeviceK1":
        """K1 arrays built on the device from a DeviceCSR (intspmv_k1_build_device)."""
        torch = _torch()
        self = cls.__new__(cls)
        n = dc.nrows
        self.eng, self.nrows, self.chunkNum = eng, n, (n + 7) // 8
        i32 = dict(dtype=torch.int32, device=eng.torch_device)
        self.permi = torch.empty((self.chunkNum + 1) * 8, **i32)
        self.nmc = torch.empty(max(self.chunkNum, 1), **i32)
        self.chp = torch.empty(max(self.chunkNum, 1), **i32)
        lenvc = _u32()
        args = (eng.h, n, _dptr(dc.row), _dptr(dc.col), _dptr(dc.val), _dptr(self.permi),
                _dptr(self.nmc), _dptr(self.chp))
        eng._check(lib().intspmv_k1_build_device(*args, None, None, 0, C.byref(lenvc)),
                   "k1_build_device")
        self.lenVC = lenvc.value
        self.val = torch.empty(max(self.lenVC, 1), dtype=torch.float64, device=eng.torch_device)
        self.col = torch.empty(max(self.lenVC, 1), **i32)
        eng._check(lib().intspmv_k1_build_device(*args, _dptr(self.val), _dptr(self.col),
                                                 self.lenVC, C.byref(lenvc)), "k1_build_device")
        nmc = np.ascontiguousarray(self.nmc[: self.chunkNum].cpu().numpy().view(np.uint32))
        h = _vp()
        eng._check(lib().intspmv_k1_plan(eng.h, n, self.chunkNum, _np_ptr(nmc), C.byref(h)),
                   "k1_plan")
        self.plan = _Plan(eng, h)
        return self

    def int_k1(self, x, res) -> None:
        """Permuted result res[chunkNum*8] (int_k1, intSpmv_main.c:943-965)."""
        self.eng._check(lib().spmv_k1(self.eng.h, self.plan.h, _dptr(self.val), _dptr(self.col),
                                      _dptr(self.chp), _dptr(self.nmc), None, _dptr(x),
                                      _dptr(res)), "spmv_k1")

    def order(self, res, y) -> None:
        """orderArrayK1 (intSpmv_main.c:969-976)."""
        self.eng._check(lib().spmv_k1_order(self.eng.h, self.nrows, _dptr(self.permi), _dptr(res),
                                            _dptr(y)), "spmv_k1_order")

    def int_k1p(self, x, y) -> None:
        """Scatter variant (int_k1p, intSpmv_main.c:1026-1066)."""
        self.eng._check(lib().spmv_k1(self.eng.h, self.plan.h, _dptr(self.val), _dptr(self.col),
                                      _dptr(self.chp), _dptr(self.nmc), _dptr(self.permi),
                                      _dptr(x), _dptr(y)), "spmv_k1p")


def _f32_round_to_odd(a: np.ndarray) -> np.ndarray:
    """FP64 -> FP32 rounded to odd: truncated toward zero, lowest mantissa bit
    set when the conversion was inexact.  Rounding that result once more to a
    format at least two bits narrower (BF16: 8 bits against FP32's 24) gives
    the correctly rounded value of the FP64 input (Boldo and Melquiond,
    "Emulation of FMA and correctly rounded sums: proved algorithms using
    rounding to odd", IEEE TC 2008)."""
    d = np.ascontiguousarray(a, dtype=np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        f = d.astype(np.float32)
        over = np.isfinite(d) & (np.abs(f.astype(np.float64)) > np.abs(d))
        f = np.where(over, np.nextafter(f, np.float32(0)), f)
        bits = f.view(np.uint32).copy()
        inexact = np.isfinite(d) & (f.astype(np.float64) != d)
    bits[inexact] |= np.uint32(1)
    return bits.view(np.float32)


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even BF16 bit patterns (uint16) of FP32 or FP64
    values.  FP64 input is rounded once (through FP32 rounded to odd), not
    FP64 -> FP32 -> BF16 to nearest twice."""
    a = np.asarray(a)
    f32 = _f32_round_to_odd(a) if a.dtype == np.float64 else np.ascontiguousarray(a, np.float32)
    f = f32.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f32)
    r = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((f[nan] >> 16) | 0x40).astype(np.uint16)  # stays a (quiet) NaN
    return r


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def axt_planes(a) -> tuple[np.ndarray, np.ndarray]:
    """Splits the interleaved ax of an AXT host object into value and x planes
    in slot order (slot = T*th*8 + R*8 + C)."""
    ax = a.ax.reshape(-1, 2, 8)
    return np.ascontiguousarray(ax[:, 0, :]).reshape(-1), np.ascontiguousarray(ax[:, 1, :]).reshape(-1)


class DeviceAXT:
    """AXT NOC on the device.  ``val_type`` / ``x_type`` select the stored
    precision: (F64, F64) is the reference layout (interleaved ax); (F32, F32)
    interleaved FP32; (BF16, F32) uses planar value / x arrays."""

    def __init__(self, eng: Engine, a, val_type: int = F64, x_type: int = F64,
                 layout: int | None = None):
        if layout is None:
            layout = PLANAR if val_type == BF16 else INTERLEAVED
        self.eng, self.nrows = eng, a.nrows
        self.val_type, self.x_type, self.layout = val_type, x_type, layout
        d = AxtDesc(nrows=a.nrows, nnz=a.nnz, tileHW=a.tileHW, tileH=a.tileH, tileN=a.tileN,
                    lenAX=a.lenAX, lenSEC=a.lenSEC, colN=a.colN, bs=a.bs, log=a.log,
                    mode=NOC if a.mode == "NOC" else COM, val_type=val_type, x_type=x_type,
                    layout=layout)
        self.desc = d
        self.xplane = None
        if layout == INTERLEAVED:
            if val_type == F64 and x_type == F64:
                self.ax = eng.upload(a.ax)
            elif val_type == F32 and x_type == F32:
                self.ax = eng.upload(a.ax.astype(np.float32))
            else:
                raise ValueError("interleaved layout needs equal value and x types")
        else:
            vp, xp = axt_planes(a)
            self.ax = eng.upload({F64: vp, F32: vp.astype(np.float32),
                                  BF16: bf16_bits(vp)}[val_type])
            self.xplane = eng.upload({F64: xp, F32: xp.astype(np.float32)}[x_type])
        self.sec = eng.upload(a.sec)
        h = _vp()
        con = np.ascontiguousarray(a.con)
        eng._check(lib().intspmv_axt_plan(eng.h, C.byref(d), _np_ptr(con), C.byref(h)),
                   "axt_plan")
        self.plan = _Plan(eng, h)
        self.ydtype = np.float64 if val_type == F64 else np.float32

    def int_axt(self, y) -> None:
        self.eng._check(lib().spmv_axt(self.eng.h, self.plan.h, _dptr(self.ax), _dptr(self.xplane),
                                       _dptr(self.sec), _dptr(y)), "spmv_axt")


PAD = 0xFFFFFFFF


def _value_plane(a, val_type: int) -> np.ndarray:
    vp = a.value_plane()
    return {F64: vp, F32: vp.astype(np.float32), BF16: bf16_bits(vp)}[val_type]


def _axt_desc(a, val_type, x_type, layout) -> AxtDesc:
    return AxtDesc(nrows=a.nrows, nnz=a.nnz, tileHW=a.tileHW, tileH=a.tileH, tileN=a.tileN,
                   lenAX=a.lenAX, lenSEC=a.lenSEC, colN=a.colN, bs=a.bs, log=a.log,
                   mode=NOC if a.mode == "NOC" else COM, val_type=val_type, x_type=x_type,
                   layout=layout)


class DeviceAXTGather:
    """AXT NOC with a column map: values and columns in slot order, x separate
    (n x k row-major, 1 <= k <= 32).  ``a`` must be built with want_acol."""

    def __init__(self, eng: Engine, a, val_type: int = F64, x_type: int = F64, vals=None,
                 placement="plain"):
        if a.mode != "NOC":
            raise ValueError("the column-map kernels take NOC layouts")
        self.eng, self.nrows, self.a = eng, a.nrows, a
        self.val_type, self.x_type = val_type, x_type
        self.nslots = a.lenAX // 2
        self.desc = _axt_desc(a, val_type, x_type, COLMAP)
        self.aval = vals if vals is not None else eng.upload(_value_plane(a, val_type))
        self.acol = eng.upload(a.acol)
        self.sec = eng.upload(a.sec)
        self.plans = {}
        self.set_placement(placement)

    def _make_plan(self, placement) -> _Plan:
        h = _vp()
        con = np.ascontiguousarray(self.a.con)
        if isinstance(placement, str) and placement == "plain":
            self.eng._check(lib().intspmv_axt_plan(self.eng.h, C.byref(self.desc), _np_ptr(con),
                                                   C.byref(h)), "axt_plan")
        elif isinstance(placement, str) and placement == "xcd":
            # eight row groups of about equal tiles, group g on XCD g
            self.eng._check(lib().intspmv_axt_plan_grouped(self.eng.h, C.byref(self.desc),
                                                           _np_ptr(con), 8, None, C.byref(h)),
                            "axt_plan_grouped")
        elif isinstance(placement, str):
            raise ValueError(f"unknown placement {placement!r}")
        else:
            # rows [placement[g], placement[g+1]) run on XCD g % 8
            g = np.ascontiguousarray(placement, dtype=np.uint32)
            self.eng._check(lib().intspmv_axt_plan_grouped(self.eng.h, C.byref(self.desc),
                                                           _np_ptr(con), len(g) - 1, _np_ptr(g),
                                                           C.byref(h)), "axt_plan_grouped")
        return _Plan(self.eng, h)

    def set_placement(self, placement) -> None:
        """Which plan the products use: "plain" (intspmv_axt_plan: long rows
        dealt round-robin over the XCDs, normal blocks in contiguous eighths),
        "xcd" (intspmv_axt_plan_grouped: eight row groups of equal tiles, each
        with its long rows on one XCD) or explicit group row starts.  Same
        bits either way."""
        key = placement if isinstance(placement, str) else "custom"
        if key == "custom" or key not in self.plans:
            self.plans[key] = self._make_plan(placement)
        self.plan = self.plans[key]
        self.placement = key

    def tune_placement_c(self, k: int = 1, reps: int = 10) -> dict:
        """The same choice made by the library (intspmv_axt_plan_tuned): keeps
        the faster of the two plans as the "tuned" placement."""
        h, g = _vp(), C.c_int()
        tp, tg = C.c_float(), C.c_float()
        con = np.ascontiguousarray(self.a.con)
        self.eng._check(lib().intspmv_axt_plan_tuned(
            self.eng.h, C.byref(self.desc), _np_ptr(con), _dptr(self.aval), _dptr(self.acol),
            _dptr(self.sec), k, reps, C.byref(h), C.byref(g), C.byref(tp), C.byref(tg)),
            "axt_plan_tuned")
        self.plans["tuned"] = _Plan(self.eng, h)
        self.plan, self.placement = self.plans["tuned"], "tuned"
        return dict(grouped=bool(g.value), plain_us=round(tp.value, 2),
                    xcd_us=round(tg.value, 2))

    def tune_placement(self, k: int = 1, reps: int = 10) -> dict:
        """Times the product with k right-hand sides under the "plain" and
        "xcd" plans (random X, the engine's stream) and keeps the faster.
        Returns both times in microseconds."""
        torch = _torch()
        dt = torch.float64 if self.x_type == F64 else torch.float32
        n = self.nrows
        X = torch.rand(max(n * k, 1), dtype=dt, device=self.eng.torch_device)
        Y = torch.empty(max(n * k, 1), dtype=dt, device=self.eng.torch_device)
        times = {}
        for pl in ("plain", "xcd"):
            self.set_placement(pl)
            for _ in range(2):
                self.matvec(X, Y, k)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(self.eng.stream)
            for _ in range(reps):
                self.matvec(X, Y, k)
            e1.record(self.eng.stream)
            e1.synchronize()
            times[pl] = e0.elapsed_time(e1) * 1e3 / reps
        self.set_placement(min(times, key=times.get))
        return dict(placement=self.placement, plain_us=round(times["plain"], 2),
                    xcd_us=round(times["xcd"], 2))

    def matvec(self, x, y, k: int = 1, ldx: int | None = None, ldy: int | None = None,
               vals=None, kc: int = 0) -> None:
        """y = A x for k right-hand sides (x: n*ldx, y: n*ldy elements); kc =
        right-hand sides per pass (0: automatic)."""
        self.eng._check(lib().spmv_axt_gather(
            self.eng.h, self.plan.h, _dptr(vals if vals is not None else self.aval),
            _dptr(self.acol), _dptr(self.sec), _dptr(x), ldx or k, _dptr(y), ldy or k, k, kc),
            "spmv_axt_gather")

    def grad_vals(self, x, ybar, vbar) -> None:
        """vbar[s] = ybar[row(s)] * x[acol[s]] over the slots."""
        self.eng._check(lib().spmv_axt_grad_vals(self.eng.h, self.plan.h, _dptr(self.acol),
                                                 _dptr(self.sec), _dptr(x), _dptr(ybar),
                                                 _dptr(vbar)), "spmv_axt_grad_vals")


class AxtAdjoint:
    """The transposed tile map of an AXT matrix, built once on the host by
    intspmv_axt_transpose_map: the AXT layout of A^T (same tile height) plus
    the slot permutation from A's slot order.  ``adjoint`` computes
    X_bar = A^T Y_bar with the column-map kernel; the values of A^T follow
    A's values through ``spmv_permute_vals``."""

    def __init__(self, eng: Engine, m, fwd: DeviceAXTGather):
        from . import host as H
        self.eng, self.fwd = eng, fwd
        self.tmap = H.AxtTransposeMap(m, fwd.a)
        at = self.at = self.tmap.at
        self.permT = eng.upload(self.tmap.permT)
        self.nslots_t = at.lenAX // 2
        self.gT = DeviceAXTGather(eng, at, fwd.val_type, fwd.x_type)
        self.refresh(fwd.aval)

    def refresh(self, vals) -> None:
        """Moves A's slot values (device) into A^T's slot order."""
        self.eng._check(lib().spmv_permute_vals(self.eng.h, self.nslots_t, self.fwd.val_type,
                                                _dptr(self.permT), _dptr(vals),
                                                _dptr(self.gT.aval)), "spmv_permute_vals")

    # how adjoint(vals=...) takes A's new values: "permuted" reads them
    # through the slot map inside the product; "refresh" first copies them into
    # A^T's slot order (refresh) and then runs the plain product.  The two give
    # the same bits; which is faster depends on the matrix (tune_values_path).
    values_path = "permuted"

    def adjoint(self, ybar, xbar, k: int = 1, vals=None) -> None:
        """X_bar = A^T Y_bar.  With vals (A's slot values, device) the result
        follows those values, by the path values_path names (after the
        "refresh" path A^T's copy holds them too); otherwise it uses the copy
        made by the last refresh()."""
        if vals is not None and self.values_path == "refresh":
            self.refresh(vals)
            vals = None
        if vals is None:
            self.gT.matvec(ybar, xbar, k)
            return
        g = self.gT
        self.eng._check(lib().spmv_axt_gather_permuted(
            self.eng.h, g.plan.h, _dptr(vals), _dptr(self.permT), _dptr(g.acol), _dptr(g.sec),
            _dptr(ybar), k, _dptr(xbar), k, k, 0), "spmv_axt_gather_permuted")


    def adjoint_grad(self, ybar, xbar, x, vbar_t, vals_t=None) -> None:
        """X_bar = A^T Y_bar (one right-hand side) with A^T's copy of the values
        (vals_t, or the one refresh() last wrote) and, in the same pass,
        A's value gradient in A^T's slot order (vbar_t) for the forward's x."""
        g = self.gT
        self.eng._check(lib().spmv_axt_gather_grad(
            self.eng.h, g.plan.h, _dptr(vals_t if vals_t is not None else g.aval), _dptr(g.acol),
            _dptr(g.sec), _dptr(ybar), _dptr(xbar), _dptr(x), _dptr(vbar_t)),
            "spmv_axt_gather_grad")

    def tune_values_path(self, k: int = 1, reps: int = 10) -> dict:
        """Times adjoint(vals=...) by both paths on this matrix (A's current
        values, random Y_bar with k columns, on the engine's stream) and keeps
        the faster one in values_path.  Returns both times in microseconds."""
        torch = _torch()
        n_out = self.at.nrows
        n_in = self.fwd.a.nrows
        dt = torch.float64 if self.fwd.val_type == F64 else torch.float32
        yb = torch.rand(n_in * k, dtype=dt, device=self.eng.torch_device)
        xb = torch.empty(max(n_out * k, 1), dtype=dt, device=self.eng.torch_device)
        vals = self.fwd.aval
        times = {}
        for path in ("permuted", "refresh"):
            self.values_path = path
            for _ in range(2):
                self.adjoint(yb, xb, k, vals=vals)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(self.eng.stream)
            for _ in range(reps):
                self.adjoint(yb, xb, k, vals=vals)
            e1.record(self.eng.stream)
            e1.synchronize()
            times[path] = e0.elapsed_time(e1) * 1e3 / reps
        self.values_path = min(times, key=times.get)
        return dict(path=self.values_path, permuted_us=round(times["permuted"], 2),
                    refresh_us=round(times["refresh"], 2))


def H_AXT(m, x, **kw):
    from . import host as H
    return H.AXT(m, x, **kw)


def refresh_x(dev_axt: "DeviceAXT", acol_dev, x) -> None:
    """Rewrites the embedded x of a reference-layout AXT on the device."""
    dev_axt.eng._check(lib().spmv_axt_refresh_x(dev_axt.eng.h, dev_axt.plan.h, _dptr(acol_dev),
                                                _dptr(x), _dptr(dev_axt.ax)), "spmv_axt_refresh_x")


def grad_vals_embedded(dev_axt: "DeviceAXT", ybar, vbar) -> None:
    dev_axt.eng._check(lib().spmv_axt_grad_vals_embedded(dev_axt.eng.h, dev_axt.plan.h,
                                                         _dptr(dev_axt.ax), _dptr(dev_axt.sec),
                                                         _dptr(ybar), _dptr(vbar)),
                       "spmv_axt_grad_vals_embedded")


class BatchMatrix(C.Structure):
    _fields_ = [("aval", _vp), ("acol", _vp), ("sec", _vp), ("x", _vp), ("y", _vp),
                ("ldx", _u32), ("ldy", _u32), ("k", _u32), ("kc", _u32), ("nrows", _u32),
                ("tileH", _u32), ("tileN", _u32), ("lenAX", _u32), ("colN", _u32),
                ("umax", _u32), ("con_host", _vp)]


def upload_batch_arrays(eng: Engine, a, val_type: int):
    """Device copies of a column-map AXT's value plane, acol and sec."""
    return eng.upload(_value_plane(a, val_type)), eng.upload(a.acol), eng.upload(a.sec)


class DeviceBatch:
    """Many column-map AXT matrices with their own tile heights and 1..32
    right-hand sides in one launch of two concurrent persistent kernels (narrow
    and wide passes; optionally captured in a hipGraph).

    entries: list of dicts with keys a (host AXT built with want_acol), X and
    Y (device tensors, n*k row-major), k and optionally ldx (leading
    dimension of X, default k), kc, umax (units per wave block) and dev (the (aval, acol, sec) device arrays of a, from
    upload_batch_arrays, so that entries can share them)."""

    def __init__(self, eng: Engine, entries, val_type: int = F32, x_type: int = F32):
        self.eng = eng
        self._keep = []
        arr = (BatchMatrix * len(entries))()
        for i, e in enumerate(entries):
            a = e["a"]
            aval, acol, sec = e.get("dev") or upload_batch_arrays(eng, a, val_type)
            con = np.ascontiguousarray(a.con)
            self._keep += [aval, acol, sec, con, e["X"], e["Y"]]
            k = e["k"]
            arr[i] = BatchMatrix(aval=aval.data_ptr(), acol=acol.data_ptr(), sec=sec.data_ptr(),
                                 x=e["X"].data_ptr(), y=e["Y"].data_ptr(), ldx=e.get("ldx", k), ldy=k, k=k,
                                 kc=e.get("kc", 0), nrows=a.nrows, tileH=a.tileH, tileN=a.tileN,
                                 lenAX=a.lenAX, colN=a.colN, umax=e.get("umax", 0),
                                 con_host=con.ctypes.data)
        h = _vp()
        eng._check(lib().intspmv_batch_create(eng.h, val_type, x_type, len(entries), arr,
                                              C.byref(h)), "batch_create")
        self.h = h

    def info(self) -> dict:
        a, b, c = _u32(), _u32(), _u32()
        lib().intspmv_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c))
        return dict(nitems=a.value, grid=b.value, nsplit=c.value)

    def launch(self) -> None:
        self.eng._check(lib().intspmv_batch_launch(self.eng.h, self.h), "batch_launch")

    def capture(self) -> None:
        self.eng._check(lib().intspmv_batch_graph_capture(self.eng.h, self.h), "batch_capture")

    def replay(self) -> None:
        self.eng._check(lib().intspmv_batch_graph_launch(self.eng.h, self.h), "batch_replay")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None and not sys.is_finalizing():
            _lib.intspmv_batch_destroy(self.h)
            self.h = None
