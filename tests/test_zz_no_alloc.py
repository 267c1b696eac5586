"""GPU: spmv_axt_h1_noalloc, the allocation-free product (no plan, no
hipMalloc; a fixed device scratch array), through the C ABI, against the
existing entry point spmv_axt on the same arrays: two calls in a row, the
second on a smaller matrix (the scratch keeps the first call's contents),
within 1e-12 scaled per row and bit-identical on rows that lie inside one
64-tile chunk and one plan block; an oversized call is refused with nothing
written."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal, csr_sparse_gaps

pytestmark = pytest.mark.gpu


def _noalloc(engine, a, ax, sec, y):
    return D.lib().spmv_axt_h1_noalloc(engine.h, a.nrows, a.tileN, D._dptr(ax), D._dptr(sec),
                                       D._dptr(y))


def _case(engine, m, x):
    import torch
    a = H.AXT(m, x, th=1)
    d = D.DeviceAXT(engine, a)  # spmv_axt with its plan: the existing entry point
    y_plan = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device=engine.torch_device)
    d.int_axt(y_plan)
    y = torch.full_like(y_plan, float("nan"))
    print(f"[noalloc] n={m.nrows} tileN={a.tileN}", flush=True)
    assert _noalloc(engine, a, d.ax, d.sec, y) == 0
    torch.cuda.synchronize()
    g, gp = y.cpu().numpy(), y_plan.cpu().numpy()
    yr = O.cf_csr(m.row, m.col, m.val, x)
    assert O.scaled_err(m.row, m.col, m.val, x, gp, g)[0] <= 1e-12
    assert O.scaled_err(m.row, m.col, m.val, x, yr, g)[0] <= 1e-12
    # rows inside one 64-tile chunk and inside one plan block: the oracle's bits
    tiles = (np.diff(m.row.astype(np.int64)) + 7) // 8
    start = np.concatenate([[0], np.cumsum(tiles)[:-1]])
    inside = (tiles == 0) | ((start // 64 == (start + tiles - 1) // 64) & (tiles <= 48))
    assert bitwise_equal(g[inside], gp[inside])
    assert bitwise_equal(g[inside], O.int_axt(a)[inside])
    return a, d


def test_noalloc_two_calls_then_smaller(engine):
    big = H.CSR.powerlaw(n=1 << 15, target_nnz=1_200_000, maxlen=8192, seed=3)  # rows over many chunks
    small = csr_sparse_gaps()  # 60 % empty rows, long gaps, a 900-entry row
    for m in (big, small):
        _case(engine, m, H.uniform(m.nrows, seed=4))


def test_noalloc_degenerate_and_oversize(engine):
    import torch
    # one row, one tile; every row empty (no tiles)
    one = H.CSR.from_arrays(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([2.5]))
    _case(engine, one, np.array([3.0]))
    y = torch.full((5,), float("nan"), dtype=torch.float64, device=engine.torch_device)
    assert D.lib().spmv_axt_h1_noalloc(engine.h, 5, 0, None, None, D._dptr(y)) == 0
    torch.cuda.synchronize()
    assert np.all(y.cpu().numpy() == 0)
    # beyond the scratch: refused before any launch, y untouched
    y.fill_(float("nan"))
    dummy = torch.zeros(16, dtype=torch.float64, device=engine.torch_device)
    sec = torch.zeros(1, dtype=torch.int32, device=engine.torch_device)
    st = D.lib().spmv_axt_h1_noalloc(engine.h, 5, (8 << 20) // 16 * 64 + 1, D._dptr(dummy),
                                     D._dptr(sec), D._dptr(y))
    assert st == -7 and "scratch" in D.lib().intspmv_strerror(st).decode()
    torch.cuda.synchronize()
    assert np.all(np.isnan(y.cpu().numpy()))
    assert D.lib().spmv_axt_h1_noalloc(None, 5, 1, None, None, None) == -1
    assert C.sizeof(C.c_double) == 8
