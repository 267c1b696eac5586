"""GPU: device-side conversions against the host builders (bit for bit) --
the K1 build (rows sorted by length on the device, chunk arrays scattered
there) and the products computed from its arrays."""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal, csr_long_rows, csr_sparse_gaps, csr_with_empty_rows

pytestmark = pytest.mark.gpu


def _cases():
    yield "lap256", H.CSR.laplace2d(256)
    yield "powerlaw16k", H.CSR.powerlaw(n=1 << 14, target_nnz=460_000, maxlen=8192, seed=1)
    yield "empty_rows", csr_with_empty_rows()
    yield "sparse_gaps", csr_sparse_gaps()
    yield "long_rows", csr_long_rows()
    yield "n1003", H.CSR.banded(1003, 1, 9, seed=4)  # nrows not a multiple of 8
    yield "one_row", H.CSR.from_arrays(np.array([0, 2], np.uint32), np.array([0, 0], np.uint32),
                                       np.array([1.0, 2.0]))


def _u32(t, n):
    return t[:n].cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("name,m", list(_cases()), ids=[c[0] for c in _cases()])
def test_k1_build_on_device_equals_host(engine, name, m):
    k = H.K1(m)
    dc = D.DeviceCSR(engine, m)
    dk = D.DeviceK1.from_device_csr(engine, dc)
    assert dk.chunkNum == k.chunkNum and dk.lenVC == k.lenVC, name
    assert np.array_equal(_u32(dk.permi, (k.chunkNum + 1) * 8), k.permi), name
    assert np.array_equal(_u32(dk.nmc, k.chunkNum), k.nmc), name
    assert np.array_equal(_u32(dk.chp, k.chunkNum), k.chp), name
    assert np.array_equal(_u32(dk.col, k.lenVC), k.col), name
    assert bitwise_equal(dk.val[: k.lenVC].cpu().numpy(), k.val), name
    # the product from the device-built arrays: the oracle's int_k1p bits
    x = H.uniform(m.nrows, seed=9)
    xd = engine.upload(x)
    import torch
    y = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device=engine.torch_device)
    dk.int_k1p(xd, y)
    torch.cuda.synchronize()
    ref = O.int_k1p(k.chunkNum, m.nrows, k.val, k.col, k.chp, k.nmc, k.permi, x)
    g = y.cpu().numpy()
    steps = np.zeros(m.nrows, np.int64)
    steps[k.permi[: m.nrows]] = np.repeat(k.nmc, 8)[: m.nrows]
    short = steps <= 32  # chunks of a wave's narrow section: the oracle's exact order
    assert bitwise_equal(g[short], ref[short]), name
    assert np.allclose(g, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), name


def test_k1_build_on_device_reports_capacity(engine):
    import ctypes as C

    import torch
    m = H.CSR.laplace2d(32)
    dc = D.DeviceCSR(engine, m)
    n, cn = m.nrows, (m.nrows + 7) // 8
    i32 = dict(dtype=torch.int32, device=engine.torch_device)
    permi, nmc, chp = torch.empty((cn + 1) * 8, **i32), torch.empty(cn, **i32), torch.empty(cn, **i32)
    val = torch.empty(8, dtype=torch.float64, device=engine.torch_device)
    col = torch.empty(8, **i32)
    lenvc = C.c_uint32()
    st = D.lib().intspmv_k1_build_device(engine.h, n, D._dptr(dc.row), D._dptr(dc.col),
                                         D._dptr(dc.val), D._dptr(permi), D._dptr(nmc),
                                         D._dptr(chp), D._dptr(val), D._dptr(col), 8,
                                         C.byref(lenvc))
    assert st == -5 and lenvc.value == H.K1(m).lenVC  # INTSPMV_ERR_OVERFLOW, size reported
