"""Vendor comparators (timed only): MKL's CSR kernels give the CSR product."""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import comparators as CMP
from int_spmv_amd import host as H


@pytest.mark.skipif(not CMP.mkl_available(), reason="libmkl_rt not present")
@pytest.mark.parametrize("threads", [1, 2])
def test_mkl_csr_matches_oracle(threads):
    m = H.CSR.powerlaw(n=4096, target_nnz=60_000, s=2.0, maxlen=512, band_frac=0.8,
                       band_half=100, seed=3)
    x = H.uniform(m.nrows, seed=3)
    yref = O.cf_csr(m.row, m.col, m.val, x)
    res = CMP.mkl_csr(m.row, m.col, m.val, x, threads, budget_s=0.05)
    assert set(res) == {"mkl_CSR", "sp_mkl_CSR", "sp_insp_mkl_CSR"}
    for name, r in res.items():
        err, _ = O.scaled_err(m.row, m.col, m.val, x, yref, r["y"])
        assert err <= 1e-12, name
        assert r["ms"] > 0 and r["calls"] >= 1


def test_rocsparse_loader_does_not_need_a_device():
    # loading the library (no handle, no device call) works or reports absence
    assert CMP.rocsparse_available() in (True, False)
