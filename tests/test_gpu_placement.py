"""GPU: column-map plans with row groups pinned to XCDs
(intspmv_axt_plan_grouped) give the bits of the ungrouped plan.

Placement only moves work blocks between workgroups: every row is still
summed by one lane (or one long-row wave) in tile order, so the products,
the adjoint on A^T's map and the fused adjoint + value gradient must equal
the ungrouped plan's output bit for bit, for narrow and wide passes, FP64 and
FP32, with empty rows, split long rows, empty groups and more groups than
XCDs.  y is pre-filled with NaN so a row no block writes fails.
"""
import numpy as np
import pytest
import torch

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import csr_long_rows, csr_sparse_gaps, csr_with_empty_rows

pytestmark = pytest.mark.gpu


def _random_groups(n, ng, seed):
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.integers(0, n + 1, size=ng - 1))
    return np.concatenate([[0], cuts, [n]]).astype(np.uint32)


def _matrices():
    return [("empty_rows", csr_with_empty_rows()), ("sparse_gaps", csr_sparse_gaps()),
            ("long_rows", csr_long_rows()),
            ("powerlaw", H.CSR.powerlaw(n=20000, target_nnz=600000, maxlen=8192,
                                        band_frac=0.8, band_half=500, seed=4))]


@pytest.fixture(scope="module")
def eng():
    return D.Engine(0)


def _run(g, X, k, dt, n):
    Y = torch.full((max(n * k, 1),), float("nan"), dtype=dt, device="cuda")
    g.matvec(X, Y, k)
    torch.cuda.synchronize()
    return Y[: n * k].cpu()


@pytest.mark.parametrize("name,m", _matrices())
@pytest.mark.parametrize("th", [1, 4])
def test_grouped_plan_same_bits(eng, name, m, th):
    n = m.nrows
    x = H.uniform(n, -1.0, 1.0, seed=2)
    a = H.AXT(m, x, th=th, want_acol=True)
    for vt, dt, ks in ((D.F64, torch.float64, (1, 3, 8, 32)), (D.F32, torch.float32, (1, 8, 32))):
        g = D.DeviceAXTGather(eng, a, vt, vt)
        for k in ks:
            if th > 1 and k > 8:
                continue
            X = torch.rand(n * k, dtype=dt, device="cuda") * 2 - 1
            g.set_placement("plain")
            ref = _run(g, X, k, dt, n)
            assert not torch.isnan(ref).any()
            for pl in ("xcd", _random_groups(n, 13, k), _random_groups(n, 3, k + 1)):
                g.set_placement(pl)
                got = _run(g, X, k, dt, n)
                assert torch.equal(got, ref), f"{name} th={th} vt={vt} k={k} placement={pl}"
        if vt == D.F64:
            # and the grouped product is the product (against the oracle)
            g.set_placement("xcd")
            y = _run(g, eng.upload(x), 1, dt, n).numpy()
            yref = O.cf_csr(m.row, m.col, m.val, x)
            err, row = O.scaled_err(m.row, m.col, m.val, x, yref, y)
            assert err <= 1e-12, f"{name} th={th}: {err} at row {row}"


@pytest.mark.parametrize("name,m", _matrices())
def test_grouped_plan_adjoint_same_bits(eng, name, m):
    """x_bar = A^T y_bar through A^T's map, and x_bar + A_bar in one pass,
    under a grouped plan of A^T's map."""
    n = m.nrows
    x = H.uniform(n, -1.0, 1.0, seed=5)
    a = H.AXT(m, x, th=1, want_acol=True)
    fwd = D.DeviceAXTGather(eng, a)
    adj = D.AxtAdjoint(eng, m, fwd)
    yb = eng.upload(H.uniform(n, -1.0, 1.0, seed=6))
    xd = eng.upload(x)
    outs = {}
    for pl in ("plain", "xcd"):
        adj.gT.set_placement(pl)
        xb = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        adj.adjoint(yb, xb)
        xb2 = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        adj.adjoint(yb, xb2, vals=fwd.aval)
        xb3 = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        vbar = torch.full((adj.nslots_t,), float("nan"), dtype=torch.float64, device="cuda")
        adj.adjoint_grad(yb, xb3, xd, vbar)
        torch.cuda.synchronize()
        outs[pl] = [t.cpu() for t in (xb, xb2, xb3, vbar)]
    for a0, a1 in zip(outs["plain"], outs["xcd"]):
        assert torch.equal(a0, a1), name


def test_tune_placement_keeps_a_plan(eng):
    m = H.CSR.powerlaw(n=20000, target_nnz=600000, maxlen=8192, band_frac=0.8, band_half=500,
                       seed=4)
    x = H.uniform(m.nrows, -1.0, 1.0, seed=2)
    g = D.DeviceAXTGather(eng, H.AXT(m, x, th=1, want_acol=True))
    r = g.tune_placement(k=1, reps=3)
    assert r["placement"] in ("plain", "xcd") and g.placement == r["placement"]
    assert r["plain_us"] > 0 and r["xcd_us"] > 0


@pytest.mark.parametrize("name,m", _matrices())
def test_csr_grouped_plan_same_bits(eng, name, m):
    """spmv_csr under intspmv_csr_plan_grouped equals the plain plan bit for
    bit, and both are within 1e-12 of the oracle."""
    n = m.nrows
    x = H.uniform(n, -1.0, 1.0, seed=8)
    dc = D.DeviceCSR(eng, m)
    xd = eng.upload(x)
    out = {}
    for pl in ("plain", "xcd"):
        dc.set_placement(pl)
        y = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        dc.spmv(xd, y)
        torch.cuda.synchronize()
        out[pl] = y.cpu()
    assert torch.equal(out["plain"], out["xcd"]), name
    yref = O.cf_csr(m.row, m.col, m.val, x)
    err, row = O.scaled_err(m.row, m.col, m.val, x, yref, out["xcd"].numpy())
    assert err <= 1e-12, f"{name}: {err} at row {row}"


@pytest.mark.parametrize("vt,dt", [(D.F64, torch.float64), (D.F32, torch.float32)])
def test_plan_tuned_in_c(eng, vt, dt):
    """intspmv_axt_plan_tuned keeps the faster of the two plans (by its own
    timings) and the kept plan gives the ungrouped plan's bits."""
    m = H.CSR.powerlaw(n=20000, target_nnz=600000, maxlen=8192, band_frac=0.8, band_half=500,
                       seed=4)
    n = m.nrows
    g = D.DeviceAXTGather(eng, H.AXT(m, H.uniform(n, -1.0, 1.0, seed=2), th=1, want_acol=True),
                          vt, vt)
    X = torch.rand(n * 3, dtype=dt, device="cuda") * 2 - 1
    ref = _run(g, X, 3, dt, n)
    r = g.tune_placement_c(k=3, reps=3)
    assert g.placement == "tuned" and r["plain_us"] > 0 and r["xcd_us"] > 0
    assert r["grouped"] == (r["xcd_us"] < r["plain_us"])
    assert torch.equal(_run(g, X, 3, dt, n), ref)
    with pytest.raises(D.IntSpmvError):  # k outside 1..32 is refused
        g.tune_placement_c(k=33, reps=1)
