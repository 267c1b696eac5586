"""GPU: degenerate matrices through every entry point -- a 1x1 matrix, a
matrix whose rows are all empty, a single empty row, a matrix whose only
entries sit in its last row, and runs of empty rows longer than a block.  y is pre-filled with NaN, so a row that no
kernel writes fails; the expected values come from the oracle (cf_CSR), and
with at most two products per row every format must give its bits."""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal

pytestmark = pytest.mark.gpu


def _cases():
    u32 = np.uint32
    return [
        ("one_by_one", H.CSR.from_arrays(np.array([0, 1], u32), np.array([0], u32),
                                         np.array([2.5]))),
        ("all_empty5", H.CSR.from_arrays(np.zeros(6, u32), np.zeros(0, u32), np.zeros(0))),
        ("one_empty", H.CSR.from_arrays(np.zeros(2, u32), np.zeros(0, u32), np.zeros(0))),
        ("last_only", H.CSR.from_arrays(np.array([0, 0, 0, 0, 2], u32), np.array([0, 3], u32),
                                        np.array([1.0, -2.0]))),
    ]


def _nan(eng, n, dtype=np.float64):
    t = eng.empty(max(n, 1), dtype)
    t.fill_(float("nan"))
    return t


def _host(t, n):
    import torch
    torch.cuda.synchronize()
    return t.reshape(-1)[:n].double().cpu().numpy()


@pytest.mark.parametrize("name,m", _cases(), ids=[c[0] for c in _cases()])
def test_every_format_on_degenerate_matrices(engine, name, m):
    n = m.nrows
    x = H.uniform(n, seed=4)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    xd = engine.upload(x)
    got = {}
    dc = D.DeviceCSR(engine, m)
    y = _nan(engine, n)
    dc.spmv(xd, y)
    got["csr"] = _host(y, n)
    y = _nan(engine, n)
    dc.cf_csr(xd, y)
    got["csr_ordered"] = _host(y, n)
    y = _nan(engine, n)
    dc.spmv_exact(xd, y)
    got["csr_exact_plan"] = _host(y, n)
    dc.set_placement("xcd")
    y = _nan(engine, n)
    dc.spmv(xd, y)
    got["csr_grouped_plan"] = _host(y, n)
    y = _nan(engine, n)
    D.DeviceAXC(engine, H.AXC(m, x)).int_axc(y)
    got["axc"] = _host(y, n)
    k = H.K1(m)
    dk = D.DeviceK1(engine, k)
    y = _nan(engine, n)
    dk.int_k1p(xd, y)
    got["k1p"] = _host(y, n)
    res = _nan(engine, k.chunkNum * 8)
    dk.int_k1(xd, res)
    y = _nan(engine, n)
    dk.order(res, y)
    got["k1"] = _host(y, n)
    for th in (1, 2, 4, 8, 32):
        y = _nan(engine, n)
        D.DeviceAXT(engine, H.AXT(m, x, th=th)).int_axt(y)
        got[f"axt_th{th}"] = _host(y, n)
    for bs in (2, 128):
        y = _nan(engine, n)
        D.DeviceAXT(engine, H.AXT(m, x, th=1, mode="COM", bs=bs)).int_axt(y)
        got[f"com_bs{bs}"] = _host(y, n)
    ag = H.AXT(m, np.zeros(n), th=1, want_acol=True)
    g = D.DeviceAXTGather(engine, ag)
    y = _nan(engine, n)
    g.matvec(xd, y)
    got["colmap"] = _host(y, n)
    for pl in ("xcd", np.array([0, n], np.uint32)):  # grouped plans: eight even groups, one group
        g.set_placement(pl)
        y = _nan(engine, n)
        g.matvec(xd, y)
        got[f"colmap_grouped_{len(pl) if not isinstance(pl, str) else pl}"] = _host(y, n)
    g.set_placement("plain")
    for name_f, got_y in got.items():
        assert bitwise_equal(got_y, yr), (name, name_f, got_y, yr)
    # multi-RHS, adjoint and a batch of the matrix
    k3 = 3
    X = np.stack([H.uniform(n, seed=10 + j) for j in range(k3)], axis=1)
    Y = _nan(engine, n * k3)
    g.matvec(engine.upload(X.reshape(-1)), Y, k=k3)
    Yg = _host(Y, n * k3).reshape(n, k3)
    for j in range(k3):
        assert bitwise_equal(Yg[:, j], O.cf_csr(m.row, m.col, m.val, X[:, j].copy())), (name, j)
    adj = D.AxtAdjoint(engine, m, g)
    ybar = H.uniform(n, seed=12)
    xb = _nan(engine, n)
    adj.adjoint(engine.upload(ybar), xb)
    ref = O.csr_transpose_product(m.row, m.col, m.val, ybar)
    assert bitwise_equal(_host(xb, n), ref), name
    import torch
    Yb = torch.full((n * k3,), float("nan"), dtype=torch.float32, device=engine.torch_device)
    Xb = torch.tensor(X.astype(np.float32).reshape(-1), device=engine.torch_device)
    b = D.DeviceBatch(engine, [dict(a=ag, X=Xb, Y=Yb, k=k3)], D.F32, D.F32)
    b.launch()
    Ybh = _host(Yb, n * k3).reshape(n, k3)
    v32 = m.val.astype(np.float32).astype(np.float64)
    for j in range(k3):
        want = O.cf_csr(m.row, m.col, v32, X[:, j].astype(np.float32).astype(np.float64))
        assert np.allclose(Ybh[:, j], want, rtol=1e-6, atol=0), (name, j)


@pytest.mark.parametrize("lead", [64, 67])
def test_long_runs_of_empty_rows(engine, lead):
    """Runs of 600 and 300 empty rows (longer than a block's 256 rows, so
    whole blocks hold no unit) after `lead` slots -- a multiple of 8 or not,
    so an empty block starts on a tile boundary or inside a tile -- and at
    the end.  Every format must write each empty row's 0 and keep its own
    bits on the rows with entries (all short)."""
    rng = np.random.default_rng(lead)
    lens = np.concatenate([np.full(lead // 4, 4), [lead % 4] if lead % 4 else [],
                           np.zeros(600), rng.integers(1, 20, 500), np.zeros(300)]).astype(np.int64)
    n = len(lens)
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, n, size=k)) for k in lens]).astype(np.uint32)
    m = H.CSR.from_arrays(row, col, rng.standard_normal(len(col)))
    x = rng.uniform(-1, 1, n)
    yr = O.cf_csr(m.row, m.col, m.val, x)
    empty = lens == 0
    xd = engine.upload(x)
    got = {}
    dc = D.DeviceCSR(engine, m)
    for name, fn in (("csr_ordered", dc.cf_csr), ("csr", dc.spmv), ("csr_exact", dc.spmv_exact)):
        y = _nan(engine, n)
        fn(xd, y)
        got[name] = (_host(y, n), yr)
    c = H.AXC(m, x)
    y = _nan(engine, n)
    D.DeviceAXC(engine, c).int_axc(y)
    got["axc"] = (_host(y, n), O.int_axc(n, c.ax, c.brp))
    k = H.K1(m)
    y = _nan(engine, n)
    D.DeviceK1(engine, k).int_k1p(xd, y)
    got["k1p"] = (_host(y, n), O.int_k1p(k.chunkNum, n, k.val, k.col, k.chp, k.nmc, k.permi, x))
    for th in (1, 4, 16):
        a = H.AXT(m, x, th=th)
        y = _nan(engine, n)
        D.DeviceAXT(engine, a).int_axt(y)
        got[f"axt_th{th}"] = (_host(y, n), O.int_axt(a))
        if th in (1, 4):
            for vt, xt in ((D.F32, D.F32), (D.BF16, D.F32)):
                y = _nan(engine, n, np.float32)
                D.DeviceAXT(engine, a, val_type=vt, x_type=xt).int_axt(y)
                g = _host(y, n)
                assert np.all(g[empty] == 0.0), (th, vt)
                assert np.all(np.isfinite(g)), (th, vt)
    for bs in (2, 128):
        a = H.AXT(m, x, th=1, mode="COM", bs=bs)
        y = _nan(engine, n)
        D.DeviceAXT(engine, a).int_axt(y)
        got[f"com_bs{bs}"] = (_host(y, n), yr)
    ag = H.AXT(m, np.zeros(n), th=1, want_acol=True)
    g = D.DeviceAXTGather(engine, ag)
    y = _nan(engine, n)
    g.matvec(xd, y)
    got["colmap"] = (_host(y, n), O.int_axt(H.AXT(m, x, th=1)))
    for name, (g, ref) in got.items():
        assert np.all(g[empty] == 0.0), name
        assert bitwise_equal(g, ref), name


def test_bad_arguments_are_refused_before_launch(engine):
    """Inconsistent sizes and unsupported blockings come back as
    INTSPMV_ERR_ARG (IntSpmvError, status -1) from the host-side checks, and
    nothing is written."""
    import torch
    m = H.CSR.banded(500, 2, 20, seed=3)
    n = m.nrows
    a = H.AXT(m, np.zeros(n), th=1, want_acol=True)
    g64 = D.DeviceAXTGather(engine, a)
    X = torch.rand(n * 40, dtype=torch.float64, device=engine.torch_device)
    Y = _nan(engine, n * 40)
    bad = [dict(k=0), dict(k=33), dict(k=4, ldx=3), dict(k=4, ldy=3), dict(k=4, kc=3),
           dict(k=8, kc=64)]
    for kw in bad:
        with pytest.raises(D.IntSpmvError) as ei:
            g64.matvec(X, Y, **kw)
        assert ei.value.status == -1, kw
    # FP64 passes of 16 or 32 are wide passes, which need tile height 1
    a4 = H.AXT(m, np.zeros(n), th=4, want_acol=True)
    g4 = D.DeviceAXTGather(engine, a4)
    for kw in (dict(k=16, kc=16), dict(k=8, kc=32)):
        with pytest.raises(D.IntSpmvError) as ei:
            g4.matvec(X, Y, **kw)
        assert ei.value.status == -1, kw
    with pytest.raises(D.IntSpmvError) as ei:
        g64.matvec(None, Y)
    assert ei.value.status == -1
    ent = dict(a=a, X=X[:n * 4], Y=Y[:n * 4], k=4)
    for kw in (dict(kc=64), dict(kc=3), dict(ldx=3), dict(umax=4), dict(umax=256),
               dict(a=a4, kc=16), dict(a=a4, kc=32)):
        with pytest.raises(D.IntSpmvError) as ei:
            D.DeviceBatch(engine, [dict(ent, **kw)], D.F64, D.F64)
        assert ei.value.status == -1, kw
    torch.cuda.synchronize()
    assert torch.isnan(Y).all()


def test_gather_rate_calibration():
    """intspmv_gather_rate (the gather ceiling bench.py reports) returns a
    positive rate for every element size and refuses bad arguments without
    launching; so does intspmv_stream_rate (the streaming ceiling)."""
    eng = D.Engine(0)
    for elem in (4, 8, 16):
        r = eng.gather_rate(1 << 20, elem, count=1 << 20, reps=2)
        assert 1e9 < r < 1e14, (elem, r)
    for span, elem, count, reps in ((1 << 20, 3, 1 << 10, 1), (4, 8, 1 << 10, 1),
                                    (1 << 20, 8, 0, 1), (1 << 20, 8, 1 << 10, 0)):
        with pytest.raises(D.IntSpmvError):
            eng.gather_rate(span, elem, count=count, reps=reps)
    # the streaming ceiling (intspmv_stream_rate): a plausible rate, bad sizes refused
    r = eng.stream_rate(64 << 20, reps=3)
    assert 1e11 < r < 2e13, r
    for nbytes, reps in ((8, 1), (1 << 20, 0), (1 << 50, 1)):  # the last: too many workgroups
        with pytest.raises(D.IntSpmvError):
            eng.stream_rate(nbytes, reps=reps)
