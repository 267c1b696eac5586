"""GPU parity at the benchmark workloads' full sizes, through the C ABI.

Workloads (int_spmv_amd/workloads.py, the seeds bench.py uses):
  powerlaw_1m    n = 2^20, 30,000,004 entries, Zipf(2) rows up to 8192, every
                 format and tile height; the column map (1 and 8 right-hand
                 sides), the adjoint, the value gradient, and the autograd
                 operator checked by central finite differences
  stencil27_128  n = 2,097,152, 55,742,968 entries \\  FP64, FP32 and BF16
  rmat_22        n = 4,194,304, 65,245,084 entries /  value storage
  batch_256      256 matrices, 11.3M rows, 383M entries, 1..32 right-hand
                 sides each, one persistent launch and its graph replays

Bar, as in the smaller parity tests: bit-exact against the oracle where it
fixes the rounding order (rows that stay in one wave block; the plans'
long-row thresholds are mirrored in _normal_rows), otherwise
|y_gpu - y_ref| <= tol * sum_j |a_ij x_j| per row against CSR-naive with
tol = 1e-12 (FP64), 1e-6 (FP32 storage), 1e-2 (BF16 storage); FP32 / BF16
results are held to the FP64 oracle of the unrounded matrix.  Outputs are
NaN-filled first, so an unwritten row fails.
"""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H
from int_spmv_amd.workloads import WORKLOADS, batch_256

from .util import bitwise_equal

pytestmark = pytest.mark.gpu

TOL64, TOL32, TOL16 = 1e-12, 1e-6, 1e-2


@pytest.fixture(scope="module")
def wl():
    """Loads a workload once per module: (m, x, cf_CSR(m, x))."""
    cache = {}

    def get(name):
        if name not in cache:
            print(f"[fullsize] generating {name}", flush=True)
            m, x = WORKLOADS[name]()
            cache[name] = (m, x, O.cf_csr(m.row, m.col, m.val, x))
        return cache[name]
    yield get
    cache.clear()


def _nan(eng, n, dt=np.float64):
    t = eng.empty(n, dt)
    t.fill_(float("nan"))
    return t


def _host(t, n):
    import torch
    torch.cuda.synchronize()
    return t.reshape(-1)[:n].double().cpu().numpy()


def _serr(m, x, yr, y):
    return O.scaled_err(m.row, m.col, m.val, x, yr, np.asarray(y, np.float64))[0]


def _many_empty(m):
    return np.count_nonzero(np.diff(m.row.astype(np.int64)) == 0) * 4 > m.nrows


def _normal_rows(m, th, colmap=False):
    """Rows the AXT NOC FP64 plan keeps in normal blocks (one lane adds their
    unit sums in order: the oracle's bits).  Mirrors intspmv_axt_plan: th = 1
    long rows above 48 tiles (64 with many empty rows or in column-map
    plans); th > 1 above min(512/th, umax/2) columns, at least 4, with umax
    256 (th 2), 64 (th 8, dense rows), 32 (th >= 16, dense rows), else 128."""
    dense = not _many_empty(m)
    if th == 1:
        lim = 48 if dense and not colmap else 64
    else:
        umax = 256 if th == 2 and not colmap else 64 if th == 8 and dense \
            else 32 if th >= 16 and dense and not colmap else 128
        lim = max(4, min(512 // th, umax // 2))
    per = 8 if th == 1 else th
    units = (np.diff(m.row.astype(np.int64)) + per - 1) // per
    return units <= lim


# ------------------------------------------------------------- powerlaw_1m


def test_powerlaw_csr(engine, wl):
    m, x, yr = wl("powerlaw_1m")
    d = D.DeviceCSR(engine, m)
    xd = engine.upload(x)
    y1, y2 = _nan(engine, m.nrows), _nan(engine, m.nrows)
    d.cf_csr(xd, y1)
    d.spmv(xd, y2)
    assert bitwise_equal(_host(y1, m.nrows), yr)
    g = _host(y2, m.nrows)
    short = np.diff(m.row.astype(np.int64)) <= 256
    assert bitwise_equal(g[short], yr[short])
    assert _serr(m, x, yr, g) <= TOL64
    y3 = _nan(engine, m.nrows)
    d.spmv_exact(xd, y3)  # exact plan: cf_CSR's bits for all rows
    assert bitwise_equal(_host(y3, m.nrows), yr)


def test_powerlaw_axc_k1(engine, wl):
    m, x, yr = wl("powerlaw_1m")
    a = H.AXC(m, x)
    d = D.DeviceAXC(engine, a)
    y = _nan(engine, m.nrows)
    d.int_axc(y)
    g = _host(y, m.nrows)
    ok = np.diff(a.brp.astype(np.int64)) // 16 <= 48  # rows in normal blocks
    assert bitwise_equal(g[ok], O.int_axc(m.nrows, a.ax, a.brp)[ok])
    assert _serr(m, x, yr, g) <= TOL64
    del d, a
    k = H.K1(m)
    dk = D.DeviceK1(engine, k)
    xd = engine.upload(x)
    res = _nan(engine, k.chunkNum * 8)
    dk.int_k1(xd, res)
    y, y2 = _nan(engine, m.nrows), _nan(engine, m.nrows)
    dk.order(res, y)
    dk.int_k1p(xd, y2)
    ref = O.int_k1(k.chunkNum, k.val, k.col, k.chp, k.nmc, x)
    narrow = ~np.repeat(k.nmc > 64, 8)  # chunks one wave slice adds in order
    assert bitwise_equal(_host(res, k.chunkNum * 8)[narrow], ref[narrow])
    assert _serr(m, x, yr, _host(y, m.nrows)) <= TOL64
    assert _serr(m, x, yr, _host(y2, m.nrows)) <= TOL64
    # the K1 conversion on the device gives the host builder's arrays
    dev = D.DeviceK1.from_device_csr(engine, D.DeviceCSR(engine, m))
    for f, n in (("permi", (k.chunkNum + 1) * 8), ("nmc", k.chunkNum), ("chp", k.chunkNum)):
        got = getattr(dev, f)[:n].cpu().numpy().view(np.uint32)
        assert np.array_equal(got, getattr(k, f)), f
    assert dev.lenVC == k.lenVC
    assert bitwise_equal(dev.val[: k.lenVC].cpu().numpy(), k.val)
    assert np.array_equal(dev.col[: k.lenVC].cpu().numpy().view(np.uint32), k.col)


@pytest.mark.parametrize("th", [1, 2, 4, 8, 16, 32, 48])
def test_powerlaw_axt_noc(engine, wl, th):
    m, x, yr = wl("powerlaw_1m")
    a = H.AXT(m, x, th=th)
    d = D.DeviceAXT(engine, a)
    y = _nan(engine, m.nrows)
    d.int_axt(y)
    g = _host(y, m.nrows)
    ok = _normal_rows(m, th)
    assert bitwise_equal(g[ok], O.int_axt(a)[ok]), th
    e = _serr(m, x, yr, g)
    print(f"[fullsize] powerlaw AXT NOC th{th}: max scaled err {e:.2e}", flush=True)
    assert e <= TOL64


@pytest.mark.parametrize("bs", [128, 256, 512])
def test_powerlaw_axt_com(engine, wl, bs):
    """COM against CSR-naive (the reference's COM kernel is defective): rows
    of up to 128 entries carry cf_CSR's bits."""
    m, x, yr = wl("powerlaw_1m")
    a = H.AXT(m, x, th=1, mode="COM", bs=bs)
    d = D.DeviceAXT(engine, a)
    y = _nan(engine, m.nrows)
    d.int_axt(y)
    g = _host(y, m.nrows)
    short = np.diff(m.row.astype(np.int64)) <= 128
    assert bitwise_equal(g[short], yr[short]), bs
    assert _serr(m, x, yr, g) <= TOL64, bs


@pytest.mark.parametrize("th", [1, 4])
def test_powerlaw_column_map_adjoint(engine, wl, th):
    """The column-map product (1 and 8 right-hand sides), the adjoint
    x_bar = A^T y_bar on the transposed map by both value paths, and the
    value gradient, at full size."""
    import torch
    m, x, yr = wl("powerlaw_1m")
    n = m.nrows
    a = H.AXT(m, x, th=th, want_acol=True)
    g = D.DeviceAXTGather(engine, a)
    y = _nan(engine, n)
    g.matvec(engine.upload(x), y)
    g1 = _host(y, n)
    e = D.DeviceAXT(engine, a)
    ye = _nan(engine, n)
    e.int_axt(ye)
    ok = _normal_rows(m, th) & _normal_rows(m, th, colmap=True)
    assert bitwise_equal(g1[ok], _host(ye, n)[ok])  # same tiles, same trees
    assert _serr(m, x, yr, g1) <= TOL64
    del e, ye
    k = 8
    X = np.random.default_rng(8).uniform(-1, 1, (n, k))
    Y = _nan(engine, n * k)
    g.matvec(engine.upload(X.reshape(-1)), Y, k=k)
    Yr = O.csr_mm(m.row, m.col, m.val, X)
    err, at_ = O.scaled_err_mm(m.row, m.col, m.val, X, Yr, _host(Y, n * k).reshape(n, k))
    assert err <= TOL64, (err, at_)
    del Y
    adj = D.AxtAdjoint(engine, m, g)
    ybar = H.normal(n, seed=12)
    yb = engine.upload(ybar)
    xb, xb2 = _nan(engine, n), _nan(engine, n)
    adj.adjoint(yb, xb)
    adj.adjoint(yb, xb2, vals=g.aval)
    got = _host(xb, n)
    at_csr, _ = m.transpose()
    ref = O.csr_transpose_product(m.row, m.col, m.val, ybar)
    e_adj = _serr(at_csr, ybar, ref, got)
    print(f"[fullsize] powerlaw adjoint th{th}: max scaled err {e_adj:.2e}", flush=True)
    assert e_adj <= TOL64
    assert bitwise_equal(_host(xb2, n), got)
    vb = _nan(engine, a.lenAX // 2)
    g.grad_vals(engine.upload(x), yb, vb)
    rows = np.repeat(np.arange(n), np.diff(m.row.astype(np.int64)))
    want = np.zeros(a.lenAX // 2)
    want[a.nzslot] = ybar[rows] * x[m.col]
    assert bitwise_equal(_host(vb, a.lenAX // 2), want)
    del adj, g, vb
    torch.cuda.empty_cache()


@pytest.mark.parametrize("vt,h,tol_adj,tol_fd", [(D.F64, 1e-6, 1e-12, 1e-6),
                                                 (D.F32, 1e-3, 1e-6, 1e-4)])
def test_powerlaw_autograd_finite_differences(engine, wl, vt, h, tol_adj, tol_fd):
    """L(v, x) = y_bar^T A(v) x on the 2^20-row matrix: x_bar against the CSR
    transpose product, and central differences along 64 random directions
    against the directional derivative from (v_bar, x_bar).  L is bilinear,
    so the differences are exact but for rounding; they are taken along the
    directions the perturbed inputs actually have after rounding to the
    storage type, and L is summed in FP64."""
    import torch

    from int_spmv_amd import autograd as AG
    m, _, _ = wl("powerlaw_1m")
    n = m.nrows
    op = AG.AxtOperator(engine, m, th=1, val_type=vt)
    dev, dt = engine.torch_device, op.dtype
    v = op.initial_values().requires_grad_(True)
    gen = torch.Generator(device=dev)
    gen.manual_seed(21)
    x = (torch.rand(n, device=dev, dtype=torch.float64, generator=gen) * 2 - 1).to(dt)
    x.requires_grad_(True)
    ybar = torch.randn(n, device=dev, dtype=torch.float64, generator=gen).to(dt)
    L = (AG.spmv(v, x, op) * ybar).sum()
    L.backward()
    gv, gx = v.grad.detach().double(), x.grad.detach().double()
    vals = m.val.astype(np.float32).astype(np.float64) if vt == D.F32 else m.val
    yb = ybar.double().cpu().numpy()
    ref = O.csr_transpose_product(m.row, m.col, vals, yb)
    at, _ = m.transpose()
    assert _serr(at, yb, ref, gx.cpu().numpy()) <= tol_adj
    occupied = (op.fwd.acol.view(torch.int32) != -1).to(torch.float64)
    worst = 0.0
    with torch.no_grad():
        yb64 = ybar.double()
        for _ in range(64):
            dv = torch.randn(len(occupied), device=dev, dtype=torch.float64, generator=gen) * occupied
            dx = torch.randn(n, device=dev, dtype=torch.float64, generator=gen)
            vp, vm = (v.double() + h * dv).to(dt), (v.double() - h * dv).to(dt)
            xp, xm = (x.double() + h * dx).to(dt), (x.double() - h * dx).to(dt)
            dv_e = (vp.double() - vm.double()) / (2 * h)
            dx_e = (xp.double() - xm.double()) / (2 * h)
            lp = (AG.spmv(vp, xp, op).double() * yb64).sum()
            lm = (AG.spmv(vm, xm, op).double() * yb64).sum()
            fd = ((lp - lm) / (2 * h)).item()
            an = ((gv * dv_e).sum() + (gx * dx_e).sum()).item()
            scale = ((gv * dv_e).abs().sum() + (gx * dx_e).abs().sum()).item()
            worst = max(worst, abs(fd - an) / scale)
    print(f"[fullsize] finite differences {'F64' if vt == D.F64 else 'F32'}: {worst:.2e}",
          flush=True)
    assert worst < tol_fd, worst
    del op
    torch.cuda.empty_cache()


# ------------------------------------------------ stencil27_128, rmat_22


@pytest.mark.parametrize("name", ["stencil27_128", "rmat_22"])
def test_precision_workloads(engine, wl, name):
    """FP64 (drop-in layout: CSR, AXC, K1p, AXT NOC th 1/4, COM), FP32
    storage (th 1 and 4) and BF16 values with FP32 x (th 1 and 4), each
    against the FP64 oracle of the unrounded matrix."""
    import torch
    m, x, yr = wl(name)
    n = m.nrows
    errs = {}
    xd = engine.upload(x)
    d = D.DeviceCSR(engine, m)
    y = _nan(engine, n)
    d.spmv(xd, y)
    errs["CSR_F64"] = _serr(m, x, yr, _host(y, n))
    y = _nan(engine, n)
    d.spmv_exact(xd, y)
    assert bitwise_equal(_host(y, n), yr), f"{name}: exact CSR plan"
    del d
    c = H.AXC(m, x)
    dc = D.DeviceAXC(engine, c)
    y = _nan(engine, n)
    dc.int_axc(y)
    g = _host(y, n)
    # rows in normal blocks (up to 48 bricks, 64 with many empty rows, whose
    # 128-brick blocks take two load rounds): the reference kernel's bits
    ok = np.diff(c.brp.astype(np.int64)) // 16 <= (64 if _many_empty(m) else 48)
    assert bitwise_equal(g[ok], O.int_axc(n, c.ax, c.brp)[ok]), f"{name}: AXC"
    errs["AXC_F64"] = _serr(m, x, yr, g)
    del dc, c
    k = H.K1(m)
    dk = D.DeviceK1(engine, k)
    y = _nan(engine, n)
    dk.int_k1p(xd, y)
    errs["K1P_F64"] = _serr(m, x, yr, _host(y, n))
    del dk, k
    for th in (1, 4):
        a = H.AXT(m, x, th=th)
        dd = D.DeviceAXT(engine, a)
        y = _nan(engine, n)
        dd.int_axt(y)
        g = _host(y, n)
        ok = _normal_rows(m, th)
        assert bitwise_equal(g[ok], O.int_axt(a)[ok]), (name, th)
        errs[f"AXT_NOC_H{th}_F64"] = _serr(m, x, yr, g)
        del dd
        for vt, xt, tag in ((D.F32, D.F32, "F32"), (D.BF16, D.F32, "BF16")):
            dd = D.DeviceAXT(engine, a, val_type=vt, x_type=xt)
            y = _nan(engine, n, np.float32)
            dd.int_axt(y)
            errs[f"AXT_NOC_H{th}_{tag}"] = _serr(m, x, yr, _host(y, n))
            del dd
        del a
    a = H.AXT(m, x, th=1, mode="COM", bs=128)
    short = np.diff(m.row.astype(np.int64)) <= 128
    for vt, tag in ((D.F64, "F64"), (D.F32, "F32")):
        dd = D.DeviceAXT(engine, a, val_type=vt, x_type=vt)
        y = _nan(engine, n, np.float64 if vt == D.F64 else np.float32)
        dd.int_axt(y)
        g = _host(y, n)
        if vt == D.F64:  # rows of up to 128 entries: cf_CSR's bits
            assert bitwise_equal(g[short], yr[short]), f"{name}: COM"
        errs[f"AXT_COM_BS128_{tag}"] = _serr(m, x, yr, g)
        del dd
    del a
    torch.cuda.empty_cache()
    print(f"[fullsize] {name} max scaled errors: "
          + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for f, e in errs.items():
        tol = TOL64 if f.endswith("F64") else TOL32 if f.endswith("F32") else TOL16
        assert e <= tol, (name, f, e)


# -------------------------------------------------------------- batch_256


def test_batch_256_full(engine):
    """All 256 matrices of the batch workload (FP32 values and x, 1..32
    right-hand sides ~ U[-1, 1], tile height and RHS blocking from the
    autotuner's byte model) in one persistent launch, captured and replayed
    twice with eager work between: every result column of every matrix
    within 1e-6 of cf_CSR of the FP32-rounded values and x, and the replays
    bit-identical to the eager launch."""
    import torch
    dev = engine.torch_device
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    ents = []
    print("[fullsize] batch_256: generating and building", flush=True)
    for i, (m, k) in enumerate(batch_256(seed=0)):
        c = H.tune_model(m, k, 4, 4)[0]
        a = H.AXT(m, np.zeros(m.nrows), th=c["tileH"], want_acol=True)
        X = torch.rand(m.nrows * k, dtype=torch.float32, device=dev, generator=gen) * 2 - 1
        Y = torch.full((m.nrows * k,), float("nan"), dtype=torch.float32, device=dev)
        ents.append(dict(a=a, X=X, Y=Y, k=k, kc=c["kc"], m=m))
    b = D.DeviceBatch(engine, ents, D.F32, D.F32)
    for e in ents:
        e.pop("a")
    b.launch()
    torch.cuda.synchronize()
    eager = [e["Y"].cpu().numpy().copy() for e in ents]
    worst, where = 0.0, None
    for i, e in enumerate(ents):
        m, k = e["m"], e["k"]
        v32 = m.val.astype(np.float32).astype(np.float64)
        X = e["X"].double().cpu().numpy().reshape(m.nrows, k)
        Yr = O.csr_mm(m.row, m.col, v32, X)
        err, pos = O.scaled_err_mm(m.row, m.col, v32, X, Yr,
                                   eager[i].astype(np.float64).reshape(m.nrows, k))
        if err > worst:
            worst, where = err, (i, m.nrows, k, pos)
    print(f"[fullsize] batch_256: {len(ents)} matrices, max scaled err {worst:.2e} at {where}",
          flush=True)
    assert worst <= TOL32, (worst, where)
    b.capture()
    for rep in range(2):
        for e in ents:
            e["Y"].fill_(float("nan"))
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for i, e in enumerate(ents):
            assert bitwise_equal(e["Y"].cpu().numpy(), eager[i]), (rep, i)
    del b, ents
    torch.cuda.empty_cache()
