"""GPU: column-map AXT (multi-RHS), the transposed tile map (adjoint), the
value gradient, device-side x refresh, and the autograd Function checked by
central finite differences.  Oracle: CSR-naive (cf_CSR) and the sequential
CSR transpose product."""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal, csr_long_rows, csr_sparse_gaps, csr_with_empty_rows, scaled

pytestmark = pytest.mark.gpu


def _host(t, n):
    import torch
    torch.cuda.synchronize()
    return t.reshape(-1)[:n].double().cpu().numpy()


def _nan(eng, n, dtype=np.float64):
    t = eng.empty(n, dtype)
    t.fill_(float("nan"))
    return t


def _mats():
    return [("powerlaw", H.CSR.powerlaw(n=1 << 13, target_nnz=200_000, maxlen=4096, seed=2)),
            ("empty_rows", csr_with_empty_rows()),
            ("long_rows", csr_long_rows()),
            ("sparse_gaps", csr_sparse_gaps())]


@pytest.fixture(scope="module")
def mats():
    return _mats()


@pytest.mark.parametrize("th", [1, 4])
def test_gather_k1_equals_embedded(engine, mats, th):
    """Same tiles, same trees: for every row inside a normal block the
    column-map product of one x has the bits of the embedded (reference
    layout) product; long rows agree within the FP64 tolerance."""
    for name, m in mats:
        x = H.uniform(m.nrows, seed=3)
        a = H.AXT(m, x, th=th, want_acol=True)
        g = D.DeviceAXTGather(engine, a)
        e = D.DeviceAXT(engine, a)
        y1, y2 = _nan(engine, m.nrows), _nan(engine, m.nrows)
        g.matvec(engine.upload(x), y1)
        e.int_axt(y2)
        g1, g2 = _host(y1, m.nrows), _host(y2, m.nrows)
        per = 8 if th == 1 else th
        units = (np.diff(m.row.astype(np.int64)) + per - 1) // per
        ok = units <= (48 if th == 1 else 64)  # rows in normal blocks of both plans
        assert bitwise_equal(g1[ok], g2[ok]), (name, th)
        assert scaled(m, x, O.cf_csr(m.row, m.col, m.val, x), g1) <= 1e-12, (name, th)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 13, 32])
@pytest.mark.parametrize("th", [1, 8])
def test_gather_multi_rhs(engine, mats, k, th):
    for name, m in mats:
        n = m.nrows
        X = np.stack([H.uniform(n, seed=100 + j) for j in range(k)], axis=1)  # n x k
        a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
        g = D.DeviceAXTGather(engine, a)
        Y = _nan(engine, n * k)
        g.matvec(engine.upload(X.reshape(-1)), Y, k=k)
        Yg = _host(Y, n * k).reshape(n, k)
        for j in range(k):
            yr = O.cf_csr(m.row, m.col, m.val, X[:, j])
            assert scaled(m, X[:, j], yr, Yg[:, j]) <= 1e-12, (name, k, th, j)


@pytest.mark.parametrize("vt", [D.F64, D.F32])
def test_gather_wide_x_plain_loads(engine, mats, vt):
    """x spanning 2 GiB or more (here through a large leading dimension) is
    gathered with plain loads instead of the buffer resource; both give the
    same bits, padding included."""
    import torch
    m = mats[0][1]
    n = m.nrows
    dt = torch.float64 if vt == D.F64 else torch.float32
    wide = (1 << 31) // (n * (8 if vt == D.F64 else 4))  # n * ldx * size == 2 GiB
    for th, k in ((1, 8), (1, 3), (4, 1), (4, 8)):
        X = np.stack([H.uniform(n, seed=60 + j) for j in range(k)], axis=1)
        a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
        g = D.DeviceAXTGather(engine, a, val_type=vt, x_type=vt)
        Xp = torch.tensor(X.reshape(-1), dtype=dt, device=engine.torch_device)
        Xw = torch.zeros((n, wide), dtype=dt, device=engine.torch_device)
        Xw[:, :k] = Xp.view(n, k)
        Y1 = torch.full((n * k,), float("nan"), dtype=dt, device=engine.torch_device)
        Y2 = torch.full_like(Y1, float("nan"))
        g.matvec(Xp, Y1, k=k)
        g.matvec(Xw.view(-1), Y2, k=k, ldx=wide)
        torch.cuda.synchronize()
        assert bitwise_equal(Y1.cpu().numpy(), Y2.cpu().numpy()), (th, k)
        del Xw
    torch.cuda.empty_cache()


@pytest.mark.parametrize("vt", [D.F32, D.BF16])
def test_gather_sixteen_per_pass(engine, mats, vt):
    """Passes of 16 right-hand sides (narrow at tile height 4, wide at tile
    height 1) against passes of 8: each right-hand side keeps its own sums in
    the same order, so the bits are equal -- at tile height 1 for the rows of
    the plan's normal blocks (a wide pass adds a long row's tiles in another
    order: within tolerance); FP64 refuses narrow passes of 16."""
    import torch
    for name, m in mats[:2]:
        n = m.nrows
        for th, k in ((1, 16), (1, 21), (4, 32)):
            X = np.stack([H.uniform(n, seed=80 + j) for j in range(k)], axis=1).astype(np.float32)
            a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
            g = D.DeviceAXTGather(engine, a, val_type=vt, x_type=D.F32)
            Xd = engine.upload(X.reshape(-1))
            Y16, Y8 = _nan(engine, n * k, np.float32), _nan(engine, n * k, np.float32)
            g.matvec(Xd, Y16, k=k)
            g.matvec(Xd, Y8, k=k, kc=8)
            torch.cuda.synchronize()
            y16, y8 = Y16.cpu().numpy().reshape(n, k), Y8.cpu().numpy().reshape(n, k)
            ok = (np.diff(m.row.astype(np.int64)) + 7) // 8 <= 64 if th == 1 else slice(None)
            assert bitwise_equal(y16[ok], y8[ok]), (name, th, k)
            Yg = _host(Y16, n * k).reshape(n, k)
            Y8g = _host(Y8, n * k).reshape(n, k)
            for j in (0, k - 1):
                yr = O.cf_csr(m.row, m.col, m.val, X[:, j].astype(np.float64))
                tol = 1e-6 if vt == D.F32 else 1e-2
                assert scaled(m, X[:, j].astype(np.float64), yr, Y8g[:, j]) <= tol, (name, th, k, j)
                yr = O.cf_csr(m.row, m.col, m.val, X[:, j].astype(np.float64))
                tol = 1e-6 if vt == D.F32 else 1e-2
                assert scaled(m, X[:, j].astype(np.float64), yr, Yg[:, j]) <= tol, (name, th, k, j)
    # FP64 has no narrow pass of 16 (tile height 4 refuses it); at tile height
    # 1 kc = 16 is a wide pass
    m = mats[0][1]
    for th, want in ((4, -1), (1, 0)):
        a = H.AXT(m, np.zeros(m.nrows), th=th, want_acol=True)
        g = D.DeviceAXTGather(engine, a)
        Y = _nan(engine, m.nrows * 16)
        st = D.lib().spmv_axt_gather(engine.h, g.plan.h, D._dptr(g.aval), D._dptr(g.acol),
                                     D._dptr(g.sec),
                                     D._dptr(engine.upload(np.zeros(m.nrows * 16))), 16,
                                     D._dptr(Y), 16, 16, 16)
        assert st == want, th


@pytest.mark.parametrize("vt,xt,tol", [(D.F32, D.F32, 1e-6), (D.BF16, D.F32, 1e-2)])
def test_gather_precision(engine, mats, vt, xt, tol):
    for name, m in mats:
        n, k = m.nrows, 4
        X = np.stack([H.uniform(n, seed=7 + j) for j in range(k)], axis=1)
        a = H.AXT(m, np.zeros(n), th=1, want_acol=True)
        g = D.DeviceAXTGather(engine, a, val_type=vt, x_type=xt)
        Y = _nan(engine, n * k, np.float32)
        g.matvec(engine.upload(X.astype(np.float32).reshape(-1)), Y, k=k)
        Yg = _host(Y, n * k).reshape(n, k)
        for j in range(k):
            yr = O.cf_csr(m.row, m.col, m.val, X[:, j])
            assert scaled(m, X[:, j], yr, Yg[:, j]) <= tol, (name, vt, j)


@pytest.mark.parametrize("th", [1, 2, 8])
def test_adjoint_and_grad_vals(engine, mats, th):
    for name, m in mats:
        n = m.nrows
        x = H.uniform(n, seed=11)
        ybar = H.normal(n, seed=12)
        a = H.AXT(m, x, th=th, want_acol=True)
        fwd = D.DeviceAXTGather(engine, a)
        adj = D.AxtAdjoint(engine, m, fwd)
        xb = _nan(engine, n)
        adj.adjoint(engine.upload(ybar), xb)
        got = _host(xb, n)
        ref = O.csr_transpose_product(m.row, m.col, m.val, ybar)
        at, _ = m.transpose()
        assert scaled(at, ybar, ref, got) <= 1e-12, (name, th)
        # A's values read through the slot map: the same bits, k = 1 and 3
        xb2 = _nan(engine, n)
        adj.adjoint(engine.upload(ybar), xb2, vals=fwd.aval)
        assert bitwise_equal(_host(xb2, n), got), (name, th)
        Yb = np.stack([H.normal(n, seed=20 + j) for j in range(3)], axis=1)
        Xr, Xp = _nan(engine, 3 * n), _nan(engine, 3 * n)
        adj.adjoint(engine.upload(Yb.reshape(-1)), Xr, k=3)
        adj.adjoint(engine.upload(Yb.reshape(-1)), Xp, k=3, vals=fwd.aval)
        assert bitwise_equal(_host(Xp, 3 * n), _host(Xr, 3 * n)), (name, th)
        vb = _nan(engine, a.lenAX // 2)
        fwd.grad_vals(engine.upload(x), engine.upload(ybar), vb)
        g = _host(vb, a.lenAX // 2)
        rows = np.repeat(np.arange(n), np.diff(m.row.astype(np.int64)))
        want = np.zeros(a.lenAX // 2)
        want[a.nzslot] = ybar[rows] * x[m.col]
        assert bitwise_equal(g, want), (name, th)
        # embedded layout: the tiles already carry x[col]
        e = D.DeviceAXT(engine, a)
        vb2 = _nan(engine, a.lenAX // 2)
        D.grad_vals_embedded(e, engine.upload(ybar), vb2)
        assert bitwise_equal(_host(vb2, a.lenAX // 2), want), (name, th)


@pytest.mark.parametrize("vt", [D.F64, D.F32])
def test_adjoint_values_paths(engine, mats, vt):
    """After a value update adjoint(vals=...) follows the new values by either
    path, with the same bits: A's values doubled (an exact scaling in binary
    floating point) give exactly twice the adjoint of the stored values.
    tune_values_path keeps one of the two paths; the result does not change."""
    for name, m in mats:
        n = m.nrows
        dt = np.float64 if vt == D.F64 else np.float32
        a = H.AXT(m, np.zeros(n), th=4, want_acol=True)
        fwd = D.DeviceAXTGather(engine, a, vt, vt)
        adj = D.AxtAdjoint(engine, m, fwd)
        yb = engine.upload(H.normal(n, seed=13).astype(dt))
        base = _nan(engine, n, dt)
        adj.adjoint(yb, base)
        want = 2 * _host(base, n)
        v2 = fwd.aval * 2
        for path in ("permuted", "refresh"):
            adj.values_path = path
            adj.refresh(fwd.aval)  # A^T's copy back at the stored values
            xb = _nan(engine, n, dt)
            adj.adjoint(yb, xb, vals=v2)
            assert bitwise_equal(_host(xb, n), want), (name, vt, path)
        adj.refresh(fwd.aval)
        t = adj.tune_values_path()
        assert t["path"] in ("permuted", "refresh") and adj.values_path == t["path"], t
        xb = _nan(engine, n, dt)
        adj.adjoint(yb, xb, vals=v2)
        assert bitwise_equal(_host(xb, n), want), (name, vt, t)


@pytest.mark.parametrize("th", [1, 4, 8])
def test_refresh_x_on_device(engine, mats, th):
    """The device refresh gives the host rebuild's arrays bit for bit,
    padding included (x half 0)."""
    for name, m in mats:
        x0, x1 = H.uniform(m.nrows, seed=1), H.uniform(m.nrows, seed=2)
        a = H.AXT(m, x0, th=th, want_acol=True)
        e = D.DeviceAXT(engine, a)
        stale = a.ax.copy()
        stale.reshape(-1, 16)[:, 8:] = np.nan  # every x half must be rewritten
        e.ax.copy_(engine.upload(stale))
        D.refresh_x(e, engine.upload(a.acol), engine.upload(x1))
        y = _nan(engine, m.nrows)
        e.int_axt(y)
        b = H.AXT(m, x1, th=th)
        import torch
        torch.cuda.synchronize()
        assert bitwise_equal(e.ax.cpu().numpy(), b.ax), (name, th)
        assert scaled(m, x1, O.cf_csr(m.row, m.col, m.val, x1), _host(y, m.nrows)) <= 1e-12


@pytest.mark.parametrize("vt,h,tol_adj,tol_fd", [(D.F64, 1e-6, 1e-12, 1e-6),
                                                 (D.F32, 1e-3, 1e-6, 1e-4)])
def test_autograd_finite_differences(engine, vt, h, tol_adj, tol_fd):
    """L(v, x) = ybar^T A(v) x is bilinear, so central differences along any
    direction equal the directional derivative up to rounding (FP64 with
    h = 1e-6, FP32 with h = 1e-3; a float32 numpy restatement of this case
    stays at 3e-6 relative, so 1e-4 holds with margin)."""
    import torch

    from int_spmv_amd import autograd as AG
    m = H.CSR.powerlaw(n=1 << 12, target_nnz=80_000, maxlen=2000, seed=5)
    n = m.nrows
    op = AG.AxtOperator(engine, m, th=1, val_type=vt)
    dev, dt = engine.torch_device, op.dtype
    v = op.initial_values().requires_grad_(True)
    xh = H.uniform(n, seed=21).astype(np.float32 if vt == D.F32 else np.float64)
    x = torch.tensor(xh, device=dev, requires_grad=True)
    ybar = torch.tensor(H.normal(n, seed=22), device=dev, dtype=dt)
    L = (AG.spmv(v, x, op) * ybar).sum()
    L.backward()
    gv, gx = v.grad.detach(), x.grad.detach()
    # x_bar against the CSR transpose product (of the values as stored)
    vals = m.val.astype(np.float32).astype(np.float64) if vt == D.F32 else m.val
    yb = ybar.double().cpu().numpy()
    ref = O.csr_transpose_product(m.row, m.col, vals, yb)
    at, _ = m.transpose()
    assert scaled(at, yb, ref, gx.double().cpu().numpy()) <= tol_adj
    occupied = op.fwd.acol.cpu().numpy().view(np.uint32) != D.PAD
    rng = np.random.default_rng(0)
    worst = 0.0
    with torch.no_grad():
        for _ in range(64):
            dv = torch.tensor(rng.standard_normal(len(occupied)) * occupied, device=dev, dtype=dt)
            dx = torch.tensor(rng.standard_normal(n), device=dev, dtype=dt)
            lp = (AG.spmv(v + h * dv, x + h * dx, op) * ybar).sum()
            lm = (AG.spmv(v - h * dv, x - h * dx, op) * ybar).sum()
            fd = ((lp - lm) / (2 * h)).item()
            an = ((gv * dv).sum() + (gx * dx).sum()).item()
            scale = ((gv * dv).abs().sum() + (gx * dx).abs().sum()).item()
            worst = max(worst, abs(fd - an) / scale)
    assert worst < tol_fd, worst


def test_plan_layouts_are_checked(engine):
    """A column-map plan is refused by the embedded product, and an embedded
    plan with larger blocks than the column-map kernels stage is refused by
    the column-map product (INTSPMV_ERR_ARG), instead of running."""
    import ctypes as C

    import torch
    m = H.CSR.powerlaw(n=4096, target_nnz=60_000, maxlen=512, seed=2)
    a = H.AXT(m, H.uniform(m.nrows, seed=1), th=2, want_acol=True)
    g = D.DeviceAXTGather(engine, a)
    e = D.DeviceAXT(engine, a)  # FP64 reference layout, th = 2: 256-column blocks
    y = torch.empty(m.nrows, dtype=torch.float64, device=engine.torch_device)
    x = engine.upload(H.uniform(m.nrows, seed=3))
    st = D.lib().spmv_axt(engine.h, g.plan.h, D._dptr(e.ax), None, D._dptr(e.sec), D._dptr(y))
    assert st == -1
    st = D.lib().spmv_axt_gather(engine.h, e.plan.h, D._dptr(g.aval), D._dptr(g.acol),
                                 D._dptr(g.sec), D._dptr(x), 1, D._dptr(y), 1, 1, 0)
    assert st == -1
    assert e.plan.info()["nblocks"] > 0 and C.sizeof(D.AxtDesc) > 0


@pytest.mark.parametrize("th", [1, 4])
def test_adjoint_from_c(th):
    """A C host computes x_bar = A^T y_bar through the two C libraries alone
    (tests/c/adjoint_from_c.c: intspmv_axt_transpose_map, spmv_permute_vals,
    spmv_axt_gather and spmv_axt_gather_permuted): within 1e-12 of the
    sequential transpose product, both device paths bit-identical."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "adjoint_from_c")
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    print(f"[adjoint_from_c] th={th}", flush=True)
    r = subprocess.run([exe, "40000", "900000", str(th)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "paths_differ 0" in r.stdout, r.stdout


@pytest.mark.parametrize("vt", [D.F64, D.F32, D.BF16])
def test_wide_passes_unaligned_rows(engine, mats, vt):
    """Wide passes (tile height 1) load each x row with 16-byte loads when
    rows are aligned, one value at a time otherwise (here an odd leading
    dimension), with lanes past k idle: the same bits either way, equal to
    narrow passes of 2 for rows inside a block, and within tolerance of
    cf_CSR everywhere."""
    import torch
    m = mats[0][1]
    n = m.nrows
    xt = D.F64 if vt == D.F64 else D.F32
    dt = torch.float64 if xt == D.F64 else torch.float32
    a = H.AXT(m, np.zeros(n), th=1, want_acol=True)
    g = D.DeviceAXTGather(engine, a, vt, xt)
    for k in (7, 13, 20):
        X = np.stack([H.uniform(n, seed=90 + j) for j in range(k)], axis=1)
        X = X.astype(np.float32).astype(np.float64) if xt == D.F32 else X
        Xa = torch.tensor(X.reshape(-1), dtype=dt, device=engine.torch_device)
        Xo = torch.zeros((n, k + 1), dtype=dt, device=engine.torch_device)  # odd rows
        Xo[:, :k] = Xa.view(n, k)
        Ya, Yo, Y2 = (torch.full((n * k,), float("nan"), dtype=dt, device=engine.torch_device)
                      for _ in range(3))
        g.matvec(Xa, Ya, k=k)
        g.matvec(Xo.view(-1), Yo, k=k, ldx=k + 1)
        g.matvec(Xa, Y2, k=k, kc=2)
        torch.cuda.synchronize()
        ya, yo, y2 = (t.double().cpu().numpy().reshape(n, k) for t in (Ya, Yo, Y2))
        assert bitwise_equal(ya, yo), (vt, k)
        units = (np.diff(m.row.astype(np.int64)) + 7) // 8
        ok = units <= 64  # rows in the plan's normal blocks
        assert bitwise_equal(ya[ok], y2[ok]), (vt, k)
        vals = m.val if vt == D.F64 else (m.val.astype(np.float32).astype(np.float64) if vt == D.F32
                                          else D.bf16_to_f32(D.bf16_bits(m.val)).astype(np.float64))
        Yr = O.csr_mm(m.row, m.col, vals, X)
        tol = 1e-12 if vt == D.F64 else 1e-6
        assert O.scaled_err_mm(m.row, m.col, vals, X, Yr, ya)[0] <= tol, (vt, k)


@pytest.mark.parametrize("vt", [D.F64, D.F32])
def test_adjoint_grad_fused(engine, mats, vt):
    """spmv_axt_gather_grad: x_bar with the bits of the plain adjoint, and the
    value gradient in A^T's slot order with the bits of spmv_axt_grad_vals
    moved there by the slot map (padding 0)."""
    import torch
    for name, m in mats:
        n = m.nrows
        dt = np.float64 if vt == D.F64 else np.float32
        for th in (1, 4):
            a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
            fwd = D.DeviceAXTGather(engine, a, vt, vt)
            adj = D.AxtAdjoint(engine, m, fwd)
            xd = engine.upload(H.uniform(n, seed=41).astype(dt))
            yb = engine.upload(H.normal(n, seed=42).astype(dt))
            x1, x2 = _nan(engine, n, dt), _nan(engine, n, dt)
            gv = _nan(engine, a.lenAX // 2, dt)
            gvt = _nan(engine, adj.nslots_t, dt)
            adj.adjoint(yb, x1)
            adj.adjoint_grad(yb, x2, xd, gvt)
            fwd.grad_vals(xd, yb, gv)
            want = torch.empty_like(gvt)
            engine._check(D.lib().spmv_permute_vals(engine.h, adj.nslots_t, vt, D._dptr(adj.permT),
                                                    D._dptr(gv), D._dptr(want)), "permute")
            assert bitwise_equal(_host(x2, n), _host(x1, n)), (name, th)
            assert bitwise_equal(_host(gvt, adj.nslots_t), _host(want, adj.nslots_t)), (name, th)


@pytest.mark.parametrize("vt", [D.F64, D.F32])
def test_autograd_pair_keeps_copies_in_step(engine, vt):
    """Plain SGD on the pair operator's two leaves: after every step A^T's
    copy is A's values moved by the slot map, bit for bit, and each step's
    loss, x_bar and value gradient equal the single-copy operator's."""
    import torch

    from int_spmv_amd import autograd as AG
    m = H.CSR.powerlaw(n=1 << 12, target_nnz=80_000, maxlen=2000, seed=7)
    n = m.nrows
    op = AG.AxtOperator(engine, m, th=1, val_type=vt)
    dev, dt = engine.torch_device, op.dtype
    v, vt_ = (t.requires_grad_(True) for t in op.initial_values_pair())
    v1 = op.initial_values().requires_grad_(True)
    x = torch.tensor(H.uniform(n, seed=43), device=dev, dtype=dt, requires_grad=True)
    x1 = x.detach().clone().requires_grad_(True)
    ybar = torch.tensor(H.normal(n, seed=44), device=dev, dtype=dt)
    lr = 0.01
    for step in range(3):
        L = (AG.spmv_pair(v, vt_, x, op) * ybar).sum()
        L1 = (AG.spmv(v1, x1, op) * ybar).sum()
        L.backward()
        L1.backward()
        assert bitwise_equal(L.detach().double().cpu().numpy(), L1.detach().double().cpu().numpy())
        assert bitwise_equal(x.grad.double().cpu().numpy(), x1.grad.double().cpu().numpy()), step
        assert bitwise_equal(v.grad.double().cpu().numpy(), v1.grad.double().cpu().numpy()), step
        with torch.no_grad():
            for p in (v, vt_, x, v1, x1):
                p -= lr * p.grad
                p.grad = None
        want = torch.empty_like(vt_)
        engine._check(D.lib().spmv_permute_vals(engine.h, op.adj.nslots_t, vt, D._dptr(op.adj.permT),
                                                D._dptr(v.detach()), D._dptr(want)), "permute")
        torch.cuda.synchronize()
        assert bitwise_equal(vt_.detach().double().cpu().numpy(), want.double().cpu().numpy()), step
