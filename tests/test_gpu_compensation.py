"""FP32 accumulation of long rows is compensated (Kahan-Babuska-Neumaier).

The case: rows of 8,192 and 20,000 entries whose products are
[1, s, s, ..., s, -1] with s = 2^-25 (x = 1 on those columns), among short
random rows.  The exact row sum is (L - 2) s.  A plain FP32 running sum in
CSR order loses every s (1 + 2^-25 rounds back to 1, below half an ulp) and
ends at 0: scaled error (L - 2) s / (2 + (L - 2) s) > 1e-4.  A compensated
running sum keeps them.  Every value here is exact in FP32 and BF16, so the
storage type adds no error and each FP32 kernel path (embedded FP32 and
BF16 storage at several tile heights, COM, the column map with 1, 3 and 16
right-hand sides -- narrow and wide passes -- and the batched launch) must
meet 1e-6 on these rows.
"""
import numpy as np
import pytest

import oracle as O
from int_spmv_amd import device as D
from int_spmv_amd import host as H

S = 2.0 ** -25
LONG = {5: 8192, 700: 20000}  # row -> length (8192: one wave; 20000: split)


def _case():
    n = 21000
    rng = np.random.default_rng(17)
    lens = rng.integers(1, 20, size=n)
    for r, L in LONG.items():
        lens[r] = L
    row = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    cols, vals = [], []
    for r in range(n):
        L = lens[r]
        if r in LONG:
            c = np.arange(L)
            v = np.full(L, S)
            v[0], v[-1] = 1.0, -1.0
        else:
            c = np.sort(rng.choice(n, size=L, replace=False))
            v = rng.standard_normal(L)
        cols.append(c)
        vals.append(v)
    m = H.CSR.from_arrays(row, np.concatenate(cols).astype(np.uint32), np.concatenate(vals))
    x = rng.uniform(-1, 1, n)
    x[: max(LONG.values())] = 1.0
    return m, x


def _seq_fp32(p, compensated):
    s, c = np.float32(0), np.float32(0)
    for v in p.astype(np.float32):
        t = np.float32(s + v)
        if compensated:
            c = np.float32(c + (np.float32(s - t) + v if abs(s) >= abs(v) else np.float32(v - t) + s))
        s = t
    return float(np.float32(s + c))


def test_case_defeats_plain_fp32_sum():
    """On the CPU: a plain FP32 running sum of each long row fails 1e-6 by
    two orders of magnitude; the compensated one is exact."""
    m, x = _case()
    for r, L in LONG.items():
        a, b = m.row[r], m.row[r + 1]
        p = m.val[a:b] * x[m.col[a:b]]
        exact, scale = (L - 2) * S, float(np.abs(p).sum())
        assert abs(_seq_fp32(p, False) - exact) / scale > 1e-4
        assert _seq_fp32(p, True) == exact


def _check(m, x, y, tol_other, what):
    yr = O.cf_csr(m.row, m.col, m.val.astype(np.float32).astype(np.float64),
                  x.astype(np.float32).astype(np.float64))
    lr = np.array(sorted(LONG))
    long_rows = np.zeros(m.nrows, bool)
    long_rows[lr] = True
    scale = np.array([np.abs(m.val[m.row[r]:m.row[r + 1]]).sum() for r in lr])  # |x| = 1 there
    e_long = np.max(np.abs(y[lr] - yr[lr]) / scale)
    assert e_long <= 1e-6, (what, e_long, y[lr], yr[lr])
    e = O.scaled_err(m.row, m.col, m.val, x, O.cf_csr(m.row, m.col, m.val, x), y)[0]
    assert e <= tol_other, (what, e)


@pytest.mark.gpu
def test_fp32_long_rows_compensated(engine):
    import torch
    m, x = _case()
    n = m.nrows
    print(f"[compensation] n={n} nnz={m.nnz}", flush=True)

    def host(t, cnt=n):
        torch.cuda.synchronize()
        return t.reshape(-1)[:cnt].double().cpu().numpy()

    def nan(cnt):
        return torch.full((cnt,), float("nan"), dtype=torch.float32, device=engine.torch_device)
    for th in (1, 2, 4, 8):
        a = H.AXT(m, x, th=th)
        for vt, tag, tol in ((D.F32, "F32", 1e-6), (D.BF16, "BF16", 1e-2)):
            if vt == D.BF16 and th == 8:
                continue
            d = D.DeviceAXT(engine, a, val_type=vt, x_type=D.F32)
            y = nan(n)
            d.int_axt(y)
            _check(m, x, host(y), tol, f"AXT NOC th{th} {tag}")
    a = H.AXT(m, x, th=1, mode="COM", bs=128)
    d = D.DeviceAXT(engine, a, val_type=D.F32, x_type=D.F32)
    y = nan(n)
    d.int_axt(y)
    _check(m, x, host(y), 1e-6, "AXT COM F32")
    x32 = torch.tensor(x, dtype=torch.float32, device=engine.torch_device)
    for th in (1, 4):
        a = H.AXT(m, np.zeros(n), th=th, want_acol=True)
        g = D.DeviceAXTGather(engine, a, D.F32, D.F32)
        for k in (1, 3, 16):  # 16 at tile height 1: a wide pass
            X = x32.repeat_interleave(k)  # n x k row-major, every column x
            Y = nan(n * k)
            g.matvec(X, Y, k=k)
            Yh = host(Y, n * k).reshape(n, k)
            for j in range(k):
                _check(m, x, Yh[:, j], 1e-6, f"column map th{th} k{k} col{j}")
        for k in (3, 16):
            Y = nan(n * k)
            b = D.DeviceBatch(engine, [dict(a=a, X=x32.repeat_interleave(k), Y=Y, k=k)], D.F32,
                              D.F32)
            b.launch()
            Yh = host(Y, n * k).reshape(n, k)
            for j in range(k):
                _check(m, x, Yh[:, j], 1e-6, f"batch th{th} k{k} col{j}")
            del b
