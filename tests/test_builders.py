"""Host builders (C) against closed forms, the hand-derived fixture and the
independent numpy restatement; I/O round trips; error paths."""
import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

from int_spmv_amd import host as H
from oracle import builders_np as B

from .util import csr_with_empty_rows

# 256x256 Laplacian: (mode, th, bs) -> (tileN, lenAX, lenSEC), derived from the
# length formula of getArraysLenAXT with row lengths 4x3, 1016x4, 64516x5.
LAP256_AXT = {
    ("NOC", 1, 128): (65536, 1048576, 65536),
    ("NOC", 2, 128): (24449, 782368, 195592),
    ("NOC", 4, 128): (16257, 1040448, 130056),
    ("NOC", 8, 128): (8192, 1048576, 65536),
    ("NOC", 16, 128): (8192, 2097152, 65536),
    ("COM", 1, 128): (40832, 653312, 326656),
}


def test_laplacian_closed_forms(lap256):
    m, x = lap256
    assert (m.nrows, m.nnz) == (65536, 326656)
    rl = m.rl
    assert (np.sum(rl == 3), np.sum(rl == 4), np.sum(rl == 5)) == (4, 1016, 64516)
    for (mode, th, bs), want in LAP256_AXT.items():
        a = H.AXT(m, x, th=th, mode=mode, bs=bs)
        assert (a.tileN, a.lenAX, a.lenSEC) == want, (mode, th)
    assert H.AXC(m, x).lenAX == 1048576
    k = H.K1(m)
    assert (k.chunkNum, k.lenVC) == (8192, 326664)


def test_tiny10_fixture_arrays(tiny10):
    fx, m, x = tiny10
    a = H.AXT(m, x, th=1)
    f = fx["noc_th1"]
    assert (a.tileN, a.lenAX, a.lenSEC) == (f["tileN"], f["lenAX"], f["lenSEC"])
    assert a.con.tolist() == f["con"] and a.sec.tolist() == f["sec"]
    assert a.ax[16:32].tolist() == f["tile1"] and a.ax[32:48].tolist() == f["tile2"]
    a = H.AXT(m, x, th=2)
    f = fx["noc_th2"]
    assert (a.tileN, a.lenAX, a.lenSEC) == (f["tileN"], f["lenAX"], f["lenSEC"])
    assert a.con.tolist() == f["con"] and a.sec.tolist() == f["sec"]
    assert a.ax[0:16].tolist() == f["tile0_brickrow0"] and a.ax[16:32].tolist() == f["tile0_brickrow1"]
    a = H.AXT(m, x, th=1, mode="COM", bs=8)
    f = fx["com_bs8"]
    assert (a.tileN, a.lenAX, a.lenSEC, a.log) == (f["tileN"], f["lenAX"], f["lenSEC"], f["log"])
    assert a.sec.tolist() == f["sec"]


def _cases():
    yield "lap64", H.CSR.laplace2d(64)
    yield "powerlaw", H.CSR.powerlaw(n=3000, target_nnz=60000, maxlen=2000, seed=5)
    yield "empty_rows", csr_with_empty_rows()
    yield "banded", H.CSR.banded(2000, 1, 40, seed=6)
    yield "stencil", H.CSR.stencil27(6, 0.01, seed=1)


@pytest.mark.parametrize("name,m", list(_cases()))
def test_c_builders_equal_numpy(name, m):
    x = H.uniform(m.nrows, seed=11)
    row, col, val = m.row.copy(), m.col.copy(), m.val.copy()
    for mode, th, bs in [("NOC", 1, 128), ("NOC", 2, 128), ("NOC", 3, 128), ("NOC", 4, 128),
                         ("NOC", 8, 128), ("NOC", 16, 128), ("NOC", 32, 128), ("NOC", 48, 128),
                         ("COM", 1, 2), ("COM", 1, 128), ("COM", 1, 512)]:
        a = H.AXT(m, x, th=th, mode=mode, bs=bs, want_acol=True)
        b = B.axt(row, col, val, x, th=th, mode=mode, bs=bs)
        assert np.array_equal(a.ax, b["ax"]), (name, mode, th, bs)
        assert np.array_equal(a.sec, b["sec"]), (name, mode, th, bs)
        assert np.array_equal(a.con, b["con"]), (name, mode, th, bs)
        assert (a.tileN, a.lenAX, a.lenSEC, a.colN) == (b["tileN"], b["lenAX"], b["lenSEC"], b["colN"])
    ax, brp = B.axc(row, col, val, x)
    c = H.AXC(m, x)
    assert np.array_equal(c.ax, ax) and np.array_equal(c.brp, brp)
    k = H.K1(m)
    kb = B.k1(row, col, val)
    for f in ("permi", "nmc", "chp", "val", "col"):
        assert np.array_equal(getattr(k, f), kb[f]), f
    assert (k.chunkNum, k.lenVC) == (kb["chunkNum"], kb["lenVC"])


@settings(max_examples=400, deadline=None, derandomize=True)
@given(seed=st.integers(0, 2**31 - 1), n=st.integers(1, 160),
       th=st.sampled_from([1, 2, 3, 4, 5, 7, 8, 16, 32, 48]),
       bs=st.sampled_from([2, 4, 8, 32, 128, 512]), long_row=st.booleans())
def test_c_builders_equal_numpy_random(seed, n, th, bs, long_row):
    """Random ragged matrices (runs of empty rows, duplicate columns, one
    long row): the C builders and the numpy restatement give the same arrays
    for NOC at this tile height, COM at this block size, AXC and K1."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 14, size=n)
    lens[rng.random(n) < 0.25] = 0
    if long_row:
        lens[rng.integers(0, n)] = rng.integers(40, 600)
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, n, size=k)) for k in lens] +
                         [np.zeros(0, np.int64)]).astype(np.uint32)
    val = rng.standard_normal(len(col))
    m = H.CSR.from_arrays(row, col, val)
    x = rng.uniform(-1, 1, n)
    for mode, t in (("NOC", th), ("COM", 1)):
        a = H.AXT(m, x, th=t, mode=mode, bs=bs, want_acol=True)
        b = B.axt(row, col, val, x, th=t, mode=mode, bs=bs)
        for f in ("ax", "sec", "con"):
            assert np.array_equal(getattr(a, f), b[f]), (mode, t, bs, f)
        assert (a.tileN, a.lenAX, a.lenSEC, a.colN) == (b["tileN"], b["lenAX"], b["lenSEC"], b["colN"])
    ax, brp = B.axc(row, col, val, x)
    c = H.AXC(m, x)
    assert np.array_equal(c.ax, ax) and np.array_equal(c.brp, brp)
    k = H.K1(m)
    kb = B.k1(row, col, val)
    for f in ("permi", "nmc", "chp", "val", "col"):
        assert np.array_equal(getattr(k, f), kb[f]), f
    assert (k.chunkNum, k.lenVC) == (kb["chunkNum"], kb["lenVC"])


def test_axt_slot_invariants():
    m = H.CSR.powerlaw(n=2000, target_nnz=40000, maxlen=1500, seed=8)
    x = H.uniform(m.nrows, seed=2)
    for th in (1, 4):
        a = H.AXT(m, x, th=th, want_acol=True)
        v = a.ax.reshape(-1, 2, 8)
        vals, xs = v[:, 0, :].reshape(-1), v[:, 1, :].reshape(-1)
        acol = a.acol
        occ = acol != H.ACOL_PAD
        assert occ.sum() == m.nnz
        assert np.all(vals[~occ] == 0) and np.all(xs[~occ] == 0)
        assert np.array_equal(xs[occ], x[acol[occ]])
        assert np.array_equal(np.sort(vals[occ]), np.sort(m.val))
        if th == 1:
            assert np.all(np.diff(a.sec.astype(np.int64)) >= 0)


def test_refresh_x_equals_rebuild():
    m = H.CSR.powerlaw(n=1500, target_nnz=30000, maxlen=1000, seed=9)
    x0, x1 = H.uniform(m.nrows, seed=1), H.uniform(m.nrows, seed=2)
    for th, mode in ((1, "NOC"), (4, "NOC"), (1, "COM")):
        a = H.AXT(m, x0, th=th, mode=mode, want_acol=True)
        a.refresh_x(x1)
        b = H.AXT(m, x1, th=th, mode=mode)
        assert np.array_equal(a.ax, b.ax)


def test_invalid_arguments():
    m = H.CSR.laplace2d(8)
    x = np.zeros(m.nrows)
    with pytest.raises(H.HostError):
        H.AXT(m, x, th=0)
    with pytest.raises(H.HostError):
        H.AXT(m, x, th=2, mode="COM")
    with pytest.raises(H.HostError):
        H.AXT(m, x, th=1, mode="COM", bs=100)
    with pytest.raises(H.HostError):
        H.AXT(m, x, th=1, mode="XYZ")
    with pytest.raises(ValueError):
        H.AXT(m, x[:-1], th=1)


@pytest.mark.parametrize("ext", [".bin", ".csr", ".mtx"])
def test_io_roundtrip(tmp_path, ext):
    m = H.CSR.powerlaw(n=500, target_nnz=8000, maxlen=300, seed=4)
    p = str(tmp_path / f"m{ext}")
    m.write(p)
    r = H.CSR.read(p)
    assert r.nrows == m.nrows and r.nnz == m.nnz
    assert np.array_equal(r.row, m.row) and np.array_equal(r.col, m.col)
    assert np.array_equal(r.val, m.val)
    assert np.array_equal(r.rl, np.diff(m.row.astype(np.int64)))


def test_mtx_symmetric_and_duplicates(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 4\n"
                 "1 1 2.0\n2 1 -1.0\n3 3 5.0\n3 3 1.0\n")
    m = H.CSR.read(str(p))
    assert m.row.tolist() == [0, 2, 3, 4]
    assert m.col.tolist() == [0, 1, 0, 2]
    assert m.val.tolist() == [2.0, -1.0, -1.0, 6.0]


def test_bad_files(tmp_path):
    with pytest.raises(H.HostError):
        H.CSR.read(str(tmp_path / "missing.bin"))
    p = tmp_path / "bad.csr"
    p.write_text("3 2\n1.0 0\n2.0 7\n0 1 1 2\n")  # column 7 out of range
    with pytest.raises(H.HostError):
        H.CSR.read(str(p))
    with pytest.raises(H.HostError):
        H.CSR.read(str(tmp_path / "x.txt"))


def test_sta_file(tmp_path):
    """printMatrixStats (intSpmv_main.c:253-302): <name>.sta beside the matrix
    file, its name line the matrix file name.  The 4x4-grid Laplacian has 4
    rows of 3, 8 of 4 and 4 of 5 entries: mean 4, variance 0.5."""
    m = H.CSR.laplace2d(4)
    mat = str(tmp_path / "lap.mtx")
    s = m.stats(mat)
    assert (s["rmin"], s["rmax"], s["bw"]) == (3, 5, 4)
    rsd = 0.5 ** 0.5
    want = ("------------------------------------\n"
            "matrix's statistics\n"
            "------------------------------------\n"
            f"name:  {mat:>28s}\n"
            f"nrows: {16:28d}\n"
            f"nnz:   {64:28d}\n"
            f"rmin:  {3:28d}\n"
            f"rave:  {4.0:28.2f}\n"
            f"rmax:  {5:28d}\n"
            f"rsd:   {rsd:28.2f}\n"
            f"rsdp:  {rsd / 4 * 100:28.2f}\n"
            f"bw:    {4:28d}\n")
    assert open(str(tmp_path / "lap.sta")).read() == want


def test_generators_deterministic():
    a = H.CSR.powerlaw(n=4000, target_nnz=100000, maxlen=4000, seed=3)
    b = H.CSR.powerlaw(n=4000, target_nnz=100000, maxlen=4000, seed=3)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.val, b.val)
    assert abs(a.nnz - 100000) < 2000
    s = H.CSR.stencil27(5, 0.01, seed=0)
    assert s.nnz == (3 * 5 - 2) ** 3
    r = H.CSR.rmat(8, 8, 0.57, 0.19, 0.19, seed=1)
    assert r.nrows == 256 and r.nnz <= 8 * 256
    assert np.all(np.diff(r.row.astype(np.int64)) >= 0)
    assert np.all((r.val > 0))
    x0 = H.uniform(1000, seed=0)
    assert np.all((x0 >= -1) & (x0 < 1)) and np.array_equal(x0, H.uniform(1000, seed=0))


def _slot_rows(a):
    """Row of every slot of an AXT NOC layout (th == 1: the tile's sec word;
    th > 1: the sec word of the slot's tile column)."""
    ns, th = a.lenAX // 2, a.tileH
    s = np.arange(ns, dtype=np.int64)
    T, C = s // (8 * th), s % 8
    return a.sec[T] if th == 1 else a.sec[8 * T + C]


@pytest.mark.parametrize("th", [1, 2, 3, 4, 8, 16])
def test_axt_transpose_map(th):
    """intspmv_axt_transpose_map: slot s of A^T's layout holds the entry of
    A's slot permT[s] transposed (its column is that slot's row, its row that
    slot's column, its value that slot's value); every stored entry of A
    appears exactly once; padding maps to ACOL_PAD; and A^T's arrays are those
    the AXT builder makes from the CSR transpose."""
    from .util import csr_with_empty_rows
    PAD = 0xFFFFFFFF
    for m in (csr_with_empty_rows(), H.CSR.powerlaw(n=3000, target_nnz=60000, maxlen=900, seed=4)):
        a = H.AXT(m, np.zeros(m.nrows), th=th, want_acol=True)
        t = H.AxtTransposeMap(m, a)
        at, pt = t.at, t.permT
        assert len(pt) == at.lenAX // 2 and at.tileH == th and at.nnz == m.nnz
        real = pt != PAD
        assert np.array_equal(real, at.acol != PAD)
        assert np.count_nonzero(real) == m.nnz
        src = pt[real].astype(np.int64)
        assert len(np.unique(src)) == m.nnz and np.all(a.acol[src] != PAD)
        assert np.array_equal(at.acol[real], _slot_rows(a)[src])
        assert np.array_equal(_slot_rows(at)[real], a.acol[src])
        assert np.array_equal(at.value_plane()[real], a.value_plane()[src])
        assert np.all(at.value_plane()[~real] == 0)
        at_csr, _ = m.transpose()
        ref = H.AXT(at_csr, np.zeros(m.nrows), th=th, want_acol=True)
        for f in ("ax", "sec", "con", "acol"):
            assert np.array_equal(getattr(at, f), getattr(ref, f)), f


def test_axt_transpose_map_refuses_bad_input():
    m = H.CSR.laplace2d(8)
    with pytest.raises(H.HostError):  # no acol / nzslot
        H.AxtTransposeMap(m, H.AXT(m, np.zeros(m.nrows), th=1))
    with pytest.raises(H.HostError):  # COM layouts have no transposed map
        H.AxtTransposeMap(m, H.AXT(m, np.zeros(m.nrows), th=1, mode="COM", bs=8, want_acol=True))
    with pytest.raises(H.HostError):  # another matrix
        H.AxtTransposeMap(H.CSR.laplace2d(9), H.AXT(m, np.zeros(m.nrows), th=1, want_acol=True))
