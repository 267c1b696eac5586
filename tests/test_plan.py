"""Host-only checks of the work decomposition (intspmv_plan_layout): every
row is covered exactly once, every unit exactly once, limits hold, split
pieces are consecutive and numbered, long blocks precede normal ones."""
import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import csr_long_rows, csr_with_empty_rows


def _check(uptr, umax, rmax, lu, lmax):
    b, s, nlw = D.plan_layout(uptr, umax, rmax, lu, lmax)
    n = len(uptr) - 1
    real = b[b[:, 2] != 0xFFFFFFFF]
    assert nlw % 8 == 0 and nlw * 4 <= len(b)
    # long section first
    assert np.all(b[: nlw * 4, 5] >= 1) and np.all(b[nlw * 4:, 5] == 0)
    rows = np.zeros(n, np.int64)
    units = np.zeros(int(uptr[-1]), np.int64)
    for ub, ue, rb, re, b0, k, piece, _ in real:
        if k <= 1:
            rows[rb:re] += 1
        if k == 0:
            assert ue - ub <= umax and re - rb <= rmax
            assert ub == uptr[rb] and ue == uptr[re]
            assert np.all(np.diff(uptr[rb:re + 1].astype(np.int64)) <= lu)
        else:
            assert re == rb + 1 and ue - ub <= lmax
        units[ub:ue] += 1
    for row, p0, k, _ in s:
        pieces = real[(real[:, 5] == k) & (real[:, 2] == row)]
        assert len(pieces) == k
        assert list(pieces[:, 6]) == list(range(p0, p0 + k))
        assert pieces[0, 0] == uptr[row] and pieces[-1, 1] == uptr[row + 1]
        assert np.all(pieces[1:, 0] == pieces[:-1, 1])
        rows[row] += 1
    assert np.all(rows == 1)
    assert np.all(units == 1)
    return b, s


@pytest.mark.parametrize("m", [csr_with_empty_rows(), csr_long_rows(),
                               H.CSR.powerlaw(n=20000, target_nnz=600000, maxlen=8192, seed=1),
                               H.CSR.laplace2d(64)])
@pytest.mark.parametrize("params", [(128, 256, 64, 1024), (256, 256, 64, 1024),
                                    (512, 256, 256, 8192), (64, 16, 8, 40)])
def test_plan_layout_invariants(m, params):
    _check(m.row, *params)


def test_plan_layout_splits_and_order():
    m = csr_long_rows()
    b, s = _check(m.row, 512, 256, 256, 8192)
    # rows of 20000, 9000, 12000 and 30000 entries exceed 8192: 3, 2, 2, 4 pieces,
    # ordered by row, numbered consecutively
    assert s[:, 0].tolist() == [10, 11, 5000, 29999]
    assert s[:, 2].tolist() == [3, 2, 2, 4]
    assert s[:, 1].tolist() == [0, 3, 5, 7]


@settings(max_examples=300, deadline=None, derandomize=True)
@given(seed=st.integers(0, 2**31 - 1), n=st.integers(1, 400), umax=st.integers(2, 512),
       rmax=st.integers(1, 256), lu_frac=st.floats(0.0, 0.99), lmax_mul=st.integers(1, 64))
def test_plan_layout_invariants_random(seed, n, umax, rmax, lu_frac, lmax_mul):
    """Random row lengths (runs of empty rows, a few rows far above every
    limit) and random limits: the same invariants."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, size=n)
    lens[rng.random(n) < 0.3] = 0
    lens[rng.random(n) < 0.02] = rng.integers(200, 5000)
    uptr = np.zeros(n + 1, np.uint32)
    uptr[1:] = np.cumsum(lens)
    lu = max(1, min(umax - 1, int(lu_frac * umax)))
    _check(uptr, umax, rmax, lu, lu + lmax_mul * 8)


def _xcd_of_blocks(nblocks, nlw):
    """XCD each block of a plan runs on: hardware workgroup g goes to XCD
    g % 8 and takes blocks 4*lg .. 4*lg+3 with lg from block_index()
    (common.hpp): the identity in the long section, xcd_remap after it."""
    grid = nlw + (nblocks - nlw * 4 + 3) // 4
    xcd = np.full(nblocks, -1, np.int64)
    nn = grid - nlw
    for g in range(grid):
        if g < nlw:
            lg = g
        else:
            h = g - nlw
            q, r, x, l = nn // 8, nn % 8, h % 8, h // 8
            lg = nlw + x * q + min(x, r) + l
        for w in range(4):
            bid = lg * 4 + w
            if bid < nblocks:
                assert xcd[bid] == -1, "block handed out twice"
                xcd[bid] = g % 8
    assert np.all(xcd >= 0), "block never handed out"
    return xcd


def _check_grouped(uptr, umax, rmax, lu, lmax, gbeg):
    b, s, nlw = D.plan_layout_grouped(uptr, umax, rmax, lu, lmax, gbeg)
    n = len(uptr) - 1
    assert nlw % 8 == 0 and nlw * 4 <= len(b) and (len(b) - nlw * 4) % 32 == 0
    assert np.all(b[: nlw * 4, 5] >= 1) and np.all(b[nlw * 4:, 5] == 0)
    real = b[b[:, 2] != 0xFFFFFFFF]
    rows = np.zeros(n, np.int64)
    units = np.zeros(int(uptr[-1]), np.int64)
    group = np.searchsorted(np.asarray(gbeg, np.int64), np.arange(n), side="right") - 1
    xcd = _xcd_of_blocks(len(b), nlw)
    for i, (ub, ue, rb, re, b0, k, piece, _) in enumerate(b):
        if rb == 0xFFFFFFFF or re == rb:
            continue
        # every row of a block is in one group, and that group's XCD runs it
        assert np.all(group[rb:re] == group[rb]) and xcd[i] == group[rb] % 8
        if k <= 1:
            rows[rb:re] += 1
        if k == 0:
            assert ue - ub <= umax and re - rb <= rmax
            assert ub == uptr[rb] and ue == uptr[re]
        else:
            assert re == rb + 1 and ue - ub <= lmax
        units[ub:ue] += 1
    for row, p0, k, _ in s:
        pieces = real[(real[:, 5] == k) & (real[:, 2] == row)]
        assert len(pieces) == k
        assert list(pieces[:, 6]) == list(range(p0, p0 + k))
        assert pieces[0, 0] == uptr[row] and pieces[-1, 1] == uptr[row + 1]
        rows[row] += 1
    assert np.all(rows == 1)
    assert np.all(units == 1)


@settings(max_examples=200, deadline=None, derandomize=True)
@given(seed=st.integers(0, 2**31 - 1), n=st.integers(1, 600), ng=st.integers(1, 24),
       umax=st.integers(2, 256), lu_frac=st.floats(0.0, 0.99), lmax_mul=st.integers(1, 16))
def test_plan_layout_grouped_random(seed, n, ng, umax, lu_frac, lmax_mul):
    """Row groups pinned to XCDs (intspmv_plan_layout_grouped): the plan
    invariants above, blocks never straddle groups, and every block is
    handed to a workgroup on its group's XCD (group % 8), empty groups
    included."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, size=n)
    lens[rng.random(n) < 0.3] = 0
    lens[rng.random(n) < 0.03] = rng.integers(200, 3000)
    uptr = np.zeros(n + 1, np.uint32)
    uptr[1:] = np.cumsum(lens)
    cuts = np.sort(rng.integers(0, n + 1, size=ng - 1))
    gbeg = np.concatenate([[0], cuts, [n]]).astype(np.uint32)
    lu = max(1, min(umax - 1, int(lu_frac * umax)))
    _check_grouped(uptr, umax, 256, lu, lu + lmax_mul * 8, gbeg)


def test_plan_layout_grouped_refuses_bad_groups():
    uptr = np.arange(11, dtype=np.uint32)
    for g in ([0, 5], [1, 10], [0, 6, 5, 10], []):
        with pytest.raises(D.IntSpmvError):
            D.plan_layout_grouped(uptr, 128, 256, 64, 1024, np.asarray(g + [10] if not g else g,
                                                                         np.uint32))


@pytest.mark.parametrize("m", [csr_long_rows(), H.CSR.laplace2d(64),
                               H.CSR.powerlaw(n=20000, target_nnz=600000, maxlen=8192, seed=1)])
@pytest.mark.parametrize("ng", [1, 8, 16])
def test_plan_layout_even_groups(m, ng):
    """group_rows NULL: ngroups groups of about equal units; the invariants
    hold and every group's share of units is within one row of even."""
    uptr = m.row
    b, s, nlw = D.plan_layout_grouped(uptr, 128, 256, 64, 1024, None, ng)
    n = m.nrows
    tot = int(uptr[-1])
    # recover the cuts the library made: the first row of each group's blocks
    cuts = np.searchsorted(uptr.astype(np.int64), [tot * g // ng for g in range(1, ng)])
    gbeg = np.concatenate([[0], cuts, [n]]).astype(np.uint32)
    _check_grouped(uptr, 128, 256, 64, 1024, gbeg)
    b2, s2, nlw2 = D.plan_layout_grouped(uptr, 128, 256, 64, 1024, gbeg)
    assert np.array_equal(b, b2) and np.array_equal(s, s2) and nlw == nlw2
