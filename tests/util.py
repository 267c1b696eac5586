"""Helpers shared by the tests: matrices with awkward shapes and error checks."""
import numpy as np

import oracle as O
from int_spmv_amd import host as H


def csr_with_empty_rows(n=300, seed=3):
    """Rows 0-2, a run in the middle and the last rows are empty; one row is
    long enough to be split across waves by every plan."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, 40, size=n)
    lens[:3] = 0
    lens[100:117] = 0
    lens[-5:] = 0
    lens[50] = 3000
    lens[51] = 700
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, size=min(l, n), replace=False)) if l <= n
                          else np.sort(rng.integers(0, n, size=l)) for l in lens]).astype(np.uint32)
    # duplicates are allowed in CSR products; keep them for the long row
    val = rng.standard_normal(len(col))
    return H.CSR.from_arrays(row, col, val)


def csr_sparse_gaps(n=4000, seed=5):
    """About 60% of the rows empty, plus gaps of 300 (longer than a block's
    row cap) and 70 consecutive empty rows, and empty first and last rows:
    the empty-row zero fill of the keyed kernels across and inside blocks."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, 30, size=n)
    lens[rng.random(n) < 0.6] = 0
    lens[1000:1300] = 0
    lens[2000:2070] = 0
    lens[0] = 0
    lens[-1] = 0
    lens[1500] = 900
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens]).astype(np.uint32)
    val = rng.standard_normal(len(col))
    return H.CSR.from_arrays(row, col, val)


def scaled(m, x, yref, y):
    e, arg = O.scaled_err(m.row, m.col, m.val, x, yref, np.asarray(y, np.float64))
    return e


def bitwise_equal(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def csr_long_rows(n=30000, seed=5):
    """Mostly short rows plus rows of 9,000-30,000 entries, longer than one
    wave streams alone in every format, so split rows are exercised."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, size=n)
    lens[10], lens[11], lens[5000], lens[n - 1] = 20000, 9000, 12000, 30000
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.integers(0, n, size=l)) for l in lens]).astype(np.uint32)
    val = rng.standard_normal(len(col))
    return H.CSR.from_arrays(row, col, val)
