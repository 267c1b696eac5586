"""The benchmark and test inputs, generated deterministically on the host.

  laplace_256      2-D 5-point Laplacian on 256x256 (n = 65,536,
                   nnz = 326,656), x ~ U[-1, 1] seed 0 -- the reference-sized
                   CPU-runnable case
  powerlaw_1m      n = 2^20, row lengths ~ Zipf(s=2) clipped to [1, 8192] and
                   rescaled to ~30M entries; 80% of columns within +-2.5% of n
                   of the diagonal, 20% uniform; values N(0,1); x ~ U[-1, 1]
  stencil27_128    3-D 27-point stencil on 128^3 (n = 2,097,152,
                   nnz = 55,742,968), diag 26, off-diagonal -1 + N(0, 0.01^2)
                   (sigma = 0.01)
  rmat_22          R-MAT scale 22, edge factor 16, (a,b,c) = (.57,.19,.19),
                   duplicates summed, values U(0, 1]
"""
from __future__ import annotations

from . import host as H


def laplace_256(seed: int = 0):
    m = H.CSR.laplace2d(256)
    return m, H.uniform(m.nrows, -1.0, 1.0, seed=seed)


def powerlaw_1m(seed: int = 0):
    n = 1 << 20
    m = H.CSR.powerlaw(n=n, target_nnz=30_000_000, s=2.0, maxlen=8192, band_frac=0.8,
                       band_half=int(0.025 * n), seed=seed)
    return m, H.uniform(m.nrows, -1.0, 1.0, seed=seed)


def powerlaw_4m(seed: int = 0):
    """The power-law generator at 2^22 rows and 67M entries (the R-MAT
    workload's size): how the headline kernel's rate changes with size."""
    n = 1 << 22
    m = H.CSR.powerlaw(n=n, target_nnz=67_000_000, s=2.0, maxlen=8192, band_frac=0.8,
                       band_half=int(0.025 * n), seed=seed)
    return m, H.uniform(m.nrows, -1.0, 1.0, seed=seed)


def stencil27_128(seed: int = 0):
    m = H.CSR.stencil27(128, sigma=0.01, seed=seed)
    return m, H.uniform(m.nrows, -1.0, 1.0, seed=seed)


def rmat_22(seed: int = 0):
    m = H.CSR.rmat(22, 16, 0.57, 0.19, 0.19, seed=seed)
    return m, H.uniform(m.nrows, -1.0, 1.0, seed=seed)


def batch_256(seed: int = 0, count: int = 256, nmax: int = 200_000):
    """Many matrices: n log-uniform in [1000, nmax], 4..64 entries per row,
    70% of columns in a band of width 0.05 n, 30% uniform, values N(0,1);
    k in {1, 4, 8, 16, 32} right-hand sides each.  Returns [(m, k)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    ks = (1, 4, 8, 16, 32)
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(nmax))))
        m = H.CSR.banded(n, 4, 64, band_frac=0.7, band_w=max(2, int(0.05 * n)), seed=seed * 1000003 + i)
        out.append((m, ks[int(rng.integers(0, len(ks)))]))
    return out


WORKLOADS = {
    "laplace_256": laplace_256,
    "powerlaw_1m": powerlaw_1m,
    "powerlaw_4m": powerlaw_4m,
    "stencil27_128": stencil27_128,
    "rmat_22": rmat_22,
}
