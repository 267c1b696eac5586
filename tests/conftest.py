import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libintspmv.so")


@pytest.fixture(scope="session")
def tiny10():
    """Hand-derived 10x10 fixture (tests/golden/tiny10.json)."""
    from int_spmv_amd import host as H
    with open(os.path.join(GOLDEN, "tiny10.json")) as f:
        fx = json.load(f)
    n = fx["n"]
    row = np.array(fx["row_ptr"], np.uint32)
    col = np.array([c for r in fx["rows"] for c in r], np.uint32)
    val = np.array([10 * i + j + 1 for i, r in enumerate(fx["rows"]) for j in r], np.float64)
    x = np.arange(1, n + 1, dtype=np.float64)
    return fx, H.CSR.from_arrays(row, col, val), x


@pytest.fixture(scope="session")
def lap256():
    """The 256x256 5-point Laplacian (n = 65,536), x ~ U[-1,1] seed 0."""
    from int_spmv_amd import host as H
    m = H.CSR.laplace2d(256)
    return m, H.uniform(m.nrows, -1.0, 1.0, seed=0)


@pytest.fixture(scope="session")
def engine():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP GPU visible")
    from int_spmv_amd import device as D
    return D.Engine(0)
