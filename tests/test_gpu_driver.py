"""GPU: the C driver (int_spmv_amd/csrc/host/intspmv_main.c), which replays
the reference's main (intSpmv_main.c:1617-1761) with every kernel on the GPU,
runs to completion and prints the kernels' table with every result within
the FP64 tolerance of CSR-naive (1e-12 scaled per row)."""
import os
import subprocess

import pytest

from int_spmv_amd import build as B

pytestmark = pytest.mark.gpu

KERNELS = ["spmv_csr_ordered", "spmv_csr", "spmv_axc", "spmv_k1", "spmv_k1p"] + \
    [f"spmv_axt_noc_h{th}_hw8" for th in (1, 4, 8, 16, 32, 48)] + \
    [f"spmv_axt_com_h1_hw8_bs{bs}" for bs in (128, 256, 512)]


def _table(out: str) -> dict:
    """The reference's kernels' table (seven columns) and the scaled errors."""
    lines = out.splitlines()
    i = next(k for k, line in enumerate(lines) if line.startswith("kernels' results"))
    j = next(k for k, line in enumerate(lines) if line.startswith("kernels' scaled errors"))
    rows = {}
    for line in lines[i + 3:j - 1]:
        f = line.split()
        assert len(f) == 7, line
        rows[f[0]] = dict(et=float(f[1]), aerr=float(f[4]))
    for line in lines[j + 3:]:
        f = line.split()
        if len(f) == 2 and f[0] in rows:
            rows[f[0]]["serr"] = float(f[1])
    return rows


@pytest.mark.parametrize("gen", ["laplace:64", "powerlaw:20000:300000"])
def test_driver_kernels_table(gen):
    exe = os.path.join(B.LIBDIR, "intspmv_main")
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    print(f"[driver] {gen}", flush=True)
    r = subprocess.run([exe, "--gen", gen, "--ite", "3", "--x", "uniform"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = _table(r.stdout)
    assert sorted(rows) == sorted(KERNELS), sorted(rows)
    for name, row in rows.items():
        assert row["et"] > 0, (gen, name)
        assert row["serr"] <= 1e-12, (gen, name, row)
    assert rows["spmv_csr_ordered"]["aerr"] == 0.0  # the oracle's own order


@pytest.mark.parametrize("seed,ext", [(1, ".mtx"), (2, ".bin"), (8, ".mtx")])
def test_driver_on_random_matrix_file(tmp_path, seed, ext):
    """The driver on a random ragged matrix read from a file (runs of empty
    rows, one long row, duplicate columns): every kernel within the FP64
    tolerance, and the statistics file written beside the matrix."""
    from .test_gpu_random import _case
    exe = os.path.join(B.LIBDIR, "intspmv_main")
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    m = _case(seed)[0]
    path = tmp_path / f"rand{seed}{ext}"
    m.write(str(path))
    print(f"[driver] {path.name} n={m.nrows} nnz={m.nnz}", flush=True)
    r = subprocess.run([exe, str(path), "--ite", "2", "--x", "uniform"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = _table(r.stdout)
    assert sorted(rows) == sorted(KERNELS), sorted(rows)
    for name, row in rows.items():
        assert row["serr"] <= 1e-12, (seed, name, row)
    assert (tmp_path / f"rand{seed}.sta").exists()
