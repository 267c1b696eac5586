"""The CPU suite again with the C host library (I/O, statistics, builders,
transposed map, generators, tuner model) built with AddressSanitizer and
UndefinedBehaviorSanitizer, any report fatal (-fno-sanitize-recover=all).
Python itself is not instrumented, so the sanitizer runtimes are preloaded
and leak checking (which would report the interpreter's own allocations) is
off.  No GPU is used."""
import os
import subprocess
import sys

import pytest

from int_spmv_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHECK_LOADED = (
    "from int_spmv_amd import host as H; H.lib(); "
    "maps = open('/proc/self/maps').read(); "
    "assert 'lib_sanitized/libintspmv_host.so' in maps, 'sanitized library not loaded'; "
    "print('sanitized host library loaded')")


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True,
                          text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_cpu_suite_under_asan_ubsan():
    lib = os.path.join(B.SANITIZED_LIBDIR, "libintspmv_host.so")
    if not os.path.exists(lib):
        pytest.fail("build first: python -c 'import __graft_entry__ as g; g.build()'")
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not found")
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", INTSPMV_LIB_DIR=B.SANITIZED_LIBDIR,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", CHECK_LOADED], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "tests", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider", "--ignore=tests/test_sanitizers.py"],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr
