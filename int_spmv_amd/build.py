"""Build recipe for the native parts of int_spmv_amd.

Produces, in-tree (so the libraries travel with the source tree):
  int_spmv_amd/lib/libintspmv_host.so  C host: I/O, stats, builders, generators
  int_spmv_amd/lib/libintspmv.so       gfx950 HIP kernels behind the C ABI
  int_spmv_amd/lib/intspmv_main        C driver (the reference's main sequence)
  tests/c/adjoint_from_c               C caller of both libraries run by a GPU test
  int_spmv_amd/lib_sanitized/          the C host library under ASan + UBSan (CPU tests)

Run ``python -m int_spmv_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "int_spmv_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = "gfx950"

# Device code: gfx950 only; packed-FP32 instructions off (hipcc would otherwise
# form v_pk_* FP32 ops from scalar float code); no implicit FMA contraction so
# that every fused multiply-add in the kernels is an explicit fma().
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.c")))
    srcs = [s for s in srcs if not s.endswith("intspmv_main.c")]
    out = os.path.join(LIBDIR, "libintspmv_host.so")
    _run([CC, *HOST_FLAGS, "-shared", f"-I{INCLUDE}", *srcs, "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


# The C host library again, instrumented with AddressSanitizer and
# UndefinedBehaviorSanitizer (host code only, built by gcc), for the CPU test
# suite (tests/test_sanitizers.py); the GPU library beside it is the normal one.
SANITIZED_LIBDIR = os.path.join(PKG, "lib_sanitized")
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all",
                  "-fsanitize=address,undefined"]


def build_host_sanitized() -> str:
    os.makedirs(SANITIZED_LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.c")))
    srcs = [s for s in srcs if not s.endswith("intspmv_main.c")]
    out = os.path.join(SANITIZED_LIBDIR, "libintspmv_host.so")
    flags = [f for f in HOST_FLAGS if f != "-O2"] + SANITIZE_FLAGS
    _run([CC, *flags, "-shared", f"-I{INCLUDE}", *srcs, "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    dev = os.path.join(SANITIZED_LIBDIR, "libintspmv.so")
    if not os.path.lexists(dev):
        os.symlink(os.path.join("..", "lib", "libintspmv.so"), dev)
    return out


def build_device(outdir: str = LIBDIR, defines: tuple = ()) -> str:
    """defines: extra -D flags overriding kernel tunables (ISPV_* macros)."""
    os.makedirs(outdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "kernels", "*.hip")))
    out = os.path.join(outdir, "libintspmv.so")
    _run([HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], f"-I{INCLUDE}",
          f"-I{os.path.join(CSRC, 'kernels')}", *srcs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_driver() -> str:
    src = os.path.join(CSRC, "host", "intspmv_main.c")
    out = os.path.join(LIBDIR, "intspmv_main")
    if not os.path.exists(src):
        return ""
    _run([CC, *[f for f in HOST_FLAGS if f != "-fPIC"], f"-I{INCLUDE}", src, "-o", out,
          f"-L{LIBDIR}", "-lintspmv_host", "-lintspmv", f"-Wl,-rpath,{LIBDIR}",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64",
          "-lm"])
    return out


def build_c_tests() -> str:
    """C host programs the GPU tests run (tests/c/*.c, linked like any C
    caller of the two libraries)."""
    src = os.path.join(ROOT, "tests", "c", "adjoint_from_c.c")
    out = os.path.join(ROOT, "tests", "c", "adjoint_from_c")
    if not os.path.exists(src):
        return ""
    _run([CC, "-O2", "-std=c11", "-Wall", "-Wextra", f"-I{INCLUDE}", src, "-o", out,
          f"-L{LIBDIR}", "-lintspmv_host", "-lintspmv", f"-Wl,-rpath,{LIBDIR}",
          "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64", "-lm"])
    return out


def build_all() -> None:
    build_host()
    build_device()
    build_driver()
    build_c_tests()
    build_host_sanitized()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"host": build_host, "device": build_device, "driver": build_driver,
         "ctests": build_c_tests, "sanitized": build_host_sanitized, "all": build_all}[w]()
