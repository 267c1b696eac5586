"""Build recipe of the oracle (TEST INFRASTRUCTURE ONLY).

Compiles oracle.c into oracle/liboracle.so and the AVX-512 CPU baseline
cpu_avx512.c into oracle/liboracle_avx512.so.  -ffp-contract=off keeps the
compiler from fusing multiply-adds the restatement does not spell out with
fma().  The reference itself (intSpmv_main.c) is not built here: it includes
<mkl.h>, which this image does not have, so there is no oracle/_ref build.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build_avx512() -> str:
    """The AVX-512 CPU baseline (cpu_avx512.c), built for any host; it is
    only called where the CPU reports avx512f."""
    out = os.path.join(HERE, "liboracle_avx512.so")
    cmd = [os.environ.get("CC", "gcc"), "-O3", "-std=gnu11", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-mavx512f", "-mavx512vl", "-mfma", "-Wall",
           os.path.join(HERE, "cpu_avx512.c"), "-o", out + ".tmp", "-lm"]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build() -> str:
    build_avx512()
    out = os.path.join(HERE, "liboracle.so")
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-march=x86-64-v3", "-Wall", os.path.join(HERE, "oracle.c"),
           "-o", out + ".tmp", "-lm"]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
