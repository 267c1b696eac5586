"""int_spmv_amd -- MI355X-native SpMV engine for the int_spmv formats.

Host side (C, ctypes-bound in :mod:`int_spmv_amd.host`): matrix I/O, the
CSR -> AXC / K1 / AXT builders and synthetic generators.  Device side
(:mod:`int_spmv_amd.device`): hand-written gfx950 HIP kernels behind the C
ABI declared in include/intspmv.h.
"""
__version__ = "0.1.0"
