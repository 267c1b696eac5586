"""Locates and loads the in-tree native libraries; fails loudly if absent."""
from __future__ import annotations

import ctypes
import os

# INTSPMV_LIB_DIR points at another build of the libraries (e.g. a tuning variant)
LIBDIR = os.environ.get("INTSPMV_LIB_DIR") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "lib")


class NativeLibraryMissing(RuntimeError):
    pass


def lib_path(name: str) -> str:
    return os.path.join(LIBDIR, name)


def load_library(name: str) -> ctypes.CDLL:
    path = lib_path(name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -m int_spmv_amd.build` (or "
            f"__graft_entry__.build()) first")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
