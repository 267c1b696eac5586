"""The C-ABI libraries load and export every symbol their headers declare.
No compute is launched (this suite runs without a GPU)."""
import ctypes
import os
import re

import pytest

from int_spmv_amd import _loader

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", txt,
                           flags=re.M))
    return sorted(n for n in names if not n.startswith(("if", "sizeof")))


@pytest.mark.parametrize("header,lib", [("intspmv.h", "libintspmv.so"),
                                        ("intspmv_host.h", "libintspmv_host.so")])
def test_exports(header, lib):
    names = _declared(header)
    assert len(names) > 10
    L = _loader.load_library(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from int_spmv_amd import device as D
    bound = {s[0] for s in D.SIGNATURES}
    assert set(_declared("intspmv.h")) <= bound


def test_strerror_and_no_device_path():
    from int_spmv_amd import device as D
    L = D.lib()
    for st in (0, -1, -2, -3, -4, -5, -6, -99):
        assert L.intspmv_strerror(st)
    if D.device_count() == 0:
        h = ctypes.c_void_p()
        assert L.intspmv_create(ctypes.byref(h), 0) == -6  # INTSPMV_ERR_NODEVICE
        assert not h.value


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_loader, "LIBDIR", str(tmp_path))
    with pytest.raises(_loader.NativeLibraryMissing):
        _loader.load_library("libintspmv.so")
