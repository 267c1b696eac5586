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


def test_headers_compile_and_link_as_c(tmp_path):
    """A C11 host that includes both headers and takes the address of every
    declared entry point compiles warning-free and links against the two
    libraries: the boundary a C caller (the reference's own language) binds."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    names = _declared("intspmv.h") + _declared("intspmv_host.h")
    src = tmp_path / "use.c"
    src.write_text('#include "intspmv.h"\n#include "intspmv_host.h"\n#include <stdio.h>\n'
                   "int main(void) {\n    void *p[] = {\n"
                   + "".join(f"        (void *)&{n},\n" for n in names)
                   + '    };\n    printf("%zu\\n", sizeof p / sizeof p[0]);\n    return 0;\n}\n')
    lib = os.path.join(ROOT, "int_spmv_amd", "lib")
    r = subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-Wno-pedantic",
                        f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(tmp_path / "use"),
                        f"-L{lib}", "-lintspmv", "-lintspmv_host", "-L/opt/rocm/lib",
                        "-Wl,--allow-shlib-undefined"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
