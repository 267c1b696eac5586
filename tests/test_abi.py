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


def test_null_context_is_refused_before_any_hip_call():
    """Every entry point that takes a context refuses a NULL one (and NULL
    arrays) with INTSPMV_ERR_ARG before any HIP call, so it runs on a host
    without a GPU; the destroy calls accept NULL like free()."""
    from int_spmv_amd import device as D
    L = D.lib()
    skip = {"intspmv_create", "intspmv_strerror", "intspmv_device_count", "intspmv_last_hip_error"}
    checked = 0
    for name, _, argtypes in D.SIGNATURES:
        if name in skip or not argtypes:
            continue
        args = [None if (t is ctypes.c_void_p or t is ctypes.c_char_p or
                         (isinstance(t, type) and issubclass(t, ctypes._Pointer))) else t(0)
                for t in argtypes]
        st = getattr(L, name)(*args)
        assert st == (0 if name.endswith("_destroy") else -1), (name, st)
        checked += 1
    assert checked >= 30
