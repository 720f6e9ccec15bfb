"""CPU-side checks of the drop-in boundary: libfemdr.so loads, exports every
entry point include/femdr.h declares, and the host mirror agrees with it.
No compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "femdr.h")
LIB = os.path.join(ROOT, "femtools.jl_amd", "femtools_amd", "libfemdr.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(fem_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "fem_stokes_solve" in names and "fem_ctx_create" in names
    assert len(names) >= 25


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfemdr.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfemdr.so not built")
def test_python_mirror_binds_every_declared_symbol():
    import femtools_amd as fa
    assert sorted(fa.EXPORTED) == declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfemdr.so not built")
def test_struct_layouts_match_header():
    import femtools_amd as fa
    # sizes derived from the header's field lists (LP64)
    assert ctypes.sizeof(fa._Material) == 8 + 8 * 64 * 6 + 8 * 3 + 8
    assert ctypes.sizeof(fa._DP) == 8 + 8 * 64 * 6
    assert ctypes.sizeof(fa._SolveParams) == 8 * 3 + 8 * 4 + 8 + 4 * 8
    assert ctypes.sizeof(fa._BC) == 24
