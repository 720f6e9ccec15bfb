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


STRUCTS = {"fem_mesh_info": "_MeshInfo", "fem_material": "_Material", "fem_drucker_prager": "_DP",
           "fem_state_opts": "_StateOpts", "fem_bc": "_BC", "fem_solve_params": "_SolveParams",
           "fem_solve_stats": "_SolveStats", "fem_thermal_material": "_ThermalMaterial",
           "fem_thermal_opts": "_ThermalOpts", "fem_thermal_params": "_ThermalParams",
           "fem_thermal_stats": "_ThermalStats"}


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfemdr.so not built")
def test_struct_layouts_match_header(tmp_path):
    """sizeof and every field offset of the header's structs, as gcc lays
    them out, against the ctypes mirror."""
    import subprocess
    import femtools_amd as fa
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in getattr(fa, pyname)._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, pyname in STRUCTS.items():
        cls = getattr(fa, pyname)
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
