"""write_vtk / write_stokes_vtk (postprocess.jl:257-602) against the
reference's own assertions on the written text (test/test_vtk.jl), and the
oracle's strain-rate/stress diagnostics (postprocess.jl:17-100) against
closed forms: a linear velocity field has the exact, uniform strain rate, a
uniform point stress its own average, and the invariants their formulas."""
import math
import os
import tempfile

import numpy as np
import pytest

import pyoracle as O

fa = pytest.importorskip("femtools_amd")


def _text(write):
    fd, path = tempfile.mkstemp(suffix=".vtk")
    os.close(fd)
    try:
        write(path)
        with open(path) as fh:
            return fh.read()
    finally:
        os.remove(path)


TRI = fa.VtkMesh([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], [[0, 1, 2]])


def test_legacy_ascii_unstructured_grid():
    # test_vtk.jl:17-41
    text = _text(lambda p: fa.write_vtk(p, TRI, point_data={"T": [1.0, 2.0, 3.0]}, cell_data={"Q": [4.0]},
                                        title="tiny"))
    for s in ("# vtk DataFile Version 3.0", "tiny", "DATASET UNSTRUCTURED_GRID", "POINTS 3 float", "CELLS 1 4",
              "CELL_TYPES 1\n5", "POINT_DATA 3", "SCALARS T float 1", "CELL_DATA 1", "SCALARS Q float 1"):
        assert s in text


def test_high_order_triangles_are_linearised():
    # test_vtk.jl:43-61
    t6 = fa.VtkMesh([[0, 0], [1, 0], [0, 1], [0.5, 0], [0.5, 0.5], [0, 0.5]], [[0, 1, 2, 3, 4, 5]])
    text = _text(lambda p: fa.write_vtk(p, t6, point_data={"marker": np.arange(10.0, 16.0)}))
    assert "POINTS 3 float" in text
    assert "CELLS 1 4\n3 0 1 2" in text
    assert "SCALARS marker float 1\nLOOKUP_TABLE default\n10.0\n11.0\n12.0" in text


def test_hex27_cells_are_written_by_their_corners():
    # test_vtk.jl:63-78 (the reference's Hex27 node order)
    mesh = fa.VtkMesh(*O.hex27_mesh(1, 1, 1))
    text = _text(lambda p: fa.write_vtk(p, mesh))
    assert "POINTS 8 float" in text and "CELLS 1 9\n8 " in text and "CELL_TYPES 1\n12" in text
    # the eight corners are the cube's
    pts = text.split("POINTS 8 float\n")[1].split("\nCELLS")[0].split("\n")
    assert sorted(tuple(map(float, r.split())) for r in pts) == sorted(
        (x, y, z) for x in (0.0, 1.0) for y in (0.0, 1.0) for z in (0.0, 1.0))


def test_tuple_fields_are_vectors_and_tensors():
    # test_vtk.jl:80-124
    text = _text(lambda p: fa.write_vtk(p, TRI, point_data={"v": ([1.0, 2.0, 3.0], [4.0, 5.0, 6.0])},
                                        cell_data={"f": ([1.0], [2.0], [3.0])}))
    assert "VECTORS v float\n1.0 4.0 0.0\n2.0 5.0 0.0\n3.0 6.0 0.0" in text
    assert "VECTORS f float\n1.0 2.0 3.0" in text
    text = _text(lambda p: fa.write_vtk(p, TRI, cell_data={"t": ([1.0], [2.0], [3.0], [4.0], [5.0], [6.0])}))
    assert "TENSORS t float\n1.0 4.0 5.0\n4.0 2.0 6.0\n5.0 6.0 3.0" in text
    with pytest.raises(ValueError, match="must hold 2 or 3"):
        _text(lambda p: fa.write_vtk(p, TRI, point_data={"v": ([1.0, 2.0, 3.0],)}))
    with pytest.raises(fa.DimensionMismatch):
        _text(lambda p: fa.write_vtk(p, TRI, point_data={"v": ([1.0, 2.0, 3.0], [1.0, 2.0])}))


def test_field_lengths_are_validated():
    # test_vtk.jl:126-143
    with pytest.raises(fa.DimensionMismatch):
        _text(lambda p: fa.write_vtk(p, TRI, point_data={"bad": [1.0, 2.0]}))
    with pytest.raises(fa.DimensionMismatch):
        _text(lambda p: fa.write_vtk(p, TRI, cell_data={"bad": [1.0, 2.0]}))
    with pytest.raises(ValueError, match="cannot write VTK 2D cells with 5 local nodes"):
        _text(lambda p: fa.write_vtk(p, fa.VtkMesh(np.zeros((5, 2)), [[0, 1, 2, 3, 4]])))


def test_write_stokes_vtk_2d():
    # test_vtk.jl:145-186: advected coordinates, P averaged to the corners, Q to cells
    t6 = fa.VtkMesh([[0, 0], [1, 0], [0, 1], [0.5, 0], [0.5, 0.5], [0, 0.5]], [[0, 1, 2, 3, 4, 5]], el2nP=[[0, 1, 2]])
    post = dict(zip(("exx", "eyy", "ezz", "exy", "eII", "txx", "tyy", "tzz", "txy", "tII"),
                    ([float(k)] for k in range(1, 11))))
    adv = [[2.0, 0.0], [0.0, 2.0], [0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0]]
    text = _text(lambda p: fa.write_stokes_vtk(p, t6, adv, [[0, 1, 2]], [[0, 1, 2]], [10.0, 20.0, 30.0],
                                               (np.arange(1.0, 7.0), np.arange(2.0, 8.0)), post,
                                               Q=[4.0, 5.0, 6.0]))
    for s in ("POINT_DATA 3", "2.0 0.0", "SCALARS P float 1", "SCALARS Vx float 1", "VECTORS V float",
              "CELL_DATA 1", "SCALARS strain_xx float 1", "SCALARS tau_II float 1", "SCALARS Q float 1"):
        assert s in text
    assert "SCALARS P float 1\nLOOKUP_TABLE default\n10.0\n20.0\n30.0" in text
    assert "SCALARS Q float 1\nLOOKUP_TABLE default\n5.0" in text


def test_write_stokes_vtk_3d():
    # test_vtk.jl:188-220
    mesh = fa.VtkMesh(*O.hex27_mesh(1, 1, 1))
    n = mesh.nnodes
    vel = tuple(np.full(n, float(c)) for c in (1, 2, 3))
    keys = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "eII", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tII")
    post = {k: [float(i + 1)] for i, k in enumerate(keys)}
    DoFsP = np.arange(4)[None, :]
    text = _text(lambda p: fa.write_stokes_vtk(p, mesh, mesh.coords, DoFsP, DoFsP, np.ones(4), vel, post))
    assert "POINT_DATA 8" in text and "VECTORS velocity float\n1.0 2.0 3.0" in text
    assert "SCALARS P float 1" in text and "SCALARS z float 1" in text
    zs = text.split("SCALARS z float 1\nLOOKUP_TABLE default\n")[1].split("\n")[:8]
    assert sorted(map(float, zs)) == [0.0] * 4 + [1.0] * 4
    assert "TENSORS tau float\n8.0 11.0 12.0\n11.0 9.0 13.0\n12.0 13.0 10.0" in text
    assert "TENSORS strain float\n1.0 4.0 5.0\n4.0 2.0 6.0\n5.0 6.0 3.0" in text
    assert "SCALARS tau_II float 1" in text


def parse_vtk(text):
    """A legacy ASCII unstructured grid read back (test_vtk.jl:222-271)."""
    lines = [ln.strip() for ln in text.split("\n") if ln.strip()]
    out = dict(points=[], cells=[], cell_types=[], point_data={}, cell_data={})
    npts = ncells = 0
    section = None
    i = 0
    while i < len(lines):
        f = lines[i].split()
        if f[0] == "POINTS":
            npts = int(f[1])
            out["points"] = [list(map(float, lines[i + k].split())) for k in range(1, npts + 1)]
            i += npts
        elif f[0] == "CELLS":
            ncells = int(f[1])
            out["cells"] = [list(map(int, lines[i + k].split()))[1:] for k in range(1, ncells + 1)]
            i += ncells
        elif f[0] == "CELL_TYPES":
            out["cell_types"] = [int(lines[i + k]) for k in range(1, int(f[1]) + 1)]
            i += int(f[1])
        elif f[0] == "POINT_DATA":
            section = out["point_data"]
        elif f[0] == "CELL_DATA":
            section = out["cell_data"]
        elif f[0] == "SCALARS":
            n = npts if section is out["point_data"] else ncells
            section[f[1]] = [float(lines[i + 1 + k]) for k in range(1, n + 1)]
            i += 1 + n
        i += 1
    return out


def test_round_trip_of_a_two_triangle_mesh():
    # test_vtk.jl:273-300
    mesh = fa.VtkMesh([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]], [[0, 1, 2], [1, 3, 2]])
    T, Q = [1.0, 2.0, 4.0, 8.0], [7.0, 9.0]
    v = parse_vtk(_text(lambda p: fa.write_vtk(p, mesh, point_data={"T": T}, cell_data={"Q": Q}, title="pair")))
    assert [p[:2] for p in v["points"]] == [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]]
    assert all(p[2] == 0.0 for p in v["points"])
    assert v["cells"] == [[0, 1, 2], [1, 3, 2]] and v["cell_types"] == [5, 5]
    assert v["point_data"]["T"] == T and v["cell_data"]["Q"] == Q


def test_numbers_print_as_julia_prints_them():
    j = fa.vtk.julia_float
    cases = {1.0: "1.0", 0.5: "0.5", 123456.0: "123456.0", 1e6: "1.0e6", 1234567.0: "1.234567e6",
             1e-5: "1.0e-5", 1e-4: "0.0001", -2.25: "-2.25", 0.1 + 0.2: "0.30000000000000004", -0.0: "-0.0"}
    for x, s in cases.items():
        assert j(x) == s and float(s) == x
    assert j(math.nan) == "NaN" and j(-math.inf) == "-Inf"


def test_diagnostics_closed_forms():
    # postprocess.jl:17-100 on a linear velocity field and a uniform point stress
    coords, el2n = O.t7_mesh(3, 2)
    a, b, c = 0.3, -0.7, 0.4
    vx = a * coords[:, 0] + b * coords[:, 1]
    vy = c * coords[:, 0] - a * coords[:, 1]
    ne = len(el2n)
    tau = (np.full((ne, 7), 1.0), np.full((ne, 7), -2.0), np.full((ne, 7), 0.5))
    p = O.postprocess2(coords, el2n, vx, vy, tau)
    exy = (b + c) / 2
    expect = dict(exx=a, eyy=-a, ezz=0.0, exy=exy, eII=math.sqrt(a * a + exy * exy), txx=1.0, tyy=-2.0, tzz=1.0,
                  txy=0.5, tII=math.sqrt(3.0 + 0.25))
    for k, v in expect.items():
        assert np.allclose(p[k], v, rtol=1e-13, atol=1e-15), k
    # a quadratic field: the average of ε̇xx = ∂(x²)/∂x = 2x is 2·(centroid x)
    p = O.postprocess2(coords, el2n, coords[:, 0] ** 2, 0 * vy, tau)
    cen = coords[el2n[:, :3]].mean(axis=1)
    assert np.allclose(p["exx"], 2 * cen[:, 0], rtol=1e-13)


def test_diagnostics_3d_closed_forms():
    # postprocess.jl:169-212 on a linear velocity field and uniform stresses
    coords, el2n = O.hex27_mesh(2, 2, 1)
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    v = (0.2 * x - 0.5 * y + 0.1 * z, 0.3 * x + 0.4 * y, -0.6 * z + 0.7 * x)
    ne = len(el2n)
    tau = tuple(np.full((ne, 27), c) for c in (1.0, -2.0, 1.0, 0.5, -0.25, 0.75))
    p = O.postprocess3(coords, el2n, v, tau)
    exx, eyy, ezz = 0.2, 0.4, -0.6
    exy, exz, eyz = (-0.5 + 0.3) / 2, (0.1 + 0.7) / 2, 0.0
    tr = (exx + eyy + ezz) / 3
    eII = math.sqrt(((exx - tr) ** 2 + (eyy - tr) ** 2 + (ezz - tr) ** 2) / 2 + exy ** 2 + exz ** 2 + eyz ** 2)
    tII = math.sqrt((1 + 4 + 1) / 2 + 0.25 + 0.0625 + 0.5625)
    expect = dict(exx=exx, eyy=eyy, ezz=ezz, exy=exy, exz=exz, eyz=eyz, eII=eII, txx=1.0, tyy=-2.0, tzz=1.0,
                  txy=0.5, txz=-0.25, tyz=0.75, tII=tII)
    for k, val in expect.items():
        assert np.allclose(p[k], val, rtol=1e-13, atol=1e-14), k
