"""Legacy ASCII VTK output of meshes and Stokes solutions: write_vtk and
write_stokes_vtk (src/postprocess/postprocess.jl:257-602).

High-order cells are written by their corner nodes (only the corners are
renumbered into VTK points); tuple fields become VTK vectors (2 or 3
components, a plane vector padded with 0.0), symmetric tensors (6 components
in the order xx, yy, zz, xy, xz, yz, mirrored into a full 3 x 3) or full
row-major tensors (9). Numbers are printed as Julia prints a Float64, so the
files are byte-identical to the reference's for the same data.
"""
from __future__ import annotations

import math
from decimal import Decimal

import numpy as np

__all__ = ["Mesh", "write_vtk", "write_stokes_vtk", "julia_float"]


try:   # the package's exception when imported from it
    from . import DimensionMismatch
except ImportError:   # pragma: no cover
    class DimensionMismatch(ValueError):
        """Field or coordinate lengths that do not fit the mesh."""


class Mesh:
    """A bare mesh for output: coords (nnodes, dim) and el2n (nels, nloc), 0-based."""

    def __init__(self, coords, el2n, el2nP=None):
        self.coords = np.asarray(coords, dtype=float)
        if self.coords.ndim == 1:
            self.coords = self.coords[:, None]
        self.el2n = np.asarray(el2n, dtype=np.int64)
        self.el2nP = None if el2nP is None else np.asarray(el2nP, dtype=np.int64)
        self.dim = self.coords.shape[1]
        self.nels = len(self.el2n)
        self.nnodes = len(self.coords)


def julia_float(x) -> str:
    """A Float64 as Julia's print shows it: the shortest round-trip digits,
    positional for 1e-5 < |x| < 1e6 (always with a decimal point), scientific
    as d.ddde±x otherwise; NaN, Inf, -Inf."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    t = Decimal(repr(abs(x))).normalize().as_tuple()   # repr: shortest round-trip digits
    digits = "".join(map(str, t.digits))
    e10 = len(digits) - 1 + t.exponent
    if -5 < e10 < 6:
        if e10 >= 0:
            whole = digits[:e10 + 1].ljust(e10 + 1, "0")
            frac = digits[e10 + 1:] or "0"
        else:
            whole, frac = "0", "0" * (-e10 - 1) + digits
        return f"{sign}{whole}.{frac}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}e{e10}"


_CORNERS = {
    1: {2: [0, 1], 3: [0, 2]},
    2: {3: [0, 1, 2], 6: [0, 1, 2], 7: [0, 1, 2], 4: [0, 1, 2, 3], 8: [0, 1, 2, 3], 9: [0, 1, 2, 3]},
    3: {4: [0, 1, 2, 3], 10: [0, 1, 2, 3], 11: [0, 1, 2, 3], 8: list(range(8)), 27: list(range(8))},
}
_KIND = {1: "line", 2: "2D", 3: "3D"}


def _corner_rows(dim, nlocal):
    """_vtk_corner_rows (postprocess.jl:360-379)"""
    try:
        return _CORNERS[dim][nlocal]
    except KeyError:
        raise ValueError(f"cannot write VTK {_KIND.get(dim, dim)} cells with {nlocal} local nodes") from None


def _cell_type(dim, ncorner):
    return {1: 3, 2: 5 if ncorner == 3 else 9, 3: 10 if ncorner == 4 else 12}[dim]


def _mesh_arrays(mesh):
    """_vtk_mesh_arrays (postprocess.jl:320-330): 2-D mixed meshes are written by
    their pressure (vertex) connectivity, 3-D ones by the Hex27 velocity cells."""
    coords = np.asarray(mesh.coords, dtype=float)
    if coords.ndim == 1:
        coords = coords[:, None]
    dim = coords.shape[1]
    el2nP = getattr(mesh, "el2nP", None)
    if el2nP is not None and dim == 2:
        el2n = np.asarray(el2nP, dtype=np.int64)
    elif hasattr(mesh, "vertex_el2n") and dim == 2:
        el2n = np.asarray(mesh.vertex_el2n(), dtype=np.int64)
    else:
        el2n = np.asarray(mesh.el2n, dtype=np.int64)
    return coords, el2n, dim


class _Topology:
    def __init__(self, mesh, coords=None):
        mc, el2n, dim = _mesh_arrays(mesh)
        if coords is not None:
            coords = np.asarray(coords, dtype=float)
            if coords.ndim == 1:
                coords = coords[:, None]
            if len(coords) != len(mc):
                raise DimensionMismatch(f"coords has length {len(coords)}; expected {len(mc)}")
        else:
            coords = mc
        rows = _corner_rows(dim, el2n.shape[1])
        self.el2n = el2n[:, rows]
        self.nodes = np.unique(self.el2n.ravel())
        if len(self.nodes) and (self.nodes[0] < 0 or self.nodes[-1] >= len(coords)):
            raise IndexError("connectivity refers to a node outside coords")
        self.node_map = np.full(len(coords), -1, dtype=np.int64)
        self.node_map[self.nodes] = np.arange(len(self.nodes))
        self.coords = coords
        self.nels = len(el2n)
        self.cell_type = _cell_type(dim, len(rows))


def _point_values(field, topo, name):
    if isinstance(field, tuple):
        return tuple(_point_values(c, topo, name) for c in field)
    v = np.asarray(field).ravel()
    if len(v) == len(topo.nodes):
        return v
    if len(v) == len(topo.coords):
        return v[topo.nodes]
    raise DimensionMismatch(f"point_data[{name}] has length {len(v)}; expected {len(topo.nodes)} or "
                            f"{len(topo.coords)}")


def _cell_values(field, topo, name):
    if isinstance(field, tuple):
        return tuple(_cell_values(c, topo, name) for c in field)
    v = np.asarray(field).ravel()
    if len(v) == topo.nels:
        return v
    raise DimensionMismatch(f"cell_data[{name}] has length {len(v)}; expected {topo.nels}")


_SYM = (0, 3, 4, 3, 1, 5, 4, 5, 2)   # row-major 3 x 3 of (xx, yy, zz, xy, xz, yz)


def _write_field(out, name, values):
    f = julia_float
    if not isinstance(values, tuple):
        out.append(f"SCALARS {name} float 1")
        out.append("LOOKUP_TABLE default")
        out.extend(f(x) for x in values)
        return
    n = len(values)
    if len({len(c) for c in values}) > 1:
        raise DimensionMismatch(f"the components of field {name} differ in length")
    if n in (2, 3):
        out.append(f"VECTORS {name} float")
        for i in range(len(values[0])):
            third = values[2][i] if n == 3 else 0.0
            out.append(f"{f(values[0][i])} {f(values[1][i])} {f(third)}")
        return
    if n in (6, 9):
        layout = _SYM if n == 6 else tuple(range(9))
        out.append(f"TENSORS {name} float")
        for i in range(len(values[0])):
            for row in range(3):
                out.append(" ".join(f(values[layout[3 * row + c]][i]) for c in range(3)))
        return
    raise ValueError(f"field {name} has {n} components; a tuple field must hold 2 or 3 (vector), "
                     "6 (symmetric tensor), or 9 (full tensor) of them")


def write_vtk(path, mesh, point_data=None, cell_data=None, title="FEMTools", coords=None):
    """write_vtk(path, mesh; point_data, cell_data, title, coords)
    (postprocess.jl:273-318). point_data/cell_data map names to arrays or
    tuples of component arrays; a point field has one value per mesh node or
    per written VTK point, a cell field one per element. coords replaces the
    written geometry (same node count)."""
    topo = _Topology(mesh, coords)
    f = julia_float
    pts = [(n, _point_values(v, topo, n)) for n, v in (point_data or {}).items()]
    cls = [(n, _cell_values(v, topo, n)) for n, v in (cell_data or {}).items()]
    out = ["# vtk DataFile Version 3.0", str(title), "ASCII", "DATASET UNSTRUCTURED_GRID",
           f"POINTS {len(topo.nodes)} float"]
    for i in topo.nodes:
        c = topo.coords[i]
        c3 = [c[k] if k < len(c) else 0.0 for k in range(3)]
        out.append(" ".join(f(x) for x in c3))
    nc = topo.el2n.shape[1]
    out.append(f"CELLS {topo.nels} {(nc + 1) * topo.nels}")
    for row in topo.node_map[topo.el2n]:
        out.append(" ".join([str(nc)] + [str(int(k)) for k in row]))
    out.append(f"CELL_TYPES {topo.nels}")
    out.extend([str(topo.cell_type)] * topo.nels)
    if pts:
        out.append(f"POINT_DATA {len(topo.nodes)}")
        for n, v in pts:
            _write_field(out, n, v)
    if cls:
        out.append(f"CELL_DATA {topo.nels}")
        for n, v in cls:
            _write_field(out, n, v)
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_stokes_vtk(path, mesh, coords_v, el2nP, DoFsP, P, v, post, title=None, cell_data=None, Q=None):
    """write_stokes_vtk (postprocess.jl:488-602). 2-D (v = (vx, vy)): pressure
    averaged from the element DoFs to the corner points, Vx, Vy and the vector
    V, the cell diagnostics strain_xx ... tau_II of post and, when Q is given,
    its cell mean. 3-D (v = (vx, vy, vz)): the vector velocity, the point
    elevation z, the cell pressure (the centre mode of Hex27 cells, the mean
    of the four values of tetrahedra) and the strain and tau tensors.
    el2nP and DoFsP are (nels, NP), 0-based; post maps the diagnostic names
    (exx, eyy, ezz, exy[, exz, eyz], eII, txx, tyy, tzz, txy[, txz, tyz], tII)
    to per-cell arrays."""
    topo = _Topology(mesh, coords_v)
    el2nP = np.asarray(el2nP, dtype=np.int64)
    DoFsP = np.asarray(DoFsP, dtype=np.int64)
    P = np.asarray(P, dtype=float).ravel()
    cell_data = dict(cell_data or {})
    if len(v) == 2:
        nodes = topo.node_map[el2nP]
        acc = np.zeros(len(topo.nodes))
        cnt = np.zeros(len(topo.nodes))
        np.add.at(acc, nodes.ravel(), P[DoFsP].ravel())
        np.add.at(cnt, nodes.ravel(), 1)
        vtk_P = acc / cnt
        Vx = np.asarray(v[0], dtype=float)[topo.nodes]
        Vy = np.asarray(v[1], dtype=float)[topo.nodes]
        for key, name in (("exx", "strain_xx"), ("eyy", "strain_yy"), ("ezz", "strain_zz"), ("exy", "strain_xy"),
                          ("eII", "strain_II"), ("txx", "tau_xx"), ("tyy", "tau_yy"), ("tzz", "tau_zz"),
                          ("txy", "tau_xy"), ("tII", "tau_II")):
            cell_data[name] = post[key]
        if Q is not None:
            cell_data["Q"] = np.asarray(Q, dtype=float)[DoFsP].mean(axis=1)
        return write_vtk(path, mesh, point_data={"P": vtk_P, "Vx": Vx, "Vy": Vy, "V": (Vx, Vy)},
                         cell_data=cell_data, coords=coords_v, title=title or "FEMTools Stokes 2D")
    vel = tuple(np.asarray(c, dtype=float)[topo.nodes] for c in v)
    z = topo.coords[topo.nodes, 2] if topo.coords.shape[1] > 2 else np.zeros(len(topo.nodes))
    nloc = np.asarray(mesh.el2n).shape[1]
    cell_P = P[DoFsP].mean(axis=1) if nloc in (10, 11) else P[DoFsP[:, 0]]
    cell_data.update({
        "P": cell_P,
        "strain": tuple(post[k] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")),
        "strain_II": post["eII"],
        "tau": tuple(post[k] for k in ("txx", "tyy", "tzz", "txy", "txz", "tyz")),
        "tau_II": post["tII"],
    })
    return write_vtk(path, mesh, point_data={"velocity": vel, "z": z}, cell_data=cell_data, coords=coords_v,
                     title=title or "FEMTools Stokes 3D")
