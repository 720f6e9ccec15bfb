"""GPU: compute_strain_rate_stress_postprocess (postprocess.jl:17-100) through
the C ABI (fem_stokes_postprocess) against the oracle to 1e-12, and a solved
state written with write_stokes_vtk and read back."""
import os
import sys
import tempfile

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

from test_vtk import parse_vtk  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_gpu_postprocess_matches_the_oracle():
    say("[gpu] strain-rate / stress diagnostics")
    n = 6
    coords, el2n = O.t7_mesh(n, n)
    mesh = fa.t7_box_mesh(n, n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial())
    rng = np.random.default_rng(3)
    vx, vy = rng.normal(size=len(coords)), rng.normal(size=len(coords))
    tau = tuple(rng.normal(size=(len(el2n), 7)) for _ in range(3))
    dr.v.x.copyto(vx)
    dr.v.y.copyto(vy)
    for name, t in zip(("xx", "yy", "xy"), tau):
        getattr(dr.tau, name).copyto(t)
    got = fa.compute_strain_rate_stress_postprocess(dr, mesh)
    ref = O.postprocess2(coords, el2n, vx, vy, tau)
    errs = {k: np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300) for k in ref if k != "ezz"}
    say("  " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-12
    assert not got["ezz"].any()


def test_gpu_solved_state_to_vtk():
    say("[gpu] solved state written as VTK")
    from test_gpu_parity import _viscous_pair
    w, pr, dr, mesh, (bx, vx, by, vy) = _viscous_pair(8)
    st = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy),
                  dt=w.dt, pressure_factor=w.pressure_factor, tolerance=1e-8, check_interval=w.check_interval)
    assert st["converged"]
    fa.update_stokes_current_stress(dr, mesh, w.dt)
    post = fa.compute_strain_rate_stress_postprocess(dr, mesh)
    el2nP = mesh.el2n[:, :3]
    DoFsP = np.arange(3 * mesh.nels).reshape(-1, 3)
    fd, path = tempfile.mkstemp(suffix=".vtk")
    os.close(fd)
    try:
        P = dr.P.numpy()
        fa.write_stokes_vtk(path, mesh, mesh.coords, el2nP, DoFsP, P, (dr.v.x.numpy(), dr.v.y.numpy()), post)
        with open(path) as fh:
            v = parse_vtk(fh.read())
    finally:
        os.remove(path)
    nvert = (8 + 1) ** 2
    assert len(v["points"]) == nvert and len(v["cells"]) == mesh.nels and set(v["cell_types"]) == {5}
    assert np.allclose(v["cell_data"]["tau_II"], post["tII"], rtol=1e-15)
    # pressure at a corner = the mean of the element values that meet there
    corner = int(el2nP[0, 0])
    k = np.searchsorted(np.unique(el2nP), corner)
    assert v["point_data"]["P"][k] == pytest.approx(P[DoFsP][el2nP == corner].mean(), rel=1e-14)


def test_gpu_postprocess_3d_matches_the_oracle():
    say("[gpu] 3-D strain-rate / stress diagnostics")
    coords, el2n = O.hex27_mesh(3, 2, 2)
    mesh = fa.hex27_box_mesh(3, 2, 2)
    assert np.array_equal(mesh.coords, coords)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(g=(0.0, 0.0, 0.0)))
    rng = np.random.default_rng(4)
    v = tuple(rng.normal(size=len(coords)) for _ in range(3))
    tau = tuple(rng.normal(size=(len(el2n), 27)) for _ in range(6))
    for c, a in zip("xyz", v):
        getattr(dr.v, c).copyto(a)
    for name, a in zip(("xx", "yy", "zz", "xy", "xz", "yz"), tau):
        getattr(dr.tau, name).copyto(a)
    got = fa.compute_strain_rate_stress_postprocess(dr, mesh)
    ref = O.postprocess3(coords, el2n, v, tau)
    errs = {k: np.abs(got[k] - ref[k]).max() / np.abs(ref[k]).max() for k in ref}
    say("  " + " ".join(f"{k} {e:.1e}" for k, e in errs.items()))
    assert max(errs.values()) < 1e-12
