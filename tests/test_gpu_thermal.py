"""GPU parity of the ThermalDiffusionDR path (libfemdr.so through its C ABI)
against the CPU oracle: residual and Jacobian diagnostics to 1e-12 of the
field scale, converged temperatures to 1e-10, identical iteration counts;
closed forms; error paths."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


MAT2 = dict(k=(1.5, 0.4), Cp=(1.2, 0.8), rho0=(2.0, 3.0), alpha=(0.02, 0.01), K=(50.0, math.inf))


def pair(n, mat=MAT2, seed=0, structured=True, bc=None, inclusion=True):
    coords, el2n = O.t7_mesh(n, n)
    if structured:
        mesh = fa.t7_box_mesh(n, n)
    else:
        mesh = fa.t7_mesh_from_arrays(coords, el2n)
    nn = len(coords)
    ph = np.ones(nn, np.int32)
    if inclusion:
        ph[(coords[:, 0] - 0.5) ** 2 + (coords[:, 1] - 0.5) ** 2 < 0.3 ** 2] = 2
    mo = O.thermal_material(**mat)
    pr = O.ThermalProblem(coords, el2n, mo, phases=ph, bc=bc if bc is not None else (np.zeros(0, np.int32), np.zeros(0)))
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat))
    dr.phases.copyto(ph)
    rng = np.random.default_rng(seed)
    for k in ("T", "T0", "source", "P"):
        v = rng.normal(size=nn) + (300.0 if k in ("T", "T0") else 0.0)
        getattr(pr, k)[:] = v
        getattr(dr, k).copyto(v)
    return pr, dr, mesh, coords


@pytest.mark.parametrize("structured", [True, False])
def test_thermal_residual_and_jacobian_parity(structured):
    say(f"[gpu] thermal residual parity structured={structured}")
    pr, dr, mesh, _ = pair(7, structured=structured)
    pr.residual(0.3, Tref=290.0, jacobian=True)
    fa.thermal_residual(dr, 0.3, Tref=290.0, compute_jacobian=True)
    for name in ("R", "dRdT", "PC"):
        e = rel(getattr(dr, name).numpy(), getattr(pr, name))
        say(f"  {name}: {e:.2e}")
        assert e < 1e-12


@pytest.mark.parametrize("dt", [0.05, 1.0])
def test_thermal_solve_parity_inclusion(dt):
    # two-phase disc, thermal expansion and compressibility, smooth initial
    # field, Dirichlet values on the box boundary
    n = 12
    say(f"[gpu] thermal solve parity n={n} dt={dt}")
    coords, _ = O.t7_mesh(n, n)
    b = O.box_boundary_nodes(coords)
    x, y = coords[:, 0], coords[:, 1]
    T0 = 300 + np.sin(3 * x) * np.cos(2 * y)
    pr, dr, mesh, coords = pair(n, bc=(b, T0[b] + 0.5))
    for k, v in (("T", T0), ("T0", T0), ("source", 1 + x * y), ("P", 0.1 * y)):
        getattr(pr, k)[:] = v
        getattr(dr, k).copyto(v)
    st_o = pr.solve(dt, tolerance=1e-9, check_interval=20, Tref=300.0)
    st_g = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(b, T0[b] + 0.5), dt=dt, tolerance=1e-9,
                    check_interval=20, Tref=300.0)
    say(f"  oracle {st_o} | gpu {st_g}")
    assert st_o["converged"] and st_g["converged"]
    assert st_o["iterations"] == st_g["iterations"]
    assert rel(dr.T.numpy(), pr.T) < 1e-10


def test_thermal_uniform_source_closed_form():
    mat = dict(k=(1.0,), Cp=(2.0,), rho0=(1.5,), alpha=(0.0,), K=(math.inf,))
    mesh = fa.t7_box_mesh(8, 8)
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat))
    dr.source.fill(3.0)
    dr.T0.fill(1.0)
    dr.T.fill(1.0)
    st = fa.solve(dr, mesh, None, dt=0.5, tolerance=1e-12, check_interval=10)
    assert st["converged"]
    assert np.allclose(dr.T.numpy(), 1.0 + 0.5 * 3.0 / (1.5 * 2.0), rtol=1e-10)


def test_thermal_error_paths():
    mat = dict(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    mesh = fa.t7_box_mesh(2, 2)
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat), eps=0.0)
    dr.source.fill(1.0)
    with pytest.raises(fa.FemError, match="Thermal diffusion DR solver did not converge"):
        fa.solve(dr, mesh, None, dt=1.0, max_iterations=3, check_interval=1)
    st = fa.solve(dr, mesh, None, dt=1.0, max_iterations=3, check_interval=1, throw_on_failure=False)
    assert not st["converged"] and st["iterations"] == 3
    # a node in no element has a zero preconditioner
    coords, el2n = O.t7_mesh(1, 1)
    coords = np.vstack([coords, [[2.0, 2.0]]])
    m2 = fa.t7_mesh_from_arrays(coords, el2n)
    d2 = fa.ThermalDiffusionDR(m2, fa.ThermalMaterial(**mat))
    with pytest.raises(fa.FemError, match="thermal diffusion preconditioner produced invalid λmax"):
        fa.solve(d2, m2, None, dt=1.0)
    # phase ids are checked against the material
    d3 = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat))
    d3.phases.fill(2)
    with pytest.raises(ValueError, match="phase id"):
        fa.solve(d3, mesh, None, dt=1.0)
