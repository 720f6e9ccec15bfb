"""GPU parity of ThermalDiffusionDR on Q9 boxes (k_thermal_q9 through the C
ABI) against the oracle's Q9 restatement (itself pinned by closed forms in
test_oracle_thermal_q9.py): the generated mesh equal to the oracle's,
residual and Jacobian diagnostics to 1e-12 of the field scale, and DR solves
with identical iteration counts and temperatures to 1e-10, up to the
48 x 48 one-step case."""
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


def pair(n, mat=MAT2, seed=0, bc=None, box=((0.0, 1.0), (0.0, 1.0))):
    mesh = fa.q9_box_mesh(*n, domain=box)
    coords, el2n = mesh.coords, mesh.el2n
    nn = len(coords)
    ph = np.ones(nn, np.int32)
    ph[((coords - 0.5) ** 2).sum(axis=1) < 0.3 ** 2] = 2
    pr = O.ThermalProblemQ9(coords, el2n, O.thermal_material(**mat), phases=ph,
                            bc=bc if bc is not None else (np.zeros(0, np.int32), np.zeros(0)))
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat))
    dr.phases.copyto(ph)
    rng = np.random.default_rng(seed)
    for k in ("T", "T0", "source", "P"):
        v = rng.normal(size=nn) + (300.0 if k in ("T", "T0") else 0.0)
        getattr(pr, k)[:] = v
        getattr(dr, k).copyto(v)
    return pr, dr, mesh, coords


def test_q9_mesh_matches_the_oracle_mesh():
    say("[gpu] Q9 mesh layout")
    box = ((0.5, 2.5), (-1.0, 0.5))
    mesh = fa.q9_box_mesh(3, 2, domain=box)
    coords, el2n = O.q9_mesh(3, 2, (0.5, 2.5, -1.0, 0.5))
    assert np.array_equal(mesh.el2n, el2n)
    assert np.array_equal(mesh.coords, coords)
    assert mesh.nv_local == 9 and mesh.nq == 9 and mesh.nnodesP == 0 and mesh.np_local == 0
    assert np.array_equal(np.sort(mesh.boundary_nodes), O.box_boundary_nodes(coords, (0.5, 2.5, -1.0, 0.5)))
    with pytest.raises(ValueError, match="Q9"):
        fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0,)))


@pytest.mark.parametrize("n", [(2, 2), (5, 3), (16, 12)])
def test_q9_thermal_residual_and_jacobian_parity(n):
    say(f"[gpu] Q9 thermal residual parity {n}")
    pr, dr, mesh, _ = pair(n, box=((0.0, 1.3), (0.0, 0.7)))
    pr.residual(0.3, Tref=290.0, jacobian=True)
    fa.thermal_residual(dr, 0.3, Tref=290.0, compute_jacobian=True)
    for name in ("R", "dRdT", "PC"):
        e = rel(getattr(dr, name).numpy(), getattr(pr, name))
        say(f"  {name}: {e:.2e}")
        assert e < 1e-12, name


def test_q9_thermal_solve_parity():
    say("[gpu] Q9 thermal solve parity")
    mesh = fa.q9_box_mesh(6, 5)
    x = mesh.coords[:, 0]
    bx = np.nonzero((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12))[0].astype(np.int32)
    bcv = 1.0 - x[bx]
    pr, dr, mesh, coords = pair((6, 5), bc=(bx, bcv))
    for a in (pr.T, pr.T0):
        a[:] = 0.0
    dr.T.fill(0.0)
    dr.T0.fill(0.0)
    kw = dict(tolerance=1e-9, max_iterations=50_000, check_interval=50)
    so = pr.solve(0.05, Tref=0.0, **kw)
    sd = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(bx, bcv), dt=0.05, Tref=0.0, **kw)
    say(f"  iterations device {sd['iterations']} oracle {so['iterations']}")
    assert sd["converged"] and so["converged"] and sd["iterations"] == so["iterations"]
    assert rel(dr.T.numpy(), pr.T) < 1e-10


def test_q9_thermal_step48_parity():
    say("[gpu] Q9 48x48 thermal step")
    from femtools_amd import problems
    mesh = fa.q9_box_mesh(48, 48)
    coords, el2n = mesh.coords, mesh.el2n
    spec = problems.q9_thermal_step(coords)
    assert len(spec["bc_dofs"]) == 4 * 96
    tm = spec["material"]
    pr = O.ThermalProblemQ9(coords, el2n, O.thermal_material(**tm), bc=(spec["bc_dofs"], spec["bc_vals"]))
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**tm))
    for k in ("T", "T0", "source"):
        getattr(pr, k)[:] = spec[k]
        getattr(dr, k).copyto(spec[k])
    kw = spec["solve"]
    so = pr.solve(**kw)
    sd = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(spec["bc_dofs"], spec["bc_vals"]), **kw)
    say(f"  iterations device {sd['iterations']} oracle {so['iterations']}, {sd['seconds'] * 1e3:.1f} ms on the device")
    assert sd["converged"] and so["converged"] and sd["iterations"] == so["iterations"]
    assert rel(dr.T.numpy(), pr.T) < 1e-10
