"""GPU parity of ThermalDiffusionDR on Hex27 boxes (k_thermal_hex27 through
the C ABI) against the oracle's Hex27 restatement (itself pinned by closed
forms in test_oracle_thermal3.py): residual and Jacobian diagnostics to 1e-12
of the field scale, and DR solves with identical iteration counts and
temperatures to 1e-10."""
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


def pair(n, mat=MAT2, seed=0, bc=None):
    mesh = fa.hex27_box_mesh(*n)
    coords, el2n = mesh.coords, mesh.el2n
    nn = len(coords)
    ph = np.ones(nn, np.int32)
    ph[((coords - 0.5) ** 2).sum(axis=1) < 0.3 ** 2] = 2
    pr = O.ThermalProblem3(coords, el2n, O.thermal_material(**mat), phases=ph,
                           bc=bc if bc is not None else (np.zeros(0, np.int32), np.zeros(0)))
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat))
    dr.phases.copyto(ph)
    rng = np.random.default_rng(seed)
    for k in ("T", "T0", "source", "P"):
        v = rng.normal(size=nn) + (300.0 if k in ("T", "T0") else 0.0)
        getattr(pr, k)[:] = v
        getattr(dr, k).copyto(v)
    return pr, dr, mesh, coords


@pytest.mark.parametrize("n", [(2, 2, 2), (4, 3, 2)])
def test_hex27_thermal_residual_and_jacobian_parity(n):
    say(f"[gpu] Hex27 thermal residual parity {n}")
    pr, dr, mesh, _ = pair(n)
    pr.residual(0.3, Tref=290.0, jacobian=True)
    fa.thermal_residual(dr, 0.3, Tref=290.0, compute_jacobian=True)
    for name in ("R", "dRdT", "PC"):
        e = rel(getattr(dr, name).numpy(), getattr(pr, name))
        say(f"  {name}: {e:.2e}")
        assert e < 1e-12, name


def test_hex27_thermal_solve_parity():
    say("[gpu] Hex27 thermal solve parity")
    mesh = fa.hex27_box_mesh(3, 3, 2)
    x = mesh.coords[:, 0]
    bx = np.nonzero((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12))[0].astype(np.int32)
    bcv = 1.0 - x[bx]
    pr, dr, mesh, coords = pair((3, 3, 2), bc=(bx, bcv))
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
