"""GPU: the coupled thermal--Stokes step on Hex27 states (solve_coupled!,
src/stokes/solvers/DR.jl:44-165, 514-661), the 3-D shear dissipation
(shear_heating.jl:91-103) and the temperature terms of the 3-D residuals,
through the C ABI against the oracle (itself pinned in
test_oracle_coupled3.py): Φ and residuals to 1e-12 of their scale, coupled
solves with identical iteration counts and fields to 1e-10."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

from test_gpu_hex27 import make_pair3, rand_tau, DP2  # noqa: E402
from test_oracle_coupled3 import buoyancy3  # noqa: E402

INF = math.inf


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _deviatoric(tau):
    tr = (tau[0] + tau[1] + tau[2]) / 3
    return [t - tr if k < 3 else t for k, t in enumerate(tau)]


@pytest.mark.parametrize("plastic", [False, True])
def test_gpu_shear_heating3_parity(plastic):
    n = (3, 2, 2)
    ne = 12
    ph = (np.arange(ne) % 2 + 1).astype(np.int32)
    tau = _deviatoric(rand_tau(ne, 13, 0.2))
    say(f"[gpu] hex27 shear heating plastic={plastic}")
    pr, dr, mesh, dp_o, dp_g = make_pair3(n, eta=(1.3, 0.4), G=(0.8, 2.0), etab=(10.0, 10.0), phases=ph,
                                          tau_old=tau, dp=DP2 if plastic else None, seed=7)
    dt = 0.4
    phi_o = pr.shear_heating(dt)
    phi_g = fa.assemble_shear_heating(dr, mesh, dt, plastic=dp_g)
    assert phi_g.shape == (ne, 27)
    say("  rel", rel(phi_g, phi_o))
    assert rel(phi_g, phi_o) < 1e-12


def test_gpu_temperature_terms_of_the_3d_residual():
    n = (3, 2, 2)
    ne = 12
    ph = (np.arange(ne) % 2 + 1).astype(np.int32)
    say("[gpu] hex27 residual with a temperature field")
    pr, dr, mesh, dp_o, dp_g = make_pair3(n, eta=(1.3, 0.2), G=(0.8, 2.0), K=(50.0, 30.0), alpha=(0.01, 0.02),
                                          Tref=1.5, rho0=(1.0, 2.0), g=(0.0, 0.0, -1.0), phases=ph,
                                          tau_old=_deviatoric(rand_tau(ne, 11)), seed=3)
    rng = np.random.default_rng(5)
    T = 2.0 + rng.normal(size=4 * ne)
    T0 = T - 0.3 * rng.normal(size=4 * ne)
    pr.T, pr.T0 = T, T0
    dr.T.copyto(T)
    dr.T0.copyto(T0)
    dt = 0.4
    pr.pressure_scaling(20.0, dt, True)
    fa.assemble_viscosity_weighted_pressure_scaling(dr, 20.0, dt, True)
    pr.pressure_residual(dt, True)
    pr.momentum_residual(dt)
    fa.residual(dr, dt, finite_K=True)
    assert rel(dr.RP.numpy(), pr.RP) < 1e-12
    for c in range(3):
        assert rel(dr.Rv[c].numpy(), pr.Rv[c]) < 1e-12
    # the temperature changes the result (it is not read as zero)
    pr.T = pr.T0 = None
    pr.pressure_residual(dt, True)
    assert rel(dr.RP.numpy(), pr.RP) > 1e-6


def _gpu_buoyancy3(n):
    th_o, pr_o, kw, el2n, tmat = buoyancy3(n)
    mesh = fa.hex27_box_mesh(*n)
    assert np.array_equal(mesh.el2n, el2n)
    coords = mesh.coords
    gam = O.box_boundary_nodes3(coords)
    th = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,)),
                               eps=1e-9)
    th.source.copyto(th_o.source)
    th.T0.copyto(th_o.T0)
    th.T.copyto(th_o.T)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0,), etab=(1.0,), G=(INF,), alpha=(0.05,), rho0=(1.0,),
                                             K=(INF,), g=(0.0, 0.0, -10.0)))
    zero = fa.DirichletBoundaryCondition(gam, np.zeros(len(gam)))
    return th_o, pr_o, kw, th, dr, mesh, fa.DirichletBoundaryCondition(gam, np.zeros(len(gam))), (zero, zero, zero)


def test_gpu_coupled3_buoyancy_parity():
    n = (2, 2, 2)
    say(f"[gpu] hex27 coupled buoyancy {n}")
    th_o, pr_o, kw, th, dr, mesh, bc_T, bcs = _gpu_buoyancy3(n)
    dt = 0.01
    st_o = O.coupled_solve3(th_o, pr_o, dt, tolerance_T=1e-9, **kw)
    kw_g = dict(kw)
    kw_g["iterMax"] = kw_g.pop("iter_max")
    st_g = fa.solve_coupled(th, dr, mesh, mesh, bc_T, *bcs, dt=dt, **kw_g)
    say(f"  oracle iter {st_o['iter']} itPH {st_o['itPH']} thermal {st_o['thermal_iterations']} | gpu iter "
        f"{st_g['iter']} itPH {st_g['itPH']} thermal {st_g['thermal_iterations']}")
    assert st_o["converged"] and st_g["converged"] and st_g["thermal_converged"]
    assert st_g["iter"] == st_o["iter"] and st_g["itPH"] == st_o["itPH"]
    assert st_g["thermal_iterations"] == st_o["thermal_iterations"]
    assert rel(th.T.numpy(), th_o.T) < 1e-10
    assert rel(dr.T.numpy(), pr_o.T) < 1e-10
    for c in range(3):
        e = rel(dr.v[c].numpy(), pr_o.v[c])
        say("  v", c, e)
        assert e < 1e-10
    assert rel(dr.P.numpy(), pr_o.P) < 1e-10


def test_gpu_coupled3_shear_heating_closed_form():
    # test_oracle_coupled3.py's simple shear of a Hex27 box with the wall
    # temperature held at Δt Φ: the whole field reaches Δt Φ = Δt η γ̇²
    coords_o, el2n = O.hex27_mesh(2, 1, 2)
    mesh = fa.hex27_box_mesh(2, 1, 2)
    coords = mesh.coords
    gam = O.box_boundary_nodes3(coords)
    eta, rate, dt = 2.0, 0.5, 1.0
    vx = rate * coords[:, 2]
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=(eta,), etab=(1.0,), G=(INF,), K=(INF,), g=(0.0, 0.0, 0.0)))
    dr.v.x.copyto(vx + 0.1 * np.sin(np.pi * coords[:, 0]) * np.sin(np.pi * coords[:, 2]))
    bcs = (fa.DirichletBoundaryCondition(gam, vx[gam]), fa.DirichletBoundaryCondition(gam, np.zeros(len(gam))),
           fa.DirichletBoundaryCondition(gam, np.zeros(len(gam))))
    phi = eta * rate ** 2
    th = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,)),
                               eps=1e-10)
    st = fa.solve_coupled(th, dr, mesh, mesh, fa.DirichletBoundaryCondition(gam, np.full(len(gam), dt * phi)),
                          *bcs, dt=dt, shear_heating=True, tolerance=1e-10, check_interval=10, max_iterations=20000,
                          iterMax=20000, pressure_factor=1.0)
    say(f"[gpu] hex27 coupled shear heating: iter {st['iter']} err_T {st['err_T']:.2e}")
    assert st["converged"] and st["thermal_converged"]
    assert np.abs(th.T.numpy() - dt * phi).max() < 1e-9
    assert np.all(th.source.numpy() == 0.0)


def test_gpu_coupled3_argument_checks():
    mesh3 = fa.hex27_box_mesh(1, 1, 1)
    mesh2 = fa.t7_box_mesh(2, 2)
    dr = fa.StokesDR(mesh3, fa.StokesMaterial(eta=(1.0,), g=(0.0, 0.0, 0.0)))
    th2 = fa.ThermalDiffusionDR(mesh2, fa.ThermalMaterial(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,)))
    b = fa.DirichletBoundaryCondition(np.zeros(0, np.int32), np.zeros(0))
    with pytest.raises((fa.DimensionMismatch, ValueError)):
        fa.solve_coupled(th2, dr, mesh2, mesh3, b, b, b, b, dt=1.0)
    th3 = fa.ThermalDiffusionDR(mesh3, fa.ThermalMaterial(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,)))
    with pytest.raises(fa.DimensionMismatch):
        fa.solve_coupled(th3, dr, mesh3, mesh3, b, b, b, dt=1.0)   # two velocity conditions on a 3-D state
    with pytest.raises(ValueError):
        fa.solve_coupled(th3, dr, mesh3, mesh3, b, b, b, b, dt=1.0, precision="fp32")
