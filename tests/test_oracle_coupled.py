"""The oracle's coupled thermal--Stokes step (solve_coupled!, src/stokes/
solvers/DR.jl:44-165, 514-661) and shear dissipation (assemble_shear_heating!,
src/stokes/assemblers/shear_heating.jl), pinned by the reference's own tests
(test/test_coupled_solver.jl) and by identities of the Maxwell model."""
import math

import numpy as np
import pytest

import pyoracle as O


def _thermal(coords, el2n, bc, k=1.0, alpha=0.0, eps=1e-8):
    m = O.thermal_material(k=(k,), Cp=(1.0,), rho0=(1.0,), alpha=(alpha,), K=(math.inf,))
    return O.ThermalProblem(coords, el2n, m, bc=bc)


def test_coupled_reference_case():
    # test/test_coupled_solver.jl:8-56: one cell, zero Dirichlet T and v, unit
    # source; the coupled step converges with the standalone thermal answer
    coords, el2n = O.t7_mesh(1, 1)
    gam = O.box_boundary_nodes(coords)
    zero = np.zeros(len(gam))
    th = _thermal(coords, el2n, (gam, zero))
    th.source[:] = 1.0
    ref = _thermal(coords, el2n, (gam, zero))
    ref.source[:] = 1.0
    ref.solve(1.0, tolerance=1e-8, max_iterations=2000, check_interval=1)
    pr = O.Problem(coords, el2n, O.material(eta=(1.0,), etab=(1.0,), G=(math.inf,), alpha=(0.0,)),
                   bc=((gam, zero), (gam, zero)))
    st = O.coupled_solve(th, pr, 1.0, tolerance_T=1e-8, pressure_factor=1.0, check_interval=1, iter_max=2000,
                         max_iterations=2000, max_ph_iterations=5, tolerance=1e-8)
    assert st["converged"]
    assert st["err_T"] < 1e-8
    assert st["thermal_iterations"] == st["iter"]
    assert np.allclose(th.T, ref.T, rtol=1e-12, atol=0)
    assert np.array_equal(pr.T.reshape(-1, 3), th.T[el2n[:, :3]])
    assert np.max(th.T) > 0


def _simple_shear(n, dt, eta, G, gd, shear):
    # test/test_coupled_solver.jl:59-106
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    eve = 1 / (1 / eta + 1 / (G * dt))
    Texp = dt * (eve * gd) ** 2 / eta
    th = _thermal(coords, el2n, (gam, np.full(len(gam), Texp if shear else 0.0)))
    pr = O.Problem(coords, el2n, O.material(eta=(eta,), etab=(1.0,), G=(G,), alpha=(0.0,)),
                   bc=((gam, gd * coords[gam, 1]), (gam, np.zeros(len(gam)))))
    pr.vx[:] = gd * coords[:, 1] + 0.1 * np.sin(np.pi * coords[:, 0]) * np.sin(np.pi * coords[:, 1])
    st = O.coupled_solve(th, pr, dt, tolerance_T=1e-10, shear_heating=shear, pressure_factor=1.0, check_interval=10,
                         iter_max=20000, max_iterations=20000, tolerance=1e-10)
    return st, th, pr, Texp


def test_coupled_shear_heating_closed_form():
    # uniform Maxwell stress τxy = ηve γ̇ dissipates (ηve γ̇)²/η; with the
    # boundary held at Δt Φ the whole field is Δt Φ
    st, th, pr, Texp = _simple_shear(2, 0.5, 2.0, 1.0, 1.0, True)
    assert st["converged"] and st["thermal_converged"]
    assert np.all(np.abs(th.T - Texp) < 1e-8)
    assert np.all(th.source == 0.0)
    st, th, pr, _ = _simple_shear(2, 0.5, 2.0, 1.0, 1.0, False)
    assert st["converged"]
    assert np.all(np.abs(th.T) < 1e-12)


def _random_state(n, seed, G, dp=None, vscale=1.0):
    rng = np.random.default_rng(seed)
    coords, el2n = O.t7_mesh(n, n)
    ne = len(el2n)
    ph = rng.integers(1, 3, ne).astype(np.int32)
    tau_old = tuple(vscale * rng.normal(size=(ne, 7)) for _ in range(3))
    pr = O.Problem(coords, el2n, O.material(eta=(1.0, 7.0), etab=(1.0, 1.0), G=G, alpha=(0.0, 0.0)),
                   dp=dp, phases_v=ph, phases_P=ph, tau_old=tau_old)
    pr.vx[:] = vscale * rng.normal(size=len(coords))
    pr.vy[:] = vscale * rng.normal(size=len(coords))
    pr.P[:] = rng.normal(size=3 * ne)
    return pr, ph, tau_old


@pytest.mark.parametrize("G", [(math.inf, math.inf), (0.7, 3.0)])
def test_shear_heating_is_the_viscous_dissipation_of_the_maxwell_stress(G):
    # the visco-elastic update τ = 2ηve(ε̇' + τ_old/(2GΔt)) rearranges to
    # ε̇' = τ/(2η) + (τ − τ_old)/(2GΔt), so τ : (ε̇ − ε̇ᵉˡ) = τ : τ/(2η) for any
    # τ_old, with the out-of-plane deviator in τ : τ
    dt = 0.3
    pr, ph, _ = _random_state(3, 5, G)
    tau = [t.copy() for t in pr.update_current_stress(dt)]
    phi = pr.shear_heating(dt)
    eta = np.array([1.0, 7.0])[ph - 1][:, None]
    txx, tyy, txy = tau
    tzz = -(txx + tyy)
    expect = (txx ** 2 + tyy ** 2 + tzz ** 2 + 2 * txy ** 2) / (2 * eta)
    assert np.allclose(phi, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())


def test_shear_heating_plastic_adds_the_plastic_work():
    # with a Drucker-Prager return τ = τᵗʳ − 2ηve λ ∂Q/∂τ, the same algebra
    # gives τ : (ε̇ − ε̇ᵉˡ) = τ : τ/(2η) + τ : λ∂Q/∂τ, λ∂Q/∂τ = (τᵗʳ − τ)/(2ηve)
    # (its zz component −(xx + yy)), zero at points that do not yield. (The
    # sign of the plastic term is not fixed: the return of rheology.jl:272-334
    # scales ∂Q/∂τ by the trial stress and can step past the yield surface.)
    dt = 0.3
    eta_p, G_p = np.array([1.0, 7.0]), np.array([0.7, 3.0])
    dp = O.drucker_prager(phi=(math.radians(30),) * 2, psi=(0.0, 0.0), C_=(0.3, 0.3), eta_reg=(0.0, 0.0),
                          Kb=(1.0, 1.0))
    pr, ph, _ = _random_state(3, 6, tuple(G_p), dp=dp, vscale=0.1)
    pr.P[:] = np.abs(pr.P)
    tau = [t.copy() for t in pr.update_current_stress(dt)]
    phi = pr.shear_heating(dt)
    pr_ve, _, _ = _random_state(3, 6, tuple(G_p), vscale=0.1)
    pr_ve.P[:] = pr.P
    trial = [t.copy() for t in pr_ve.update_current_stress(dt)]
    eta, G = eta_p[ph - 1][:, None], G_p[ph - 1][:, None]
    eve = 1 / (1 / eta + 1 / (G * dt))
    txx, tyy, txy = tau
    tzz = -(txx + tyy)
    visc = (txx ** 2 + tyy ** 2 + tzz ** 2 + 2 * txy ** 2) / (2 * eta)
    lq = [(a - b) / (2 * eve) for a, b in zip(trial, tau)]
    plastic = txx * lq[0] + tyy * lq[1] + tzz * -(lq[0] + lq[1]) + 2 * txy * lq[2]
    assert np.allclose(phi, visc + plastic, rtol=1e-12, atol=1e-12 * np.abs(phi).max())
    yielded = ~np.isclose(tau[2], trial[2], rtol=0, atol=1e-14)
    assert yielded.any() and (~yielded).any()
    assert np.all(plastic[yielded] != 0)
    assert np.all(plastic[~yielded] == 0)


def _buoyancy(n, ncheck=10):
    # a hot blob under gravity: T drives the density and the thermal-expansion
    # term of the continuity residual, so the coupling acts on the flow
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    x, y = coords[:, 0], coords[:, 1]
    th = O.ThermalProblem(coords, el2n, O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,),
                                                           K=(math.inf,)),
                          bc=(gam, np.zeros(len(gam))))
    th.source[:] = 50.0 * np.exp(-((x - 0.4) ** 2 + (y - 0.5) ** 2) / 0.02)
    th.T0[:] = 0.2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    th.T[:] = th.T0
    zero = np.zeros(len(gam))
    mat = O.material(eta=(1.0,), etab=(1.0,), G=(math.inf,), alpha=(0.05,), rho0=(1.0,), K=(math.inf,),
                     g=(0.0, -10.0))
    pr = O.Problem(coords, el2n, mat, bc=((gam, zero), (gam, zero)))
    kw = dict(pressure_factor=1.0, check_interval=ncheck, tolerance=1e-9, iter_max=20000, max_iterations=40000)
    return th, pr, kw, el2n


def test_coupled_buoyancy_is_consistent_with_both_standalone_solves():
    th, pr, kw, el2n = _buoyancy(6)
    dt = 0.01
    st = O.coupled_solve(th, pr, dt, tolerance_T=1e-9, **kw)
    assert st["converged"] and st["thermal_converged"] and st["err_T"] < 1e-9
    assert st["thermal_iterations"] == st["iter"]
    assert np.array_equal(pr.T.reshape(-1, 3), th.T[el2n[:, :3]])
    assert np.array_equal(pr.T0.reshape(-1, 3), th.T0[el2n[:, :3]])
    # the thermal step does not see the flow: it matches the standalone solve
    th2, _, _, _ = _buoyancy(6)
    th2.solve(dt, tolerance=1e-12, max_iterations=20000, check_interval=10)
    assert np.allclose(th.T, th2.T, rtol=0, atol=1e-7 * np.abs(th2.T).max())
    # the flow matches a Stokes solve with the final temperature held fixed
    _, pr2, _, _ = _buoyancy(6)
    pr2.T[:], pr2.T0[:] = pr.T, pr.T0
    st2 = pr2.solve(dt, **kw)
    assert st2["converged"]
    scale = max(np.abs(pr2.vx).max(), np.abs(pr2.vy).max())
    assert scale > 0
    assert np.abs(pr.vx - pr2.vx).max() < 1e-6 * scale
    assert np.abs(pr.vy - pr2.vy).max() < 1e-6 * scale


def test_coupled_node_layout_mismatch():
    coords, el2n = O.t7_mesh(1, 1)
    c2, e2 = O.t7_mesh(2, 2)
    th = _thermal(c2, e2, (np.zeros(0, np.int32), np.zeros(0)))
    pr = O.Problem(coords, el2n, O.material())
    with pytest.raises(RuntimeError, match="must match the Stokes velocity-node layout"):
        O.coupled_solve(th, pr, 1.0)
