"""Oracle: the 3-D coupled thermal--Stokes pieces (solve_coupled! on a Hex27
state, src/stokes/solvers/DR.jl:44-165; the 3-D shear dissipation,
src/stokes/assemblers/shear_heating.jl:91-103; the temperature terms of the
3-D residuals, pressure_residual.jl:132-158 and momentum_residuals.jl:11-16),
pinned by closed forms and by the standalone thermal and Stokes solves."""
import math

import numpy as np
import pytest

import pyoracle as O

INF = math.inf


def _random_state3(n, seed, G, dp=None, vscale=1.0):
    rng = np.random.default_rng(seed)
    coords, el2n = O.hex27_mesh(*n)
    ne = len(el2n)
    ph = rng.integers(1, 3, ne).astype(np.int32)
    tau_old = [vscale * rng.normal(size=(ne, 27)) for _ in range(6)]
    tr = (tau_old[0] + tau_old[1] + tau_old[2]) / 3   # a deviatoric stress history
    for k in range(3):
        tau_old[k] -= tr
    pr = O.Problem3(coords, el2n, O.material(eta=(1.0, 7.0), etab=(1.0, 1.0), G=G, alpha=(0.0, 0.0)),
                    dp=dp, phases=ph, tau_old=tau_old)
    for c in range(3):
        pr.v[c][:] = vscale * rng.normal(size=len(coords))
    pr.P[:] = rng.normal(size=4 * ne)
    return pr, ph


def _tau_store(pr, dt):
    pr.update_current_stress(dt)
    return [t.copy() for t in pr.tau_store]


@pytest.mark.parametrize("G", [(INF, INF), (0.7, 3.0)])
def test_shear_heating3_is_the_viscous_dissipation_of_the_maxwell_stress(G):
    # τ = 2ηve(ε̇' + τ_old/(2GΔt)) gives ε̇' = τ/(2η) + (τ − τ_old)/(2GΔt), so
    # τ : (ε̇ − ε̇ᵉˡ) = τ : τ/(2η) for any τ_old (τ deviatoric, shear terms twice)
    dt = 0.3
    pr, ph = _random_state3((2, 2, 2), 5, G)
    tau = _tau_store(pr, dt)
    phi = pr.shear_heating(dt)
    eta = np.array([1.0, 7.0])[ph - 1][:, None]
    txx, tyy, tzz, txy, txz, tyz = tau
    expect = (txx ** 2 + tyy ** 2 + tzz ** 2 + 2 * (txy ** 2 + txz ** 2 + tyz ** 2)) / (2 * eta)
    assert np.allclose(phi, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())


def test_shear_heating3_plastic_adds_the_plastic_work():
    # with the Drucker-Prager return τ = τᵗʳ − 2ηve λ ∂Q/∂τ the same algebra
    # adds τ : (τᵗʳ − τ)/(2ηve), zero at points that do not yield
    dt = 0.3
    eta_p, G_p = np.array([1.0, 7.0]), np.array([0.7, 3.0])
    dp = O.drucker_prager(phi=(math.radians(30),) * 2, psi=(0.0, 0.0), C_=(0.3, 0.3), eta_reg=(0.0, 0.0),
                          Kb=(1.0, 1.0))
    pr, ph = _random_state3((2, 2, 2), 6, tuple(G_p), dp=dp, vscale=0.1)
    pr.P[:] = np.abs(pr.P)
    tau = _tau_store(pr, dt)
    phi = pr.shear_heating(dt)
    pr_ve, _ = _random_state3((2, 2, 2), 6, tuple(G_p), vscale=0.1)
    pr_ve.P[:] = pr.P
    trial = _tau_store(pr_ve, dt)
    eta, G = eta_p[ph - 1][:, None], G_p[ph - 1][:, None]
    eve = 1 / (1 / eta + 1 / (G * dt))
    visc = (sum(t ** 2 for t in tau[:3]) + 2 * sum(t ** 2 for t in tau[3:])) / (2 * eta)
    lq = [(a - b) / (2 * eve) for a, b in zip(trial, tau)]
    plastic = sum(t * l for t, l in zip(tau[:3], lq[:3])) + 2 * sum(t * l for t, l in zip(tau[3:], lq[3:]))
    assert np.allclose(phi, visc + plastic, rtol=1e-11, atol=1e-11 * np.abs(phi).max())
    yielded = ~np.isclose(tau[3], trial[3], rtol=0, atol=1e-14)
    assert yielded.any() and (~yielded).any()


def test_shear_heating3_uniform_pure_shear():
    # v = ε̇ (x − ½, 0, −(z − ½)) in a viscous material: τ = 2η ε̇', Φ = 4η ε̇²
    coords, el2n = O.hex27_mesh(2, 1, 2)
    pr = O.Problem3(coords, el2n, O.material(eta=(3.0,), G=(INF,)))
    pr.v[0][:] = 0.4 * (coords[:, 0] - 0.5)
    pr.v[2][:] = -0.4 * (coords[:, 2] - 0.5)
    phi = pr.shear_heating(1.0)
    assert np.allclose(phi, 4 * 3.0 * 0.4 ** 2, rtol=1e-12)


def test_temperature_terms_of_the_3d_residuals():
    # v = 0, P = 0, T − T0 = ΔT everywhere: the continuity residual is
    # α ΔT/Δt ∫ N^P_i dΩ (the cell volume for the constant-sum mode, zero for
    # the odd modes ξ, η, ζ); the momentum residual sums to the weight of the
    # expanded column, ΣR_z = −ρ0 (1 − α (T − Tref)) g_z V
    n = (2, 2, 1)
    coords, el2n = O.hex27_mesh(*n)
    ne = len(el2n)
    al, rho0, gz, Tref, T, dT, dt = 0.02, 1.3, -2.0, 0.5, 3.0, 0.7, 0.25
    mat = O.material(eta=(1.0,), etab=(INF,), G=(INF,), rho0=(rho0,), K=(INF,), alpha=(al,), Tref=Tref)
    pr = O.Problem3(coords, el2n, mat, g=(0.0, 0.0, gz))
    pr.T, pr.T0 = np.full(4 * ne, T), np.full(4 * ne, T - dT)
    pr.pressure_residual(dt)
    RP = pr.RP.reshape(ne, 4)
    vol = 1.0 / ne
    assert np.allclose(RP[:, 0], al * dT / dt * vol, rtol=1e-12)
    assert np.allclose(RP[:, 1:], 0.0, atol=1e-14)
    pr.momentum_residual(dt)
    assert abs(pr.Rv[2].sum() - (-rho0 * (1 - al * (T - Tref)) * gz)) < 1e-12
    assert np.abs(pr.Rv[0]).max() < 1e-13 and np.abs(pr.Rv[1]).max() < 1e-13
    # without T the density is taken at T = 0
    pr.T = pr.T0 = None
    pr.momentum_residual(dt)
    assert abs(pr.Rv[2].sum() - (-rho0 * (1 - al * (0.0 - Tref)) * gz)) < 1e-12


def buoyancy3(n, ncheck=10):
    # a hot blob under gravity in a closed box: T drives the density and the
    # thermal-expansion term of the continuity residual
    coords, el2n = O.hex27_mesh(*n)
    gam = O.box_boundary_nodes3(coords)
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    tmat = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,))
    th = O.ThermalProblem3(coords, el2n, tmat, bc=(gam, np.zeros(len(gam))))
    th.source[:] = 50.0 * np.exp(-((x - 0.4) ** 2 + (y - 0.5) ** 2 + (z - 0.45) ** 2) / 0.05)
    th.T0[:] = 0.2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    th.T[:] = th.T0
    zero = (gam, np.zeros(len(gam)))
    mat = O.material(eta=(1.0,), etab=(1.0,), G=(INF,), alpha=(0.05,), rho0=(1.0,), K=(INF,))
    pr = O.Problem3(coords, el2n, mat, g=(0.0, 0.0, -10.0), bc=(zero, zero, zero))
    kw = dict(pressure_factor=1.0, check_interval=ncheck, tolerance=1e-9, iter_max=20000, max_iterations=40000)
    return th, pr, kw, el2n, tmat


def test_coupled3_buoyancy_is_consistent_with_both_standalone_solves():
    th, pr, kw, el2n, tmat = buoyancy3((2, 2, 2))
    dt = 0.01
    st = O.coupled_solve3(th, pr, dt, tolerance_T=1e-9, **kw)
    assert st["converged"] and st["thermal_converged"] and st["err_T"] < 1e-9
    assert st["thermal_iterations"] == st["iter"]
    rows = [26, 22, 23, 25]
    assert np.array_equal(pr.T.reshape(-1, 4), th.T[el2n[:, rows]])
    assert np.array_equal(pr.T0.reshape(-1, 4), th.T0[el2n[:, rows]])
    # the thermal step does not see the flow: it matches the standalone solve
    th2, _, _, _, _ = buoyancy3((2, 2, 2))
    th2.solve(dt, tolerance=1e-12, max_iterations=20000, check_interval=10)
    assert np.allclose(th.T, th2.T, rtol=0, atol=1e-7 * np.abs(th2.T).max())
    # the flow matches a Stokes solve with the final temperature held fixed
    _, pr2, _, _, _ = buoyancy3((2, 2, 2))
    pr2.T, pr2.T0 = pr.T.copy(), pr.T0.copy()
    st2 = pr2.solve(dt, **kw)
    assert st2["converged"]
    scale = max(np.abs(v).max() for v in pr2.v)
    assert scale > 0
    for c in range(3):
        assert np.abs(pr.v[c] - pr2.v[c]).max() < 1e-6 * scale


def test_coupled3_shear_heating_reaches_the_dissipation():
    # the 3-D form of test/test_coupled_solver.jl:59-106: simple shear on the
    # z walls of a viscous box (a perturbed start), the wall temperature held
    # at Δt Φ: the dissipation of the converged flow, Φ = η γ̇² everywhere,
    # heats the whole field to Δt Φ (the thermal residual is a pure source)
    coords, el2n = O.hex27_mesh(2, 1, 2)
    gam = O.box_boundary_nodes3(coords)
    eta, rate, dt = 2.0, 0.5, 1.0
    vx = rate * coords[:, 2]
    bc = ((gam, vx[gam]), (gam, np.zeros(len(gam))), (gam, np.zeros(len(gam))))
    pr = O.Problem3(coords, el2n, O.material(eta=(eta,), etab=(1.0,), G=(INF,), K=(INF,)), bc=bc)
    pr.v[0][:] = vx + 0.1 * np.sin(np.pi * coords[:, 0]) * np.sin(np.pi * coords[:, 2])
    phi = eta * rate ** 2
    tmat = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,))
    th = O.ThermalProblem3(coords, el2n, tmat, bc=(gam, np.full(len(gam), dt * phi)))
    st = O.coupled_solve3(th, pr, dt, tolerance_T=1e-10, shear_heating=True, tolerance=1e-10, check_interval=10,
                          max_iterations=20000, iter_max=20000, pressure_factor=1.0)
    assert st["converged"] and st["thermal_converged"]
    assert np.allclose(th.T, dt * phi, rtol=0, atol=1e-9)
    assert np.all(th.source == 0.0)


def test_coupled3_node_layout_mismatch():
    c1, e1 = O.hex27_mesh(1, 1, 1)
    c2, e2 = O.hex27_mesh(2, 1, 1)
    tmat = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(INF,))
    th = O.ThermalProblem3(c2, e2, tmat)
    pr = O.Problem3(c1, e1, O.material(eta=(1.0,)))
    with pytest.raises(RuntimeError, match="thermal mesh nodes"):
        O.coupled_solve3(th, pr, 1.0)
