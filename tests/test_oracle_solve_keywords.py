"""The oracle's restatement of solve!'s array keywords (DR.jl:183-226,
302-418, 451-457, 556-561, 722-727): a viscosity sampled at the quadrature
points, assembled nodal loads, the previous-step pressure at the points, the
pressure source Qq/Q2D and collect_history. Pinned by the reference's tests
(test_stokes_quadrature_viscosity.jl:4-113, test_stokes.jl:295-340, 506-580,
test_solver_convergence_api.jl:300-328) and by closed forms: the hydrostatic
pressure P = y + C, a source that relaxes P to K, ∫N_i dΩ = area/3, and the
tensile cap's material-point recurrence."""
import math

import numpy as np
import pytest

import pyoracle as O


def _box(n):
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    return coords, el2n, gam


def _quadrature_problem(**kw):
    # test_stokes_quadrature_viscosity.jl:5-16: η = 2, ηb = G = Inf, α = ρ0 = 0, v = (x², xy)
    coords, el2n, gam = _box(2)
    pr = O.Problem(coords, el2n, O.material(eta=(2.0,), etab=(math.inf,), G=(math.inf,), rho0=(0.0,)), **kw)
    pr.vx[:] = coords[:, 0] ** 2
    pr.vy[:] = coords[:, 0] * coords[:, 1]
    return pr


def _momentum(pr, **kw):
    for k, v in kw.items():
        setattr(pr, k, v)
    pr.momentum_residual(1.0)
    return pr.Rvx.copy(), pr.Rvy.copy()


def test_quadrature_viscosity_residual_and_scale():
    # test_stokes_quadrature_viscosity.jl:17-36
    pr = _quadrature_problem()
    ne = pr.nels
    constant = np.full((ne, 7), 2.0)
    ref = _momentum(pr, eta_ip=None)
    got = _momentum(pr, eta_ip=constant)
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14 * np.abs(b).max())
    # a spatially varying field is not replaced by one element value
    varying = constant.copy()
    varying[:, 0] = 4.0
    vx = _momentum(pr, eta_ip=varying)[0]
    assert not np.allclose(vx, ref[0])
    # the pressure scale of a sampled viscosity: γ = factor·η·(gphy = gnum)/(2 gnum) = 10
    pr.eta_ip = None
    pr.eta_scale_ip = constant
    pr.pressure_scaling(10.0, 1.0)
    assert np.allclose(pr.gammaP, 10.0, rtol=1e-14)
    pr.eta_scale_ip = varying
    pr.pressure_scaling(10.0, 1.0)
    assert pr.gammaP.max() > 10.0


def _gravity_load(coords, el2n):
    """load = −R of the gravitational assembly (g = (0, 1), ρ0 = 1) at v = P = 0
    (test_stokes_quadrature_viscosity.jl:38-50)."""
    g = O.Problem(coords, el2n, O.material(eta=(2.0,), etab=(math.inf,), G=(math.inf,), rho0=(1.0,), g=(0.0, 1.0)))
    g.momentum_residual(1.0)
    return -g.Rvx.copy(), -g.Rvy.copy()


def test_body_force_gives_the_hydrostatic_pressure():
    # test_stokes_quadrature_viscosity.jl:51-60: P − mean = y − 1/2, v = 0
    coords, el2n, gam = _box(2)
    load = _gravity_load(coords, el2n)
    assert np.abs(load[0]).max() < 1e-14 * np.abs(load[1]).max()
    assert load[1].sum() == pytest.approx(1.0, rel=1e-13)   # ∫ρ g dΩ over the unit square
    zero = np.zeros(len(gam))
    pr = O.Problem(coords, el2n, O.material(eta=(2.0,), etab=(math.inf,), G=(math.inf,), rho0=(0.0,)),
                   bc=((gam, zero), (gam, zero)), eta_ip=np.full((len(el2n), 7), 2.0), body_force=load)
    st = pr.solve(1.0, pressure_factor=50.0, check_interval=10, tolerance=1e-11)
    assert st["converged"]
    Pc = coords[el2n[:, :3]].reshape(-1, 2)
    P = pr.P - (pr.P * pr.MP).sum() / pr.MP.sum()
    assert np.allclose(P, Pc[:, 1] - 0.5, atol=1e-7)
    assert np.abs(pr.vx).max() < 1e-7 and np.abs(pr.vy).max() < 1e-7


def test_body_force_is_the_gravity_load():
    # the load solve and the gravitational solve reach the same state
    coords, el2n, gam = _box(4)
    load = _gravity_load(coords, el2n)
    zero = np.zeros(len(gam))
    mats = dict(eta=(2.0,), etab=(math.inf,), G=(math.inf,))
    a = O.Problem(coords, el2n, O.material(rho0=(0.0,), **mats), bc=((gam, zero), (gam, zero)), body_force=load)
    b = O.Problem(coords, el2n, O.material(rho0=(1.0,), g=(0.0, 1.0), **mats), bc=((gam, zero), (gam, zero)))
    sa = a.solve(1.0, check_interval=10, tolerance=1e-10)
    sb = b.solve(1.0, check_interval=10, tolerance=1e-10)
    assert sa["converged"] and sb["converged"]
    assert abs(sa["iter"] - sb["iter"]) <= 10
    assert np.allclose(a.P, b.P, atol=1e-7)


def test_point_pressure_history_enters_the_continuity_rate():
    # test_stokes.jl:295-312: with P = 3, K = 2, Δt = 1 the rate is −(P − P_old)/(KΔt)
    coords, el2n, gam = _box(2)
    mat = O.material(eta=(1.0,), etab=(2.0,), G=(math.inf,), K=(2.0,))
    pr = O.Problem(coords, el2n, mat)
    pr.P[:] = 3.0
    pr.P0[:] = 0.0
    pr.pressure_residual(1.0, finite_K=True)
    nodal = pr.RP.copy()
    area = 1.0 / len(el2n)
    assert np.allclose(nodal, -1.5 * area / 3, rtol=1e-13)
    pr.Pold_ip = np.ones((len(el2n), 7))
    pr.pressure_residual(1.0, finite_K=True)
    assert np.allclose(pr.RP, -1.0 * area / 3, rtol=1e-13)
    # a point field sampled from a P1 field is the nodal field
    rng = np.random.default_rng(0)
    P0 = rng.normal(size=3 * len(el2n))
    pr.Pold_ip = P0.reshape(-1, 3) @ O.t7_tables()[1].T
    pr.P[:] = rng.normal(size=3 * len(el2n))
    pr.pressure_residual(1.0, finite_K=True)
    pts = pr.RP.copy()
    pr.Pold_ip = None
    pr.P0[:] = P0
    pr.pressure_residual(1.0, finite_K=True)
    assert np.allclose(pts, pr.RP, rtol=1e-12, atol=1e-15)


def test_pressure_source_normalises_the_area_rate():
    # test_stokes.jl:315-340
    coords, el2n, gam = _box(1)
    pr = O.Problem(coords, el2n, O.material())
    Qq = np.linspace(0.5, 1.5, pr.nnodesP)
    pr.set_pressure_source(Qq, 2.0)
    NqP = O.t7_tables()[1]
    integrated = ((pr.Q.reshape(-1, 3) @ NqP.T) * pr.dOmegaP).sum()
    assert integrated == pytest.approx(2.0, rel=1e-14)
    assert np.allclose(pr.Q / Qq, pr.Q[0] / Qq[0])
    pr.set_pressure_source(np.zeros(pr.nnodesP), 0.0)
    assert not pr.Q.any()
    with pytest.raises(ValueError, match="nonempty positive Qq support"):
        pr.set_pressure_source(np.zeros(pr.nnodesP), 1.0)
    with pytest.raises(ValueError, match="finite and nonnegative"):
        pr.set_pressure_source(-Qq, 1.0)
    with pytest.raises(ValueError, match="Qq must be finite"):
        pr.set_pressure_source(np.full(pr.nnodesP, np.nan))
    pr.set_pressure_source(Qq)
    assert np.array_equal(pr.Q, Qq)


def _source_problem():
    # test_solver_convergence_api.jl:300-328: K = 10, γP = 10, Qq = 1, Q2D = 1
    coords, el2n, gam = _box(4)
    K = 10.0
    zero = np.zeros(len(gam))
    pr = O.Problem(coords, el2n, O.material(eta=(1.0,), etab=(math.inf,), rho0=(1.0,), K=(K,)),
                   bc=((gam, zero), (gam, zero)))
    pr.pressure_scaling(1.0, 1.0, finite_K=True)
    pr.gammaP[:] = 10.0
    pr.set_pressure_source(np.ones(pr.nnodesP), 1.0)
    return pr, K


def test_pressure_source_relaxes_the_pressure_to_K():
    pr, K = _source_problem()
    st = pr.solve(1.0, finite_K=True, skip_scaling=True, check_interval=1, iter_max=1, max_iterations=1,
                  max_ph_iterations=2)
    assert st["converged"]
    assert np.allclose(pr.Q, 1.0, rtol=1e-13)
    assert np.allclose(pr.P, K, rtol=1e-12)


def test_collect_history_records_every_check():
    # test_stokes_quadrature_viscosity.jl:100-113 (history layout)
    coords, el2n, gam = _box(2)
    zero = np.zeros(len(gam))
    pr = O.Problem(coords, el2n, O.material(eta=(2.0,), etab=(math.inf,), rho0=(1.0,), g=(0.0, 1.0)),
                   bc=((gam, zero), (gam, zero)))
    st = pr.solve(1.0, check_interval=10, tolerance=1e-11, collect_history=True)
    h = st["history"]
    assert st["converged"] and len(h) > 2
    assert h[0, 0] == 0 and h[-1, 0] == st["iter"]
    assert h[-1, 5] == pytest.approx(st["err_P"])
    assert np.allclose(np.maximum(h[:, 2], h[:, 3]), h[:, 1], rtol=1e-14)
    assert np.all(np.diff(h[:, 0]) >= 0)


def test_cap_carries_the_corrected_pressure_across_steps():
    # test_stokes.jl:506-580: uniaxial extension imposed on every boundary node,
    # P_old = the accepted corrected point pressure; the field follows the
    # material-point recurrence of the cap's local update
    coords, el2n, gam = _box(2)
    ne = len(el2n)
    dt, K, edot = 1.0, 4.0, 0.1
    phi, psi, C_, pT, ereg, C_min, H_C = math.radians(30), math.radians(5), 1.0, -0.5, 0.1, 0.5, -1.0
    cap = O.drucker_prager_cap((phi,), (psi,), (C_,), (pT,), (ereg,), (K,), C_min=(C_min,), H_C=(H_C,))
    mat = O.material(eta=(1.0,), etab=(K,), G=(1.0,), K=(K,))
    zeros = tuple(np.zeros((ne, 7)) for _ in range(3))
    bc = ((gam, edot * coords[gam, 0]), (gam, np.zeros(len(gam))))
    pr = O.Problem(coords, el2n, mat, tau_old=zeros, cap=cap, history=True, bc=bc)
    pr.vx[:] = edot * coords[:, 0]
    pr.Pold_ip = np.zeros((ne, 7))
    tau_mp, P_mp, g_mp, th_mp = np.zeros(3), 0.0, 0.0, 0.0
    eve = 1.0 / (1.0 / 1.0 + 1.0 / dt)
    for step in range(3):
        st = pr.solve(dt, pressure_factor=20.0, tolerance=1e-10, finite_K=True)
        assert st["converged"]
        pr.update_current_stress(dt)
        pr.commit_plastic_history(dt)
        P_corr = pr.P_store.copy()
        pr.Pold_ip = P_corr
        pr.tau_old = tuple(t.copy() for t in pr.tau_store)
        # the material point: trial stress from ε̇' and τ_old, trial P = P − KΔt ε̇
        P_tr = P_mp - K * dt * edot
        dev = np.array([edot - edot / 3, -edot / 3, 0.0])
        tau_tr = 2 * eve * (dev + tau_mp / (2 * 1.0 * dt))
        c = math.cos(phi) * min(max(C_ + H_C * g_mp, C_min), C_)
        mp = O.cap_local_update(tau_tr, P_tr, eve, K * dt, math.sin(phi), math.sin(psi), c, pT, ereg)
        tau_mp, P_mp = mp["tau"], mp["P"]
        g_mp += dt * mp["gdot"]
        th_mp += (P_mp - P_tr) / K
        assert np.allclose(P_corr, P_mp, atol=1e-8)
        assert np.allclose(pr.tau_store[0], tau_mp[0], atol=1e-8)
        assert np.allclose(pr.gamma_hist, g_mp, atol=1e-8)
        assert np.allclose(pr.theta_hist, th_mp, atol=1e-8)
    # elastic first step, then two tensile steps: the corrected pressure stays
    # at the cap, the trial pressure does not
    assert th_mp > 0
    assert -0.51 < P_mp < -0.5
    assert np.all(pr.P < -0.9)
