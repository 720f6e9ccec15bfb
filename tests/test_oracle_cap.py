"""The oracle's Drucker-Prager tensile cap (src/stokes/assemblers/
rheology.jl:510-787, Popov, Berlie & Kaus 2025) pinned by the reference's own
tests (test/test_drucker_prager_cap.jl, test/test_popov_2d_material_point.jl):
the sixteen JustRelax fixtures, the geometry identities, the closed forms of
the return map, and finite differences of the assembled Jacobian."""
import math

import numpy as np
import pytest

import pyoracle as O

PHI, PSI = math.radians(30.0), math.radians(10.0)
K, KQ, CC, PT = math.sin(PHI), math.sin(PSI), 10.0 * math.cos(PHI), -5.0

# test_drucker_prager_cap.jl:377-397 (JustRelax 14ffc40c, C=η=G=Δt=1, K=4,
# ϕ=30°, pT=-0.5, ηve=0.5): ψ[°], η_reg, s_trial, p_trial -> τII, P, λ, θdot
JUSTRELAX = (
    (0.0, 0.0, 0.0, -1.0, 0.0, -0.5, 0.125, 0.125),
    (0.0, 0.0, 0.0, -0.8, 0.0, -0.5, 0.07500000000000002, 0.07500000000000002),
    (0.0, 0.0, 1.0, -1.0, 0.7138809973230174, -0.1988680769707268, 0.6062751909426288, 0.2002829807573183),
    (0.0, 0.0, 1.0, -0.8, 0.7393352034857215, -0.1717513436572687, 0.5444750388719386, 0.15706216408568285),
    (0.0, 0.1, 0.0, -1.0, 0.0, -0.5109358077913947, 0.12226604805215133, 0.12226604805215133),
    (0.0, 0.1, 0.0, -0.8, 0.0, -0.5065614846748367, 0.07335962883129082, 0.07335962883129082),
    (0.0, 0.1, 1.0, -1.0, 0.7478011906745448, -0.2332127198290175, 0.5395967276728152, 0.19169682004274563),
    (0.0, 0.1, 1.0, -0.8, 0.770966851846044, -0.201037213621191, 0.4819201261930389, 0.14974069659470224),
    (5.0, 0.0, 0.0, -1.0, 0.0, -0.5, 0.1245279300120768, 0.12499999999999997),
    (5.0, 0.0, 0.0, -0.8, 0.0, -0.5, 0.07471675800724613, 0.07500000000000002),
    (5.0, 0.0, 1.0, -1.0, 0.7415662149212339, -0.1692756291438945, 0.5549273268904352, 0.20768109271402638),
    (5.0, 0.0, 1.0, -0.8, 0.7660214876335598, -0.1409973422964033, 0.49423784294486456, 0.16475066442589917),
    (5.0, 0.1, 0.0, -1.0, 0.0, -0.510895408013067, 0.12181436479906836, 0.12227614799673324),
    (5.0, 0.1, 0.0, -0.8, 0.0, -0.5065372448078401, 0.07308861887944106, 0.07336568879803998),
    (5.0, 0.1, 1.0, -1.0, 0.7700437546899073, -0.2043524795101405, 0.499191884741223, 0.19891188012246488),
    (5.0, 0.1, 1.0, -0.8, 0.792336244059665, -0.171203715822286, 0.4424045245893892, 0.1571990710444285),
)


@pytest.mark.parametrize("row", range(len(JUSTRELAX)))
def test_justrelax_fixtures(row):
    psi, ereg, s_tr, p_tr, s, p, lam, thdot = JUSTRELAX[row]
    k, c = math.sin(math.pi / 6), math.cos(math.pi / 6)
    r = O.cap_return_map(s_tr, p_tr, 0.5, 4.0, k, math.sin(math.radians(psi)), c, -0.5, ereg)
    assert r["converged"]
    got = np.array([r["tII"], r["P"], r["lam"], r["lam"] * r["Ap"]])
    assert np.allclose(got, [s, p, lam, thdot], rtol=1e-12, atol=1e-14)


def test_cap_geometry_identities():
    # test_drucker_prager_cap.jl:108-137
    g = O.cap_geometry(K, KQ, CC, PT)
    assert g["a"] == pytest.approx(math.sqrt(1 + K * K))
    assert g["b"] == pytest.approx(math.sqrt(1 + KQ * KQ))
    assert g["R_y"] == pytest.approx(g["p_y"] - PT)
    assert g["R_y"] == pytest.approx((K * g["p_y"] + CC) / g["a"])
    assert g["tau_d"] == pytest.approx(K * g["p_d"] + CC)
    assert g["tau_d"] == pytest.approx(g["R_y"] / g["a"])
    assert math.hypot(g["tau_d"], g["p_d"] - g["p_y"]) == pytest.approx(g["R_y"])
    assert g["p_q"] == pytest.approx(g["p_d"] + KQ * g["tau_d"])
    assert PT < g["p_d"] < g["p_y"] and g["tau_d"] > 0


def test_yield_surface_is_the_zero_level_set():
    # test_drucker_prager_cap.jl:139-165
    g = O.cap_geometry(K, KQ, CC, PT)
    F = lambda t, p: O.cap_yield_function(t, p, K, KQ, CC, PT)   # noqa: E731
    assert abs(F(0.0, PT)) < 1e-12
    th_d = math.atan2(1 / g["a"], -K / g["a"])
    for th in np.linspace(th_d, math.pi, 9):
        P, t = g["p_y"] + g["R_y"] * math.cos(th), g["R_y"] * math.sin(th)
        assert abs(F(t, P)) < 1e-10
    for P in np.linspace(g["p_d"], g["p_d"] + 20, 9):
        assert abs(F(K * P + CC, P)) < 1e-10
    assert F(0.0, g["p_y"]) < 0 and F(0.0, PT - 1.0) > 0 and F(2 * g["tau_d"], g["p_d"] + 20) < 0
    # gradient magnitude a on both branches (finite differences)
    h = 1e-5
    for t, P in ((2 * g["tau_d"], g["p_d"] + 10.0), (0.2 * g["tau_d"], PT - 0.5)):
        dt_ = (F(t + h, P) - F(t - h, P)) / (2 * h)
        dp_ = (F(t, P + h) - F(t, P - h)) / (2 * h)
        assert math.hypot(dt_, dp_) == pytest.approx(g["a"], rel=1e-6)
    # the shear branch is the Drucker-Prager surface bit for bit
    g2 = O.cap_geometry(K, KQ, CC, -50.0)
    assert 60.0 + K * 1.0 >= g2["p_y"]
    assert O.cap_yield_function(1.0, 60.0, K, KQ, CC, -50.0) == 1.0 - CC - 60.0 * K


def test_invariants_branches_and_derivatives():
    # test_drucker_prager_cap.jl:26-101: potential branch selection, the
    # closed forms of Aτ and Ap, and Aτ = (∂Q/∂s)/2, Ap = −∂Q/∂p by differences
    g = O.cap_geometry(K, KQ, CC, PT)
    for s, p in ((1.0, -6.0), (30.0, 20.0)):
        F, At, Ap = O.cap_invariants(s, p, K, KQ, CC, PT)
        if p + KQ * s >= g["p_q"]:
            assert (At, Ap) == (0.5, KQ)
        else:
            Rq = math.hypot(s, p - g["p_q"])
            assert 2 * At == pytest.approx(g["b"] * s / Rq)
            assert Ap == pytest.approx(-g["b"] * (p - g["p_q"]) / Rq)
    assert O.cap_invariants(0.0, -6.0, K, KQ, CC, PT)[1:] == pytest.approx((0.0, g["b"]))
    assert O.cap_invariants(0.0, g["p_q"], K, KQ, CC, PT)[1:] == (0.5, KQ)
    Q = lambda s, p: g["b"] * math.hypot(s, p - g["p_q"])   # noqa: E731
    h = 1e-6
    _, At, Ap = O.cap_invariants(1.0, -6.0, K, KQ, CC, PT)
    assert 2 * At == pytest.approx((Q(1 + h, -6) - Q(1 - h, -6)) / (2 * h), rel=1e-8)
    assert Ap == pytest.approx(-(Q(1, -6 + h) - Q(1, -6 - h)) / (2 * h), rel=1e-8)


def _residual(r, s, p, kq, ereg, eve=1.0, KDt=4.0, k=math.sin(math.radians(30)), c=math.cos(math.radians(30)),
              pT=-0.5):
    F, At, Ap = O.cap_invariants(r["tII"], r["P"], k, kq, c, pT)
    return (r["tII"] - s + 2 * eve * r["lam"] * At, r["P"] - p - KDt * r["lam"] * Ap, F - ereg * r["lam"])


@pytest.mark.parametrize("psi", [0.0, 5.0])
@pytest.mark.parametrize("ereg", [0.1, 0.0])
def test_coupled_return_map_closed_forms(psi, ereg):
    # test_drucker_prager_cap.jl:282-318
    k, c, pT, eve, KDt = math.sin(math.radians(30)), math.cos(math.radians(30)), -0.5, 1.0, 4.0
    a = math.sqrt(1 + k * k)
    kq = math.sin(math.radians(psi))
    for s, p in ((0.0, -1.0), (0.0, -0.8), (1.0, -1.0), (1.0, -0.8), (3.0, 2.0)):
        r = O.cap_return_map(s, p, eve, KDt, k, kq, c, pT, ereg)
        assert r["converged"] and r["lam"] > 0
        assert max(abs(x) for x in _residual(r, s, p, kq, ereg)) <= 1e-12
        if p < 0:
            assert r["Ap"] > 0
    if kq == 0:
        r = O.cap_return_map(0.0, -1.0, eve, KDt, k, kq, c, pT, ereg)
        assert r["lam"] == pytest.approx(a * (pT + 1.0) / (ereg + a * KDt))
        assert r["tII"] == 0
    s, p = 3.0, 2.0
    lam = (s - k * p - c) / (eve + ereg + KDt * k * kq)
    r = O.cap_return_map(s, p, eve, KDt, k, kq, c, pT, ereg)
    assert r["lam"] == pytest.approx(lam)
    assert r["tII"] == pytest.approx(s - eve * lam)
    assert r["P"] == pytest.approx(p + KDt * kq * lam)


def test_return_map_elastic_switching_and_failures():
    k, c, pT, eve, KDt = math.sin(math.radians(30)), math.cos(math.radians(30)), -0.5, 1.0, 4.0
    r = O.cap_return_map(0.2, 0.0, eve, KDt, k, 0.0, c, pT, 0.1)
    assert r["converged"] and (r["tII"], r["P"], r["lam"]) == (0.2, 0.0, 0.0)
    kq = math.sin(math.radians(20))
    g = O.cap_geometry(k, kq, c, pT)
    for s, p in ((g["tau_d"] / 2 + 1, g["p_y"] - k * g["tau_d"] / 2 + 0.01),
                 (2 * g["tau_d"], g["p_y"] - 2 * k * g["tau_d"] - 0.01)):
        r = O.cap_return_map(s, p, eve, KDt, k, kq, c, pT, 0.1)
        assert r["converged"]
        assert max(abs(x) for x in _residual(r, s, p, kq, 0.1)) <= 1e-12
    for e_, K_ in ((0.0, KDt), (eve, 0.0), (eve, math.inf)):
        assert not O.cap_return_map(0.0, -1.0, e_, K_, k, 0.0, c, pT, 0.1)["converged"]
    assert not O.cap_return_map(0.0, -1.0, eve, KDt, k, 0.0, c, pT, 0.1, maxiter=0)["converged"]
    failed = O.cap_local_update((0.0, 0.0, 0.0), -1.0, 0.0, KDt, k, 0.0, c, pT, 0.1)
    assert np.all(np.isnan(failed["tau"])) and math.isnan(failed["P"]) and math.isnan(failed["gdot"])


def test_local_update_is_radial_and_rotation_invariant():
    # test_drucker_prager_cap.jl:331-351
    k, c, pT, eve, KDt = math.sin(math.radians(30)), math.cos(math.radians(30)), -0.5, 1.0, 4.0
    tau = (1.2, -0.3, 0.4)

    def s2(t):
        tzz = -t[0] - t[1]
        return math.sqrt((t[0] ** 2 + t[1] ** 2 + tzz ** 2) / 2 + t[2] ** 2)

    out = O.cap_local_update(tau, -0.6, eve, KDt, k, 0.1, c, pT, 0.1)
    r = O.cap_return_map(s2(tau), -0.6, eve, KDt, k, 0.1, c, pT, 0.1)
    assert s2(out["tau"]) == pytest.approx(r["tII"])
    assert out["P"] == pytest.approx(r["P"])
    assert out["gdot"] == pytest.approx(r["lam"] * r["At"])
    assert out["thdot"] == pytest.approx(r["lam"] * r["Ap"])

    def rot(t, th):
        R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        M = R @ np.array([[t[0], t[2]], [t[2], t[1]]]) @ R.T
        return (M[0, 0], M[1, 1], M[0, 1])

    o2 = O.cap_local_update(rot(tau, 0.7), -0.6, eve, KDt, k, 0.1, c, pT, 0.1)
    assert np.allclose(o2["tau"], rot(out["tau"], 0.7))
    assert o2["P"] == pytest.approx(out["P"])
    z = O.cap_local_update((0.0, 0.0, 0.0), -1.0, eve, KDt, k, 0.0, c, pT, 0.1)
    assert np.array_equal(z["tau"], [0.0, 0.0, 0.0])
    assert z["P"] == pytest.approx(O.cap_return_map(0.0, -1.0, eve, KDt, k, 0.0, c, pT, 0.1)["P"])


def test_popov_material_point():
    # test_popov_2d_material_point.jl:8-33
    k, kq, c, pT = math.sin(math.radians(30)), 0.0, 20.0 * math.cos(math.radians(30)), -1.0
    assert abs(O.cap_yield_function(0.0, pT, k, kq, c, pT)) < 1e-12
    el = O.cap_return_map(0.0, -0.5, 1.0, 64.0, k, kq, c, pT, 0.1)
    assert el["converged"] and el["lam"] == 0.0 and el["P"] == -0.5
    y = O.cap_return_map(0.0, -1.2, 1.0, 64.0, k, kq, c, pT, 0.1)
    assert y["converged"] and y["lam"] > 0 and -1.2 < y["P"] < pT


def _cap_problem(n=3, seed=0, history=False, H_C=(0.0, 0.0)):
    rng = np.random.default_rng(seed)
    coords, el2n = O.t7_mesh(n, n)
    ne = len(el2n)
    ph = rng.integers(1, 3, ne).astype(np.int32)
    cap = O.drucker_prager_cap((math.radians(30), math.radians(25)), (math.radians(5), 0.0), (1.0, 1.5),
                               (-0.5, -0.3), (0.1, 0.05), (4.0, 6.0), C_min=(0.6, 1.0), H_C=H_C)
    mat = O.material(eta=(1.0, 3.0), etab=(1.0, 1.0), G=(1.0, 2.0), K=(4.0, 6.0))
    tau_old = tuple(0.3 * rng.normal(size=(ne, 7)) for _ in range(3))
    pr = O.Problem(coords, el2n, mat, phases_v=ph, phases_P=ph, tau_old=tau_old, cap=cap, history=history)
    pr.vx[:] = 0.6 * rng.normal(size=len(coords))
    pr.vy[:] = 0.6 * rng.normal(size=len(coords))
    pr.P[:] = rng.normal(size=3 * ne) - 0.3
    pr.pressure_scaling(1.0, 0.5, finite_K=True)
    return pr, ph


def test_cap_residual_uses_the_corrected_pressure():
    # the cap corrects the pressure in the momentum balance: the residual
    # differs from the visco-elastic one, P_store holds P_trial + KΔt θdot
    pr, ph = _cap_problem()
    pr.update_current_stress(0.5)
    Pstore = pr.P_store.copy()
    NqP = O.t7_tables()[1]
    Ptr = pr.P.reshape(-1, 3) @ NqP.T
    assert not np.allclose(Pstore, Ptr)
    assert np.isfinite(Pstore).all()


def test_cap_jacobian_matches_finite_differences():
    # the augmented Jacobian diagnostics carry the implicit derivative of the
    # return (stress and corrected pressure): |J| row sums and |J_ii| from
    # central differences of the assembled residual with Pnum(v) inline
    pr, ph = _cap_problem(n=2, seed=3, history=True, H_C=(-0.5, -0.2))
    pr.gamma_hist[:] = 0.3
    dt = 0.5
    pr.augmented_jacobian(dt, finite_K=True)
    Jrow = (pr.dRdvx.copy(), pr.dRdvy.copy())
    Jdiag = (pr.PCx.copy(), pr.PCy.copy())
    nn = pr.nnodes
    v0 = (pr.vx.copy(), pr.vy.copy())

    def residual():
        pr.pressure_residual(dt, finite_K=True)
        Pn = pr.gammaP * pr.RP / pr.MP
        pr.momentum_residual(dt, Pnum=Pn)
        return np.concatenate([pr.Rvx.copy(), pr.Rvy.copy()])

    h = 1e-6
    J = np.zeros((2 * nn, 2 * nn))
    for col in range(2 * nn):
        comp, node = divmod(col, nn)
        for sgn in (1, -1):
            pr.vx[:], pr.vy[:] = v0[0], v0[1]
            (pr.vx if comp == 0 else pr.vy)[node] += sgn * h
            J[:, col] += sgn * residual() / (2 * h)
    pr.vx[:], pr.vy[:] = v0[0], v0[1]
    # per-element row sums cannot be read from the assembled matrix; the
    # diagonal can, and the row sums bound the assembled |J| row sums
    diag = np.abs(np.diag(J))
    scale = np.abs(J).max()
    # the oracle assembles element |J_ii| sums; for nodes of one element
    # (bubbles) they equal the assembled diagonal
    bub = pr.el2n[:, 6]
    assert np.allclose(Jdiag[0][bub], diag[bub], rtol=1e-5, atol=1e-7 * scale)
    assert np.allclose(Jdiag[1][bub], diag[nn + bub], rtol=1e-5, atol=1e-7 * scale)
    row = np.abs(J).sum(axis=1)
    assert np.all(Jrow[0] + 1e-6 * scale >= row[:nn])
    assert np.allclose(Jrow[0][bub], row[bub], rtol=1e-5, atol=1e-7 * scale)
    assert np.allclose(Jrow[1][bub], row[nn + bub], rtol=1e-5, atol=1e-7 * scale)


def test_commit_plastic_history():
    # DR.jl:899-958: γ += Δt γdot, θ += Δt θdot with the rates of the local
    # update at the current state, softened by the accepted history. Without
    # softening the rates do not depend on γ: a second commit adds the same
    # increment; with H_C < 0 the weaker cohesion yields faster.
    dt = 0.5
    pr, _ = _cap_problem(history=True)
    pr.commit_plastic_history(dt)
    g1, t1 = pr.gamma_hist.copy(), pr.theta_hist.copy()
    assert (g1 > 0).any() and (g1 >= 0).all() and (g1 == 0).any()
    pr.commit_plastic_history(dt)
    assert np.allclose(pr.gamma_hist, 2 * g1, rtol=1e-14, atol=0)
    assert np.allclose(pr.theta_hist, 2 * t1, rtol=1e-14, atol=0)
    ps, _ = _cap_problem(history=True, H_C=(-0.5, -0.2))
    ps.commit_plastic_history(dt)
    s1 = ps.gamma_hist.copy()
    assert np.array_equal(s1, g1)   # γ = 0 before the first commit: no softening yet
    ps.commit_plastic_history(dt)
    inc = ps.gamma_hist - s1
    y = s1 > 0
    assert np.all(inc[y] >= s1[y] * (1 - 1e-12)) and np.any(inc[y] > s1[y] * (1 + 1e-6))
