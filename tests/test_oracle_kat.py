"""Pins the CPU oracle against the reference's own known-answer tests and
against closed forms the mathematics fixes. Each test names the reference
test or formula it reproduces (paths relative to the FEMTools.jl root)."""
import math

import numpy as np
import pytest

import pyoracle as O

EPS = np.finfo(float).eps


# --------------------------------------------------------------------------
# reference elements
# --------------------------------------------------------------------------
T7_NODES = [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5), (1 / 3, 1 / 3)]


def test_t7_kronecker_partition_of_unity():
    # shape_functions.jl:92-113 docstring: Kronecker delta at the 7 nodes,
    # partition of unity and zero gradient sum everywhere.
    for j, (x, y) in enumerate(T7_NODES):
        N, _ = O.t7_basis(x, y)
        expect = np.zeros(7)
        expect[j] = 1
        np.testing.assert_allclose(N, expect, atol=1e-14)
    rng = np.random.default_rng(0)
    for _ in range(20):
        x, y = rng.random(2) * 0.5
        N, dN = O.t7_basis(x, y)
        assert abs(N.sum() - 1) < 1e-14
        np.testing.assert_allclose(dN.sum(axis=0), 0, atol=1e-13)


def test_t7_gradients_match_finite_differences():
    h = 1e-6
    for x, y in [(0.2, 0.3), (0.05, 0.7), (1 / 3, 1 / 3)]:
        _, dN = O.t7_basis(x, y)
        fx = (O.t7_basis(x + h, y)[0] - O.t7_basis(x - h, y)[0]) / (2 * h)
        fy = (O.t7_basis(x, y + h)[0] - O.t7_basis(x, y - h)[0]) / (2 * h)
        np.testing.assert_allclose(dN[:, 0], fx, atol=1e-8)
        np.testing.assert_allclose(dN[:, 1], fy, atol=1e-8)


def test_t7_basis_spans_cubic_bubble_and_quadratics():
    # Interpolating a polynomial in the T7 space (P2 + bubble) is exact.
    f = lambda x, y: 1 + 2 * x - 3 * y + 0.5 * x * x + x * y - 2 * y * y + 5 * x * y * (1 - x - y)
    nodal = np.array([f(x, y) for x, y in T7_NODES])
    for x, y in [(0.1, 0.2), (0.6, 0.1), (0.25, 0.25)]:
        N, _ = O.t7_basis(x, y)
        assert abs(N @ nodal - f(x, y)) < 1e-13


def test_dunavant_rule_degree5():
    # integration_points.jl:188-202; weights sum to 1/2 and degree <= 5
    # monomials integrate to a!b!/(a+b+2)! (constants carry 15 digits).
    xi, eta, w = O.t7_rule()
    assert abs(w.sum() - 0.5) < 1e-14
    for a in range(6):
        for b in range(6 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * xi**a * eta**b) - exact) < 2e-14
    # degree 6 is NOT exact (pins that this is the 7-point degree-5 rule)
    assert abs(np.sum(w * xi**6) - math.factorial(6) / math.factorial(8)) > 1e-7


# --------------------------------------------------------------------------
# mesh, geometry, node->element map
# --------------------------------------------------------------------------
def test_t7_mesh_counts_and_single_cell_connectivity():
    coords, el2n = O.t7_mesh(1, 1)
    # generate_element2node(T7) connectivity.jl:111-140 for ex=ey=0 (even):
    # (BL,BR,TR,bot,rgt,ctr,cA), (BL,TR,TL,ctr,top,lft,cB), 1-based 1..9 + 10,11
    expect = np.array([[1, 3, 9, 2, 6, 5, 10], [1, 9, 7, 5, 8, 4, 11]]) - 1
    np.testing.assert_array_equal(el2n, expect)
    np.testing.assert_allclose(coords[9], [2 / 3, 1 / 3])
    np.testing.assert_allclose(coords[10], [1 / 3, 2 / 3])
    for n in (2, 3, 7):
        c, e = O.t7_mesh(n, n + 1)
        assert len(e) == 2 * n * (n + 1)
        assert len(c) == (2 * n + 1) * (2 * n + 3) + 2 * n * (n + 1)


def test_t7_mesh_16_matches_reference_dof_count():
    # API_SIMPLIFICATION_BASELINE.md:124 -- 16x16 T7/P1 has 4738 DoFs.
    c, e = O.t7_mesh(16, 16)
    assert 2 * len(c) + 3 * len(e) == 4738


def test_geometry_affine_closed_form():
    coords, el2n = O.t7_mesh(4, 3, (0.0, 2.0, -1.0, 0.5))
    invJ, dO, dOP = O.t7_geometry(coords, el2n)
    # sum |det J| w = area ; velocity and pressure weights agree (affine)
    assert abs(dO.sum() - 2.0 * 1.5) < 1e-13
    np.testing.assert_allclose(dO, dOP, rtol=1e-13)
    # closed-form J from the three corners
    for e in (0, 5, len(el2n) - 1):
        p = coords[el2n[e, :3]]
        J = np.array([[p[1, 0] - p[0, 0], p[2, 0] - p[0, 0]], [p[1, 1] - p[0, 1], p[2, 1] - p[0, 1]]])
        Ji = np.linalg.inv(J)
        for q in range(7):
            np.testing.assert_allclose(invJ[e, q].reshape(2, 2), Ji, rtol=1e-12, atol=1e-12)


def test_node_valence_structured_t7():
    # Valence: vertices 4 or 8 (alternating diagonals, connectivity.jl:129-135)
    # in the interior, mid-edge nodes 2, bubbles 1.
    coords, el2n = O.t7_mesh(4, 4)
    val = np.bincount(el2n.ravel(), minlength=len(coords))
    bubbles = el2n[:, 6]
    assert np.all(val[bubbles] == 1)
    assert set(np.unique(val)) <= {1, 2, 3, 4, 6, 8}
    interior_vertices = [iy * 9 + ix for iy in range(2, 7, 2) for ix in range(2, 7, 2)]
    assert set(val[interior_vertices]) == {4, 8}


# --------------------------------------------------------------------------
# element-level KATs (test/test_stokes.jl)
# --------------------------------------------------------------------------
def _pressure_el(dNdx, dO, Nq, vx, vy, P, P0, T, T0, alpha, bulk, dt, Q=None, phase=None):
    NV = dNdx.shape[1]
    NP = Nq.shape[1]
    nq = dNdx.shape[0]
    out = np.zeros(NP)
    ph = np.ones(NP, np.int32) if phase is None else np.asarray(phase, np.int32)
    L = O.lib()
    L.orc_pressure_element(NV, NP, nq, O.ptr(O.f64(dNdx)), O.ptr(O.f64(dO)), O.ptr(O.f64(dO)), O.ptr(O.f64(Nq)),
                           O.ptr(O.f64(vx)), O.ptr(O.f64(vy)), O.ptr(O.f64(P)), O.ptr(O.f64(P0)), O.ptr(O.f64(T)),
                           O.ptr(O.f64(T0)), O.ptr(O.f64(Q)) if Q is not None else O.ptr(None), O.ptr(ph, O.PI),
                           O.ptr(O.f64(alpha)), O.ptr(O.f64(bulk)), dt, O.ptr(out))
    return out


def test_pressure_rate_kat():
    # test/test_stokes.jl:723-744 "pressure rates interpolate nodal increments"
    Nv = np.array([[0.2, 0.3, 0.5]])
    P, P0 = np.array([3.0, 5, 8]), np.array([1.0, 2, 3])
    T, T0 = np.array([7.0, 4, 2]), np.array([2.0, 1, 1])
    dt, etab, alpha, dO = 2.0, 4.0, 0.25, 0.5
    r = _pressure_el(np.zeros((1, 3, 2)), [dO], Nv, np.zeros(3), np.zeros(3), P, P0, T, T0, [alpha], [etab], dt)
    rate = -np.sum(Nv[0] * (P - P0)) / (etab * dt) + alpha * np.sum(Nv[0] * (T - T0)) / dt
    np.testing.assert_allclose(r, Nv[0] * rate * dO, rtol=1e-14)
    # finite K replaces etab (test_stokes.jl:746-767): bulk = K = 8
    r = _pressure_el(np.zeros((1, 3, 2)), [0.5], Nv, np.zeros(3), np.zeros(3), P, P0, np.zeros(3), np.zeros(3),
                     [0.0], [8.0], 2.0)
    np.testing.assert_allclose(r, Nv[0] * (-np.dot(Nv[0], P - P0) / (8.0 * 2.0) * 0.5), rtol=1e-14)


def test_pressure_divergence_and_source_kat():
    # 2-D counterpart of test/test_stokes.jl:274-295: one node, grad N=(0.2,0.3),
    # v=(1,2), dOmega=2: RP = -(0.2+0.6)*2 = -1.6; with source Q=3: -1.6+6 = 4.4
    dNdx = np.array([[[0.2, 0.3]]])
    r = _pressure_el(dNdx, [2.0], np.array([[1.0]]), [1.0], [2.0], [0.0], [0.0], [0.0], [0.0], [0.0], [math.inf], 1.0)
    np.testing.assert_allclose(r, [-1.6], rtol=1e-14)
    r = _pressure_el(dNdx, [2.0], np.array([[1.0]]), [1.0], [2.0], [0.0], [0.0], [0.0], [0.0], [0.0], [math.inf], 1.0,
                     Q=[3.0])
    np.testing.assert_allclose(r, [4.4], rtol=1e-14)
    # P_old / finite-K storage term: P=3, bulk K=2, dt=1: -3/2*2 = -3  (test_stokes.jl:297-313)
    r = _pressure_el(np.zeros((1, 1, 2)), [2.0], np.array([[1.0]]), [0.0], [0.0], [3.0], [0.0], [0.0], [0.0], [0.0],
                     [2.0], 1.0)
    np.testing.assert_allclose(r, [-3.0], rtol=1e-14)


def _momentum_el(dNdx, dO, Nq, NqP, vx, vy, P, T, mat, Pnum=None, phase=None, dp=None, dt=1.0, tau_old=None):
    NV = dNdx.shape[1]
    NP = NqP.shape[1]
    nq = dNdx.shape[0]
    Rx, Ry = np.zeros(NV), np.zeros(NV)
    ph = np.ones(NV, np.int32) if phase is None else np.asarray(phase, np.int32)
    L = O.lib()
    import ctypes as C
    L.orc_momentum_element(NV, NP, nq, O.ptr(O.f64(dNdx)), O.ptr(O.f64(dO)), O.ptr(O.f64(Nq)), O.ptr(O.f64(NqP)),
                           O.ptr(O.f64(vx)), O.ptr(O.f64(vy)), O.ptr(O.f64(P)),
                           O.ptr(O.f64(Pnum)) if Pnum is not None else O.ptr(None), O.ptr(O.f64(T)),
                           O.ptr(ph, O.PI), C.byref(mat), O.ptr(None),
                           C.byref(dp) if dp is not None else C.cast(None, C.POINTER(O.DP)), dt,
                           O.ptr(O.f64(tau_old)) if tau_old is not None else O.ptr(None),
                           O.ptr(Rx), O.ptr(Ry), O.ptr(None))
    return Rx, Ry


third = np.full((1, 3), 1 / 3)


@pytest.mark.parametrize("kw,expect_x,expect_y", [
    # test/test_stokes.jl:777-867 body-force KATs (1-point linear triangle, dOmega = 0.5)
    (dict(rho0=(3.0,), g=(0.0, -10.0)), 0.0, 5.0),
    (dict(rho0=(2.0,), g=(5.0, 0.0)), -5 / 3, 0.0),
    (dict(rho0=(4.0,), alpha=(0.1,), g=(0.0, -10.0), T=1.0), 0.0, 6.0),
    (dict(rho0=(2.0,), K=(4.0,), g=(0.0, -10.0), P=2.0), 0.0, 5.0),
    (dict(rho0=(1.0,), g=(0.0, 0.0)), 0.0, 0.0),
])
def test_momentum_body_force_kats(kw, expect_x, expect_y):
    T = kw.pop("T", 0.0)
    P = kw.pop("P", 0.0)
    mat = O.material(eta=(1.0,), G=(math.inf,), **kw)
    Rx, Ry = _momentum_el(np.zeros((1, 3, 2)), [0.5], third, third, np.zeros(3), np.zeros(3), np.full(3, P),
                          np.full(3, T), mat)
    np.testing.assert_allclose(Rx, expect_x, atol=1e-14)
    np.testing.assert_allclose(Ry, expect_y, atol=1e-14)


def test_momentum_two_phase_body_force_kat():
    # test_stokes.jl:854-867: phase_loc=[1,1,2], rho0=(1,3), g=(0,-6) -> Ry = 5/3
    mat = O.material(eta=(1.0, 1.0), rho0=(1.0, 3.0), g=(0.0, -6.0))
    Rx, Ry = _momentum_el(np.zeros((1, 3, 2)), [0.5], third, third, np.zeros(3), np.zeros(3), np.zeros(3),
                          np.zeros(3), mat, phase=[1, 1, 2])
    np.testing.assert_allclose(Ry, 5 / 3, rtol=1e-14)
    np.testing.assert_allclose(Rx, 0, atol=1e-14)


def test_momentum_pressure_term_closed_form():
    # Constant pressure P on a reference-gradient triangle: R_x = -dN/dx*P*dOmega
    dNdx = np.array([[[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]]])
    mat = O.material(eta=(1.0,))
    Rx, Ry = _momentum_el(dNdx, [1.0], third, third, np.zeros(3), np.zeros(3), np.ones(3), np.zeros(3), mat)
    np.testing.assert_allclose(Rx, [1.0, -1.0, 0.0], atol=1e-14)
    np.testing.assert_allclose(Ry, [1.0, 0.0, -1.0], atol=1e-14)
    # Pnum adds to the pressure (test_stokes.jl:869-893): Pnum = 0.5 -> x1.5
    Rx2, _ = _momentum_el(dNdx, [1.0], third, third, np.zeros(3), np.zeros(3), np.ones(3), np.zeros(3), mat,
                          Pnum=np.full(3, 0.5))
    np.testing.assert_allclose(Rx2, 1.5 * Rx, atol=1e-14)


def test_momentum_viscous_closed_form():
    # Linear velocity on a reference-gradient triangle: uniform strain rate.
    dNdx = np.array([[[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]]])
    vx = np.array([0.2, -0.1, 0.4])
    vy = np.array([-0.3, 0.5, 0.1])
    eta = 2.0
    mat = O.material(eta=(eta,))
    Rx, Ry = _momentum_el(dNdx, [1.0], third, third, vx, vy, np.zeros(3), np.zeros(3), mat)
    gx = dNdx[0].T @ vx
    gy = dNdx[0].T @ vy
    exx, eyy, exy = gx[0], gy[1], 0.5 * (gx[1] + gy[0])
    tr = (exx + eyy) / 3
    txx, tyy, txy = 2 * eta * (exx - tr), 2 * eta * (eyy - tr), 2 * eta * exy
    np.testing.assert_allclose(Rx, dNdx[0, :, 0] * txx + dNdx[0, :, 1] * txy, rtol=1e-13)
    np.testing.assert_allclose(Ry, dNdx[0, :, 0] * txy + dNdx[0, :, 1] * tyy, rtol=1e-13)


def test_momentum_elastic_history_closed_form():
    # finite G with stress history, zero velocity: tau = 2 etave tau_old/(2 G dt)
    dNdx = np.array([[[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]]])
    eta, G, dt = 2.0, 4.0, 0.25
    to = np.array([[0.3], [-0.1], [0.2]])
    mat = O.material(eta=(eta,), G=(G,))
    Rx, Ry = _momentum_el(dNdx, [1.0], third, third, np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), mat,
                          dt=dt, tau_old=to)
    etave = 1 / (1 / eta + 1 / (G * dt))
    t = 2 * etave * to[:, 0] / (2 * G * dt)
    np.testing.assert_allclose(Rx, dNdx[0, :, 0] * t[0] + dNdx[0, :, 1] * t[2], rtol=1e-13)
    np.testing.assert_allclose(Ry, dNdx[0, :, 0] * t[2] + dNdx[0, :, 1] * t[1], rtol=1e-13)


def _dp_stress(dp, eta=1.0, G=math.inf, dt=1.0, Pq=0.0, vx0=2.0):
    mat = O.material(eta=(eta,), G=(G,))
    gvx = np.array([vx0, 0.0])
    gvy = np.zeros(2)
    tau = np.zeros(3)
    import ctypes as C
    lam = O.lib().orc_stress_at_ip(O.ptr(gvx), O.ptr(gvy), O.ptr(np.array([1.0, 0, 0])), 3,
                                   O.ptr(np.ones(3, np.int32), O.PI), C.byref(mat), C.byref(dp), dt,
                                   O.ptr(np.zeros(3)), Pq, O.ptr(tau))
    return tau, lam


def test_dp_return_map_kat_associated_zero_friction():
    # test/test_stokes.jl:367-393: phi=psi=C=0, eta_reg=1, Kb=1 -> lambda = tauII/2
    dp = O.drucker_prager([0.0], [0.0], [0.0], [1.0], [1.0])
    tau, lam = _dp_stress(dp)
    txx, tyy = 8 / 3, -4 / 3
    tII = math.sqrt((txx**2 + tyy**2 + (-txx - tyy) ** 2) / 2)
    lam_e = tII / 2
    assert abs(lam - lam_e) < 1e-14
    np.testing.assert_allclose(tau[0], txx - 2 * lam_e * (2 * txx + tyy) / (2 * tII), rtol=1e-13)
    np.testing.assert_allclose(tau[1], tyy - 2 * lam_e * (txx + 2 * tyy) / (2 * tII), rtol=1e-13)
    assert tau[2] == 0


def test_dp_return_map_kat_dilatant_denominator():
    # test/test_stokes.jl:395-419: phi=psi=30deg, C=0, eta_reg=3, Kb=2:
    # lambda = tauII/(1 + 3 + 2 sin^2(30deg))
    dp = O.drucker_prager([math.pi / 6], [math.pi / 6], [0.0], [3.0], [2.0])
    tau, lam = _dp_stress(dp)
    txx, tyy = 8 / 3, -4 / 3
    tII = math.sqrt((txx**2 + tyy**2 + (-txx - tyy) ** 2) / 2)
    lam_e = tII / (1.0 + 3.0 + 2.0 * math.sin(math.pi / 6) ** 2)
    assert abs(lam - lam_e) < 1e-14
    np.testing.assert_allclose(tau[0], txx - 2 * lam_e * (2 * txx + tyy) / (2 * tII), rtol=1e-13)


def test_dp_dilation_stiffens_return():
    # test/test_stokes.jl:691-715
    tII = lambda t: math.sqrt((t[0] ** 2 + t[1] ** 2 + (t[0] + t[1]) ** 2) / 2 + t[2] ** 2)
    mk = lambda psi: O.drucker_prager([math.radians(30)], [psi], [0.1], [1e-3], [100.0])
    t3, _ = _dp_stress(mk(math.radians(3)), G=1.0, dt=0.1)
    t0, _ = _dp_stress(mk(0.0), G=1.0, dt=0.1)
    big = O.drucker_prager([math.radians(30)], [0.0], [1e9], [1e-3], [100.0])
    ttrial, lam = _dp_stress(big, G=1.0, dt=0.1)
    assert lam == 0
    assert tII(t3) < tII(ttrial)
    assert tII(t3) > tII(t0)


# --------------------------------------------------------------------------
# assemblers on whole meshes (test/test_stokes.jl:988-1139)
# --------------------------------------------------------------------------
def test_pressure_scaling_homogeneous_and_finite_bulk():
    coords, el2n = O.t7_mesh(2, 2)
    pr = O.Problem(coords, el2n, O.material(eta=(6.0, 99.0), etab=(math.inf, math.inf)))
    pr.pressure_scaling(2.0, 1.0)
    assert abs(pr.MP.sum() - 1.0) < 1e-12 and np.all(pr.MP > 0)
    np.testing.assert_allclose(pr.gammaP, 6.0, atol=1e-12)          # gamma*eta/2
    pr = O.Problem(coords, el2n, O.material(eta=(6.0, 99.0), etab=(1.0, 1.0), K=(4.0, 99.0)))
    pr.pressure_scaling(2.0, 0.5, finite_K=True)
    np.testing.assert_allclose(pr.gammaP, 12 / 7, atol=1e-12)       # test_stokes.jl:1013-1030


def test_zero_velocity_zero_residual_and_body_force_sum():
    coords, el2n = O.t7_mesh(3, 2)
    pr = O.Problem(coords, el2n, O.material(eta=(1.0,), etab=(math.inf,)))
    pr.momentum_residual(1.0)
    assert np.abs(pr.Rvx).max() < 1e-14 and np.abs(pr.Rvy).max() < 1e-14
    rho0, gy = 1.0, -9.81
    pr = O.Problem(coords, el2n, O.material(eta=(1.0,), rho0=(rho0,), g=(0.0, gy)))
    pr.momentum_residual(1.0)
    assert np.abs(pr.Rvx).max() < 1e-14
    assert abs(pr.Rvy.sum() - (-rho0 * gy * 1.0)) < 1e-10        # test_stokes.jl:1110-1139


def test_rigid_rotation_zero_residual_and_divergence():
    # A rigid rotation has zero strain rate and zero divergence.
    coords, el2n = O.t7_mesh(3, 3)
    pr = O.Problem(coords, el2n, O.material(eta=(3.0,), etab=(math.inf,)))
    pr.vx[:] = -(coords[:, 1] - 0.5)
    pr.vy[:] = coords[:, 0] - 0.5
    pr.momentum_residual(1.0)
    pr.pressure_residual(1.0)
    assert np.abs(pr.Rvx).max() < 1e-13 and np.abs(pr.Rvy).max() < 1e-13
    assert np.abs(pr.RP).max() < 1e-14


def test_pressure_residual_uniform_extension_closed_form():
    # v = (x, 0): div = 1 -> RP_a = -int N_a dOmega = -MP_a
    coords, el2n = O.t7_mesh(3, 2)
    pr = O.Problem(coords, el2n, O.material(eta=(1.0,), etab=(math.inf,)))
    pr.vx[:] = coords[:, 0]
    pr.pressure_scaling(1.0, 1.0)
    pr.pressure_residual(1.0)
    np.testing.assert_allclose(pr.RP, -pr.MP, rtol=1e-12)


def _rand_problem(nx=2, ny=2, seed=0, dp=None, G=math.inf, nph=2, single=False):
    rng = np.random.default_rng(seed)
    coords, el2n = O.t7_mesh(nx, ny)
    if single:
        nodes = el2n[0]
        coords = coords[nodes]
        el2n = np.arange(7, dtype=np.int32)[None, :]
    ne = len(el2n)
    mat = O.material(eta=tuple(10 ** rng.uniform(-1, 1, nph)), etab=(math.inf,) * nph, G=(G,) * nph,
                     rho0=tuple(rng.uniform(0.5, 1.5, nph)), g=(0.0, -1.0))
    ph = rng.integers(1, nph + 1, ne).astype(np.int32)
    tau_old = None
    if G != math.inf:
        tau_old = tuple(rng.normal(size=(ne, 7)) * 0.3 for _ in range(3))
    pr = O.Problem(coords, el2n, mat, dp=dp, phases_v=ph, phases_P=ph, tau_old=tau_old)
    pr.vx[:] = np.sin(coords[:, 0]) * coords[:, 1] + 0.1 * rng.normal(size=len(coords))
    pr.vy[:] = coords[:, 0] * np.cos(coords[:, 1]) + 0.1 * rng.normal(size=len(coords))
    pr.P[:] = rng.normal(size=3 * ne)
    return pr


def _fd_jacobian(pr, dt, with_pnum, h=1e-6):
    n = pr.nnodes
    J = np.zeros((2 * n, 2 * n))
    base = (pr.vx.copy(), pr.vy.copy())

    def resid():
        if with_pnum:
            pr.pressure_residual(dt)
            Pn = pr.gammaP * pr.RP / pr.MP
            pr.momentum_residual(dt, Pn)
        else:
            pr.momentum_residual(dt)
        return np.concatenate([pr.Rvx.copy(), pr.Rvy.copy()])

    for comp in range(2):
        arr = pr.vx if comp == 0 else pr.vy
        for j in range(n):
            arr[j] = base[comp][j] + h
            rp = resid()
            arr[j] = base[comp][j] - h
            rm = resid()
            arr[j] = base[comp][j]
            J[:, comp * n + j] = (rp - rm) / (2 * h)
    return J


@pytest.mark.parametrize("dp_on", [False, True])
def test_augmented_jacobian_matches_finite_differences_single_element(dp_on):
    # test/test_stokes.jl:1265-1299 (single T7 element, rtol 1e-6), extended to
    # the augmented operator and to the plastic branch. With a single element
    # the folded absolute row sums equal the FD row sums exactly. All phases
    # equal so the element-max viscosity equals the residual's viscosity.
    dp = O.drucker_prager([math.radians(30)], [0.0], [0.05], [1e-2], [10.0]) if dp_on else None
    pr = _rand_problem(single=True, dp=dp, G=(1.0 if dp_on else math.inf), nph=1, seed=3)
    dt = 0.5
    pr.pressure_scaling(2.0, dt)
    pr.augmented_jacobian(dt)
    J = _fd_jacobian(pr, dt, with_pnum=True)
    n = pr.nnodes
    rs = np.abs(J).sum(axis=1)
    np.testing.assert_allclose(pr.dRdvx, rs[:n], rtol=1e-6)
    np.testing.assert_allclose(pr.dRdvy, rs[n:], rtol=1e-6)
    np.testing.assert_allclose(pr.PCx, np.abs(np.diag(J))[:n], rtol=1e-6)
    np.testing.assert_allclose(pr.PCy, np.abs(np.diag(J))[n:], rtol=1e-6)


def test_augmented_jacobian_gamma_zero_is_plain_and_bounds():
    # test/test_stokes.jl:1301-1328: gamma_eff = 0 gives the plain Jacobian,
    # and row sums bound the diagonal.
    pr = _rand_problem(nx=2, ny=2, nph=1, seed=5)
    pr.MP[:] = 1.0
    pr.gammaP[:] = 0.0
    pr.augmented_jacobian(1.0)
    assert np.all(pr.dRdvx >= pr.PCx) and np.all(pr.dRdvy >= pr.PCy) and np.all(pr.PCx > 0)
    # compare with the FD of the plain residual on a single element
    pr1 = _rand_problem(single=True, nph=1, seed=6)
    pr1.MP[:] = 1.0
    pr1.gammaP[:] = 0.0
    pr1.augmented_jacobian(1.0)
    J = _fd_jacobian(pr1, 1.0, with_pnum=False)
    n = pr1.nnodes
    np.testing.assert_allclose(pr1.dRdvx, np.abs(J).sum(axis=1)[:n], rtol=1e-6)


# --------------------------------------------------------------------------
# full solves with closed-form answers
# --------------------------------------------------------------------------
def test_solve_uniform_extension_bulk_viscous():
    # test/test_solver_convergence_api.jl:330-409: P = -etab dt edot, iter <= 100
    Lx, Ly, dt, edot, etab = 1.0, 0.7, 1.0, 1.0e-5, 6.4e3
    coords, el2n = O.t7_mesh(6, 4, (0.0, Lx, 0.0, Ly))
    gam = O.box_boundary_nodes(coords, (0.0, Lx, 0.0, Ly))
    tol = 32 * EPS
    left = gam[np.abs(coords[gam, 0]) <= tol]
    right = gam[np.abs(coords[gam, 0] - Lx) <= tol]
    bcx = (np.concatenate([left, right]),
           np.concatenate([np.full(len(left), -0.5 * edot * Lx), np.full(len(right), 0.5 * edot * Lx)]))
    bcy = (gam, np.zeros(len(gam)))
    mat = O.material(eta=(1e30,), etab=(etab,), G=(4e3,), K=(6.4e3,))
    ne = len(el2n)
    pr = O.Problem(coords, el2n, mat, tau_old=tuple(np.zeros((ne, 7)) for _ in range(3)), bc=(bcx, bcy))
    pr.vx[:] = edot * (coords[:, 0] - Lx / 2)
    pr.vx[bcx[0]] = bcx[1]
    pr.pressure_scaling(20.0, dt)
    st = pr.solve(dt, tolerance=1e-6, check_interval=25, iter_max=500, skip_scaling=True)
    assert st["converged"]
    assert st["iter"] <= 4 * 25
    np.testing.assert_allclose(pr.P, -etab * dt * edot, rtol=1e-6)
    assert np.abs(pr.vy).max() < 1e-12


def test_solve_hydrostatic_pressure():
    # test/test_stokes_quadrature_viscosity.jl:4-87: P = y - 1/2 (mean removed),
    # v ~ 0. Gravity enters through the material (rho0=1, g=(0,1), K=Inf),
    # which equals the reference's assembled body-force load.
    coords, el2n = O.t7_mesh(2, 2)
    gam = O.box_boundary_nodes(coords)
    mat = O.material(eta=(2.0,), etab=(math.inf,), rho0=(1.0,), g=(0.0, 1.0))
    bc = ((gam, np.zeros(len(gam))), (gam, np.zeros(len(gam))))
    pr = O.Problem(coords, el2n, mat, bc=bc)
    st = pr.solve(1.0, pressure_factor=50.0, tolerance=1e-11, check_interval=10)
    assert st["converged"]
    pc = coords[el2n[:, :3].ravel()]
    P = pr.P - np.sum(pr.P * pr.MP) / np.sum(pr.MP)
    np.testing.assert_allclose(P, pc[:, 1] - 0.5, atol=1e-7)
    assert np.abs(pr.vx).max() < 1e-7 and np.abs(pr.vy).max() < 1e-7


def test_solve_iteration_cap_semantics():
    # test/test_solver_convergence_api.jl:245-294: converged implies the outer
    # test passed; a cap is reported as reached_total_iter.
    coords, el2n = O.t7_mesh(3, 3)
    ne = len(el2n)
    cen = coords[el2n[:, :3]].mean(axis=1)
    ph = np.where(cen[:, 1] < 0.5, 2, 1).astype(np.int32)
    (nx, _), (ny, _) = O.pure_shear_bc(coords)
    bc = ((nx, np.zeros(len(nx))), (ny, np.zeros(len(ny))))
    mat = O.material(eta=(1.0, 1.0), etab=(math.inf,) * 2, rho0=(1.0, 2.0), g=(0.0, -1.0))
    capped = 0
    for tol in (1e-4, 1e-6):
        for cap in range(300, 1801, 300):
            pr = O.Problem(coords, el2n, mat, phases_v=ph, phases_P=ph, bc=bc)
            pr.pressure_scaling(20.0, 1.0)
            st = pr.solve(1.0, tolerance=tol, max_iterations=cap, iter_max=cap, check_interval=100,
                          rel_drop0=0.1, CFL_v=0.9, c_fact=0.7, skip_scaling=True)
            capped += st["reached_total_iter"]
            assert (not st["converged"]) or (not st["reached_total_iter"] and min(st["err_abs"], st["err_rel"]) < tol)
    assert capped > 0
    pr = O.Problem(coords, el2n, mat, phases_v=ph, phases_P=ph, bc=bc)
    pr.pressure_scaling(20.0, 1.0)
    st = pr.solve(1.0, tolerance=1e-6, check_interval=100, rel_drop0=0.1, CFL_v=0.9, c_fact=0.7, skip_scaling=True)
    assert st["converged"] and not st["reached_total_iter"]


def test_solve_invalid_preconditioner_message():
    # utils.jl:152-158: an orphan node has PC = 0 -> invalid lambda max
    coords, el2n = O.t7_mesh(1, 1)
    coords = np.vstack([coords, [[2.0, 2.0]]])
    pr = O.Problem(coords, el2n, O.material(eta=(1.0,), etab=(1.0,)))
    with pytest.raises(RuntimeError, match="stokes vx preconditioner produced invalid"):
        pr.solve(1.0, max_iterations=2)
