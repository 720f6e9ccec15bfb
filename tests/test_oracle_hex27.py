"""The oracle's 3-D Hex27 / P1-disc restatement pinned by the reference's own
known-answer tests and by closed forms (basis, rule, mesh numbering, geometry,
3-D momentum KAT, pressure scaling, residual identities, Jacobian by finite
differences, full solves with closed-form answers)."""
import itertools
import math

import numpy as np
import pytest

import pyoracle as O

EPS = np.finfo(float).eps
HEX_NODES = np.array([
    [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
    [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
    [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0], [0, 0, -1], [0, -1, 0], [1, 0, 0], [0, 1, 0],
    [-1, 0, 0], [0, 0, 1], [0, 0, 0]], float)


def test_hex27_kronecker_partition_and_quadratic_reproduction():
    for a, x in enumerate(HEX_NODES):
        N, _ = O.hex27_basis(*x)
        e = np.zeros(27)
        e[a] = 1.0
        np.testing.assert_allclose(N, e, atol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(5):
        p = rng.uniform(-1, 1, 3)
        N, dN = O.hex27_basis(*p)
        assert abs(N.sum() - 1) < 1e-14 and np.abs(dN.sum(axis=0)).max() < 1e-14
        # triquadratic field x² y z² - 3 x y + z is reproduced with its gradient
        f = lambda x: x[0] ** 2 * x[1] * x[2] ** 2 - 3 * x[0] * x[1] + x[2]
        df = lambda x: np.array([2 * x[0] * x[1] * x[2] ** 2 - 3 * x[1], x[0] ** 2 * x[2] ** 2 - 3 * x[0],
                                 2 * x[0] ** 2 * x[1] * x[2] + 1])
        vals = np.array([f(x) for x in HEX_NODES])
        assert abs(N @ vals - f(p)) < 1e-13
        np.testing.assert_allclose(dN.T @ vals, df(p), atol=1e-13)


def test_gauss_rule_3x3x3_exactness():
    xi, eta, zeta, w = O.hex27_rule()
    assert abs(w.sum() - 8.0) < 1e-14
    # ∫_{-1}^{1} x^k = 2/(k+1) for even k, 0 odd; exact up to degree 5 per axis
    mono = lambda k: 0.0 if k % 2 else 2.0 / (k + 1)
    for a, b, c in itertools.product(range(6), repeat=3):
        q = np.sum(w * xi ** a * eta ** b * zeta ** c)
        assert abs(q - mono(a) * mono(b) * mono(c)) < 1e-13
    assert abs(np.sum(w * xi ** 6) - 4 * mono(6)) > 1e-3   # degree 6 is not exact


def test_hex27_single_cell_connectivity_kat():
    # test/test_mesh.jl:84-121
    coords, el2n = O.hex27_mesh(1, 1, 1)
    kat = [1, 3, 9, 7, 19, 21, 27, 25, 2, 6, 8, 4, 20, 24, 26, 22, 10, 12, 18, 16, 5, 11, 15, 17, 13, 23, 14]
    assert list(el2n[0] + 1) == kat
    assert len(coords) == 27 and len(O.box_boundary_nodes3(coords)) == 26
    # every local node sits at its reference position
    np.testing.assert_allclose(coords[el2n[0]], (HEX_NODES + 1) / 2, atol=1e-15)


def test_hex27_mesh_counts_sharing_and_geometry():
    nx, ny, nz = 3, 2, 2
    box = (0.0, 1.5, -1.0, 1.0, 0.0, 0.4)
    coords, el2n = O.hex27_mesh(nx, ny, nz, box)
    assert len(coords) == 7 * 5 * 5 and len(el2n) == 12
    # the centre node is private to its element, corners shared by up to 8
    counts = np.bincount(el2n.ravel(), minlength=len(coords))
    assert np.all(counts[el2n[:, 26]] == 1)
    assert counts.max() == 8
    invJ, dOm = O.hex27_geometry(coords, el2n)
    hx, hy, hz = 1.5 / nx, 2.0 / ny, 0.4 / nz
    np.testing.assert_allclose(invJ[:, :, 0], 2 / hx, rtol=1e-14)
    np.testing.assert_allclose(invJ[:, :, 4], 2 / hy, rtol=1e-14)
    np.testing.assert_allclose(invJ[:, :, 8], 2 / hz, rtol=1e-14)
    assert np.abs(invJ[:, :, [1, 2, 3, 5, 6, 7]]).max() < 1e-12   # rounding of Σ x_a ∂N_a/∂ξ
    assert abs(dOm.sum() - 1.5 * 2.0 * 0.4) < 1e-13


def test_momentum_element_3d_viscous_kat():
    # test/test_stokes.jl:237-256
    mat = O.material(eta=(4.0,), etab=(math.inf,), G=(math.inf,), rho0=(3.0,), K=(math.inf,))
    dNdx = np.array([[[0.2, 0.3, 0.4]]])
    R = [np.zeros(1) for _ in range(3)]
    g = np.array([0.0, -2.0, 0.0])
    O.lib().orc_momentum_element3(1, 1, 1, O.ptr(dNdx), O.ptr(np.array([2.0])), O.ptr(np.array([[1.0]])),
                                  O.ptr(np.array([[1.0]])), O.ptr(np.array([1.0])), O.ptr(np.array([2.0])),
                                  O.ptr(np.array([3.0])), O.ptr(np.array([5.0])), O.ptr(np.array([0.5])),
                                  O.ptr(np.array([1], np.int32), O.PI), mat, O.ptr(g), O.ptr(None, O.C.POINTER(O.DP)), 1.0, O.ptr(None),
                                  O.ptr(R[0]), O.ptr(R[1]), O.ptr(R[2]), O.ptr(None))
    gv = np.array([[0.2, 0.3, 0.4], [0.4, 0.6, 0.8], [0.6, 0.9, 1.2]])
    strain = (gv + gv.T) / 2
    stress = 8 * (strain - np.trace(strain) / 3 * np.eye(3))
    d = dNdx[0, 0]
    for i in range(3):
        assert R[i][0] == pytest.approx((d @ stress[:, i] - d[i] * 5.5 - 3 * g[i]) * 2, rel=1e-14)


def _unit_problem(n=(2, 2, 2), mat=None, g=(0.0, 0.0, 0.0), dp=None, seed=None, bc=None, tau_old=None):
    coords, el2n = O.hex27_mesh(*n)
    mat = mat or O.material(eta=(1.0,), etab=(math.inf,), G=(math.inf,), rho0=(1.0,), K=(math.inf,))
    pr = O.Problem3(coords, el2n, mat, g=g, dp=dp, tau_old=tau_old, bc=bc)
    if seed is not None:
        rng = np.random.default_rng(seed)
        for c in range(3):
            pr.v[c][:] = rng.normal(size=pr.nnodes)
        pr.P[:] = rng.normal(size=pr.nnodesP)
    return pr, coords, el2n


def test_pressure_scaling_positive_modal_mass_and_closed_forms():
    # test/test_mixed_mesh.jl:175-197 (positive modal mass), and the N² weights
    # in closed form
    pr, coords, el2n = _unit_problem((1, 1, 1))
    pr.pressure_scaling(2.0, 1.0)
    assert np.all(pr.MP > 0) and np.all(np.isfinite(pr.gammaP)) and np.all(pr.gammaP > 0)
    # reference cube: ∫(1-ξ-η-ζ)² = 8 + 3·8/3 = 16, ∫ξ² = 8/3; the cell maps
    # the cube onto [0,1]³ with |det J| = 1/8
    np.testing.assert_allclose(pr.MP, np.array([16.0, 8 / 3, 8 / 3, 8 / 3]) / 8, rtol=1e-13)
    # K = Inf: γ = γfact η (γ_num = γ_phy, harmonic mean half of it: γfact η / 2)
    np.testing.assert_allclose(pr.gammaP, 2.0 * 1.0 / 2, rtol=1e-13)


def test_rigid_rotation_and_zero_fields_give_zero_residuals():
    pr, coords, _ = _unit_problem((2, 2, 1))
    w = np.array([0.3, -0.2, 0.5])
    v = np.cross(w, coords)
    for c in range(3):
        pr.v[c][:] = v[:, c]
    pr.pressure_residual(1.0)
    pr.momentum_residual(1.0)
    assert np.abs(pr.RP).max() < 1e-14
    assert max(np.abs(r).max() for r in pr.Rv) < 1e-14


def test_body_force_sums_to_total_weight():
    # Σ_nodes R_z = -ρ g_z · volume with v = P = 0 (partition of unity)
    mat = O.material(eta=(1.0,), etab=(math.inf,), G=(math.inf,), rho0=(2.5,), K=(math.inf,))
    pr, _, _ = _unit_problem((2, 1, 2), mat=mat, g=(0.0, 0.0, -1.5))
    pr.momentum_residual(1.0)
    assert abs(pr.Rv[2].sum() - 2.5 * 1.5) < 1e-13
    assert abs(pr.Rv[0].sum()) < 1e-14 and abs(pr.Rv[1].sum()) < 1e-14


def test_uniform_extension_pressure_residual_closed_form():
    # v = ε̇ x: RP_a = -ε̇ ∫ N_a dΩ; on the reference cube ∫(1-ξ-η-ζ) = 8 and
    # ∫ξ = ∫η = ∫ζ = 0, times |det J| = (h/2)³
    pr, coords, _ = _unit_problem((2, 2, 2))
    pr.v[0][:] = 0.7 * coords[:, 0]
    pr.pressure_residual(1.0)
    h3 = (0.5 ** 3) / 8
    expect = np.tile([-0.7 * 8 * h3, 0.0, 0.0, 0.0], pr.nels)
    np.testing.assert_allclose(pr.RP, expect, atol=1e-15)


def _fd_jacobian3(pr, dt, h=1e-7):
    n = pr.nnodes

    def resid(vec):
        for c in range(3):
            pr.v[c][:] = vec[c * n:(c + 1) * n]
        pr.pressure_residual(dt)
        Pnum = pr.gammaP * pr.RP / pr.MP
        pr.momentum_residual(dt, Pnum)
        return np.concatenate([r.copy() for r in pr.Rv])

    v0 = np.concatenate([c.copy() for c in pr.v])
    J = np.zeros((3 * n, 3 * n))
    for k in range(3 * n):
        vp, vm = v0.copy(), v0.copy()
        vp[k] += h
        vm[k] -= h
        J[:, k] = (resid(vp) - resid(vm)) / (2 * h)
    resid(v0)
    return J


@pytest.mark.parametrize("dp_on", [False, True])
def test_augmented_jacobian_3d_matches_finite_differences_single_element(dp_on):
    dp = O.drucker_prager([math.radians(30)], [0.0], [0.05], [1e-2], [10.0]) if dp_on else None
    mat = O.material(eta=(1.3,), etab=(50.0,), G=(1.0 if dp_on else math.inf,), rho0=(1.0,), K=(math.inf,))
    ne = 1
    tau = [np.full((ne, 27), 0.2 * (k + 1) * (-1) ** k) for k in range(6)] if dp_on else None
    pr, _, _ = _unit_problem((1, 1, 1), mat=mat, dp=dp, seed=4, tau_old=tau)
    dt = 0.5
    pr.pressure_scaling(2.0, dt)
    pr.augmented_jacobian(dt)
    J = _fd_jacobian3(pr, dt)
    n = pr.nnodes
    rs = np.abs(J).sum(axis=1)
    dg = np.abs(np.diag(J))
    for c in range(3):
        np.testing.assert_allclose(pr.dRdv[c], rs[c * n:(c + 1) * n], rtol=1e-6)
        np.testing.assert_allclose(pr.PC[c], dg[c * n:(c + 1) * n], rtol=1e-6)


def test_solve_3d_uniform_extension_bulk_viscous():
    # 3-D counterpart of test/test_solver_convergence_api.jl:330-409:
    # vx = ε̇ (x - Lx/2) on the x-walls, vy = vz = 0 on all walls gives
    # P = -ηb Δt ε̇ everywhere
    L, dt, edot, etab = (1.0, 0.7, 0.5), 1.0, 1.0e-5, 6.4e3
    coords, el2n = O.hex27_mesh(3, 2, 2, (0.0, L[0], 0.0, L[1], 0.0, L[2]))
    gam = O.box_boundary_nodes3(coords, (0.0, L[0], 0.0, L[1], 0.0, L[2]))
    xw = gam[(np.abs(coords[gam, 0]) < 1e-12) | (np.abs(coords[gam, 0] - L[0]) < 1e-12)]
    bc = ((xw, edot * (coords[xw, 0] - L[0] / 2)), (gam, np.zeros(len(gam))), (gam, np.zeros(len(gam))))
    mat = O.material(eta=(1e30,), etab=(etab,), G=(4e3,), K=(6.4e3,))
    ne = len(el2n)
    pr = O.Problem3(coords, el2n, mat, tau_old=[np.zeros((ne, 27)) for _ in range(6)], bc=bc)
    pr.v[0][:] = edot * (coords[:, 0] - L[0] / 2)
    pr.pressure_scaling(20.0, dt)
    # the residual scale is ε̇ = 1e-5, so the tolerance is taken well below it
    st = pr.solve(dt, tolerance=1e-13, check_interval=25, iter_max=20_000, skip_scaling=True)
    assert st["converged"]
    # a uniform pressure has the same value in all four modes (nodal basis)
    np.testing.assert_allclose(pr.P, -etab * dt * edot, rtol=1e-6)
    assert max(np.abs(pr.v[1]).max(), np.abs(pr.v[2]).max()) < 1e-12


def test_solve_3d_hydrostatic_pressure():
    # 3-D counterpart of test/test_stokes_quadrature_viscosity.jl:4-87:
    # ρ0 = 1, g = (0, 0, 1), no-slip walls: P = z - 1/2 (mean removed), v ~ 0
    mat = O.material(eta=(2.0,), etab=(math.inf,), rho0=(1.0,), K=(math.inf,))
    coords, el2n = O.hex27_mesh(2, 2, 2)
    gam = O.box_boundary_nodes3(coords)
    z = (gam, np.zeros(len(gam)))
    pr = O.Problem3(coords, el2n, mat, g=(0.0, 0.0, 1.0), bc=(z, z, z))
    st = pr.solve(1.0, pressure_factor=50.0, tolerance=1e-11, check_interval=10)
    assert st["converged"]
    # pressure DoFs are nodal at the cell centre and the +x, +y, +z face centres
    pc = coords[el2n[:, [26, 22, 23, 25]].ravel()]
    P = pr.P - np.sum(pr.P[0::4] * pr.MP[0::4]) / np.sum(pr.MP[0::4])
    np.testing.assert_allclose(P, pc[:, 2] - 0.5, atol=1e-7)
    assert max(np.abs(v).max() for v in pr.v) < 1e-7
