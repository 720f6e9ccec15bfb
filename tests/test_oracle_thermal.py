"""The oracle's thermal-diffusion restatement (src/heat_diffusion/**,
src/dynamic_relaxation/solver.jl) pinned by the reference's own known-answer
tests and by closed forms."""
import math

import numpy as np
import pytest

import pyoracle as O


def test_integrate_residual_tref_kat():
    # test/test_heat_diffusion_residual.jl:1-24
    m = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(2.0,), alpha=(0.01,), K=(100.0,))
    N, dN = [[0.5, 0.5]], np.zeros((1, 2, 2))
    args = ([1.0], [300.0, 300.0], [300.0, 300.0], [1.0, 1.0], [1, 1], m, [0.0, 0.0], 1.0)
    r1 = O.thermal_integrate(N, dN, *args, 300.0)
    r2 = O.thermal_integrate(N, dN, *args, 250.0)
    assert np.allclose(r1, [0.25, 0.25], rtol=1e-14)
    assert np.allclose(r2, 2 * r1, rtol=1e-14)


def test_integrate_residual_incompressible_limit():
    # test/test_heat_diffusion_residual.jl:26-58: a negative shape-function
    # weight with K=Inf stays finite because 1/K is what is interpolated
    N = [[-0.25, 1.25]]
    dN = np.array([[[0.0, 1.0], [0.0, -1.0]]])
    T, T0, s, P = [300.0, 300.0], [300.0, 300.0], [1.0, 1.0], [1.0e6, 1.0e6]
    inc = O.thermal_integrate(N, dN, [1.0], T, T0, s, [1, 1],
                              O.thermal_material((1.0,), (1.0,), (2.0,), (0.01,), (math.inf,)), P, 1.0, 300.0)
    assert np.all(np.isfinite(inc))
    two = O.thermal_integrate(N, dN, [1.0], T, T0, s, [1, 2],
                              O.thermal_material((1.0, 1.0), (1.0, 1.0), (2.0, 2.0), (0.01, 0.01),
                                                 (math.inf, math.inf)), P, 1.0, 300.0)
    assert np.allclose(two, inc, rtol=1e-14)
    stiff = O.thermal_integrate(N, dN, [1.0], T, T0, s, [1, 1],
                                O.thermal_material((1.0,), (1.0,), (2.0,), (0.01,), (1.0e30,)), P, 1.0, 300.0)
    assert np.allclose(stiff, inc, rtol=1e-12)


def _t7_element_tables(seed):
    rng = np.random.default_rng(seed)
    Nq, _, dNdxi, w = O.t7_tables()
    # a random affine triangle
    X = np.array([[0.0, 0.0], [1.0, 0.1], [0.2, 0.9]]) + 0.1 * rng.normal(size=(3, 2))
    Jm = np.array([X[1] - X[0], X[2] - X[0]]).T          # columns: d x / d xi, d x / d eta
    iJ = np.linalg.inv(Jm)
    dNdx = np.einsum("qak,kd->qad", dNdxi, iJ)
    dOm = w * abs(np.linalg.det(Jm))
    return rng, Nq, dNdx, dOm


@pytest.mark.parametrize("alpha", [0.0, 0.02])
def test_element_jacobian_matches_finite_differences(alpha):
    rng, Nq, dNdx, dOm = _t7_element_tables(1)
    m = O.thermal_material(k=(1.5, 0.7), Cp=(1.2, 0.9), rho0=(2.0, 3.0), alpha=(alpha, 0.5 * alpha), K=(50.0, 80.0))
    ph = rng.integers(1, 3, 7)
    T = 300 + rng.normal(size=7)
    T0 = 300 + rng.normal(size=7)
    s, P = rng.normal(size=7), rng.normal(size=7)
    R, J = O.thermal_integrate(Nq, dNdx, dOm, T, T0, s, ph, m, P, 0.3, 290.0, jacobian=True)
    h = 1e-6
    Jfd = np.zeros((7, 7))
    for j in range(7):
        Tp, Tm = T.copy(), T.copy()
        Tp[j] += h
        Tm[j] -= h
        Jfd[:, j] = (O.thermal_integrate(Nq, dNdx, dOm, Tp, T0, s, ph, m, P, 0.3, 290.0)
                     - O.thermal_integrate(Nq, dNdx, dOm, Tm, T0, s, ph, m, P, 0.3, 290.0)) / (2 * h)
    assert np.allclose(J, Jfd, rtol=1e-6, atol=1e-9 * np.abs(J).max())
    if alpha == 0.0:
        # mass and stiffness are symmetric: a transposed index would break this
        assert np.allclose(J, J.T, rtol=1e-13, atol=1e-15)


def _problem(n, mat, box=(0.0, 1.0, 0.0, 1.0), phases=None, bc=None):
    coords, el2n = O.t7_mesh(n, n, box)
    pr = O.ThermalProblem(coords, el2n, mat, phases=phases,
                          bc=bc if bc is not None else (np.zeros(0, np.int32), np.zeros(0)))
    return pr


def test_uniform_source_without_boundary_conditions_closed_form():
    # T = T0 + Δt s / (ρ Cp), ∇T = 0
    mat = O.thermal_material(k=(1.0,), Cp=(2.0,), rho0=(1.5,), alpha=(0.0,), K=(math.inf,))
    pr = _problem(4, mat)
    pr.source[:] = 3.0
    pr.T0[:] = 1.0
    pr.T[:] = pr.T0
    st = pr.solve(0.5, tolerance=1e-12, check_interval=10)
    assert st["converged"]
    assert np.allclose(pr.T, 1.0 + 0.5 * 3.0 / (1.5 * 2.0), rtol=1e-10)


def test_uniform_source_with_thermal_expansion_fixed_point():
    # ρ = ρ0 (1 - α (T - Tref)) makes the uniform step a scalar fixed point:
    # T - T0 = Δt s / (ρ0 (1 - α (T - Tref)) Cp)
    a, r0, cp, dt, s, T0, Tref = 0.05, 1.0, 1.0, 0.4, 2.0, 1.0, 0.0
    mat = O.thermal_material(k=(1.0,), Cp=(cp,), rho0=(r0,), alpha=(a,), K=(math.inf,))
    pr = _problem(3, mat)
    pr.source[:] = s
    pr.T0[:] = T0
    pr.T[:] = T0
    st = pr.solve(dt, tolerance=1e-13, check_interval=10, Tref=Tref, max_iterations=50_000)
    assert st["converged"]
    # (T - T0)(1 - a (T - Tref)) r0 cp = dt s -> quadratic in T
    A, B, C = -a * r0 * cp, r0 * cp * (1 + a * Tref + a * T0), -r0 * cp * T0 * (1 + a * Tref) - dt * s
    roots = np.roots([A, B, C])
    T = roots[np.argmin(abs(roots - T0))].real
    assert np.allclose(pr.T, T, rtol=1e-9)


def test_linear_profile_is_reproduced():
    # T0 linear, Dirichlet values of the same linear field: the T7 space holds
    # it exactly and it is the discrete solution for any Δt
    mat = O.thermal_material(k=(1.3,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    coords, _ = O.t7_mesh(6, 6)
    lin = 2.0 * coords[:, 0] - 0.5 * coords[:, 1] + 1.0
    b = O.box_boundary_nodes(coords)
    pr = _problem(6, mat, bc=(b, lin[b]))
    pr.T0[:] = lin
    pr.T[:] = 0.0
    pr.T[b] = lin[b]
    st = pr.solve(10.0, tolerance=1e-11, check_interval=20, max_iterations=100_000)
    assert st["converged"]
    assert np.allclose(pr.T, lin, atol=1e-8)


def test_non_convergence_reports_statistics():
    # test/test_solver_convergence_api.jl:426-433 (ϵ = 0)
    mat = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    pr = _problem(2, mat)
    pr.source[:] = 1.0
    st = pr.solve(1.0, tolerance=0.0, max_iterations=3, check_interval=1)
    assert not st["converged"] and st["iterations"] == 3


def test_orphan_node_invalid_lambda_max():
    # test/test_solver_convergence_api.jl:131-145: a node in no element has a
    # zero preconditioner
    mat = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    coords, el2n = O.t7_mesh(1, 1)
    coords = np.vstack([coords, [[2.0, 2.0]]])
    pr = O.ThermalProblem(coords, el2n, mat)
    with pytest.raises(RuntimeError, match="thermal diffusion preconditioner produced invalid λmax"):
        pr.solve(1.0)
