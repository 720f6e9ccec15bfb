"""The oracle's thermal-diffusion restatement on Q9 meshes (heat
residual.jl:58-121 with QuadraticElement{2, 9}, shape_functions.jl:259-283,
and its 3x3 Gauss rule, integration_points.jl:218-241; the mesh of
mesh.jl:340-398 and connectivity.jl:193-227), pinned by closed forms:

* the basis is nodal, sums to one and reproduces every polynomial of the
  tensor space (1, x, y, x², xy, y², x²y, xy², x²y²); its gradients match
  central differences of the basis;
* the rule integrates ξ^i η^j exactly for i, j <= 5 (the 1-D three-point
  Gauss rule) and its weights sum to the reference area 4;
* with no diffusion the residual is (T0 - T + Δt s/(ρ Cp)) ∫N_i: the
  consistent-mass integrals of a rectangle are A/36 (corners), A/9 (edges)
  and 4A/9 (centre);
* a linear field has zero interior residual, and T = x² gives the interior
  residual 2 k Δt/(ρ Cp) ∫N_i (integration by parts; Q9 holds x² exactly);
* the assembled Jacobian diagnostics equal the row sums of |J| and |J_ii| of
  the central finite-difference Jacobian (a curved element, two phases,
  thermal expansion on);
* the quadratic profile of k T'' = -s is the solve's fixed point and the
  solve from T = 0 returns it.
"""
import math

import numpy as np
import pytest

import pyoracle as O

NODES = np.array([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)], float)


def test_basis_is_nodal_and_reproduces_the_tensor_space():
    for b, (xi, eta) in enumerate(NODES):
        N, _ = O.q9_basis(xi, eta)
        assert np.allclose(N, np.eye(9)[b], atol=1e-15)
    rng = np.random.default_rng(0)
    for xi, eta in rng.uniform(-1, 1, (20, 2)):
        N, dN = O.q9_basis(xi, eta)
        assert abs(N.sum() - 1) < 1e-14
        for i in range(3):
            for j in range(3):
                f = NODES[:, 0] ** i * NODES[:, 1] ** j
                assert abs(N @ f - xi ** i * eta ** j) < 1e-14
        h = 1e-6
        Np, _ = O.q9_basis(xi + h, eta)
        Nm, _ = O.q9_basis(xi - h, eta)
        assert np.allclose(dN[:, 0], (Np - Nm) / (2 * h), atol=1e-8)
        Np, _ = O.q9_basis(xi, eta + h)
        Nm, _ = O.q9_basis(xi, eta - h)
        assert np.allclose(dN[:, 1], (Np - Nm) / (2 * h), atol=1e-8)


def test_rule_is_the_tensor_gauss_rule():
    xi, eta, w = O.q9_rule()
    assert abs(w.sum() - 4) < 1e-14
    assert np.allclose(xi[:3], [-math.sqrt(0.6), 0, math.sqrt(0.6)], rtol=0, atol=1e-16)
    assert np.array_equal(eta[:3], np.full(3, eta[0])) and eta[0] < 0   # ξ fastest
    for i in range(6):
        for j in range(6):
            exact = (0 if i % 2 else 2 / (i + 1)) * (0 if j % 2 else 2 / (j + 1))
            assert abs(np.sum(w * xi ** i * eta ** j) - exact) < 1e-14


def test_mesh_places_the_local_nodes_on_their_reference_positions():
    box = (0.5, 2.5, -1.0, 0.5)
    nx, ny = 3, 2
    coords, el2n = O.q9_mesh(nx, ny, box)
    assert coords.shape == ((2 * nx + 1) * (2 * ny + 1), 2) and el2n.shape == (nx * ny, 9)
    hx, hy = (box[1] - box[0]) / nx, (box[3] - box[2]) / ny
    for e in range(nx * ny):
        ex, ey = e % nx, e // nx          # elements x fastest
        cx, cy = box[0] + (ex + 0.5) * hx, box[2] + (ey + 0.5) * hy
        expect = np.stack([cx + NODES[:, 0] * hx / 2, cy + NODES[:, 1] * hy / 2], axis=1)
        assert np.allclose(coords[el2n[e]], expect, rtol=0, atol=1e-14)
    # every node belongs to an element, centres to exactly one
    counts = np.bincount(el2n.ravel(), minlength=len(coords))
    assert counts.min() >= 1 and np.all(counts[el2n[:, 8]] == 1)
    invJ, dOm = O.q9_geometry(coords, el2n)
    assert abs(dOm.sum() - (box[1] - box[0]) * (box[3] - box[2])) < 1e-13
    assert np.allclose(invJ[..., 0], 2 / hx) and np.allclose(invJ[..., 3], 2 / hy)
    assert np.abs(invJ[..., 1]).max() < 1e-13 and np.abs(invJ[..., 2]).max() < 1e-13


def test_consistent_mass_integrals_without_diffusion():
    # k = 0: R_i = (T0 - T + Δt s/(ρ Cp)) Σ_e ∫N_i on a uniform rectangle mesh
    coords, el2n = O.q9_mesh(2, 3, (0.0, 1.0, 0.0, 1.5))
    A = 0.5 * 0.5   # cell area
    m = O.thermal_material(k=(0.0,), Cp=(2.0,), rho0=(1.5,), alpha=(0.0,), K=(math.inf,))
    t = O.ThermalProblemQ9(coords, el2n, m)
    t.T0[:] = 0.7
    t.T[:] = 0.2
    t.source[:] = 3.0
    dt = 0.4
    t.residual(dt, Tref=0.0)
    rate = 0.7 - 0.2 + dt * 3.0 / (1.5 * 2.0)
    mass = np.zeros(len(coords))
    per = np.array([1, 1, 1, 1, 4, 4, 4, 4, 16]) * A / 36
    for e in range(len(el2n)):
        np.add.at(mass, el2n[e], per)
    assert np.allclose(t.R, rate * mass, rtol=1e-13, atol=1e-15)


def test_linear_and_quadratic_fields_interior_residuals():
    coords, el2n = O.q9_mesh(3, 3)
    k, rho, cp, dt = 1.7, 1.3, 0.8, 0.1
    m = O.thermal_material(k=(k,), Cp=(cp,), rho0=(rho,), alpha=(0.0,), K=(math.inf,))
    inner = np.setdiff1d(np.arange(len(coords)), O.box_boundary_nodes(coords))
    t = O.ThermalProblemQ9(coords, el2n, m)
    lin = 1.0 + 0.3 * coords[:, 0] - 0.7 * coords[:, 1]
    t.T[:] = lin
    t.T0[:] = lin
    t.residual(dt, Tref=0.0)
    assert np.abs(t.R[inner]).max() < 1e-14
    # T = x²: -Δt/(ρCp) ∫ k ∇N_i·∇T = Δt/(ρCp) k ΔT ∫N_i = 2 k Δt/(ρCp) ∫N_i
    quad = coords[:, 0] ** 2
    t.T[:] = quad
    t.T0[:] = quad
    t.residual(dt, Tref=0.0)
    mass = np.zeros(len(coords))
    per = np.array([1, 1, 1, 1, 4, 4, 4, 4, 16]) / 36 / 9
    for e in range(len(el2n)):
        np.add.at(mass, el2n[e], per)
    expect = 2 * k * dt / (rho * cp) * mass
    assert np.allclose(t.R[inner], expect[inner], rtol=1e-12, atol=1e-15)


def test_jacobian_diagnostics_match_finite_differences():
    coords, el2n = O.q9_mesh(1, 1)
    coords = coords + 0.04 * np.sin(5.0 * coords[:, ::-1])        # a curved element
    m = O.thermal_material(k=(1.5, 0.7), Cp=(1.2, 0.9), rho0=(2.0, 3.0), alpha=(0.02, 0.01), K=(50.0, 80.0))
    rng = np.random.default_rng(4)
    ph = rng.integers(1, 3, len(coords))
    t = O.ThermalProblemQ9(coords, el2n, m, phases=ph)
    t.T[:] = 300 + rng.normal(size=len(coords))
    t.T0[:] = 300 + rng.normal(size=len(coords))
    t.source[:] = rng.normal(size=len(coords))
    t.P[:] = rng.normal(size=len(coords))
    dt, Tref = 0.3, 273.0
    t.residual(dt, Tref, jacobian=True)
    rs, pc = t.dRdT.copy(), t.PC.copy()
    n = len(coords)
    J = np.zeros((n, n))
    h = 1e-6
    T = t.T.copy()
    for j in range(n):
        t.T[:] = T
        t.T[j] += h
        t.residual(dt, Tref)
        rp = t.R.copy()
        t.T[j] -= 2 * h
        t.residual(dt, Tref)
        J[:, j] = (rp - t.R) / (2 * h)
    assert np.allclose(np.abs(J).sum(axis=1), rs, rtol=1e-6, atol=1e-9)
    assert np.allclose(np.abs(np.diag(J)), pc, rtol=1e-6, atol=1e-9)


def test_quadratic_steady_profile_is_the_discrete_solution():
    # k T'' = -s with T = 1 - x + s x (1 - x)/(2k): from T0 = that profile the
    # implicit step keeps it (the y faces are insulated, with no normal flux),
    # and the solve from T = 0 with the x faces held returns it
    k, s_, dt = 0.8, 2.5, 0.05
    m = O.thermal_material(k=(k,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    coords, el2n = O.q9_mesh(3, 2)
    x = coords[:, 0]
    prof = 1 - x + s_ * x * (1 - x) / (2 * k)
    bx = np.nonzero((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12))[0].astype(np.int32)
    t = O.ThermalProblemQ9(coords, el2n, m, bc=(bx, prof[bx]))
    t.source[:] = s_
    t.T0[:] = prof
    t.T[:] = prof
    t.residual(dt, Tref=0.0)
    free = np.setdiff1d(np.arange(len(coords)), bx)
    assert np.abs(t.R[free]).max() < 1e-13
    t.T[:] = 0.0
    st = t.solve(dt, tolerance=1e-12, max_iterations=200000, check_interval=50, Tref=0.0)
    assert st["converged"]
    assert np.allclose(t.T, prof, rtol=0, atol=1e-9)


def test_coupled_solve_rejects_a_q9_thermal_mesh():
    # a 5x1 Q9 grid and a 2x2 T7 mesh both have 33 nodes: the element types
    # differ, so the Stokes points of the shear heating and the node layout do not match
    coords, el2n = O.q9_mesh(5, 1)
    tm = O.thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    th = O.ThermalProblemQ9(coords, el2n, tm)
    c7, e7 = O.t7_mesh(2, 2)
    pr = O.Problem(c7, e7, O.material(eta=(1.0,)))
    assert th.nnodes == pr.nnodes == 33
    with pytest.raises(RuntimeError, match="thermal mesh nodes"):
        O.coupled_solve(th, pr, 1.0)
