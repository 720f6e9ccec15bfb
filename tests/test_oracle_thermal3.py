"""The oracle's thermal-diffusion restatement on Hex27 meshes (heat
residual.jl:58-121 with the 27-node element and the 3x3x3 Gauss rule,
dynamic_relaxation/solver.jl:61-122), pinned by closed forms:

* a uniform source in an insulated box: one backward-Euler step from T0 = 0
  gives T = Δt s / (ρ Cp) everywhere (the consistent-mass rate vanishes);
* a linear temperature with T0 = T has zero residual at the interior nodes
  (the element reproduces it and a constant-coefficient flux of a linear
  field has no divergence);
* the assembled Jacobian diagnostics equal the row sums of |J| and |J_ii| of
  the central finite-difference Jacobian of the assembled residual
  (one-element mesh, thermal expansion on, two phases);
* the quadratic profile solving k T'' = -s is the discrete solution of an
  implicit step from itself (zero residual off the held x faces) and the
  solve from T = 0 returns it (a wrong source weight, a lost axis or a
  transposed gradient breaks it).
"""
import math

import numpy as np
import pytest

import pyoracle as O


def _box3(n, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    return O.hex27_mesh(n, n, n, box)


def test_uniform_source_one_step():
    coords, el2n = _box3(2)
    m = O.thermal_material(k=(1.3,), Cp=(2.0,), rho0=(1.5,), alpha=(0.0,), K=(math.inf,))
    t = O.ThermalProblem3(coords, el2n, m)
    t.source[:] = 3.0
    dt = 0.4
    st = t.solve(dt, tolerance=1e-12, max_iterations=20000, check_interval=50, Tref=0.0)
    assert st["converged"]
    assert np.allclose(t.T, dt * 3.0 / (1.5 * 2.0), rtol=1e-9)


def test_linear_field_has_zero_interior_residual():
    coords, el2n = _box3(2)
    m = O.thermal_material(k=(1.7,), Cp=(1.0,), rho0=(1.3,), alpha=(0.0,), K=(math.inf,))
    lin = 1.0 + 0.3 * coords[:, 0] - 0.7 * coords[:, 1] + 0.2 * coords[:, 2]
    t = O.ThermalProblem3(coords, el2n, m)
    t.T[:] = lin
    t.T0[:] = lin
    t.residual(0.1, Tref=0.0)
    inner = np.setdiff1d(np.arange(len(coords)), O.box_boundary_nodes3(coords))
    assert np.abs(t.R[inner]).max() < 1e-13


def test_jacobian_diagnostics_match_finite_differences():
    coords, el2n = _box3(1)
    coords = coords + 0.03 * np.sin(7.0 * coords)          # a curved (non-affine) element
    m = O.thermal_material(k=(1.5, 0.7), Cp=(1.2, 0.9), rho0=(2.0, 3.0), alpha=(0.02, 0.01), K=(50.0, 80.0))
    rng = np.random.default_rng(3)
    ph = rng.integers(1, 3, len(coords))
    t = O.ThermalProblem3(coords, el2n, m, phases=ph)
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
    # k T'' = -s with T = 1 - x + s x (1 - x) / (2 k): from T0 = that profile the
    # implicit step keeps it (rate = Δt s / (ρ Cp) balances the diffusion;
    # the y and z faces are insulated and the profile has no normal flux
    # there), and a solve from T = 0 with the x faces held returns it
    k, s_, dt = 0.8, 2.5, 0.05
    m = O.thermal_material(k=(k,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))
    c3, e3 = _box3(2)
    x = c3[:, 0]
    prof = 1 - x + s_ * x * (1 - x) / (2 * k)
    bx = np.nonzero((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12))[0].astype(np.int32)
    t = O.ThermalProblem3(c3, e3, m, bc=(bx, prof[bx]))
    t.source[:] = s_
    t.T0[:] = prof
    t.T[:] = prof
    t.residual(dt, Tref=0.0)
    free = np.setdiff1d(np.arange(len(c3)), bx)
    assert np.abs(t.R[free]).max() < 1e-13
    t.T[:] = 0.0
    st = t.solve(dt, tolerance=1e-12, max_iterations=200000, check_interval=50, Tref=0.0)
    assert st["converged"]
    assert np.allclose(t.T, prof, rtol=0, atol=1e-9)
