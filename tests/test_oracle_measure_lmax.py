"""The oracle's measure_λmax option of the forward solve (power iteration on
the Jacobi-scaled velocity operator, DR.jl:210-211, 375-378, 459-490, 728-760;
estimate_velocity_λmax, adjoint_operator.jl:240-281), pinned by the
reference's own tests (test_solver_convergence_api.jl:190-238,
test_adjoint_operator.jl:56-67) and by the spectrum of the assembled operator:

* with and without it the buoyancy case converges to the same velocity
  (rtol 1e-5), the Gershgorin bound is the same, the measured λmax never
  exceeds it and the power steps are counted; a viscous solve assembles the
  Jacobian once;
* the measured value is the largest eigenvalue of P^-1/2 A P^-1/2 on the free
  rows (dense eigenvalues of the operator assembled from the element blocks),
  from below and within the power iteration's tolerance.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from test_oracle_adjoint import buoyancy_case

INF = math.inf
FWD = dict(pressure_factor=20.0, tolerance=1e-9, check_interval=100, rel_drop0=0.1, CFL_v=0.9, c_fact=0.7)


def problem(case):
    coords, el2n, ph, xw, yw, obs = case
    mat = O.material(eta=(1.0, 1.0), etab=(INF, INF), alpha=(0.0, 0.0), rho0=(1.0, 2.0), K=(INF, INF),
                     g=(0.0, -1.0), Tref=0.0)
    return O.Problem(coords, el2n, mat, phases_v=ph, phases_P=ph,
                     bc=((xw, np.zeros(len(xw))), (yw, np.zeros(len(yw)))))


def test_measured_lambda_max_converges_to_the_same_solution():
    case = buoyancy_case()
    pb = problem(case)
    bound = pb.solve(1.0, **FWD)
    assert bound["converged"] and bound["lambda_max_iterations"] == 0
    assert bound["lambda_max"] == bound["lambda_max_gershgorin"]
    pp = problem(case)
    power = pp.solve(1.0, **FWD, measure_lmax=True)
    assert power["converged"] and power["lambda_max_iterations"] > 0
    assert power["lambda_max"] <= power["lambda_max_gershgorin"]
    assert power["lambda_max_gershgorin"] == pytest.approx(bound["lambda_max_gershgorin"], rel=1e-14)
    assert power["jacobian_assemblies"] == 1
    assert np.allclose(pp.vy, pb.vy, rtol=1e-5, atol=1e-5 * np.abs(pb.vy).max())
    assert np.allclose(pp.vx, pb.vx, rtol=1e-5, atol=1e-5 * np.abs(pb.vx).max())
    assert np.any(pp.vy != 0)


def _dense_scaled_operator(pr, dt):
    """P^-1/2 A P^-1/2 restricted to the free rows, assembled from the element blocks."""
    blocks = pr.adjoint_blocks(dt)   # A and the Jacobian diagnostics of the state
    nn = pr.nnodes
    M = np.zeros((2 * nn, 2 * nn))
    for e in range(pr.nels):
        dofs = np.concatenate([pr.el2n[e], nn + pr.el2n[e]])
        M[np.ix_(dofs, dofs)] += blocks["A"][e]
    pc = np.concatenate([pr.PCx, pr.PCy])
    M = M / np.sqrt(pc)[:, None] / np.sqrt(pc)[None, :]
    fixed = np.concatenate([pr.bc[0][0], nn + pr.bc[1][0]])
    free = np.setdiff1d(np.arange(2 * nn), fixed)
    return M[np.ix_(free, free)]


@pytest.mark.parametrize("rtol,gap", [(1e-2, 0.1), (1e-12, 1e-8)])
def test_measured_value_is_the_top_of_the_scaled_spectrum(rtol, gap):
    # the power estimate approaches the top eigenvalue from below; the stopping
    # test bounds the last increment, not the error, hence the looser gap
    case = buoyancy_case()
    pr = problem(case)
    st = pr.solve(1.0, **FWD, measure_lmax=True, lmax_power_rtol=rtol, lmax_power_iterations=20000, lmax_safety=1.0)
    assert st["converged"]
    lam_true = np.abs(np.linalg.eigvals(_dense_scaled_operator(pr, 1.0))).max()
    meas = st["lambda_max"]
    assert meas <= lam_true * (1 + 1e-12)
    assert meas >= lam_true * (1 - gap)
    assert meas < st["lambda_max_gershgorin"]
