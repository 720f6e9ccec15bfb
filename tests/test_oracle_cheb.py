"""Closed-form checks of the oracle's Chebyshev scalars (_stokes_cheb,
DR.jl:39-42) and spectral estimate (_stokes_λmin, DR.jl:32-37), the two
scalar pieces of the DYREL iteration that whole-solve comparisons only pin
indirectly.

* At the start of a solve λmin = 0 and the reference states the step as
  α = Δτ², β = 1 (DR.jl:491-493).
* The inner update rate ← β rate + R/PC, v ← v − α rate applied to one
  eigenmode R = λ v (PC = 1) is the 2x2 recurrence [[β, λ], [−αβ, 1 − αλ]].
  With Δτ = 2 CFL/√λmax it must contract every mode λ in (0, λmax] for
  CFL < 1 and, with damping (λmin > 0), blow up above λ = 4/Δτ² — the
  stability edge of the scheme. A wrong factor or sign in α or β moves that
  edge.
* λmin is the Rayleigh quotient of the operator on the last step: for
  R − R0 = −α PC Λ rate (one update of a diagonal operator Λ) it returns λ_k
  exactly when rate is the k-th eigenvector and a value between the smallest
  and largest λ otherwise; a zero step gives 0.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import pyoracle as O  # noqa: E402


def spectral_radius(alpha, beta, lam):
    M = np.array([[beta, lam], [-alpha * beta, 1.0 - alpha * lam]])
    return max(abs(np.linalg.eigvals(M)))


def test_cheb_start_of_solve():
    for dtau in (1e-3, 0.07, 1.3):
        a, b = O.stokes_cheb(dtau, 0.0, 0.9)
        assert a == pytest.approx(dtau * dtau, rel=1e-15) and b == 1.0


@pytest.mark.parametrize("lmax,lmin,cfl", [(10.0, 0.1, 0.99), (2.5e4, 3.0, 0.9), (1.0, 1e-4, 0.5)])
def test_cheb_contracts_up_to_lmax_and_not_beyond(lmax, lmin, cfl):
    dtau = 2 / np.sqrt(lmax) * cfl
    a, b = O.stokes_cheb(dtau, lmin, 0.9)
    for lam in np.geomspace(lmin, lmax, 60):
        assert spectral_radius(a, b, lam) < 1.0, lam
    edge = 4 / dtau ** 2           # = lmax / cfl^2 > lmax
    assert spectral_radius(a, b, 1.02 * edge) > 1.0
    assert spectral_radius(a, b, 0.98 * edge) < 1.0


def test_lmin_rayleigh_quotient():
    rng = np.random.default_rng(7)
    n = 50
    lam = rng.uniform(0.5, 20.0, n)
    PC = rng.uniform(0.1, 3.0, n)
    R0 = rng.standard_normal(n)
    step = 0.013
    for k in (0, 17, 49):
        rate = np.zeros(n)
        rate[k] = -2.5
        R = R0 - step * PC * lam * rate
        assert O.stokes_lmin(step, rate, R, R0, PC) == pytest.approx(lam[k], rel=1e-13)
    rate = rng.standard_normal(n)
    R = R0 - step * PC * lam * rate
    got = O.stokes_lmin(step, rate, R, R0, PC)
    assert got == pytest.approx(np.sum(lam * rate ** 2) / np.sum(rate ** 2), rel=1e-12)
    assert lam.min() < got < lam.max()
    assert O.stokes_lmin(0.0, rate, R, R0, PC) == 0.0
    assert O.stokes_lmin(step, np.zeros(n), R, R0, PC) == 0.0
