"""Kernel-level entry points of the inner DR step against the oracle:
fem_dp_return_map_batch (the VEP Drucker-Prager stress at independent points,
rheology.jl:226-245, 272-334) against the oracle's point stress
(orc_stress_at_ip, pinned by the reference's return-map KATs), and
fem_dr_velocity_update (helpers.jl:31-41) against the oracle's update
(orc_dr_update) on the same arrays."""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def _oracle_point(gvx, gvy, tau_old, Pq, eta, G, Cc, phi, psi, ereg, Kb, dt):
    mat = O.material(eta=(eta,), G=(G,))
    dp = O.drucker_prager([phi], [psi], [Cc], [ereg], [Kb])
    tau = np.zeros(3)
    O.lib().orc_stress_at_ip(O.ptr(np.array(gvx)), O.ptr(np.array(gvy)), O.ptr(np.array([1.0, 0, 0])), 3,
                             O.ptr(np.ones(3, np.int32), O.PI), C.byref(mat), C.byref(dp), dt,
                             O.ptr(np.array(tau_old, float)), Pq, O.ptr(tau))
    return tau


def test_dp_return_map_batch_matches_oracle():
    import femtools_amd as fa
    rng = np.random.default_rng(9)
    n = 400
    gvx = rng.standard_normal((n, 2))
    gvy = rng.standard_normal((n, 2))
    tau_old = rng.standard_normal((3, n)) * 0.5
    P = rng.uniform(-0.5, 1.5, n)
    eta = 10 ** rng.uniform(-2, 2, n)
    G = np.where(rng.uniform(size=n) < 0.3, math.inf, 10 ** rng.uniform(-1, 1, n))
    Cc = rng.uniform(0.0, 1.5, n)
    phi = np.radians(rng.uniform(0, 40, n))
    psi = np.radians(rng.uniform(0, 10, n))
    ereg = rng.uniform(1e-3, 0.2, n)
    Kb = rng.uniform(1.0, 100.0, n)
    dt = 0.1
    exx, eyy, exy = gvx[:, 0], gvy[:, 1], 0.5 * (gvx[:, 1] + gvy[:, 0])
    txx, tyy, txy, y = fa.dp_return_map_batch(fa.default_context(), dt, exx=exx, eyy=eyy, exy=exy, tau_old=tau_old,
                                              P=P, eta=eta, G=G, C_=Cc, phi=phi, psi=psi, eta_reg=ereg, Kb=Kb)
    ref = np.array([_oracle_point(gvx[i], gvy[i], tau_old[:, i], P[i], eta[i], G[i], Cc[i], phi[i], psi[i], ereg[i],
                                  Kb[i], dt) for i in range(n)])
    got = np.stack([txx, tyy, txy], axis=1)
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-300
    assert np.all(np.abs(got - ref) <= 1e-13 * scale)
    assert 0.2 < y.mean() < 0.9           # both branches exercised


def test_dr_velocity_update_matches_oracle():
    import femtools_amd as fa
    ctx = fa.default_context()
    coords, el2n = O.t7_mesh(3, 3)
    pr = O.Problem(coords, el2n, O.material())
    n = pr.nnodes
    rng = np.random.default_rng(2)
    rate, v, R, PC = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n), rng.uniform(0.5, 2, n)
    alpha, beta = 0.013, 0.87
    pr.rate_x[:], pr.vx[:], pr.Rvx[:], pr.PCx[:] = rate, v, R, PC
    pr.PCy[:] = 1.0
    O.lib().orc_dr_update(C.byref(pr.struct()), O.ptr(np.array([alpha, alpha])), O.ptr(np.array([beta, beta])))
    d = [fa.device_array(ctx, a) for a in (rate, v, R, PC)]
    fa.dr_velocity_update(ctx, d[0], d[1], d[2], d[3], alpha, beta)
    ctx.sync()
    # the device contracts a*b + c into one fused multiply-add: equal to the last bit or one rounding apart
    for got, ref in ((d[0].numpy(), pr.rate_x), (d[1].numpy(), pr.vx)):
        assert np.all(np.abs(got - ref) <= 4e-16 * np.abs(ref).max())
