"""The linearised Drucker-Prager return of the precision-2 correction operator
(femdr_kernels.hpp: stress2d's tangent output, the EF_TANGENT branch of the
element body) against central differences of the oracle's point stress
(rheology.jl:226-245, 272-334 restated in orc_stress_at_ip).

The kernel stores per yielding point (h, kq, pq, q = M t) for the trial
stress t and applies
    τ(x + δ) - τ(x) ≈ u - h M u - q (kq q·u + pq δP),   u = 2 ηve ε̇'(δ),
M = [[2,1,0],[1,2,0],[0,0,2]]. This test evaluates that expression from the
stored quantities' definitions and differentiates the oracle's stress in the
velocity gradient and in the pressure; a wrong sign or factor in h, kq, pq or
q fails it."""
import ctypes as C
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

EPS2 = 4.930380657631324e-32


def oracle_tau(gvx, gvy, tau_old, P, eta, G, Cc, phi, psi, ereg, Kb, dt):
    mat = O.material(eta=(eta,), G=(G,))
    dp = O.drucker_prager([phi], [psi], [Cc], [ereg], [Kb])
    tau = np.zeros(3)
    O.lib().orc_stress_at_ip(O.ptr(np.array(gvx, float)), O.ptr(np.array(gvy, float)), O.ptr(np.array([1.0, 0, 0])),
                             3, O.ptr(np.ones(3, np.int32), O.PI), C.byref(mat), C.byref(dp), dt,
                             O.ptr(np.array(tau_old, float)), P, O.ptr(tau))
    return tau


def tangent(t, P, eve, y0, sphi, invden):
    """(h, kq, pq, q) of the kernel's tangent output for trial stress t."""
    txx, tyy, txy = t
    tzz = -txx - tyy
    tII = math.sqrt((txx * txx + tyy * tyy + tzz * tzz) * 0.5 + txy * txy + EPS2)
    lam = ((tII - y0) - sphi * P) * invden
    tIIs = tII + EPS2
    h = 2 * eve * lam / (tIIs * 2)
    kq = (eve / tIIs) * (invden - lam / tIIs) / (2 * tII)
    pq = -(eve * invden * sphi) / tIIs
    return h, kq, pq, np.array([2 * txx + tyy, txx + 2 * tyy, 2 * txy])


def test_linearised_return_matches_oracle_differences():
    rng = np.random.default_rng(4)
    dt, checked = 0.1, 0
    for _ in range(200):
        eta, G = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-0.5, 1)
        Cc, phi, psi = rng.uniform(0.1, 0.6), math.radians(rng.uniform(10, 40)), math.radians(rng.uniform(0, 10))
        ereg, Kb = rng.uniform(1e-2, 0.3), rng.uniform(1.0, 50.0)
        gvx, gvy = rng.standard_normal(2) * 3, rng.standard_normal(2) * 3
        tau_old, P = rng.standard_normal(3) * 0.3, rng.uniform(-0.3, 0.5)
        eve = 1 / (1 / eta + 1 / (G * dt))
        inv2Gdt = 1 / (2 * G * dt)
        invden = 1 / ((eve + ereg) + ((Kb * dt) * (-math.sin(psi))) * (-math.sin(phi)))
        y0, sphi = math.cos(phi) * Cc, math.sin(phi)
        exx, eyy, exy = gvx[0], gvy[1], 0.5 * (gvx[1] + gvy[0])
        tr = (exx + eyy) / 3
        t = 2 * eve * (np.array([exx - tr, eyy - tr, exy]) + tau_old * inv2Gdt)
        tz = -t[0] - t[1]
        tII = math.sqrt((t[0] ** 2 + t[1] ** 2 + tz ** 2) * 0.5 + t[2] ** 2)
        F = tII - y0 - sphi * P
        if not (F > 0.05 * (abs(y0) + 1) and F * invden > 0):
            continue   # the differences must not cross the yield surface
        h, kq, pq, q = tangent(t, P, eve, y0, sphi, invden)
        M = np.array([[2.0, 1, 0], [1, 2, 0], [0, 0, 2]])
        args = (eta, G, Cc, phi, psi, ereg, Kb, dt)
        e = 1e-6
        for _k in range(3):
            dgx, dgy = rng.standard_normal(2), rng.standard_normal(2)
            fd = (oracle_tau(gvx + e * dgx, gvy + e * dgy, tau_old, P, *args)
                  - oracle_tau(gvx - e * dgx, gvy - e * dgy, tau_old, P, *args)) / (2 * e)
            dxx, dyy, dxy = dgx[0], dgy[1], 0.5 * (dgx[1] + dgy[0])
            dtr = (dxx + dyy) / 3
            u = 2 * eve * np.array([dxx - dtr, dyy - dtr, dxy])
            lin = u - h * (M @ u) - q * (kq * (q @ u))
            assert np.abs(lin - fd).max() <= 1e-6 * (np.abs(fd).max() + 1), (lin, fd)
        fdP = (oracle_tau(gvx, gvy, tau_old, P + e, *args) - oracle_tau(gvx, gvy, tau_old, P - e, *args)) / (2 * e)
        assert np.abs(-q * pq - fdP).max() <= 1e-6 * (np.abs(fdP).max() + 1), (-q * pq, fdP)
        checked += 1
    assert checked > 50
