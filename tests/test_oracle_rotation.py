"""The oracle's rotate_stress! (src/stokes/stress_rotation.jl:16-61) and its
restatement of GeoParams.rotate_elastic_stress (GeoParams 0.7.21, not
present here), pinned by the reference's own tests
(test/test_stress_rotation.jl) and by scipy's rotation matrices. Linear
velocity fields are reproduced exactly by T7 and Hex27, so every quadrature
point sees the analytic vorticity."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as O


def _rot2(theta):
    return np.array([[math.cos(theta), -math.sin(theta)], [math.sin(theta), math.cos(theta)]])


def _sym2(t):
    return np.array([[t[0], t[2]], [t[2], t[1]]])


def _sym3(t):   # assembler order xx, yy, zz, xy, xz, yz
    return np.array([[t[0], t[3], t[4]], [t[3], t[1], t[5]], [t[4], t[5], t[2]]])


def _fixture2(vfun, n=2, tau0=(0.7, -0.4, 0.3)):
    coords, el2n = O.t7_mesh(n, n)
    pr = O.Problem(coords, el2n, O.material())
    v = np.array([vfun(c) for c in coords])
    pr.vx[:], pr.vy[:] = v[:, 0], v[:, 1]
    ne = len(el2n)
    tau = [np.full((ne, 7), t) for t in tau0]
    return pr, tau


def test_2d_no_vorticity_leaves_stress():
    # test_stress_rotation.jl:48-58: pure shear has ω = 0
    pr, tau = _fixture2(lambda c: (1.5 * c[0], -1.5 * c[1]))
    out = pr.rotate_stress(0.25, tau)
    for o, t in zip(out, tau):
        assert np.allclose(o, t, rtol=1e-14)


def test_2d_solid_body_rotation():
    # test_stress_rotation.jl:60-77: vx = -Ωy, vy = Ωx gives ω = -Ω and the
    # stress turns by θ = -ΩΔt; trace and determinant survive
    Om, dt = 0.8, 0.25
    pr, tau = _fixture2(lambda c: (-Om * c[1], Om * c[0]))
    out = pr.rotate_stress(dt, tau)
    R = _rot2(-Om * dt)
    E = R @ _sym2((0.7, -0.4, 0.3)) @ R.T
    for o, e in zip(out, (E[0, 0], E[1, 1], E[0, 1])):
        assert np.allclose(o, e, rtol=1e-13)
    xx, yy, xy = (o[0, 0] for o in out)
    assert xx + yy == pytest.approx(0.3, rel=1e-14)
    assert xx * yy - xy ** 2 == pytest.approx(0.7 * -0.4 - 0.09, rel=1e-13)


def test_2d_vanishing_step_is_the_identity():
    # test_stress_rotation.jl:79-88
    pr, tau = _fixture2(lambda c: (-0.8 * c[1], 0.8 * c[0]))
    out = pr.rotate_stress(0.0, tau)
    for o, t in zip(out, tau):
        assert np.array_equal(o, t)


def test_2d_general_linear_field_uses_the_antisymmetric_part():
    # a linear field = strain rate + spin: only ∂vx/∂y − ∂vy/∂x turns the stress
    A = np.array([[0.3, 1.1], [-0.4, -0.2]])   # ∇v
    dt = 0.6
    pr, tau = _fixture2(lambda c: tuple(A @ c), n=3, tau0=(0.2, 0.5, -0.7))
    out = pr.rotate_stress(dt, tau)
    R = _rot2(0.5 * (A[0, 1] - A[1, 0]) * dt)
    E = R @ _sym2((0.2, 0.5, -0.7)) @ R.T
    for o, e in zip(out, (E[0, 0], E[1, 1], E[0, 1])):
        assert np.allclose(o, e, rtol=1e-12)


def _fixture3(vfun, n=(2, 1, 1), tau0=(0.7, -0.4, -0.3, 0.3, -0.1, 0.2)):
    coords, el2n = O.hex27_mesh(*n)
    pr = O.Problem3(coords, el2n, O.material())
    v = np.array([vfun(c) for c in coords])
    for c in range(3):
        pr.v[c][:] = v[:, c]
    ne = len(el2n)
    tau = [np.full((ne, 27), t) for t in tau0]
    return pr, tau


TAU3 = (0.7, -0.4, -0.3, 0.3, -0.1, 0.2)   # (0.7, -0.4, -0.3, yz 0.2, xz -0.1, xy 0.3) of test_stress_rotation.jl:108


def test_3d_no_vorticity_leaves_stress():
    # test_stress_rotation.jl:112-120
    pr, tau = _fixture3(lambda c: (1.5 * c[0], -3.0 * c[1], 1.5 * c[2]))
    out = pr.rotate_stress(0.25, tau)
    for o, t in zip(out, tau):
        assert np.allclose(o, t, rtol=1e-14)


def test_3d_plane_flow_turns_opposite_to_2d():
    # test_stress_rotation.jl:122-140: curl (0, 0, 2Ω) turns the stress by +ΩΔt
    Om, dt = 0.8, 0.25
    pr, tau = _fixture3(lambda c: (-Om * c[1], Om * c[0], 0.0))
    out = pr.rotate_stress(dt, tau)
    R = np.array([[math.cos(Om * dt), -math.sin(Om * dt), 0], [math.sin(Om * dt), math.cos(Om * dt), 0], [0, 0, 1]])
    E = R @ _sym3(TAU3) @ R.T
    for o, e in zip(out, (E[0, 0], E[1, 1], E[2, 2], E[0, 1], E[0, 2], E[1, 2])):
        assert np.allclose(o, e, rtol=1e-13, atol=1e-15)
    # the 2-D convention for the same flow turns by -ΩΔt
    r2 = O.rotate_elastic_stress(-Om, (0.7, -0.4, 0.3), dt)
    assert not np.allclose(r2, [E[0, 0], E[1, 1], E[0, 1]], rtol=1e-13)
    E2 = _rot2(-Om * dt) @ _sym2((0.7, -0.4, 0.3)) @ _rot2(-Om * dt).T
    assert np.allclose(r2, [E2[0, 0], E2[1, 1], E2[0, 1]], rtol=1e-13)


def test_3d_rigid_rotation_about_an_oblique_axis():
    # test_stress_rotation.jl:142-160: v = a × x turns the stress by |a|Δt
    # about a (curl v = 2a); scipy's rotation vector gives the same matrix
    a = np.array([0.3, -0.5, 0.7])
    dt = 0.25
    pr, tau = _fixture3(lambda c: tuple(np.cross(a, c)))
    out = pr.rotate_stress(dt, tau)
    after = _sym3([o[0, 0] for o in out])
    before = _sym3(TAU3)
    assert not np.allclose(after, before)
    assert np.trace(after) == pytest.approx(np.trace(before), abs=1e-14)
    assert np.sum(after ** 2) == pytest.approx(np.sum(before ** 2), rel=1e-13)
    assert np.linalg.det(after) == pytest.approx(np.linalg.det(before), rel=1e-12)
    R = Rotation.from_rotvec(a * dt).as_matrix()
    assert np.allclose(after, R @ before @ R.T, rtol=1e-13, atol=1e-15)
    for o in out:
        assert np.allclose(o, o[0, 0], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("seed", [0, 1])
def test_3d_element_rotation_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    w = rng.normal(size=3)
    t = rng.normal(size=6)
    dt = rng.uniform(0.1, 2.0)
    R = Rotation.from_rotvec(0.5 * dt * w).as_matrix()
    E = R @ _sym3(t) @ R.T
    got = O.rotate_elastic_stress(w, t, dt)
    assert np.allclose(got, [E[0, 0], E[1, 1], E[2, 2], E[0, 1], E[0, 2], E[1, 2]], rtol=1e-13, atol=1e-15)
