"""GPU: rotate_stress! (src/stokes/stress_rotation.jl:16-61) through the C
ABI against the oracle (random fields, 1e-13) and the reference's closed
forms (test/test_stress_rotation.jl), in 2-D (T7, structured and uploaded
meshes) and 3-D (Hex27)."""
import math
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

TAU2 = ("xx", "yy", "xy")
TAU3 = ("xx", "yy", "zz", "xy", "xz", "yz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("structured", [True, False])
def test_gpu_rotate_stress_2d_parity(structured):
    say(f"[gpu] rotate_stress 2-D parity structured={structured}")
    rng = np.random.default_rng(7)
    n = 9
    coords, el2n = O.t7_mesh(n, n)
    if not structured:   # a sheared, stretched copy of the box mesh, uploaded
        coords = coords @ np.array([[1.2, 0.1], [0.3, 0.9]])
    mesh = fa.t7_box_mesh(n, n) if structured else fa.t7_mesh_from_arrays(coords, el2n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial())
    pr = O.Problem(coords, el2n, O.material())
    ne = len(el2n)
    pr.vx[:] = np.sin(3 * coords[:, 0]) * coords[:, 1] + 0.1 * rng.normal(size=len(coords))
    pr.vy[:] = np.cos(2 * coords[:, 1]) * coords[:, 0] + 0.1 * rng.normal(size=len(coords))
    dr.v.x.copyto(pr.vx)
    dr.v.y.copyto(pr.vy)
    tau = [rng.normal(size=(ne, 7)) for _ in range(3)]
    for name, t in zip(TAU2, tau):
        getattr(dr.tau, name).copyto(t)
    dt = 0.7
    fa.rotate_stress(dr, mesh, dt)
    ref = pr.rotate_stress(dt, tau)
    errs = [rel(getattr(dr.tau_old, name).numpy().reshape(ne, 7), r) for name, r in zip(TAU2, ref)]
    say("  " + " ".join(f"{nm} {e:.2e}" for nm, e in zip(TAU2, errs)))
    assert max(errs) < 1e-13
    # dr.tau is read, not written
    for name, t in zip(TAU2, tau):
        assert np.array_equal(getattr(dr.tau, name).numpy().reshape(ne, 7), t)


def test_gpu_rotate_stress_2d_solid_body():
    # test_stress_rotation.jl:60-88
    say("[gpu] rotate_stress 2-D solid-body rotation")
    Om, dt = 0.8, 0.25
    mesh = fa.t7_box_mesh(3, 2)
    dr = fa.StokesDR(mesh, fa.StokesMaterial())
    c = mesh.coords
    dr.v.x.copyto(-Om * c[:, 1])
    dr.v.y.copyto(Om * c[:, 0])
    for name, v in zip(TAU2, (0.7, -0.4, 0.3)):
        getattr(dr.tau, name).fill(v)
    fa.rotate_stress(dr, mesh, dt)
    th = -Om * dt
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    E = R @ np.array([[0.7, 0.3], [0.3, -0.4]]) @ R.T
    for name, e in zip(TAU2, (E[0, 0], E[1, 1], E[0, 1])):
        assert np.allclose(getattr(dr.tau_old, name).numpy(), e, rtol=1e-13)
    fa.rotate_stress(dr, mesh, 0.0)
    for name, v in zip(TAU2, (0.7, -0.4, 0.3)):
        assert np.array_equal(getattr(dr.tau_old, name).numpy(), np.full((mesh.nels, 7), v))


def test_gpu_rotate_stress_3d_parity_and_rigid_rotation():
    say("[gpu] rotate_stress 3-D parity")
    rng = np.random.default_rng(8)
    nx, ny, nz = 3, 2, 2
    mesh = fa.hex27_box_mesh(nx, ny, nz)
    coords, el2n = O.hex27_mesh(nx, ny, nz)
    assert np.array_equal(mesh.coords, coords)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(g=(0.0, 0.0, 0.0)))
    pr = O.Problem3(coords, el2n, O.material())
    ne = len(el2n)
    for k, comp in enumerate("xyz"):
        pr.v[k][:] = np.sin((k + 2) * coords[:, k]) * coords[:, (k + 1) % 3] + 0.1 * rng.normal(size=len(coords))
        getattr(dr.v, comp).copyto(pr.v[k])
    tau = [rng.normal(size=(ne, 27)) for _ in range(6)]
    for name, t in zip(TAU3, tau):
        getattr(dr.tau, name).copyto(t)
    dt = 0.9
    fa.rotate_stress(dr, mesh, dt)
    ref = pr.rotate_stress(dt, tau)
    errs = [rel(getattr(dr.tau_old, name).numpy().reshape(ne, 27), r) for name, r in zip(TAU3, ref)]
    say("  " + " ".join(f"{nm} {e:.2e}" for nm, e in zip(TAU3, errs)))
    assert max(errs) < 1e-13
    # v = a × x: rotation by the vector aΔt (test_stress_rotation.jl:142-160)
    a = np.array([0.3, -0.5, 0.7])
    v = np.cross(a, coords)
    for k, comp in enumerate("xyz"):
        getattr(dr.v, comp).copyto(v[:, k])
    t0 = (0.7, -0.4, -0.3, 0.3, -0.1, 0.2)
    for name, t in zip(TAU3, t0):
        getattr(dr.tau, name).fill(t)
    fa.rotate_stress(dr, mesh, 0.25)
    T = np.array([[t0[0], t0[3], t0[4]], [t0[3], t0[1], t0[5]], [t0[4], t0[5], t0[2]]])
    R = Rotation.from_rotvec(a * 0.25).as_matrix()
    E = R @ T @ R.T
    for name, e in zip(TAU3, (E[0, 0], E[1, 1], E[2, 2], E[0, 1], E[0, 2], E[1, 2])):
        assert np.allclose(getattr(dr.tau_old, name).numpy(), e, rtol=1e-12, atol=1e-14)
    # no vorticity: untouched
    for k, comp in enumerate("xyz"):
        getattr(dr.v, comp).copyto((1.5, -3.0, 1.5)[k] * coords[:, k])
    fa.rotate_stress(dr, mesh, 0.25)
    for name, t in zip(TAU3, t0):
        assert np.allclose(getattr(dr.tau_old, name).numpy(), t, rtol=1e-14)


def test_gpu_time_step_handoff_on_device():
    # the step boundary of a visco-elastic series: τ at the points from the
    # converged state (update_stokes_current_stress!, DR.jl:820-897), then
    # copyto!(τ_old, τ) (symmetric_tensor.jl:133-150) and P0 <- P, all on the
    # device; the next step's oracle solve from the same history agrees
    say("[gpu] time-step handoff on the device")
    from test_gpu_parity import make_pair
    n = 8
    coords, el2n = O.t7_mesh(n, n)
    ne = len(el2n)
    (bx, vx), (by, vy) = O.pure_shear_bc(coords)
    zero = tuple(np.zeros((ne, 7)) for _ in range(3))
    cen = coords[el2n[:, :3]].mean(axis=1)
    ph = np.where(((cen - 0.5) ** 2).sum(axis=1) < 0.2 ** 2, 2, 1).astype(np.int32)
    pr, dr, mesh, _, _ = make_pair(n, n, (1.0, 5.0), G=(1.0, 2.0), etab=(10.0, 10.0), phases=ph, tau_old=zero,
                                   bc=((bx, vx), (by, vy)), random_state=False)
    # a start that meets the boundary values (pure shear)
    pr.vx[:], pr.vy[:] = coords[:, 0] - 0.5, -(coords[:, 1] - 0.5)
    dr.v.x.copyto(pr.vx)
    dr.v.y.copyto(pr.vy)
    bcs = (fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy))
    dt = 0.3
    kw = dict(dt=dt, tolerance=1e-10, check_interval=20)
    for step in range(2):
        st = fa.solve(dr, mesh, *bcs, **kw)
        so = pr.solve(dt, tolerance=1e-10, check_interval=20)
        say(f"  step {step}: gpu iter {st['iter']} oracle iter {so['iter']}")
        assert st["converged"] and st["iter"] == so["iter"] > 0
        assert rel(dr.v.x.numpy(), pr.vx) < 1e-10
        fa.update_stokes_current_stress(dr, mesh, dt)
        tau = [t.copy() for t in pr.update_current_stress(dt)]
        fa.copyto(dr.tau_old, dr.tau)
        fa.copyto(dr.P0, dr.P)
        for name, t in zip(TAU2, tau):
            assert np.array_equal(getattr(dr.tau_old, name).numpy(), getattr(dr.tau, name).numpy())
            assert rel(getattr(dr.tau_old, name).numpy(), t) < 1e-10
        assert np.array_equal(dr.P0.numpy(), dr.P.numpy())
        pr.tau_old = tuple(tau)
        pr.arrays["P0"][:] = pr.P
    with pytest.raises(fa.DimensionMismatch):
        fa.copyto(dr.P0, dr.v.x)
