"""GPU: the mixed-precision solve (FP32 DR iterations on corrections of the
FP64 solution, FP64 true residuals and convergence test).

The bar for this form is the reference's: the solution must reach the same
residual tolerance. Each test re-evaluates the FP64 residual of the returned
fields with the CPU oracle and applies the reference's criterion to it, and
prints the per-field error against the oracle's FP64 solution of the same
problem (both converged to the same tolerance)."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")
from femtools_amd import problems as PB  # noqa: E402

from test_gpu_parity import make_pair, rel_err  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_errors(pr, vx, vy, P, dt):
    """err_v, err_P of the reference's outer iteration (DR.jl:535-557: residuals
    without Pnum, Rv pinned to zero on Dirichlet nodes) for the fields
    (vx, vy, P), evaluated by the oracle in FP64."""
    pr.vx[:], pr.vy[:], pr.P[:] = vx, vy, P
    pr.pressure_scaling(1.0, dt)
    pr.pressure_residual(dt)
    pr.momentum_residual(dt)
    Rx, Ry = pr.Rvx.copy(), pr.Rvy.copy()
    Rx[pr.bc[0][0]] = 0.0
    Ry[pr.bc[1][0]] = 0.0
    err_P = np.linalg.norm(pr.RP / pr.MP) / math.sqrt(len(pr.RP))
    err_v = max(np.linalg.norm(Rx), np.linalg.norm(Ry)) / (2 * math.sqrt(len(Rx)))
    return err_v, err_P


def _viscous(n):
    from test_gpu_parity import _viscous_pair
    return _viscous_pair(n)


@pytest.mark.parametrize("n", [16, 24, 48])   # 24: not a multiple of the 16-cell tile width
def test_mixed_viscous_inclusion_reaches_tolerance(n):
    tol = 1e-8
    say(f"[gpu] mixed precision viscous inclusion n={n}")
    w, pr, dr, mesh, (bx, vx, by, vy) = _viscous(n)
    st_o = pr.solve(w.dt, pressure_factor=w.pressure_factor, tolerance=tol, check_interval=w.check_interval)
    assert st_o["converged"] and st_o["iter"] > 0
    ref = (pr.vx.copy(), pr.vy.copy(), pr.P.copy())
    st = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy),
                  dt=w.dt, pressure_factor=w.pressure_factor, tolerance=tol, check_interval=w.check_interval,
                  precision="fp32")
    got = (dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy())
    errs = [rel_err(a, b) for a, b in zip(got, ref)]
    say(f"  oracle FP64: iter {st_o['iter']} err_abs {st_o['err_abs']:.3e} | mixed: corrections {st['itPH']} "
        f"FP32 iter {st['iter']} err_abs {st['err_abs']:.3e} err_rel {st['err_rel']:.3e}")
    say(f"  mixed vs oracle FP64 (max rel): vx {errs[0]:.2e} vy {errs[1]:.2e} P {errs[2]:.2e}")
    assert st["converged"]
    # the oracle's FP64 residual of the returned fields meets the criterion
    ev, eP = oracle_errors(pr, *got, w.dt)
    assert max(ev, eP) == pytest.approx(st["err_abs"], rel=1e-6)
    assert min(st["err_abs"], st["err_rel"]) < tol
    # both solutions lie within the tolerance of the discrete solution
    assert max(errs) < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp32-tangent"])
@pytest.mark.parametrize("n", [12, 16])
def test_mixed_visco_elastic_plastic_time_steps(n, precision):
    # the reduced VEP series of the parity tests (4 inclusions, C=0.5, three
    # steps with the stress-history hand-off); 16 x 16 runs the FP32 inner
    # iterations in 16 x 8-cell tiles, 12 x 12 with the element-slot gather
    tol = 1e-8
    say(f"[gpu] mixed precision VEP n={n} {precision}")
    w = PB.vep_inclusions(n, ninc=4, seed=3, nsteps=3, C=0.5)
    coords, el2n = O.t7_mesh(n, n)
    ph = PB.inclusion_phases(coords, el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(coords)
    ne = len(el2n)
    zero = tuple(np.zeros((ne, 7)) for _ in range(3))
    dp = dict(phi=w.plastic["phi"], psi=w.plastic["psi"], C_=w.plastic["C"], eta_reg=w.plastic["eta_reg"],
              Kb=w.plastic["Kb"])
    pr, dr, mesh, dp_o, dp_g = make_pair(n, n, w.eta, G=w.G, rho0=w.rho0, etab=w.etab, K=w.K, g=w.g, dp=dp,
                                         phases=ph, tau_old=zero, bc=((bx, vx), (by, vy)), random_state=False,
                                         CFL_v=w.CFL_v)
    v0x, v0y = PB.pure_shear_field(coords)
    pr.vx[:], pr.vy[:] = v0x, v0y
    dr.v.x.copyto(v0x)
    dr.v.y.copyto(v0y)
    bcx, bcy = fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)
    kw = w.solve_kwargs()
    kw["tolerance"] = tol
    for step in range(w.nsteps):
        # both solvers start the step from the mixed solution's state
        tau_old = tuple(getattr(dr.tau_old, c).numpy().reshape(ne, 7) for c in ("xx", "yy", "xy"))
        P0 = dr.P.numpy()
        pr.vx[:], pr.vy[:], pr.P[:] = dr.v.x.numpy(), dr.v.y.numpy(), P0
        pr.arrays["P0"][:] = P0
        pr.tau_old = tuple(t.copy() for t in tau_old)
        dr.P0.copyto(P0)
        st_o = pr.solve(w.dt, CFL_v=w.CFL_v, **kw)
        ref = (pr.vx.copy(), pr.vy.copy(), pr.P.copy())
        st = fa.solve(dr, mesh, bcx, bcy, dt=w.dt, plastic=dp_g, precision=precision, max_iterations=200_000,
                      iterMax=200_000, **kw)
        got = (dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy())
        errs = [rel_err(a, b) for a, b in zip(got, ref)]
        say(f"  step {step}: oracle FP64 iter {st_o['iter']} | mixed corrections {st['itPH']} FP32 iter {st['iter']} "
            f"err_abs {st['err_abs']:.3e}; vs oracle: vx {errs[0]:.2e} vy {errs[1]:.2e} P {errs[2]:.2e}")
        assert st["converged"] and st_o["converged"]
        ev, eP = oracle_errors(pr, *got, w.dt)
        assert max(ev, eP) == pytest.approx(st["err_abs"], rel=1e-6)
        assert min(st["err_abs"], st["err_rel"]) < tol
        # two solutions meeting 1e-8 differ by a few times the distance of
        # one of them to the exact discrete solution: the oracle's own step-0
        # solution at 1e-8 lies 2.9e-6 (max rel) from its solution at 1e-11
        assert max(errs) < 3e-5
        fa.update_stokes_current_stress(dr, mesh, w.dt, plastic=dp_g)
        for c in ("xx", "yy", "xy"):
            getattr(dr.tau_old, c).copyto(getattr(dr.tau, c).numpy())


def test_mixed_precision_argument_checks():
    w, pr, dr, mesh, (bx, vx, by, vy) = _viscous(4)
    with pytest.raises(ValueError):
        fa.solve(dr, mesh, fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy), dt=1.0,
                 precision="fp16")
