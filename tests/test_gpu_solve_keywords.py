"""GPU: solve!'s array keywords through the C ABI (fem_stokes_set_solve_inputs,
fem_stokes_set_pressure_source) against the oracle: residuals, Jacobian
diagnostics and pressure scale with a viscosity at the points, nodal loads and
a previous-step pressure at the points to 1e-12; solves with the same
iteration counts and history rows; the closed forms of the reference's tests
(test_stokes_quadrature_viscosity.jl, test_stokes.jl:506-580,
test_solver_convergence_api.jl:300-328) and their argument errors."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


VISC = dict(eta=(2.0,), etab=(math.inf,), G=(math.inf,))


def _pair(n, rho0=0.0, g=(0.0, 0.0), mat=None, **okw):
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    mat = mat or dict(VISC, rho0=(rho0,))
    pr = O.Problem(coords, el2n, O.material(g=g, **mat), **okw)
    mesh = fa.t7_box_mesh(n, n)
    assert np.array_equal(mesh.coords, coords)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(g=g, **mat))
    return pr, dr, mesh, gam


def _gravity_load(n):
    pr, dr, mesh, gam = _pair(n, rho0=1.0, g=(0.0, 1.0))
    fa.residual(dr, 1.0)
    return tuple(-getattr(dr.Rv, c).numpy() for c in "xy")


def test_gpu_keyword_residuals_jacobian_and_scale():
    say("[gpu] residual / Jacobian / scale with point viscosity, loads and point P_old")
    n = 4
    rng = np.random.default_rng(0)
    pr, dr, mesh, gam = _pair(n, mat=dict(eta=(2.0,), etab=(3.0,), G=(math.inf,), K=(3.0,), rho0=(1.0,)),
                              g=(0.2, -1.0))
    ne = pr.nels
    eta = 1.0 + rng.random((ne, 7))
    bf = (rng.normal(size=pr.nnodes), rng.normal(size=pr.nnodes))
    Pold = rng.normal(size=(ne, 7))
    pr.vx[:] = rng.normal(size=pr.nnodes)
    pr.vy[:] = rng.normal(size=pr.nnodes)
    pr.P[:] = rng.normal(size=3 * ne)
    dr.v.x.copyto(pr.vx)
    dr.v.y.copyto(pr.vy)
    dr.P.copyto(pr.P)
    pr.eta_ip, pr.body_force, pr.Pold_ip = eta, bf, Pold
    pr.pressure_scaling(5.0, 0.5, finite_K=True)
    fa.assemble_viscosity_weighted_pressure_scaling(dr, 5.0, 0.5, finite_K=True, eta=eta)
    errs = dict(gammaP=rel(dr.gammaP.numpy(), pr.gammaP))
    pr.pressure_residual(0.5, finite_K=True)
    Pn = pr.gammaP * pr.RP / pr.MP
    pr.momentum_residual(0.5, Pnum=Pn)
    fa.residual(dr, 0.5, with_pnum=True, finite_K=True, viscosity=eta, body_force=bf, P_old=Pold)
    errs.update(RP=rel(dr.RP.numpy(), pr.RP), Rx=rel(dr.Rv.x.numpy(), pr.Rvx), Ry=rel(dr.Rv.y.numpy(), pr.Rvy))
    pr.augmented_jacobian(0.5, finite_K=True)
    fa.jacobian_diagnostics(dr, 0.5, finite_K=True, viscosity=eta)
    errs.update(Jx=rel(dr.dRvdv.x.numpy(), pr.dRdvx), Dx=rel(dr.PC_v.x.numpy(), pr.PCx),
                Jy=rel(dr.dRvdv.y.numpy(), pr.dRdvy), Dy=rel(dr.PC_v.y.numpy(), pr.PCy))
    # nodal P_old in place of dr.P0
    P0 = rng.normal(size=3 * ne)
    pr.Pold_ip = None
    pr.P0[:] = P0
    pr.pressure_residual(0.5, finite_K=True)
    fa.residual(dr, 0.5, finite_K=True, P_old=P0)
    errs["RP_nodal"] = rel(dr.RP.numpy(), pr.RP)
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-12
    # the inputs were cleared: a plain residual is the plain one again
    pr.eta_ip, pr.body_force, pr.P0[:] = None, None, 0.0
    pr.pressure_residual(0.5, finite_K=True)
    pr.momentum_residual(0.5)
    fa.residual(dr, 0.5, finite_K=True)
    assert rel(dr.Rv.x.numpy(), pr.Rvx) < 1e-12 and rel(dr.RP.numpy(), pr.RP) < 1e-12


def test_gpu_body_force_hydrostatic_solve_and_history():
    # test_stokes_quadrature_viscosity.jl:51-60 and the history layout of :100-113
    say("[gpu] hydrostatic solve with a point viscosity and nodal loads")
    n = 8
    load = _gravity_load(n)
    pr, dr, mesh, gam = _pair(n)
    zero = np.zeros(len(gam))
    eta = np.full((pr.nels, 7), 2.0)
    bcs = (fa.DirichletBoundaryCondition(gam, zero), fa.DirichletBoundaryCondition(gam, zero))
    st = fa.solve(dr, mesh, *bcs, dt=1.0, pressure_factor=50.0, check_interval=10, tolerance=1e-11,
                  viscosity=eta, body_force=load, collect_history=True)
    P = dr.P.numpy()
    MP = dr.M_P.numpy()
    Pc = pr.coords[pr.el2n[:, :3]].reshape(-1, 2)
    assert st["converged"]
    assert np.allclose(P - (P * MP).sum() / MP.sum(), Pc[:, 1] - 0.5, atol=1e-7)
    assert np.abs(dr.v.x.numpy()).max() < 1e-7 and np.abs(dr.v.y.numpy()).max() < 1e-7
    # the history layout: every outer and inner check, in order (this v = 0
    # solution ends at the rounding floor, where iterate trajectories of two
    # implementations part; the oracle comparison is the inclusion test below)
    h = st["history"]
    assert h[0]["iter"] == 0 and h[-1]["iter"] == st["iter"]
    assert h[-1]["err_P"] == pytest.approx(st["err_P"])
    assert all(a["iter"] <= b["iter"] for a, b in zip(h, h[1:]))
    assert all(r["iter"] % 10 == 0 for r in h)
    for r in h:
        assert max(r["err_v_components"]) == pytest.approx(r["err_v"], rel=1e-14)


def test_gpu_keywords_on_a_viscous_inclusion_with_history():
    # the viscous-inclusion parity case with the phase viscosity sampled at the
    # points: iteration counts, history rows and fields. (A nodal load on this
    # case turns it into a ~16k-iteration solve whose iteration count moves by
    # hundreds under rounding-level changes, in the oracle alone; loads are
    # pinned above at element level and by the closed forms.)
    from test_gpu_parity import _viscous_pair
    say("[gpu] viscous inclusion with a point viscosity and history")
    n = 16
    w, pr, dr, mesh, (bx, vx, by, vy) = _viscous_pair(n)
    ph = pr.phases_v - 1
    eta = np.repeat(np.asarray(w.eta)[ph][:, None], 7, axis=1)
    pr.eta_ip = eta
    kw = dict(pressure_factor=w.pressure_factor, tolerance=1e-8, check_interval=w.check_interval)
    so = pr.solve(w.dt, collect_history=True, **kw)
    st = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy),
                  dt=w.dt, viscosity=eta, collect_history=True, **kw)
    h, ho = st["history"], so["history"]
    say(f"  oracle iter {so['iter']} | gpu iter {st['iter']} history rows {len(h)} / {len(ho)}")
    assert st["converged"] and so["converged"]
    assert st["iter"] == so["iter"] and st["itPH"] == so["itPH"] and len(h) == len(ho)
    assert [r["iter"] for r in h] == [int(x) for x in ho[:, 0]]
    for a, b in ((dr.v.x, pr.vx), (dr.v.y, pr.vy), (dr.P, pr.P)):
        assert rel(a.numpy(), b) < 1e-10
    assert h[0]["err_v"] == pytest.approx(ho[0, 1], rel=1e-12)
    # later rows: rounding of the last checks (residuals down to 1e-11) gives
    # up to ~2e-5 relative on this case
    # (the start field is divergence-free: err_P of row 0 is rounding, ~1e-15,
    # so each norm is compared relative to itself or to 1e-10 of its peak)
    sv, sP = 1e-10 * ho[:, 1].max(), 1e-10 * ho[:, 5].max()
    dv = max(abs(r["err_v"] - ro[1]) / max(ro[1], sv) for r, ro in zip(h, ho))
    dP = max(abs(r["err_P"] - ro[5]) / max(ro[5], sP) for r, ro in zip(h, ho))
    say(f"  history max rel diff err_v {dv:.2e} err_P {dP:.2e}")
    assert dv < 1e-4 and dP < 1e-4


def test_gpu_body_force_matches_gravity_in_fp64_and_mixed_precision():
    say("[gpu] nodal loads equal the gravity load (FP64 and FP32 corrections)")
    n = 16
    load = _gravity_load(n)
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    zero = np.zeros(len(gam))
    bcs = (fa.DirichletBoundaryCondition(gam, zero), fa.DirichletBoundaryCondition(gam, zero))
    out = {}
    for tag, rho, g, kw in (("gravity", 1.0, (0.0, 1.0), {}), ("load64", 0.0, (0.0, 0.0), dict(body_force=load)),
                            ("load32", 0.0, (0.0, 0.0), dict(body_force=load, precision="fp32"))):
        mesh = fa.t7_box_mesh(n, n)
        dr = fa.StokesDR(mesh, fa.StokesMaterial(g=g, rho0=(rho,), **VISC))
        st = fa.solve(dr, mesh, *bcs, dt=1.0, check_interval=10, tolerance=1e-9, **kw)
        assert st["converged"], tag
        out[tag] = (st["iter"], dr.P.numpy(), dr.v.y.numpy())
    say("  iterations " + " ".join(f"{k} {v[0]}" for k, v in out.items()))
    assert rel(out["load64"][1], out["gravity"][1]) < 1e-8
    assert rel(out["load32"][1], out["gravity"][1]) < 1e-5


def test_gpu_pressure_source():
    # test_stokes.jl:315-340, test_solver_convergence_api.jl:300-328
    say("[gpu] pressure source Qq/Q2D")
    n = 4
    K = 10.0
    mat = dict(eta=(1.0,), etab=(math.inf,), rho0=(1.0,), K=(K,))
    pr, dr, mesh, gam = _pair(n, mat=mat)
    Qq = np.linspace(0.5, 1.5, pr.nnodesP)
    pr.set_pressure_source(Qq, 2.0)
    fa.set_pressure_source(dr, Qq, 2.0)
    assert rel(dr.Q.numpy(), pr.Q) < 1e-14
    with pytest.raises(ValueError, match="nonempty positive Qq support"):
        fa.set_pressure_source(dr, np.zeros(pr.nnodesP), 1.0)
    with pytest.raises(ValueError, match="finite and nonnegative"):
        fa.set_pressure_source(dr, -Qq, 1.0)
    fa.set_pressure_source(dr, np.zeros(pr.nnodesP), 0.0)
    assert not dr.Q.numpy().any()
    # one pseudo-step relaxes P to K; the second outer pass records convergence
    zero = np.zeros(len(gam))
    bcs = (fa.DirichletBoundaryCondition(gam, zero), fa.DirichletBoundaryCondition(gam, zero))
    fa.assemble_viscosity_weighted_pressure_scaling(dr, 1.0, 1.0, finite_K=True)
    dr.gammaP.fill(10.0)
    st = fa.solve(dr, mesh, *bcs, dt=1.0, finite_K=True, skip_scaling=True, Qq=np.ones(pr.nnodesP), Q2D=1.0,
                  check_interval=1, iterMax=1, max_iterations=1, max_ph_iterations=2)
    assert st["converged"]
    assert np.allclose(dr.Q.numpy(), 1.0, rtol=1e-13)
    assert np.allclose(dr.P.numpy(), K, rtol=1e-12)


def test_gpu_cap_carries_the_corrected_pressure_across_steps():
    # test_stokes.jl:506-580 on the device, step by step against the oracle
    say("[gpu] tensile cap with P_old at the points over three steps")
    n = 2
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    ne = len(el2n)
    dt, K, edot = 1.0, 4.0, 0.1
    spec = dict(phi=(math.radians(30),), psi=(math.radians(5),), C_=(1.0,), pT=(-0.5,), eta_reg=(0.1,), Kb=(K,),
                C_min=(0.5,), H_C=(-1.0,))
    mat = dict(eta=(1.0,), etab=(K,), G=(1.0,), K=(K,))
    bc = ((gam, edot * coords[gam, 0]), (gam, np.zeros(len(gam))))
    pr = O.Problem(coords, el2n, O.material(**mat), tau_old=tuple(np.zeros((ne, 7)) for _ in range(3)),
                   cap=O.drucker_prager_cap(**spec), history=True, bc=bc)
    pr.vx[:] = edot * coords[:, 0]
    pr.Pold_ip = np.zeros((ne, 7))
    mesh = fa.t7_box_mesh(n, n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**mat))
    dr.v.x.copyto(edot * coords[:, 0])
    cap = fa.DruckerPragerCap(**spec)
    bcs = tuple(fa.DirichletBoundaryCondition(d, v) for d, v in bc)
    P_old = np.zeros((ne, 7))
    for step in range(3):
        so = pr.solve(dt, pressure_factor=20.0, tolerance=1e-10, finite_K=True)
        st = fa.solve(dr, mesh, *bcs, dt=dt, plastic=cap, pressure_factor=20.0, tolerance=1e-10, P_old=P_old)
        assert st["converged"] and so["converged"] and st["iter"] == so["iter"]
        pr.update_current_stress(dt)
        pr.commit_plastic_history(dt)
        fa.update_stokes_current_stress(dr, mesh, dt, plastic=cap)
        fa.commit_stokes_plastic_history(dr, mesh, dt, plastic=cap)
        P_old = dr.P_ip.numpy()
        errs = dict(P_ip=rel(P_old, pr.P_store), gamma=rel(dr.plastic_history.gamma.numpy(), pr.gamma_hist),
                    theta=rel(dr.plastic_history.theta.numpy(), pr.theta_hist), P=rel(dr.P.numpy(), pr.P))
        say(f"  step {step} iter {st['iter']}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        assert max(errs.values()) < 1e-9
        pr.Pold_ip = pr.P_store.copy()
        pr.tau_old = tuple(t.copy() for t in pr.tau_store)
        fa.copyto(dr.tau_old, dr.tau)
    P_mp = P_old.mean()
    assert -0.51 < P_mp < -0.5
    assert np.all(dr.P.numpy() < -0.9)
    assert dr.plastic_history.theta.numpy().min() > 0


def test_gpu_keyword_argument_errors():
    # test_stokes_quadrature_viscosity.jl:62-65 and DR.jl:388-418
    say("[gpu] keyword argument errors")
    n = 2
    pr, dr, mesh, gam = _pair(n)
    zero = np.zeros(len(gam))
    bcs = (fa.DirichletBoundaryCondition(gam, zero), fa.DirichletBoundaryCondition(gam, zero))
    eta = np.full((pr.nels, 7), 2.0)
    with pytest.raises(fa.DimensionMismatch, match="viscosity must be nq × nels"):
        fa.solve(dr, mesh, *bcs, dt=1.0, viscosity=np.zeros((1, 1)))
    with pytest.raises(ValueError, match="finite positive"):
        fa.solve(dr, mesh, *bcs, dt=1.0, viscosity=-eta)
    with pytest.raises(ValueError, match="body_force must contain finite values"):
        fa.solve(dr, mesh, *bcs, dt=1.0, body_force=(np.full(pr.nnodes, np.nan), np.zeros(pr.nnodes)))
    with pytest.raises(fa.DimensionMismatch, match="body_force must match velocity nodes"):
        fa.solve(dr, mesh, *bcs, dt=1.0, body_force=(np.zeros(1), np.zeros(1)))
    with pytest.raises(fa.DimensionMismatch, match="integration-point P_old"):
        fa.solve(dr, mesh, *bcs, dt=1.0, P_old=np.zeros((pr.nels, 8)))
    with pytest.raises(ValueError, match="FP64"):
        fa.solve(dr, mesh, *bcs, dt=1.0, viscosity=eta, precision="fp32")
    with pytest.raises(ValueError, match="plastic=nothing"):
        fa.solve(dr, mesh, *bcs, dt=1.0, viscosity=eta, plastic=fa.DruckerPrager((0.5,), (0.1,), (1.0,), (0.1,),
                                                                                 (1.0,)))
    ve = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0,), G=(1.0,)))
    with pytest.raises(ValueError, match="purely viscous 2-D flow"):
        fa.solve(ve, mesh, *bcs, dt=1.0, viscosity=eta)
    with pytest.raises(ValueError, match="Qq is required"):
        fa.solve(dr, mesh, *bcs, dt=1.0, Q2D=1.0)
    # a failed call leaves no inputs attached
    pr.momentum_residual(1.0)
    fa.residual(dr, 1.0)
    assert rel(dr.Rv.y.numpy(), pr.Rvy) < 1e-12


def test_gpu_body_force_3d():
    # nodal loads on the Hex27 path: the load solve equals the gravity solve
    say("[gpu] 3-D nodal loads")
    n = (3, 3, 3)
    mesh = fa.hex27_box_mesh(*n)
    gam = mesh.boundary_nodes
    zero = np.zeros(len(gam))
    bcs = tuple(fa.DirichletBoundaryCondition(gam, zero) for _ in range(3))
    grav = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0,), etab=(math.inf,), rho0=(1.0,), g=(0.0, 0.0, -1.0)))
    fa.residual(grav, 1.0)
    load = tuple(-getattr(grav.Rv, c).numpy() for c in "xyz")
    kw = dict(dt=1.0, check_interval=20, tolerance=1e-9)
    sg = fa.solve(grav, mesh, *bcs, **kw)
    dl = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0,), etab=(math.inf,), rho0=(0.0,), g=(0.0, 0.0, 0.0)))
    sl = fa.solve(dl, mesh, *bcs, body_force=load, **kw)
    say(f"  gravity iter {sg['iter']} | load iter {sl['iter']}")
    assert sg["converged"] and sl["converged"]
    assert rel(dl.P.numpy(), grav.P.numpy()) < 1e-8
