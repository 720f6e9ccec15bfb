"""GPU: the frozen-operator adjoint of StokesDR (solve_adjoint! with operator =
:blocks, DR_adjoint.jl:1-531; adjoint_operator.jl:1-660) through the C ABI
against the oracle (itself pinned in test_oracle_adjoint.py by the
reference's finite-difference gradient test and a dot-product test of the
operator): the element blocks to 1e-12 of their scale, and adjoint solves from
the same forward state with identical iteration counts (inner, outer and power
iteration) and multipliers to 1e-10; the reference's gradient check on the
GPU's own forward solves."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

from test_oracle_adjoint import buoyancy_case, FWD, ADJ, _state, vep_adjoint_case  # noqa: E402

INF = math.inf


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def gpu_twin(pr, mat_kw, phases):
    """A GPU state holding the oracle problem's mesh, material, phases and
    current fields (v, P, P0, T, T0, τ_old, γP, M_P)."""
    n = int(round(math.sqrt(pr.nels / 2)))
    mesh = fa.t7_box_mesh(n, n)
    assert np.array_equal(mesh.el2n, pr.el2n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**mat_kw), CFL_v=0.9, c_fact=0.7)
    dr.set_phases("v", phases)
    dr.set_phases("P", phases)
    for dst, src in ((dr.v.x, pr.vx), (dr.v.y, pr.vy), (dr.P, pr.P), (dr.P0, pr.P0), (dr.T, pr.T), (dr.T0, pr.T0),
                     (dr.gammaP, pr.gammaP), (dr.M_P, pr.MP)):
        dst.copyto(src)
    if pr.tau_old is not None:
        dr.tau_old.xx.copyto(pr.tau_old[0])
        dr.tau_old.yy.copyto(pr.tau_old[1])
        dr.tau_old.xy.copyto(pr.tau_old[2])
    return dr, mesh


STATES = {
    "viscous": ((INF, INF), (INF, INF), (INF, INF)),
    "ve_storage": ((INF, INF), (3.0, 8.0), (0.7, 2.0)),
    "compressible": ((40.0, 90.0), (3.0, INF), (INF, INF)),
}


@pytest.mark.parametrize("name", list(STATES))
def test_gpu_adjoint_blocks_parity(name):
    K, etab, G = STATES[name]
    say(f"[gpu] adjoint blocks {name}")
    pr, rng = _state(K, etab, G, seed=2)
    mat_kw = dict(eta=(1.0, 4.0), etab=etab, G=G, alpha=(0.01, 0.02), rho0=(1.0, 1.5), K=K, g=(0.3, -1.0),
                  Tref=0.2)
    dr, mesh = gpu_twin(pr, mat_kw, pr.phases_v)
    dt = 0.5
    bo = pr.adjoint_blocks(dt)
    bg = fa.adjoint_blocks(dr, dt)
    for k in "ABCDW":
        scale = max(np.abs(bo[k]).max(), 1e-300)
        e = np.abs(bg[k] - bo[k]).max() / scale if np.abs(bo[k]).max() > 0 else np.abs(bg[k]).max()
        say(f"  {k}: {e:.2e}")
        assert e < 1e-12, k
    for a, b in ((dr.dRvdv.x, pr.dRdvx), (dr.dRvdv.y, pr.dRdvy), (dr.PC_v.x, pr.PCx), (dr.PC_v.y, pr.PCy)):
        assert rel(a.numpy(), b) < 1e-12


def _adjoint_pair(case, rho2=2.0):
    coords, el2n, ph, xw, yw, obs = case
    mat_kw = dict(eta=(1.0, 1.0), etab=(INF, INF), alpha=(0.0, 0.0), rho0=(1.0, rho2), K=(INF, INF), g=(0.0, -1.0),
                  Tref=0.0)
    pr = O.Problem(coords, el2n, O.material(**mat_kw), phases_v=ph, phases_P=ph,
                   bc=((xw, np.zeros(len(xw))), (yw, np.zeros(len(yw)))))
    st = pr.solve(1.0, **FWD)
    assert st["converged"]
    dr, mesh = gpu_twin(pr, mat_kw, ph)
    bcs = (fa.DirichletBoundaryCondition(xw, np.zeros(len(xw))), fa.DirichletBoundaryCondition(yw, np.zeros(len(yw))))
    return pr, dr, mesh, bcs


def _gpu_kw():
    kw = dict(ADJ)
    kw.pop("CFL_v"), kw.pop("c_fact")
    kw["iterMax"] = kw.pop("iter_max")
    return kw


@pytest.mark.parametrize("measure", [True, False])
def test_gpu_adjoint_solve_parity(measure):
    case = buoyancy_case()
    obs = case[5]
    say(f"[gpu] adjoint solve parity measure_lmax={measure}")
    pr, dr, mesh, bcs = _adjoint_pair(case)
    nn, npp = pr.nnodes, pr.nnodesP
    obj = (np.zeros(nn), np.zeros(nn))
    obj[1][obs] = -1.0
    lam = [np.zeros(nn), np.zeros(nn), np.zeros(npp)]
    so = pr.solve_adjoint(1.0, obj, lam, **ADJ, measure_lmax=measure)
    sg = fa.solve_adjoint(dr, mesh, *bcs, dt=1.0, objective_vx=obj[0], objective_vy=obj[1], lam_vx=np.zeros(nn),
                          lam_vy=np.zeros(nn), lam_P=np.zeros(npp), measure_lmax=measure, **_gpu_kw())
    say("  oracle", {k: so[k] for k in ("iter", "itPH", "lmax_iterations", "lmax", "err")},
        "gpu", {k: sg[k] for k in ("iter", "itPH", "lmax_iterations", "lmax", "err")})
    assert sg["converged"] and so["converged"]
    assert sg["iter"] == so["iter"] and sg["itPH"] == so["itPH"]
    assert sg["lmax_iterations"] == so["lmax_iterations"]
    assert sg["lmax"] == pytest.approx(so["lmax"], rel=1e-12)
    assert sg["lmax_gershgorin"] == pytest.approx(so["lmax_gershgorin"], rel=1e-12)
    for a, b in zip((sg["lam_vx"], sg["lam_vy"], sg["lam_P"]), lam):
        assert rel(a, b) < 1e-10


def test_gpu_adjoint_gradient_and_warm_start():
    # the reference's gradient test on the GPU's own forward solves
    case = buoyancy_case()
    coords, el2n, ph, xw, yw, obs = case
    say("[gpu] adjoint gradient vs finite differences")

    def forward(rho2):
        mesh = fa.t7_box_mesh(3, 3)
        dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0, 1.0), etab=(INF, INF), rho0=(1.0, rho2), g=(0.0, -1.0)),
                         CFL_v=0.9, c_fact=0.7)
        bcs = (fa.DirichletBoundaryCondition(xw, np.zeros(len(xw))),
               fa.DirichletBoundaryCondition(yw, np.zeros(len(yw))))
        kw = dict(FWD)
        kw.pop("CFL_v"), kw.pop("c_fact")
        kw["iterMax"] = kw.pop("iter_max")
        st = fa.solve(dr, mesh, *bcs, dt=1.0, phases_v=ph, phases_P=ph, **kw)
        assert st["converged"]
        return dr, mesh, bcs

    dr, mesh, bcs = forward(2.0)
    nn, npp = mesh.nnodes, mesh.nnodesP
    obj_y = np.zeros(nn)
    obj_y[obs] = -1.0
    lx, ly, lp = (fa.device_array(dr.ctx, np.zeros(n)) for n in (nn, nn, npp))
    st = fa.solve_adjoint(dr, mesh, *bcs, dt=1.0, objective_vx=np.zeros(nn), objective_vy=obj_y, lam_vx=lx,
                          lam_vy=ly, lam_P=lp, **_gpu_kw())
    assert st["converged"] and st["lmax_iterations"] > 0 and st["lmax"] < st["lmax_gershgorin"]
    lam = (lx.numpy(), ly.numpy(), lp.numpy())
    assert np.any(lam[1] != 0)
    warm = fa.solve_adjoint(dr, mesh, *bcs, dt=1.0, objective_vx=np.zeros(nn), objective_vy=obj_y, lam_vx=lx,
                            lam_vy=ly, lam_P=lp, **{**_gpu_kw(), "tolerance": 1e-9}, measure_lmax=False)
    assert warm["iter"] == 0 and warm["lmax_iterations"] == 0
    assert all(np.array_equal(a.numpy(), b) for a, b in zip((lx, ly, lp), lam))
    # ∂R/∂ρ₂ at the frozen forward state (the oracle's residual of the GPU state)
    pr = O.Problem(coords, el2n, O.material(eta=(1.0, 1.0), etab=(INF, INF), rho0=(0.0, 1.0), g=(0.0, -1.0)),
                   phases_v=ph, phases_P=ph)
    pr.vx[:], pr.vy[:], pr.P[:] = dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy()
    pr.momentum_residual(1.0)
    r1 = (pr.Rvx.copy(), pr.Rvy.copy())
    pr.mat.rho0[1] = 0.0
    pr.momentum_residual(1.0)
    grad = lam[0] @ (r1[0] - pr.Rvx) + lam[1] @ (r1[1] - pr.Rvy)
    d = 1e-4
    jp = -forward(2.0 + d)[0].v.y.numpy()[obs].sum()
    jm = -forward(2.0 - d)[0].v.y.numpy()[obs].sum()
    say(f"  adjoint {grad:.8e} finite differences {(jp - jm) / (2 * d):.8e}")
    assert grad == pytest.approx((jp - jm) / (2 * d), rel=1e-4)


def test_gpu_adjoint_storage_and_compressible_solve_parity():
    # a visco-elastic state with finite ηb (D, W stored) and finite K (the
    # full-matrix layout), converged by the oracle's forward solve
    say("[gpu] adjoint solve parity, finite bulk viscosity and bulk modulus")
    coords, el2n = O.t7_mesh(4, 4)
    ne = len(el2n)
    cen = coords[el2n[:, :3]].mean(axis=1)
    ph = np.where((cen[:, 0] - 0.5) ** 2 + (cen[:, 1] - 0.5) ** 2 < 0.06, 2, 1).astype(np.int32)
    gam = O.box_boundary_nodes(coords)
    rng = np.random.default_rng(1)
    tau = [0.2 * rng.normal(size=(ne, 7)) for _ in range(3)]
    mat_kw = dict(eta=(1.0, 10.0), etab=(50.0, 80.0), G=(2.0, 5.0), alpha=(0.0, 0.0), rho0=(1.0, 1.2),
                  K=(200.0, 300.0), g=(0.0, -1.0), Tref=0.0)
    pr = O.Problem(coords, el2n, O.material(**mat_kw), phases_v=ph, phases_P=ph, tau_old=tau,
                   bc=((gam, np.zeros(len(gam))), (gam, np.zeros(len(gam)))))
    st = pr.solve(0.5, **FWD)
    assert st["converged"]
    dr, mesh = gpu_twin(pr, mat_kw, ph)
    bcs = (fa.DirichletBoundaryCondition(gam, np.zeros(len(gam))), fa.DirichletBoundaryCondition(gam, np.zeros(len(gam))))
    nn, npp = pr.nnodes, pr.nnodesP
    obj = (np.cos(3 * coords[:, 1]), np.sin(2 * coords[:, 0]))
    lam = [np.zeros(nn), np.zeros(nn), np.zeros(npp)]
    so = pr.solve_adjoint(0.5, obj, lam, **ADJ)
    sg = fa.solve_adjoint(dr, mesh, *bcs, dt=0.5, objective_vx=obj[0], objective_vy=obj[1], lam_vx=np.zeros(nn),
                          lam_vy=np.zeros(nn), lam_P=np.zeros(npp), **_gpu_kw())
    say("  oracle", {k: so[k] for k in ("iter", "itPH", "lmax_iterations", "err")},
        "gpu", {k: sg[k] for k in ("iter", "itPH", "lmax_iterations", "err")})
    assert so["converged"] and sg["converged"]
    assert sg["iter"] == so["iter"] and sg["itPH"] == so["itPH"] and sg["lmax_iterations"] == so["lmax_iterations"]
    for a, b in zip((sg["lam_vx"], sg["lam_vy"], sg["lam_P"]), lam):
        assert rel(a, b) < 1e-10


def test_gpu_adjoint_errors():
    case = buoyancy_case()
    coords, el2n, ph, xw, yw, obs = case
    mesh = fa.t7_box_mesh(3, 3)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0, 1.0), rho0=(1.0, 2.0), g=(0.0, -1.0)))
    bcs = (fa.DirichletBoundaryCondition(xw, np.zeros(len(xw))), fa.DirichletBoundaryCondition(yw, np.zeros(len(yw))))
    nn, npp = mesh.nnodes, mesh.nnodesP
    z = dict(objective_vx=np.zeros(nn), objective_vy=np.ones(nn), lam_vx=np.zeros(nn), lam_vy=np.zeros(nn),
             lam_P=np.zeros(npp))
    with pytest.raises(ValueError, match="unassembled"):
        fa.solve_adjoint(dr, mesh, *bcs, dt=1.0, phases_v=ph, phases_P=ph, **z)
    with pytest.raises(ValueError, match="Drucker-Prager model"):
        fa.solve_adjoint(dr, mesh, *bcs, dt=1.0,
                         plastic=fa.DruckerPragerCap((0.5,) * 2, (0.0,) * 2, (1.0,) * 2, (-0.1,) * 2, (0.0,) * 2,
                                                     (1.0,) * 2), **z)
    fa.assemble_viscosity_weighted_pressure_scaling(dr, 20.0, 1.0)
    budget = dict(max_iterations=2, check_interval=1, iterMax=2, max_ph_iterations=1, tolerance=1e-14)
    with pytest.raises(fa.FemError, match="did not converge after 2 iterations"):
        fa.solve_adjoint(dr, mesh, *bcs, dt=1.0, **budget, **z)
    st = fa.solve_adjoint(dr, mesh, *bcs, dt=1.0, throw_on_failure=False, **budget, **z)
    assert not st["converged"] and st["iter"] == 2


def test_gpu_forward_measure_lmax_parity():
    # measure_λmax of the forward solve (DR.jl:375-378, 459-490, 728-760): the
    # same velocity blocks drive a power iteration at every Jacobian assembly
    case = buoyancy_case()
    coords, el2n, ph, xw, yw, obs = case
    say("[gpu] forward solve with measure_lmax")
    mat_kw = dict(eta=(1.0, 1.0), etab=(INF, INF), alpha=(0.0, 0.0), rho0=(1.0, 2.0), K=(INF, INF), g=(0.0, -1.0),
                  Tref=0.0)
    pr = O.Problem(coords, el2n, O.material(**mat_kw), phases_v=ph, phases_P=ph,
                   bc=((xw, np.zeros(len(xw))), (yw, np.zeros(len(yw)))))
    kw = dict(pressure_factor=20.0, tolerance=1e-9, check_interval=100, rel_drop0=0.1)
    so = pr.solve(1.0, CFL_v=0.9, c_fact=0.7, measure_lmax=True, **kw)
    mesh = fa.t7_box_mesh(3, 3)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**mat_kw), CFL_v=0.9, c_fact=0.7)
    bcs = (fa.DirichletBoundaryCondition(xw, np.zeros(len(xw))), fa.DirichletBoundaryCondition(yw, np.zeros(len(yw))))
    sg = fa.solve(dr, mesh, *bcs, dt=1.0, phases_v=ph, phases_P=ph, measure_lmax=True, **kw)
    say("  oracle", {k: so[k] for k in ("iter", "itPH", "lambda_max", "lambda_max_iterations")},
        "gpu", {k: sg[k] for k in ("iter", "itPH", "lambda_max", "lambda_max_iterations")})
    assert so["converged"] and sg["converged"]
    assert sg["iter"] == so["iter"] and sg["itPH"] == so["itPH"]
    assert sg["lambda_max_iterations"] == so["lambda_max_iterations"] > 0
    assert sg["lambda_max"] == pytest.approx(so["lambda_max"], rel=1e-12)
    assert sg["lambda_max"] < sg["lambda_max_gershgorin"]
    for a, b in ((dr.v.x, pr.vx), (dr.v.y, pr.vy), (dr.P, pr.P)):
        assert rel(a.numpy(), b) < 1e-10
    with pytest.raises(ValueError, match="precision 0"):
        fa.solve(dr, mesh, *bcs, dt=1.0, measure_lmax=True, precision="fp32", **kw)


def test_gpu_forward_measure_lmax_plastic_parity():
    # with a yield model the Jacobian and the estimate are renewed at every
    # check (warm-started power vectors, safety ≥ 1.5)
    from femtools_amd import problems as PB
    say("[gpu] forward VEP solve with measure_lmax")
    n = 10
    w = PB.vep_inclusions(n, ninc=3, seed=5, nsteps=1, C=0.5)
    coords, el2n = O.t7_mesh(n, n)
    ph = PB.inclusion_phases(coords, el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(coords)
    ne = len(el2n)
    dp = dict(phi=w.plastic["phi"], psi=w.plastic["psi"], C_=w.plastic["C"], eta_reg=w.plastic["eta_reg"],
              Kb=w.plastic["Kb"])
    mat_kw = dict(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g)
    pr = O.Problem(coords, el2n, O.material(**mat_kw), dp=O.drucker_prager(**dp), phases_v=ph, phases_P=ph,
                   tau_old=tuple(np.zeros((ne, 7)) for _ in range(3)), bc=((bx, vx), (by, vy)))
    v0x, v0y = PB.pure_shear_field(coords)
    pr.vx[:], pr.vy[:] = v0x, v0y
    kw = dict(pressure_factor=w.pressure_factor, tolerance=1e-8, check_interval=w.check_interval,
              rel_drop0=w.rel_drop0)
    so = pr.solve(w.dt, CFL_v=w.CFL_v, measure_lmax=True, **kw)
    mesh = fa.t7_box_mesh(n, n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**mat_kw), CFL_v=w.CFL_v)
    dr.v.x.copyto(v0x)
    dr.v.y.copyto(v0y)
    sg = fa.solve(dr, mesh, fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy), dt=w.dt,
                  phases_v=ph, phases_P=ph, measure_lmax=True,
                  plastic=fa.DruckerPrager(dp["phi"], dp["psi"], dp["C_"], dp["eta_reg"], dp["Kb"]), **kw)
    say("  oracle", {k: so[k] for k in ("iter", "itPH", "jacobian_assemblies", "lambda_max_iterations")},
        "gpu", {k: sg[k] for k in ("iter", "itPH", "jacobian_assemblies", "lambda_max_iterations")})
    assert so["converged"] and sg["converged"]
    assert sg["iter"] == so["iter"] and sg["jacobian_assemblies"] == so["jacobian_assemblies"] > 1
    assert sg["lambda_max_iterations"] == so["lambda_max_iterations"]
    for a, b in ((dr.v.x, pr.vx), (dr.v.y, pr.vy), (dr.P, pr.P)):
        assert rel(a.numpy(), b) < 1e-9



def test_gpu_adjoint_blocks_parity_drucker_prager():
    say("[gpu] adjoint blocks with the Drucker-Prager return")
    K, etab, G = (INF, INF), (3.0, 8.0), (0.7, 2.0)
    pr, rng = _state(K, etab, G, seed=5)
    dpk = dict(phi=(0.5, 0.6), psi=(0.0, 0.1), C_=(0.05, 0.08), eta_reg=(0.01, 0.02), Kb=(5.0, 5.0))
    pr.dp = O.drucker_prager(**dpk)
    mat_kw = dict(eta=(1.0, 4.0), etab=etab, G=G, alpha=(0.01, 0.02), rho0=(1.0, 1.5), K=K, g=(0.3, -1.0), Tref=0.2)
    dr, mesh = gpu_twin(pr, mat_kw, pr.phases_v)
    dt = 0.5
    bo = pr.adjoint_blocks(dt)
    bg = fa.adjoint_blocks(dr, dt, plastic=fa.DruckerPrager(dpk["phi"], dpk["psi"], dpk["C_"], dpk["eta_reg"],
                                                            dpk["Kb"]))
    for k in "ABCDW":
        e = np.abs(bg[k] - bo[k]).max() / max(np.abs(bo[k]).max(), 1e-300)
        say(f"  {k}: {e:.2e}")
        assert e < 1e-12, k


def test_gpu_plastic_adjoint_solve_parity():
    # the yielding VEP step of test_oracle_adjoint.py: the oracle's forward
    # state on the device, both adjoints from zero
    say("[gpu] plastic adjoint solve parity")
    pr, w, obs, adj = vep_adjoint_case()
    mat_kw = dict(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g)
    n = int(round(math.sqrt(pr.nels / 2)))
    mesh = fa.t7_box_mesh(n, n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**mat_kw), CFL_v=w.CFL_v, c_fact=0.9)
    dr.set_phases("v", pr.phases_v)
    dr.set_phases("P", pr.phases_v)
    for dst, src in ((dr.v.x, pr.vx), (dr.v.y, pr.vy), (dr.P, pr.P), (dr.P0, pr.P0), (dr.gammaP, pr.gammaP),
                     (dr.M_P, pr.MP)):
        dst.copyto(src)
    for dst, src in zip((dr.tau_old.xx, dr.tau_old.yy, dr.tau_old.xy), pr.tau_old):
        dst.copyto(src)
    nn, npp = pr.nnodes, pr.nnodesP
    obj = (np.zeros(nn), np.zeros(nn))
    obj[1][obs] = -1.0
    lam = [np.zeros(nn), np.zeros(nn), np.zeros(npp)]
    so = pr.solve_adjoint(w.dt, obj, lam, **adj)
    bx, by = pr.bc[0][0], pr.bc[1][0]
    bcs = (fa.DirichletBoundaryCondition(bx, np.zeros(len(bx))), fa.DirichletBoundaryCondition(by, np.zeros(len(by))))
    kw = dict(adj)
    kw.pop("CFL_v"), kw.pop("c_fact")
    kw["iterMax"] = kw.pop("iter_max")
    plastic = fa.DruckerPrager(w.plastic["phi"], w.plastic["psi"], w.plastic["C"], w.plastic["eta_reg"],
                               w.plastic["Kb"])
    sg = fa.solve_adjoint(dr, mesh, *bcs, dt=w.dt, objective_vx=obj[0], objective_vy=obj[1], lam_vx=np.zeros(nn),
                          lam_vy=np.zeros(nn), lam_P=np.zeros(npp), plastic=plastic, **kw)
    say("  oracle", {k: so[k] for k in ("iter", "itPH", "lmax_iterations", "err")},
        "gpu", {k: sg[k] for k in ("iter", "itPH", "lmax_iterations", "err")})
    assert so["converged"] and sg["converged"]
    # the non-normal operator's relaxation is rounding-sensitive near the
    # stopping test: the device may take one more outer step of the check
    # cadence; the multipliers then agree to the tolerance
    assert sg["lmax_iterations"] == so["lmax_iterations"]
    assert abs(sg["itPH"] - so["itPH"]) <= 1 and abs(sg["iter"] - so["iter"]) <= adj["check_interval"]
    for a, b in zip((sg["lam_vx"], sg["lam_vy"], sg["lam_P"]), lam):
        say("  lambda rel", rel(a, b))
        assert rel(a, b) < 1e-8
    # the gradient with respect to the matrix density from the device's
    # multipliers (∂R/∂ρ from the oracle's residual at the frozen state)
    base = tuple(pr.mat.rho0[i] for i in range(pr.mat.nphases))

    def residual(rho):
        for i, r in enumerate(rho):
            pr.mat.rho0[i] = r
        pr.momentum_residual(w.dt)
        return pr.Rvx.copy(), pr.Rvy.copy()
    r1 = residual((1.0,) + (0.0,) * (len(base) - 1))
    r0 = residual((0.0,) * len(base))
    residual(base)
    g_dev = sg["lam_vx"] @ (r1[0] - r0[0]) + sg["lam_vy"] @ (r1[1] - r0[1])
    g_orc = lam[0] @ (r1[0] - r0[0]) + lam[1] @ (r1[1] - r0[1])
    say(f"  gradient device {g_dev:.10e} oracle {g_orc:.10e}")
    assert g_dev == pytest.approx(g_orc, rel=1e-9)
