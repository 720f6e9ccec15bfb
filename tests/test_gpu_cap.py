"""GPU: the Drucker-Prager tensile cap (src/stokes/assemblers/rheology.jl:
510-787) through the C ABI against the oracle: residuals, stress and
corrected pressure, Jacobian diagnostics (the return's implicit derivative)
and the plastic-history commit to 1e-12 of the field scale; converged solves
to 1e-10 with the iteration counts compared."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

TAU2 = ("xx", "yy", "xy")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CAP = dict(phi=(math.radians(30), math.radians(25)), psi=(math.radians(5), 0.0), C_=(1.0, 1.5), pT=(-0.5, -0.3),
           eta_reg=(0.1, 0.05), Kb=(4.0, 6.0), C_min=(0.6, 1.0), H_C=(-0.5, -0.2))
MAT = dict(eta=(1.0, 3.0), etab=(1.0, 1.0), G=(1.0, 2.0), K=(4.0, 6.0))


def cap_pair(n=6, seed=0, nodal=False, history=True, pf=False):
    rng = np.random.default_rng(seed)
    coords, el2n = O.t7_mesh(n, n)
    ne = len(el2n)
    ph = rng.integers(1, 3, ne).astype(np.int32)
    cap_o = O.drucker_prager_cap(CAP["phi"], CAP["psi"], CAP["C_"], CAP["pT"], CAP["eta_reg"], CAP["Kb"],
                                 C_min=CAP["C_min"], H_C=CAP["H_C"])
    cap_g = fa.DruckerPragerCap(CAP["phi"], CAP["psi"], CAP["C_"], CAP["pT"], CAP["eta_reg"], CAP["Kb"],
                                C_min=CAP["C_min"], H_C=CAP["H_C"])
    tau_old = tuple(0.3 * rng.normal(size=(ne, 7)) for _ in range(3))
    mesh = fa.t7_box_mesh(n, n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**MAT))
    if nodal:
        php = rng.integers(1, 3, (ne, 7)).astype(np.int32)
        pr = O.Problem(coords, el2n, O.material(**MAT), phases_v=php, phase_mode_v=1,
                       phases_P=php[:, :3].copy(), phase_mode_P=1, tau_old=tau_old, cap=cap_o, history=history)
        dr.set_phases("v", php)
        dr.set_phases("P", php[:, :3].copy())
    else:
        pr = O.Problem(coords, el2n, O.material(**MAT), phases_v=ph, phases_P=ph, tau_old=tau_old, cap=cap_o,
                       history=history)
        dr.set_phases("v", ph)
        dr.set_phases("P", ph)
    for name, t in zip(TAU2, tau_old):
        getattr(dr.tau_old, name).copyto(t)
    pr.vx[:] = 0.6 * rng.normal(size=len(coords))
    pr.vy[:] = 0.6 * rng.normal(size=len(coords))
    pr.P[:] = rng.normal(size=3 * ne) - 0.3
    dr.v.x.copyto(pr.vx)
    dr.v.y.copyto(pr.vy)
    dr.P.copyto(pr.P)
    if history:
        g = np.abs(0.4 * rng.normal(size=(ne, 7)))
        pr.gamma_hist[:] = g
        dr.plastic_history.gamma.copyto(g)
    if pf:
        Pf = 0.2 * rng.normal(size=3 * ne)
        pr.Pf = Pf
        dr.Pf.copyto(Pf)
    pr.pressure_scaling(1.0, 0.5, finite_K=True)
    fa.assemble_viscosity_weighted_pressure_scaling(dr, 1.0, 0.5, finite_K=True)
    return pr, dr, mesh, cap_g


@pytest.mark.parametrize("nodal", [False, True])
@pytest.mark.parametrize("with_pnum", [False, True])
def test_gpu_cap_residual_stress_and_pressure(nodal, with_pnum):
    say(f"[gpu] cap residual parity nodal={nodal} with_pnum={with_pnum}")
    pr, dr, mesh, cap = cap_pair(nodal=nodal)
    dt = 0.5
    pr.pressure_residual(dt, finite_K=True)
    Pn = pr.gammaP * pr.RP / pr.MP if with_pnum else None
    pr.momentum_residual(dt, Pnum=Pn)
    Rx, Ry, RP = pr.Rvx.copy(), pr.Rvy.copy(), pr.RP.copy()
    fa.residual(dr, dt, plastic=cap, with_pnum=with_pnum, finite_K=True)
    errs = dict(RP=rel(dr.RP.numpy(), RP), Rx=rel(dr.Rv.x.numpy(), Rx), Ry=rel(dr.Rv.y.numpy(), Ry))
    pr.update_current_stress(dt)
    fa.update_stokes_current_stress(dr, mesh, dt, plastic=cap)
    for name, t in zip(TAU2, pr.tau_store):
        errs["tau." + name] = rel(getattr(dr.tau, name).numpy(), t)
    errs["P_ip"] = rel(dr.P_ip.numpy(), pr.P_store)
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-12


@pytest.mark.parametrize("nodal", [False, True])
def test_gpu_cap_jacobian_parity(nodal):
    say(f"[gpu] cap Jacobian parity nodal={nodal}")
    pr, dr, mesh, cap = cap_pair(nodal=nodal, seed=1)
    dt = 0.5
    pr.augmented_jacobian(dt, finite_K=True)
    fa.jacobian_diagnostics(dr, dt, plastic=cap, finite_K=True)
    errs = dict(dRdvx=rel(dr.dRvdv.x.numpy(), pr.dRdvx), dRdvy=rel(dr.dRvdv.y.numpy(), pr.dRdvy),
                PCx=rel(dr.PC_v.x.numpy(), pr.PCx), PCy=rel(dr.PC_v.y.numpy(), pr.PCy))
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-12


def test_gpu_cap_commit_plastic_history():
    say("[gpu] cap plastic-history commit")
    pr, dr, mesh, cap = cap_pair(seed=2)
    dt = 0.5
    for _ in range(2):
        pr.commit_plastic_history(dt)
        fa.commit_stokes_plastic_history(dr, mesh, dt, plastic=cap)
        eg = rel(dr.plastic_history.gamma.numpy(), pr.gamma_hist)
        et = rel(dr.plastic_history.theta.numpy(), pr.theta_hist)
        say(f"  gamma {eg:.2e} theta {et:.2e}")
        assert eg < 1e-12 and et < 1e-12
    # a no-op for other yield models
    g = dr.plastic_history.gamma.numpy()
    fa.commit_stokes_plastic_history(dr, mesh, dt, plastic=None)
    assert np.array_equal(dr.plastic_history.gamma.numpy(), g)


def _extension(n, biax):
    coords, el2n = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    x, y = coords[:, 0], coords[:, 1]
    nx = gam[(np.abs(x[gam]) < 1e-12) | (np.abs(x[gam] - 1) < 1e-12)]
    ny = gam[(np.abs(y[gam]) < 1e-12) | (np.abs(y[gam] - 1) < 1e-12)]
    sy = 1.0 if biax else -1.0
    bc = ((nx, x[nx] - 0.5), (ny, sy * (y[ny] - 0.5)))
    ne = len(el2n)
    cen = coords[el2n[:, :3]].mean(axis=1)
    ph = np.where(((cen - 0.5) ** 2).sum(axis=1) < 0.15 ** 2, 2, 1).astype(np.int32)
    spec = dict(phi=(math.radians(30),) * 2, psi=(math.radians(5),) * 2, C_=(1.0, 0.5), pT=(-0.3, -0.15),
                eta_reg=(0.1, 0.1), Kb=(4.0, 4.0))
    mat = dict(eta=(1.0, 0.1), etab=(1.0, 1.0), G=(1.0, 1.0), K=(4.0, 4.0))
    zero = tuple(np.zeros((ne, 7)) for _ in range(3))
    pr = O.Problem(coords, el2n, O.material(**mat), phases_v=ph, phases_P=ph, tau_old=zero,
                   cap=O.drucker_prager_cap(**spec), bc=bc)
    mesh = fa.t7_box_mesh(n, n)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(**mat))
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    pr.vx[:], pr.vy[:] = x - 0.5, sy * (y - 0.5)
    dr.v.x.copyto(pr.vx)
    dr.v.y.copyto(pr.vy)
    bcs = tuple(fa.DirichletBoundaryCondition(d, v) for d, v in bc)
    return pr, dr, mesh, bcs, fa.DruckerPragerCap(**spec)


@pytest.mark.parametrize("n,biax", [(6, False), (8, False), (4, True)])
def test_gpu_cap_solve_parity(n, biax):
    say(f"[gpu] cap solve parity n={n} biaxial={biax}")
    pr, dr, mesh, bcs, cap = _extension(n, biax)
    kw = dict(pressure_factor=1.0, tolerance=1e-8, check_interval=20, max_iterations=100_000)
    so = pr.solve(0.5, finite_K=True, iter_max=100_000, **kw)
    st = fa.solve(dr, mesh, *bcs, dt=0.5, plastic=cap, iterMax=100_000, **kw)
    errs = dict(vx=rel(dr.v.x.numpy(), pr.vx), vy=rel(dr.v.y.numpy(), pr.vy), P=rel(dr.P.numpy(), pr.P))
    say(f"  oracle iter {so['iter']} itPH {so['itPH']} | gpu iter {st['iter']} itPH {st['itPH']} "
        f"jacobians {st['jacobian_assemblies']}; " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert st["converged"] and so["converged"]
    assert st["iter"] == so["iter"] and st["itPH"] == so["itPH"]
    assert max(errs.values()) < 1e-10
    # the stress update reports the corrected pressure where points fail
    fa.update_stokes_current_stress(dr, mesh, 0.5, plastic=cap)
    pr.update_current_stress(0.5)
    assert rel(dr.P_ip.numpy(), pr.P_store) < 1e-10


def test_gpu_cap_rejections():
    say("[gpu] cap argument checks")
    pr, dr, mesh, bcs, cap = _extension(4, False)
    with pytest.raises(ValueError, match="FP64"):
        fa.solve(dr, mesh, *bcs, dt=0.5, plastic=cap, precision="fp32")
    with pytest.raises(ValueError):
        fa.DruckerPragerCap((0.5,), (0.1,), (1.0,), (0.2,), (0.1,), (4.0,))     # pT > 0
    with pytest.raises(ValueError, match="cap radius"):
        fa.DruckerPragerCap((0.5,), (0.1,), (0.1,), (-5.0,), (0.1,), (4.0,))
    with pytest.raises(ValueError, match="bulk modulus"):
        fa.DruckerPragerCap((0.5,), (0.1,), (1.0,), (-0.5,), (0.1,), (0.0,))
    with pytest.raises(ValueError, match="must all be finite"):
        fa.DruckerPragerCap((0.5,), (0.1,), (1.0,), (-0.5,), (0.1,), (math.inf,))
    m3 = fa.hex27_box_mesh(1, 1, 1)
    d3 = fa.StokesDR(m3, fa.StokesMaterial(g=(0.0, 0.0, 0.0)))
    with pytest.raises(ValueError, match="2-D"):
        fa.residual(d3, 0.5, plastic=fa.DruckerPragerCap((0.5,), (0.1,), (1.0,), (-0.5,), (0.1,), (4.0,)))
