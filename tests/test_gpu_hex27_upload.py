"""GPU parity of the 3-D Stokes and thermal paths on uploaded Hex27 meshes
(fem_mesh_hex27_upload: graded box elements, shuffled nodes and elements)
against the oracle on the same arrays, whose uploaded-mesh geometry is pinned
in test_oracle_hex27_upload.py. Tolerances as for generated meshes: element
quantities 1e-12 of the field scale, converged solves 1e-10 with identical
iteration counts."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O
from test_oracle_hex27_upload import mesh_graded_shuffled

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

INF = math.inf
BOX = (0.0, 2.0, -1.0, 0.5, 0.0, 1.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def upload(n, box=BOX, grade=0.3, seed=0):
    coords, el2n, _, _ = mesh_graded_shuffled(n, box, grade, seed)
    mesh = fa.hex27_mesh_from_arrays(coords, el2n)
    return coords, el2n, mesh


def test_upload_layout_and_boundary():
    say("[gpu] hex27 upload 3x2x2 layout")
    coords, el2n, mesh = upload((3, 2, 2))
    assert mesh.dim == 3 and mesh.nv_local == 27 and mesh.np_local == 4 and mesh.nx == 0
    assert np.array_equal(mesh.coords, coords) and np.array_equal(mesh.el2n, el2n)
    assert mesh.nnodesP == 4 * mesh.nels and mesh.n_shared == mesh.nnodes - mesh.nels
    assert np.array_equal(np.sort(mesh.boundary_nodes), np.sort(O.box_boundary_nodes3(coords, BOX)))


def test_upload_rejects_what_the_element_kernels_cannot_take():
    say("[gpu] hex27 upload errors")
    coords, el2n = O.hex27_mesh(2, 1, 1)
    bad = el2n.copy()
    pos = coords[el2n[0]]
    mir = pos.copy()
    mir[:, 0] = pos[:, 0].min() + pos[:, 0].max() - pos[:, 0]
    bad[0] = el2n[0][[int(np.nonzero((np.abs(pos - m) < 1e-12).all(axis=1))[0][0]) for m in mir]]
    with pytest.raises(ValueError, match="inverted"):
        fa.hex27_mesh_from_arrays(coords, bad)
    moved = coords.copy()
    moved[el2n[0][8]] += (0.05, 0.0, 0.0)
    with pytest.raises(ValueError, match="axis-aligned box"):
        fa.hex27_mesh_from_arrays(moved, el2n)
    bad = el2n.copy()
    bad[1, 0] = len(coords)
    with pytest.raises(ValueError, match="out of range"):
        fa.hex27_mesh_from_arrays(coords, bad)
    with pytest.raises(ValueError):
        fa.hex27_mesh_from_arrays(coords[:, :2], el2n)


def pair(n, eta=(1.3,), etab=None, G=None, rho0=None, K=None, g=(0.0, 0.0, 0.0), dp=None, phases=None,
         tau=None, bc=None, seed=None, grade=0.3):
    nph = len(eta)
    coords, el2n, mesh = upload(n, grade=grade, seed=1)
    kw = dict(eta=eta, etab=etab or (INF,) * nph, G=G or (INF,) * nph, rho0=rho0 or (1.0,) * nph,
              K=K or (INF,) * nph, alpha=(0.0,) * nph, Tref=0.0)
    pr = O.Problem3(coords, el2n, O.material(**kw), g=g, dp=O.drucker_prager(**dp) if dp else None,
                    phases=phases, tau_old=tau, bc=bc)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(g=g, **kw))
    dp_g = fa.DruckerPrager(dp["phi"], dp["psi"], dp["C_"], dp["eta_reg"], dp["Kb"]) if dp else None
    if phases is not None:
        dr.set_phases("v", phases)
        dr.set_phases("P", phases)
    if tau is not None:
        for dst, src in zip(dr.tau_old, tau):
            dst.copyto(src)
    if seed is not None:
        rng = np.random.default_rng(seed)
        for c in range(3):
            pr.v[c][:] = np.sin((c + 2) * coords[:, c]) * coords[:, (c + 1) % 3] + 0.1 * rng.normal(size=len(coords))
            dr.v[c].copyto(pr.v[c])
        pr.P[:] = rng.normal(size=4 * len(el2n))
        dr.P.copyto(pr.P)
    return pr, dr, mesh, coords, dp_g


DP2 = dict(phi=[math.radians(30), math.radians(20)], psi=[0.0, math.radians(3)], C_=[0.4, 0.2],
           eta_reg=[1e-2, 5e-3], Kb=[20.0, INF])


@pytest.mark.parametrize("case", ["bulk_gravity", "vep_2ph"])
def test_upload_residual_and_jacobian_parity(case):
    n = (3, 2, 2)
    ne = 12
    rng = np.random.default_rng(11)
    if case == "bulk_gravity":
        args = dict(eta=(2.0,), etab=(5.0,), rho0=(1.7,), g=(0.1, -0.3, -1.0))
    else:
        args = dict(eta=(1.3, 0.2), G=(0.8, 2.0), K=(50.0, 30.0), rho0=(1.0, 2.0), g=(0.0, 0.0, -1.0),
                    dp=DP2, phases=(np.arange(ne) % 2 + 1).astype(np.int32),
                    tau=[0.3 * rng.normal(size=(ne, 27)) for _ in range(6)])
    say("[gpu] hex27 upload residual", case)
    pr, dr, mesh, _, dp_g = pair(n, seed=3, **args)
    dt, finite_K = 0.4, "K" in args
    pr.pressure_scaling(20.0, dt, finite_K)
    fa.assemble_viscosity_weighted_pressure_scaling(dr, 20.0, dt, finite_K)
    assert rel_err(dr.M_P.numpy(), pr.MP) < 1e-13 and rel_err(dr.gammaP.numpy(), pr.gammaP) < 1e-12
    pr.pressure_residual(dt, finite_K)
    Pnum = pr.gammaP * pr.RP / pr.MP
    pr.momentum_residual(dt, Pnum)
    fa.residual(dr, dt, plastic=dp_g, with_pnum=True, finite_K=finite_K)
    assert rel_err(dr.RP.numpy(), pr.RP) < 1e-12
    for c in range(3):
        e = rel_err(dr.Rv[c].numpy(), pr.Rv[c])
        say("  Rv", c, e)
        assert e < 1e-12
    pr.augmented_jacobian(dt, finite_K)
    fa.jacobian_diagnostics(dr, dt, plastic=dp_g, finite_K=finite_K)
    for c in range(3):
        assert rel_err(dr.dRvdv[c].numpy(), pr.dRdv[c]) < 1e-12
        assert rel_err(dr.PC_v[c].numpy(), pr.PC[c]) < 1e-12
    if args.get("tau") is not None:
        pr.update_current_stress(dt)
        fa.update_stokes_current_stress(dr, mesh, dt, plastic=dp_g)
        for k, t in enumerate(dr.tau):
            assert rel_err(t.numpy(), pr.tau_store[k]) < 1e-12


def _shear_bc(coords, box=BOX, rate=1.0):
    gam = O.box_boundary_nodes3(coords, box)
    on = lambda d, k: gam[np.abs(coords[gam, d] - box[2 * d + k]) < 1e-12]
    zw = np.union1d(on(2, 0), on(2, 1))
    bx = np.union1d(zw, np.union1d(on(0, 0), on(0, 1)))
    by = np.union1d(zw, np.union1d(on(1, 0), on(1, 1)))
    return ((bx, rate * (coords[bx, 2] - box[4])), (by, np.zeros(len(by))), (zw, np.zeros(len(zw))))


def test_upload_solve_parity():
    n = (3, 3, 2)
    ne = 18
    ph = np.ones(ne, np.int32)
    ph[[4, 13]] = 2
    coords, _, _ = upload(n, seed=1)
    bc = _shear_bc(coords)
    say("[gpu] hex27 upload solve 3x3x2, graded and shuffled")
    pr, dr, mesh, coords, _ = pair(n, eta=(1.0, 5.0), rho0=(1.0, 1.5), g=(0.0, 0.0, -0.5), phases=ph, bc=bc)
    kw = dict(tolerance=1e-9, check_interval=20, max_iterations=60_000)
    st_o = pr.solve(0.5, iter_max=20_000, **kw)
    st_g = fa.solve(dr, mesh, *[fa.DirichletBoundaryCondition(d, v) for d, v in bc], dt=0.5, iterMax=20_000,
                    throw_on_failure=False, **kw)
    say("  iters", st_o["iter"], st_g["iter"], "itPH", st_o["itPH"], st_g["itPH"])
    assert st_o["converged"] and st_g["converged"]
    assert st_o["iter"] == st_g["iter"] and st_o["itPH"] == st_g["itPH"]
    assert rel_err(dr.P.numpy(), pr.P) < 1e-10
    for c in range(3):
        assert rel_err(dr.v[c].numpy(), pr.v[c]) < 1e-10


def test_upload_thermal_residual_parity():
    say("[gpu] hex27 upload thermal residual")
    coords, el2n, mesh = upload((3, 2, 2), seed=2)
    mat = dict(k=(1.5, 0.4), Cp=(1.2, 0.8), rho0=(2.0, 3.0), alpha=(0.02, 0.01), K=(50.0, math.inf))
    nn = len(coords)
    ph = np.ones(nn, np.int32)
    ph[((coords - 0.6) ** 2).sum(axis=1) < 0.4 ** 2] = 2
    pr = O.ThermalProblem3(coords, el2n, O.thermal_material(**mat), phases=ph,
                           bc=(np.zeros(0, np.int32), np.zeros(0)))
    dr = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(**mat))
    dr.phases.copyto(ph)
    rng = np.random.default_rng(0)
    for k in ("T", "T0", "source", "P"):
        v = rng.normal(size=nn) + (300.0 if k in ("T", "T0") else 0.0)
        getattr(pr, k)[:] = v
        getattr(dr, k).copyto(v)
    pr.residual(0.3, Tref=290.0, jacobian=True)
    fa.thermal_residual(dr, 0.3, Tref=290.0, compute_jacobian=True)
    for name in ("R", "dRdT", "PC"):
        e = rel_err(getattr(dr, name).numpy(), getattr(pr, name))
        say(f"  {name}: {e:.2e}")
        assert e < 1e-12, name
