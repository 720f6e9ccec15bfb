"""GPU: CellPressureStokesDR (the viscous Hex27 solver with the cell pressure
modes (1, ξ, η, ζ); stokes_types.jl:545-613, DR.jl:960-1057) through the C
ABI against the oracle: preconditioner and residuals to 1e-12, solves with
the same iteration count and fields to 1e-10; the reference's
boundary-value and constructor tests (test_stokes_3d_boundary_values.jl)."""
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

from test_oracle_cell_pressure import block_problem  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def pair(n, seed=0, g=(0.0, 0.0, -1.0), bc=True):
    pr, ph = block_problem(n, g=g, bc=bc)
    mesh = fa.hex27_box_mesh(*n)
    assert np.array_equal(mesh.coords, pr.coords)
    dr = fa.CellPressureStokesDR(mesh, fa.StokesMaterial(eta=(1.0, 50.0), rho0=(1.0, 2.0), g=g), phases=ph)
    bcs = None
    if bc:
        bcs = tuple(fa.DirichletBoundaryCondition(d, v) for d, v in pr.bc)
    return pr, dr, mesh, bcs


def test_gpu_cell_pressure_preconditioner_and_residuals():
    say("[gpu] cell-pressure preconditioner and residuals")
    pr, dr, mesh, _ = pair((3, 2, 2), g=(0.3, -0.2, -1.0), bc=False)
    rng = np.random.default_rng(1)
    for c, comp in enumerate("xyz"):
        pr.v[c][:] = rng.normal(size=pr.nnodes)
        getattr(dr.v, comp).copyto(pr.v[c])
    pr.P[:] = rng.normal(size=4 * pr.nels)
    dr.P.copyto(pr.P.reshape(-1, 4))
    dg, M = pr.cp3_preconditioner()
    fa.stokes_preconditioner_3d(dr)
    errs = {f"D{c}": rel(getattr(dr.diagonal, c).numpy(), dg[k]) for k, c in enumerate("xyz")}
    errs["M"] = rel(dr.pressure_mass.numpy(), M.reshape(-1, 4))
    fa.assemble_stokes_pressure_residual_3d(dr)
    errs["RP"] = rel(dr.RP.numpy(), pr.cp3_pressure_residual().reshape(-1, 4))
    fa.assemble_stokes_momentum_residual_3d(dr)
    R = pr.cp3_momentum_residual()
    for k, c in enumerate("xyz"):
        errs["R" + c] = rel(getattr(dr.Rv, c).numpy(), R[k])
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-12


@pytest.mark.parametrize("n", [(3, 3, 3), (4, 4, 4), (6, 4, 5)])
def test_gpu_cell_pressure_solve_parity(n):
    say(f"[gpu] cell-pressure solve parity n={n}")
    pr, dr, mesh, bcs = pair(n)
    kw = dict(tolerance=1e-8, max_iterations=60_000, check_interval=100)
    so = pr.cp3_solve(**kw)
    st = fa.solve(dr, mesh, *bcs, collect_history=True, **kw)
    errs = {f"v{c}": rel(getattr(dr.v, c).numpy(), pr.v[k]) for k, c in enumerate("xyz")}
    errs["P"] = rel(dr.P.numpy(), pr.P.reshape(-1, 4))
    say(f"  oracle iter {so['iterations']} err {so['residual']:.3e} | gpu iter {st['iterations']} "
        f"err {st['residual']:.3e} ({st['seconds']:.2f} s); " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert so["converged"] and st["converged"]
    assert st["iterations"] == so["iterations"] == st["iter"]
    assert len(st["history"]) == st["iterations"] // 100
    assert max(errs.values()) < 1e-10


def _expanding():
    mesh = fa.hex27_box_mesh(1, 1, 1)
    gam = mesh.boundary_nodes
    vals = [mesh.coords[gam, 0] if c == 0 else np.zeros(len(gam)) for c in range(3)]
    bcs = tuple(fa.DirichletBoundaryCondition(gam, v) for v in vals)
    mat = fa.StokesMaterial(eta=1.0, rho0=1.0, g=(0.0, 0.0, 0.0))
    return mesh, gam, vals, bcs, mat


def test_gpu_cell_pressure_first_residual_sees_boundary_values():
    # test_stokes_3d_boundary_values.jl:7-51
    say("[gpu] cell-pressure boundary values")
    mesh, gam, vals, bcs, mat = _expanding()
    dr = fa.CellPressureStokesDR(mesh, mat)
    with pytest.raises(fa.FemError, match="3-D Stokes solve did not converge"):
        fa.solve(dr, mesh, *bcs, check_interval=1, max_iterations=1)
    dr = fa.CellPressureStokesDR(mesh, mat)
    st = fa.solve(dr, mesh, *bcs, check_interval=1, max_iterations=1, collect_history=True, throw_on_failure=False)
    assert not st["converged"] and st["reached_total_iter"]
    assert st["err_P"] > 1e-5
    assert len(st["history"]) == 1 and st["history"][0]["err_P"] == st["err_P"]
    for c, comp in enumerate("xyz"):
        assert np.array_equal(getattr(dr.v, comp).numpy()[gam], vals[c])
    cons = fa.CellPressureStokesDR(mesh, mat)
    for c, comp in enumerate("xyz"):
        v = np.zeros(mesh.nnodes)
        v[gam] = vals[c]
        getattr(cons.v, comp).copyto(v)
    sc = fa.solve(cons, mesh, *bcs, check_interval=1, max_iterations=1, throw_on_failure=False)
    assert sc["err_P"] == pytest.approx(st["err_P"])
    assert np.allclose(cons.P.numpy(), dr.P.numpy())
    for comp in "xyz":
        assert np.allclose(getattr(cons.v, comp).numpy(), getattr(dr.v, comp).numpy())
    # the oracle agrees on the first iteration
    pr = O.Problem3(O.hex27_mesh(1, 1, 1)[0], O.hex27_mesh(1, 1, 1)[1], O.material(eta=(1.0,), rho0=(1.0,)),
                    g=(0.0, 0.0, 0.0), bc=tuple((gam, v) for v in vals))
    so = pr.cp3_solve(check_interval=1, max_iterations=1)
    assert rel(dr.P.numpy(), pr.P.reshape(-1, 4)) < 1e-12


def test_gpu_cell_pressure_constructor_checks():
    say("[gpu] cell-pressure constructor checks")
    mesh, gam, vals, bcs, mat = _expanding()
    with pytest.raises(fa.DimensionMismatch, match="phases has 2 entries"):
        fa.CellPressureStokesDR(mesh, mat, phases=[1, 1])
    with pytest.raises(ValueError, match="phases must lie in 1:1"):
        fa.CellPressureStokesDR(mesh, mat, phases=[2])
    with pytest.raises(ValueError, match="purely viscous"):
        fa.CellPressureStokesDR(mesh, fa.StokesMaterial(eta=1.0, G=1.0, g=(0.0, 0.0, 0.0)))
    with pytest.raises(ValueError, match="requires a Hex27 mesh"):
        fa.CellPressureStokesDR(fa.t7_box_mesh(1, 1), fa.StokesMaterial(eta=1.0))


def test_gpu_cell_pressure_adjoint_and_material_gradient():
    # solve_adjoint! / stokes_material_gradient_3d (DR_adjoint.jl:446-531)
    # against the oracle (itself pinned to the transposed system and finite
    # differences): iteration counts, λ to 1e-10, gradients to 1e-10
    say("[gpu] cell-pressure adjoint and material gradient")
    n = (4, 3, 3)
    pr, dr, mesh, bcs = pair(n)
    kw = dict(tolerance=1e-9, max_iterations=60_000, check_interval=100)
    so = pr.cp3_solve(**kw)
    st = fa.solve(dr, mesh, *bcs, **kw)
    assert so["converged"] and st["converged"] and st["iterations"] == so["iterations"]
    v_fwd = [getattr(dr.v, c).numpy() for c in "xyz"]
    P_fwd = dr.P.numpy()
    rng = np.random.default_rng(5)
    obj = [rng.normal(size=pr.nnodes) for _ in range(3)]
    lo, lPo, sao = pr.cp3_solve_adjoint(obj, **kw)
    sa = fa.solve_adjoint(dr, mesh, *bcs, objective_v=obj, lam_v=[np.zeros(pr.nnodes)] * 3,
                          lam_P=np.zeros((pr.nels, 4)), **kw)
    errs = {f"lam_{c}": rel(sa["lam_v"][k], lo[k]) for k, c in enumerate("xyz")}
    errs["lam_P"] = rel(sa["lam_P"], lPo.reshape(-1, 4))
    say(f"  adjoint oracle iter {sao['iterations']} | gpu iter {sa['iterations']}; "
        + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert sao["converged"] and sa["converged"] and sa["iterations"] == sao["iterations"]
    assert max(errs.values()) < 1e-10
    # the forward state is untouched
    for k, c in enumerate("xyz"):
        assert np.array_equal(getattr(dr.v, c).numpy(), v_fwd[k])
    assert np.array_equal(dr.P.numpy(), P_fwd)
    go = pr.cp3_material_gradient(pr.v, lo)
    gg = fa.stokes_material_gradient_3d(dr, mesh, sa["lam_v"])
    say(f"  gradients oracle {go} | gpu {gg}")
    for p in range(2):
        assert gg["density_gradient"][p] == pytest.approx(go[0][p], rel=1e-10)
        assert gg["viscosity_gradient"][p] == pytest.approx(go[1][p], rel=1e-10)
    # a warm start from the converged adjoint converges at the first check
    with pytest.raises(fa.DimensionMismatch, match="objective_v"):
        fa.solve_adjoint(dr, mesh, *bcs, objective_v=obj[:2], lam_v=[np.zeros(pr.nnodes)] * 3,
                         lam_P=np.zeros((pr.nels, 4)))
    warm = fa.solve_adjoint(dr, mesh, *bcs, objective_v=obj, lam_v=list(sa["lam_v"]), lam_P=sa["lam_P"], **kw)
    assert warm["converged"] and warm["iterations"] <= kw["check_interval"]
