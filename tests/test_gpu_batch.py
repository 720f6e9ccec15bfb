"""Batched independent solves (fem_stokes_solve_batch): each state on its
own context and stream, several host threads. Every solve is deterministic,
so the batch must reproduce the one-at-a-time solves bit for bit."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")
from femtools_amd import problems as PB  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(n, contrast, precision):
    ctx = fa.Context(0)
    w = PB.viscous_inclusion(n)
    mesh = fa.t7_box_mesh(n, n, ctx=ctx)
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=(1.0, contrast), etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    v0 = PB.pure_shear_field(mesh.coords)
    bcs = [fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)]
    kw = dict(dt=w.dt, pressure_factor=w.pressure_factor, tolerance=1e-8, check_interval=w.check_interval,
              precision=precision)
    return dr, mesh, bcs, kw, v0


def _reset(dr, v0):
    dr.v.x.copyto(v0[0])
    dr.v.y.copyto(v0[1])
    dr.P.fill(0.0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_sequential_solves(precision):
    say(f"[gpu] batched solves {precision}")
    cases = [(16, 1e3), (24, 10.0), (32, 1e2), (16, 3.0), (20, 1e3)]
    probs = [_problem(n, c, precision) for n, c in cases]
    seq = []
    for dr, mesh, bcs, kw, v0 in probs:
        _reset(dr, v0)
        st = fa.solve(dr, mesh, *bcs, **kw)
        seq.append((st, dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy()))
    for dr, mesh, bcs, kw, v0 in probs:
        _reset(dr, v0)
    out = fa.solve_batch([(dr, mesh, bcs, kw) for dr, mesh, bcs, kw, _ in probs], nthreads=3)
    for (n, c), (dr, *_r), st, (st0, vx, vy, P) in zip(cases, probs, out, seq):
        say(f"  n={n} contrast={c:g}: iter {st['iter']} (sequential {st0['iter']}) converged {st['converged']}")
        assert st["status"] == 0 and st["converged"]
        assert st["iter"] == st0["iter"] and st["itPH"] == st0["itPH"]
        assert np.array_equal(dr.v.x.numpy(), vx) and np.array_equal(dr.v.y.numpy(), vy)
        assert np.array_equal(dr.P.numpy(), P)


def test_batch_rejects_a_shared_context():
    ctx = fa.Context(0)
    w = PB.viscous_inclusion(8)
    m1, m2 = fa.t7_box_mesh(8, 8, ctx=ctx), fa.t7_box_mesh(8, 8, ctx=ctx)
    mat = fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g)
    d1, d2 = fa.StokesDR(m1, mat), fa.StokesDR(m2, mat)
    (bx, vx), (by, vy) = PB.pure_shear_bc(m1.coords)
    bcs = [fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)]
    with pytest.raises(ValueError, match="one context"):
        fa.solve_batch([(d1, m1, bcs, dict(dt=1.0)), (d2, m2, bcs, dict(dt=1.0))])
