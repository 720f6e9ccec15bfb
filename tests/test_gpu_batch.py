"""Batched independent solves (fem_stokes_solve_batch) on one context:
mixed-precision viscous problems on tile-sized structured meshes run their
inner iterations as grouped launches replayed from a graph, the rest one
after the other. Every solve is deterministic and the grouped launches run
the same kernel body per problem, so the batch must reproduce the
one-at-a-time solves bit for bit (iterations, outer iterations, fields)."""
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


def _problem(n, contrast, precision, ctx, circles=None):
    w = PB.viscous_inclusion(n)
    mesh = fa.t7_box_mesh(n, n, ctx=ctx)
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, circles or w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
    nph = int(ph.max())
    eta = (1.0,) + (contrast,) * (nph - 1)
    inf = math.inf
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=eta, etab=(inf,) * nph, G=(inf,) * nph, rho0=(1.0,) * nph,
                                             K=(inf,) * nph, g=(0.0, 0.0)))
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


def _compare(cases, probs):
    seq = []
    for dr, mesh, bcs, kw, v0 in probs:
        _reset(dr, v0)
        st = fa.solve(dr, mesh, *bcs, **kw)
        seq.append((st, dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy()))
    for dr, mesh, bcs, kw, v0 in probs:
        _reset(dr, v0)
    out = fa.solve_batch([(dr, mesh, bcs, kw) for dr, mesh, bcs, kw, _ in probs])
    for case, (dr, *_r), st, (st0, vx, vy, P) in zip(cases, probs, out, seq):
        say(f"  {case}: iter {st['iter']} (single {st0['iter']}) itPH {st['itPH']} converged {st['converged']}")
        assert st["status"] == 0 and st["converged"]
        assert st["iter"] == st0["iter"] and st["itPH"] == st0["itPH"]
        assert np.array_equal(dr.v.x.numpy(), vx) and np.array_equal(dr.v.y.numpy(), vy)
        assert np.array_equal(dr.P.numpy(), P)


def test_grouped_batch_matches_single_solves():
    say("[gpu] grouped batch, mixed precision")
    ctx = fa.default_context()
    cases = [(16, 1e3), (32, 10.0), (48, 1e2), (16, 3.0), (64, 1e3), (32, 0.1)]
    probs = [_problem(n, c, "fp32", ctx) for n, c in cases]
    _compare(cases, probs)


def test_grouped_batch_of_random_inclusion_problems():
    # the batched workload's generator (sizes in steps of 16, 1-10 inclusions,
    # log10 eta in U[-2, 4]) at small sizes and moderate contrast
    say("[gpu] grouped batch, random inclusions")
    ctx = fa.default_context()
    rng = np.random.default_rng(4)
    cases, probs = [], []
    for k in range(10):
        n = int(rng.choice([16, 32, 48]))
        ninc = int(rng.integers(1, 11))
        circles = list(zip(rng.uniform(0.1, 0.9, ninc), rng.uniform(0.1, 0.9, ninc), rng.uniform(0.02, 0.1, ninc)))
        c = float(10.0 ** rng.uniform(-1, 2))
        cases.append((n, ninc, round(c, 3)))
        probs.append(_problem(n, c, "fp32", ctx, circles))
    _compare(cases, probs)


def test_fp64_batch_runs_one_after_the_other():
    say("[gpu] FP64 batch")
    ctx = fa.default_context()
    cases = [(16, 1e3), (24, 10.0), (20, 3.0)]
    probs = [_problem(n, c, "fp64", ctx) for n, c in cases]
    _compare(cases, probs)


def test_batch_rejects_states_on_two_contexts():
    w = PB.viscous_inclusion(8)
    m1, m2 = fa.t7_box_mesh(8, 8, ctx=fa.Context(0)), fa.t7_box_mesh(8, 8, ctx=fa.Context(0))
    mat = fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g)
    d1, d2 = fa.StokesDR(m1, mat), fa.StokesDR(m2, mat)
    (bx, vx), (by, vy) = PB.pure_shear_bc(m1.coords)
    bcs = [fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)]
    with pytest.raises(ValueError, match="one context"):
        fa.solve_batch([(d1, m1, bcs, dict(dt=1.0)), (d2, m2, bcs, dict(dt=1.0))])
