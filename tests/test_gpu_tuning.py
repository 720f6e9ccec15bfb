"""GPU: the context's launch choices (fem_ctx_set_tuning) and the autotuner.

Graph replay on or off runs the same kernels in the same order: results are
bitwise equal. Other tile shapes (4 / 8 / 16 cell rows) sum a node's element
contributions in another grouping, so the FP32 iterations round differently;
each shape must reach the tolerance on the FP64 residual and agree with the
others within it."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")
from femtools_amd import problems as PB  # noqa: E402
from femtools_amd import autotune as AT  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(n=32):
    w = PB.viscous_inclusion(n)
    mesh = fa.t7_box_mesh(n, n)
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    v0 = PB.pure_shear_field(mesh.coords)
    dr.v.x.copyto(v0[0])
    dr.v.y.copyto(v0[1])
    return w, mesh, dr, (fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)), v0


def test_tuning_roundtrip_and_checks():
    ctx = fa.default_context()
    saved = ctx.tuning
    try:
        assert set(saved) == {"tiles", "tile_rows", "graphs", "fold_interval"}
        ctx.set_tuning(tile_rows=16, graphs=0, fold_interval=10)
        t = ctx.tuning
        assert (t["tile_rows"], t["graphs"], t["fold_interval"]) == (16, 0, 10)
        for bad in (dict(tile_rows=6), dict(graphs=2), dict(fold_interval=-1), dict(tiles=3)):
            with pytest.raises(ValueError):
                ctx.set_tuning(**bad)
        with pytest.raises(ValueError):
            ctx.set_tuning(unknown=1)
        assert ctx.tuning["tile_rows"] == 16
    finally:
        ctx.set_tuning(**saved)


def test_graphs_on_off_bitwise_and_tile_shapes_reach_tolerance():
    ctx = fa.default_context()
    saved = ctx.tuning
    res = {}
    try:
        for rows in (4, 8, 16):
            for graphs in (1, 0):
                ctx.set_tuning(tiles=1, tile_rows=rows, graphs=graphs)
                w, mesh, dr, bcs, v0 = _problem()
                st = fa.solve(dr, mesh, *bcs, dt=w.dt, tolerance=1e-8, check_interval=50, precision="fp32",
                              pressure_factor=w.pressure_factor)
                say(f"  rows {rows} graphs {graphs}: iter {st['iter']} err_abs {st['err_abs']:.3e}")
                assert st["converged"]
                res[(rows, graphs)] = (st["iter"], dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy())
    finally:
        ctx.set_tuning(**saved)
    for rows in (4, 8, 16):
        a, b = res[(rows, 1)], res[(rows, 0)]
        assert a[0] == b[0]
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y)
    # solutions meeting 1e-8 lie a few 1e-6 (max rel) from the discrete
    # solution on these problems; two of them within 3e-5 of each other
    ref = res[(8, 1)]
    for rows in (4, 16):
        for x, y in zip(res[(rows, 1)][1:], ref[1:]):
            assert np.abs(x - y).max() <= 3e-5 * np.abs(y).max()


def test_autotuner_picks_a_measured_candidate_and_restores_the_state():
    ctx = fa.default_context()
    saved = ctx.tuning
    try:
        w, mesh, dr, bcs, v0 = _problem(48)
        before = [f.numpy() for f in (*dr.v, dr.P)]
        out = AT.tune_solve(dr, mesh, bcs, iterations=300, workload="48x48 viscous inclusion", dt=w.dt,
                            pressure_factor=w.pressure_factor)
        after = [f.numpy() for f in (*dr.v, dr.P)]
        for a, b in zip(before, after):
            assert np.array_equal(a, b)
        cands = {(m["tiles"], m["tile_rows"], m["graphs"]) for m in out["measurements"]}
        t = out["tuning"]
        assert (t["tiles"], t["tile_rows"], t["graphs"]) in cands
        assert ctx.tuning == t
        # 48 rows admit 4, 8 and 16-row tiles: six tile candidates and two without tiles
        assert len(cands) == 8
    finally:
        ctx.set_tuning(**saved)
