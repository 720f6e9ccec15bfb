"""Autotuner of a context's launch choices (fem_tuning).

The choices that change the speed of a solve but not what it computes are
measured on the caller's own problem and kept as JSON:

* tile_rows (4, 8 or 16 cell rows per tile of 16 cells: blocks of 128, 256
  or 512 threads) and tiles on/off for the FP32 inner iterations of
  structured 2-D meshes;
* graphs on/off (check intervals replayed from HIP graphs or launched from
  the host);
* optionally fold_interval, the mixed-precision fold cadence with a yield
  model, which changes the iteration count as well as the cost of an
  iteration: it is timed to the tolerance.

Every candidate runs from the same start state; the state is restored after
tuning. The JSON holds the device name, the workload description, every
measurement and the chosen "tuning" object, which default_context() applies
when FEMDR_TUNING names the file.
"""
from __future__ import annotations

import json
import time

from . import solve, solve_batch


def _candidates(meshes, graphs=(1, 0)):
    """(tiles, tile_rows, graphs) combinations the meshes admit."""
    out = []
    for rows in (4, 8, 16):
        if all(m.dim == 2 and getattr(m, "nx", 0) and m.nx % 16 == 0 and m.ny % rows == 0 for m in meshes):
            for g in graphs:
                out.append(dict(tiles=1, tile_rows=rows, graphs=g))
    for g in graphs:
        out.append(dict(tiles=0, tile_rows=8, graphs=g))
    return out


def _snapshot(dr):
    return [f.numpy() for f in (*dr.v, dr.P)]


def _restore(dr, snap):
    for f, a in zip((*dr.v, dr.P), snap):
        f.copyto(a)


def tune_solve(dr, mesh, bcs, iterations=1000, fold_intervals=None, out=None, workload="", graphs=(1, 0),
               **solve_kwargs):
    """Times fixed-length solves of (dr, mesh, bcs) under every admissible
    (tiles, tile_rows, graphs in `graphs`) and keeps the fastest per
    iteration; with fold_intervals (a list), then times the solve to its
    tolerance under each fold interval. Returns the JSON dict (also written to
    out)."""
    ctx = dr.ctx
    saved = ctx.tuning
    snap = _snapshot(dr)
    kw = dict(solve_kwargs)
    kw.update(max_iterations=int(iterations), iterMax=int(iterations), throw_on_failure=False, tolerance=1e-300)
    kw.setdefault("precision", "fp32")
    meas = []
    best = None
    for cand in _candidates([mesh], graphs):
        ctx.set_tuning(**dict(saved, **cand))
        walls = []
        for _ in range(2):   # the first run instantiates the graphs
            _restore(dr, snap)
            ctx.sync()
            t0 = time.perf_counter()
            st = solve(dr, mesh, *bcs, **kw)
            ctx.sync()
            walls.append(time.perf_counter() - t0)
        per = walls[-1] / max(st["iter"], 1)
        meas.append(dict(cand, us_per_iteration=round(per * 1e6, 3), iterations=st["iter"]))
        if best is None or per < best[0]:
            best = (per, cand)
    chosen = dict(saved, **best[1])
    if fold_intervals:
        tol_kw = dict(solve_kwargs)
        tol_kw.update(throw_on_failure=False)
        tol_kw.setdefault("precision", "fp32")
        best_f = None
        for fi in fold_intervals:
            ctx.set_tuning(**dict(chosen, fold_interval=int(fi)))
            _restore(dr, snap)
            ctx.sync()
            t0 = time.perf_counter()
            st = solve(dr, mesh, *bcs, **tol_kw)
            ctx.sync()
            wall = time.perf_counter() - t0
            meas.append(dict(fold_interval=int(fi), seconds_to_tolerance=round(wall, 4), iterations=st["iter"],
                             converged=bool(st["converged"])))
            if st["converged"] and (best_f is None or wall < best_f[0]):
                best_f = (wall, int(fi))
        if best_f is not None:
            chosen["fold_interval"] = best_f[1]
    ctx.set_tuning(**chosen)
    _restore(dr, snap)
    res = {"device": _device_name(), "workload": workload, "measurements": meas, "tuning": chosen}
    if out:
        with open(out, "w") as f:
            json.dump(res, f, indent=1)
    return res


def tune_batch(problems, iterations=500, out=None, workload="", graphs=(1, 0)):
    """Times one fixed-length solve_batch of the problems ((dr, mesh, bcs, kw)
    tuples on one context) under every admissible (tiles, tile_rows, graphs in
    `graphs`) and keeps the fastest. Returns the JSON dict (also written to
    out)."""
    ctx = problems[0][0].ctx
    saved = ctx.tuning
    snaps = [_snapshot(dr) for dr, *_ in problems]
    meas, best = [], None
    for cand in _candidates([m for _, m, *_ in problems], graphs):
        ctx.set_tuning(**dict(saved, **cand))
        walls = []
        for _ in range(2):
            for (dr, *_r), sn in zip(problems, snaps):
                _restore(dr, sn)
            items = []
            for dr, mesh, bcs, kw in problems:
                k = dict(kw)
                k.update(tolerance=1e-300, max_iterations=int(iterations), iterMax=int(iterations),
                         throw_on_failure=False)
                items.append((dr, mesh, bcs, k))
            ctx.sync()
            t0 = time.perf_counter()
            solve_batch(items, raise_errors=False)
            ctx.sync()
            walls.append(time.perf_counter() - t0)
        meas.append(dict(cand, seconds=round(walls[-1], 4)))
        if best is None or walls[-1] < best[0]:
            best = (walls[-1], cand)
    chosen = dict(saved, **best[1])
    ctx.set_tuning(**chosen)
    for (dr, *_r), sn in zip(problems, snaps):
        _restore(dr, sn)
    res = {"device": _device_name(), "workload": workload, "iterations": int(iterations), "measurements": meas,
           "tuning": chosen}
    if out:
        with open(out, "w") as f:
            json.dump(res, f, indent=1)
    return res


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return "unknown"
