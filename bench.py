#!/usr/bin/env python
"""Benchmark of the MI355X DR Stokes hot path.

Workload: 2-D visco-elasto-plastic Stokes, 512x512 cells -> 524,288 T7/P1-disc
elements, 20 seeded random circular inclusions (r in U[0.01,0.05],
log10 eta in U[-2,4]), G=1, dt=0.1, Drucker-Prager C=1, phi=30 deg, psi=0,
K = etab = Kb = 100, densities 1 / 1.1, g=(0,-1), pure shear; solver keywords
of the reference's VEP examples (pressure_factor 20, rel_drop0 0.75,
check_interval 100, CFL_v 0.99); each physical time step solved to residual
1e-8 (FP64), stress history handed over between steps.

A "step" is one physical time step (one solve! to 1e-8 plus the stress
refresh). The warmup steps are the first steps of the same time series.
value = GDOF.iter/s = (2 N_v + N_p) x DR iterations / seconds / 1e9 over the
timed steps, inputs resident in HBM.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "femtools.jl_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(w, nv, ne, npp, wf=8, wi=4):
    """Compulsory traffic per element-kernel launch and per DR iteration
    (each array counted once; geometry as (D^2+1) words per element)."""
    D, NV, nq = 2, 7, 7
    elastic = any(math.isfinite(g) for g in w.G)
    p_old = any(math.isfinite(b) for b in w.etab)
    tau = 3 * nq * ne if elastic else 0
    elem = wf * (D * nv + npp + (npp if p_old else 0) + npp + tau + (D * D + 1) * ne + D * nv) + wi * (NV * ne + ne)
    it = wf * (2 * D * nv + 2 * D * nv + D * nv + npp + (npp if p_old else 0) + npp + tau + (D * D + 1) * ne) \
        + wi * (NV * ne + ne)
    return elem, it


def read_traffic(kernel_prefix):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    as committed in profiles/*traffic*.txt (line: '<kernel> traffic_bytes <N>')."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if "traffic" not in fn or not fn.endswith(".txt"):
            continue
        for line in open(os.path.join(d, fn)):
            m = re.match(r"\s*(\S+)\s+traffic_bytes\s+([0-9.eE+]+)", line)
            if m and m.group(1).startswith(kernel_prefix):
                return float(m.group(2))
    return None


def cpu_baseline(w, budget_s=15.0):
    """The CPU oracle (FP64 restatement of the reference's path) on the host
    cores: bare DR iterations of the same workload, bounded to ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O
    from femtools_amd import problems as PB
    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    coords, el2n = O.t7_mesh(w.nx, w.ny)
    ph = PB.inclusion_phases(coords, el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(coords)
    ne = len(el2n)
    mat = O.material(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g)
    dp = O.drucker_prager(w.plastic["phi"], w.plastic["psi"], w.plastic["C"], w.plastic["eta_reg"],
                          w.plastic["Kb"]) if w.plastic else None
    pr = O.Problem(coords, el2n, mat, dp=dp, phases_v=ph, phases_P=ph,
                   tau_old=tuple(np.zeros((ne, 7)) for _ in range(3)), bc=((bx, vx), (by, vy)))
    pr.pressure_scaling(w.pressure_factor, w.dt)
    pr.vx[:], pr.vy[:] = PB.pure_shear_field(coords)
    pr.PCx[:] = 1.0
    pr.PCy[:] = 1.0
    al, be = [1e-3, 1e-3], [0.9, 0.9]
    t0 = time.perf_counter()
    pr.dr_iterations(w.dt, al, be, 1)
    t1 = time.perf_counter() - t0
    k = int(max(2, min(200, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    pr.dr_iterations(w.dt, al, be, k)
    t = time.perf_counter() - t0
    return {"value": w.dofs() * k / t / 1e9, "unit": "GDOF.iter/s", "cores": threads, "kind": "port",
            "sample": f"{k} bare FP64 DR pseudo-iterations (pressure residual, Pnum, momentum residual with the "
                      f"DP return, Dirichlet, update) of the same {w.nx}x{w.ny} VEP workload on {threads} OpenMP "
                      f"threads, {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed physical time steps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed physical time steps before")
    ap.add_argument("--n", type=int, default=512, help="cells per side")
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--max-iterations", type=int, default=200_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md on sharding); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import femtools_amd as fa
    from femtools_amd import problems as PB

    w = PB.vep_inclusions(args.n, nsteps=args.warmup + args.steps)
    ctx = fa.default_context()
    mesh = fa.t7_box_mesh(w.nx, w.ny)
    nv, ne, npp = mesh.nnodes, mesh.nels, mesh.nnodesP
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
    bcx, bcy = fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g),
                     CFL_v=w.CFL_v)
    v0x, v0y = PB.pure_shear_field(mesh.coords)
    dr.v.x.copyto(v0x)
    dr.v.y.copyto(v0y)
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    dp = fa.DruckerPrager(w.plastic["phi"], w.plastic["psi"], w.plastic["C"], w.plastic["eta_reg"], w.plastic["Kb"])

    def time_step(timed):
        dr.P0.copyto(dr.P.numpy())
        kw = w.solve_kwargs()
        kw["tolerance"] = args.tolerance
        st = fa.solve(dr, mesh, bcx, bcy, dt=w.dt, plastic=dp, max_iterations=args.max_iterations,
                      iterMax=args.max_iterations, throw_on_failure=False, time_kernels=timed, **kw)
        fa.update_stokes_current_stress(dr, mesh, w.dt, plastic=dp)
        for c in ("xx", "yy", "xy"):
            getattr(dr.tau_old, c).copyto(getattr(dr.tau, c).numpy())
        return st

    for i in range(args.warmup):
        st = time_step(False)
        print(f"# warmup step {i}: iter {st['iter']} converged {st['converged']} {st['seconds']:.2f}s",
              file=sys.stderr, flush=True)
    ctx.sync()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        st = time_step(True)
        stats.append(st)
        print(f"# step {i}: iter {st['iter']} converged {st['converged']} err_abs {st['err_abs']:.2e} "
              f"{st['seconds']:.2f}s", file=sys.stderr, flush=True)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    iters = sum(s["iter"] for s in stats)
    value = w.dofs() * iters / elapsed / 1e9
    ms_el = sum(s["kernel_ms_element"] for s in stats)
    ms_ga = sum(s["kernel_ms_gather"] for s in stats)
    nl = max(1, sum(s["timed_launches"] for s in stats))
    elem_bytes, iter_bytes = algorithmic_bytes(w, nv, ne, npp)
    avg_el_s = ms_el / nl * 1e-3
    avg_it_s = (ms_el + ms_ga) / nl * 1e-3
    achieved = elem_bytes / avg_el_s / 1e9 if avg_el_s > 0 else None
    traffic = read_traffic("k_t7_element")
    out = {
        "metric": "GDOF.iter/s (time-to-1e-8 residual companion)",
        "value": round(value, 4),
        "unit": "GDOF.iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random inclusions, generated in-process)",
        "config": {"workload": w.name + f", dt={w.dt}, DP C=1 phi=30deg psi=0 eta_reg=8e-3, K=etab=Kb=100, "
                                        f"pure shear, tol {args.tolerance:g}, FP64",
                   "cells": [w.nx, w.ny], "elements": ne, "velocity_nodes": nv, "pressure_dofs": npp,
                   "dofs": w.dofs(), "time_steps_timed": args.steps, "parallelism": "single GPU"},
        "time_to_tolerance_s": [round(s["seconds"], 4) for s in stats],
        "iterations": [s["iter"] for s in stats],
        "converged": [bool(s["converged"]) for s in stats],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic, "kernel": "k_t7_element (fused element residual + bubble update)",
                     "algorithmic_bytes_per_launch": elem_bytes,
                     "avg_launch_us": round(avg_el_s * 1e6, 2)},
        "iteration_roofline": {"algorithmic_bytes": iter_bytes, "avg_us": round(avg_it_s * 1e6, 2),
                               "achieved_GBs": round(iter_bytes / avg_it_s / 1e9, 1) if avg_it_s > 0 else None,
                               "avg_gather_us": round(ms_ga / nl * 1e3, 2)},
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
