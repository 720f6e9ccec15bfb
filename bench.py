#!/usr/bin/env python
"""Benchmark of the MI355X DR Stokes hot path.

Workload (the 512^2 visco-elasto-plastic series): 2-D Stokes on the unit box,
512x512 cells -> 524,288 T7/P1-disc elements, 20 seeded random circular
inclusions (r in U[0.01,0.05], log10 eta in U[-2,4]), G=1, dt=0.1,
Drucker-Prager C=1, phi=30 deg, psi=0, K = etab = Kb = 100, densities
1 / 1.1, g=(0,-1), pure shear; solver keywords of the reference's VEP
examples (pressure_factor 20, rel_drop0 0.75, check_interval 100, CFL_v
0.99); ten physical time steps, each solved to residual 1e-8, with the stress
history handed over between steps. The Perzyna viscosity eta_reg is the one
the series needs to stay stable at this resolution (problems.VEP_ETA_REG).

A "step" is one physical time step of that series. Every step is solved twice
from the same state (velocity, pressure, stress history): by the plain FP64
solver (the reference's iteration) and by the fast form (FP64 state, outer
residuals and convergence test, FP32 inner iterations; at the first step the
learned warm start supplies the initial guess, later steps start from the
previous step as in the reference's time loops). The series continues from
the FP64 solution. --steps is clamped to the series' ten steps; a step whose
solve stops with an error is recorded in the line and ends the series.

value = GDOF.iter/s = (2 N_v + N_p) x DR iterations / seconds / 1e9 of the
fast form over the timed steps, inputs resident in HBM; time_to_tolerance_s
and speedup_vs_fp64 give the time-to-1e-8 comparison.

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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "femtools.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
SERIES_STEPS = 10
T_START = time.perf_counter()


def log(msg):
    print(f"# [{time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(w, nv, ne, npp, wf=8, wi=4, structured=True):
    """Compulsory traffic per element-kernel launch and per DR iteration of the
    FP64 path (each array counted once). Structured box meshes carry no
    connectivity or geometry: the kernel derives both from the element index,
    so they are not part of what it must move."""
    D, NV, nq = 2, 7, 7
    elastic = any(math.isfinite(g) for g in w.G)
    p_old = any(math.isfinite(b) for b in w.etab)
    tau = 3 * nq * ne if elastic else 0
    mesh = 0 if structured else wf * (D * D + 1) * ne + wi * NV * ne
    elem = wf * (D * nv + npp + (npp if p_old else 0) + npp + tau + D * nv) + wi * ne + mesh
    it = wf * (2 * D * nv + 2 * D * nv + D * nv + npp + (npp if p_old else 0) + npp + tau) + wi * ne + mesh
    return elem, it


def algorithmic_bytes_fp32_rm(w, nv, ne, npp):
    """Compulsory traffic of one launch of the FP32 return-map correction
    operator (structured mesh, 16 x 8-cell tiles): it reads the correction
    (velocity at every node, pressure), the FP64 state it linearises about
    (velocity, pressure), the scale gamma_P/M_P, the pressure right-hand side,
    the FP32 stress history (3 x 7 points per element) and the cell phase, and
    writes the summed residual of every grid node and the bubble update (rate
    r+w, preconditioner and right-hand side read, velocity written)."""
    D, nq = 2, 7
    ngrid = nv - ne   # grid nodes: every node except the element-private bubbles
    return (4 * (D * nv + npp) + 8 * (D * nv + npp) + 4 * (2 * npp) + 4 * 3 * nq * ne + 4 * ne
            + 4 * D * ngrid + 4 * D * ne * 5)


def algorithmic_bytes_fp32_tile(w, nv, ne, npp, nt6):
    """Compulsory traffic of one launch of the FP32 visco-elastic correction
    kernel in 16 x 8-cell tiles (structured mesh): it reads the correction's
    velocity (every node) and pressure, the scale gamma_P/M_P, the pressure
    right-hand side and the cell phase, writes the summed residual of every
    grid node, and updates the element-private bubble nodes (rate read and
    written, preconditioner and velocity right-hand side read, velocity
    written)."""
    D = 2
    return 4 * (D * nv + 3 * npp + ne + D * nt6 + D * ne * 5)


def algorithmic_bytes_fp64_tile(w, nv, ne, npp, nt6):
    """Compulsory traffic of one launch of the FP64 inner-iteration kernel in
    16 x 8-cell tiles (structured mesh): velocity (every node), pressure, the
    previous-step pressure (finite bulk modulus), the scale gamma_P/M_P, the
    stress history and the cell phase read; the summed residual of every grid
    node written; the element-private bubble nodes updated (rate read and
    written, preconditioner read, velocity written)."""
    D, nq = 2, 7
    elastic = any(math.isfinite(g) for g in w.G)
    p_old = any(math.isfinite(b) for b in w.etab)
    tau = 3 * nq * ne if elastic else 0
    return 8 * (D * nv + npp + (npp if p_old else 0) + npp + tau + D * nt6 + D * ne * 4) + 4 * ne


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
            m = re.match(r"\s*(.+?)\s+traffic_bytes\s+([0-9.eE+]+)", line)
            if m and m.group(1).startswith(kernel_prefix):
                return float(m.group(2))
    return None


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 256))


def cpu_baseline(w, budget_s=15.0):
    """The CPU oracle (the FP64 C++ restatement of the reference's algorithm)
    on all host cores this process may use: bare DR iterations of the same
    512^2 workload for about budget_s, and the whole 64^2 viscous-inclusion
    solve to 1e-8 (time to tolerance) when it fits the same budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O
    from femtools_amd import problems as PB
    threads = host_cores()
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
    k = int(max(2, min(2000, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    pr.dr_iterations(w.dt, al, be, k)
    t = time.perf_counter() - t0
    out = {"value": round(w.dofs() * k / t / 1e9, 5), "unit": "GDOF.iter/s", "cores": threads, "kind": "port",
           "note": "restatement, not the reference: the C++ FP64 oracle that restates the reference's algorithm "
                   "(the reference is Julia, not runnable here), OpenMP on every host core this process may use",
           "sample": f"{k} bare FP64 DR pseudo-iterations (pressure residual, Pnum, momentum residual with the "
                     f"DP return, Dirichlet, update) of the same {w.nx}x{w.ny} VEP workload, {t:.1f} s"}
    # time to tolerance on the 64^2 viscous inclusion (one centred inclusion, eta 1e3), bounded
    w1 = PB.viscous_inclusion(64)
    c1, e1 = O.t7_mesh(64, 64)
    ph1 = PB.inclusion_phases(c1, e1, w1.circles)
    (b1x, v1x), (b1y, v1y) = PB.pure_shear_bc(c1)
    p1 = O.Problem(c1, e1, O.material(eta=w1.eta, etab=w1.etab, G=w1.G, rho0=w1.rho0, K=w1.K, g=w1.g),
                   phases_v=ph1, phases_P=ph1, bc=((b1x, v1x), (b1y, v1y)))
    p1.vx[:], p1.vy[:] = PB.pure_shear_field(c1)
    t0 = time.perf_counter()
    cap = int(max(1000, 2 * budget_s / max(t1 * (len(e1) / ne), 1e-6)))
    st = p1.solve(1.0, tolerance=1e-8, check_interval=50, max_iterations=cap, iter_max=cap)
    t = time.perf_counter() - t0
    out["viscous64_time_to_1e-8"] = {"seconds": round(t, 3), "iterations": st["iter"], "converged": st["converged"],
                                     "gdof_iter_per_s": round(w1.dofs() * st["iter"] / t / 1e9, 5)}
    return out


def viscous64_gpu():
    """The 64^2 viscous inclusion solved to 1e-8 on the GPU in FP64 and in the
    mixed form (the same problem as the CPU baseline's time-to-tolerance)."""
    import femtools_amd as fa
    from femtools_amd import problems as PB
    w = PB.viscous_inclusion(64)
    mesh = fa.t7_box_mesh(64, 64)
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
    (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
    bcs = (fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy))
    out = {}
    for prec in ("fp64", "fp32"):
        dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
        dr.set_phases("v", ph)
        dr.set_phases("P", ph)
        v0 = PB.pure_shear_field(mesh.coords)
        dr.v.x.copyto(v0[0])
        dr.v.y.copyto(v0[1])
        mesh.ctx.sync()
        t0 = time.perf_counter()
        st = fa.solve(dr, mesh, *bcs, dt=1.0, tolerance=1e-8, check_interval=50, max_iterations=400_000,
                      iterMax=400_000, throw_on_failure=False, precision=prec)
        mesh.ctx.sync()
        t = time.perf_counter() - t0
        out[prec] = {"seconds": round(t, 4), "iterations": st["iter"], "converged": st["converged"],
                     "gdof_iter_per_s": round(w.dofs() * st["iter"] / t / 1e9, 4)}
    return out


# FP64 flops the factorised Hex27 evaluation executes per element in the
# viscous DR-step kernel: 27 active lanes x 1234 (velocity gradient at the
# lane's point 480, continuity / stress / fluxes ~106, nodal residual 648)
HEX27_FLOPS_PER_ELEMENT = 27 * 1234

# graph replay settings the autotuner may try (main sets (0,) under --no-graphs)
GRAPHS = (1, 0)


def hex27_measure(n, max_iterations=200_000):
    """The 3-D path: n^3 Hex27 cells, 50 seeded spheres (r in U[0.02, 0.08],
    log10 eta in U[-1, 3]), viscous, pure shear, from the far-field guess to
    relative residual 1e-7 in FP64 and in FP32 bulk with FP64 norms (the same
    criterion); the FP64 element kernel and the FP32 correction step timed
    alone with HIP events."""
    import femtools_amd as fa
    from femtools_amd import problems as PB
    w = PB.hex27_inclusions(n)
    mesh = fa.hex27_box_mesh(n, n, n)
    ph = PB.hex27_sphere_phases(mesh.coords, mesh.el2n, w.spheres)
    bcs = [fa.DirichletBoundaryCondition(d, v) for d, v in PB.pure_shear_bc3(mesh.coords)]
    v0 = PB.pure_shear_field3(mesh.coords)
    out = {"workload": w.name + f", far-field guess, tol {w.tolerance:g}", "elements": mesh.nels, "dofs": w.dofs()}
    fields = {}
    for prec in ("fp32", "fp64"):
        dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
        dr.set_phases("v", ph)
        dr.set_phases("P", ph)
        fa.assemble_viscosity_weighted_pressure_scaling(dr, w.pressure_factor, w.dt)
        if prec == "fp32":
            out["fp32_correction_step_ms"] = round(fa.time_correction_kernel(dr, w.dt, n=20), 4)
        else:
            out["fp64_element_kernel_ms"] = round(fa.time_element_kernel(dr, w.dt, n=20), 4)
        for f, v in zip(dr.v, v0):
            f.copyto(v)
        dr.P.fill(0.0)
        mesh.ctx.sync()
        t0 = time.perf_counter()
        st = fa.solve(dr, mesh, *bcs, dt=w.dt, max_iterations=max_iterations, iterMax=max_iterations,
                      throw_on_failure=False, precision=prec, **w.solve_kwargs())
        mesh.ctx.sync()
        wall = time.perf_counter() - t0
        log(f"hex27 {n}^3 {prec}: iter {st['iter']} itPH {st['itPH']} converged {st['converged']} "
            f"err_abs {st['err_abs']:.2e} {wall:.2f}s")
        out[prec] = {"seconds": round(wall, 3), "iterations": st["iter"], "itPH": st["itPH"],
                     "converged": bool(st["converged"]), "err_abs": st["err_abs"],
                     "gdof_iter_per_s": round(w.dofs() * st["iter"] / wall / 1e9, 4)}
        fields[prec] = [f.numpy() for f in dr.v] + [dr.P.numpy()]
    out["speedup_fp32_vs_fp64"] = round(out["fp64"]["seconds"] / out["fp32"]["seconds"], 3)
    out["fp32_vs_fp64_max_rel_diff"] = dict(zip(
        ("vx", "vy", "vz", "P"),
        (float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) for a, b in zip(fields["fp32"], fields["fp64"]))))
    flops = HEX27_FLOPS_PER_ELEMENT * mesh.nels
    tf = flops / (out["fp64_element_kernel_ms"] * 1e-3) / 1e12
    out["roofline"] = {"bound": "fp64 alu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(tf / FP64_PEAK_TFLOPS, 4), "kernel": "k_hex27_element (viscous DR step)",
                       "flops_per_launch": flops, "avg_launch_us": round(out["fp64_element_kernel_ms"] * 1e3, 2)}
    return out


def batch_measure(count, seed=0, max_iterations=400_000):
    """The batched workload: count independent viscous T7 problems, n drawn
    from 32..256 in steps of 16 cells, 1-10 inclusions (r in U[0.02, 0.1],
    log10 eta in U[-2, 4]), pure shear, FP32 bulk, each to 1e-8, solved in
    one fem_stokes_solve_batch call (grouped inner iterations replayed from
    graphs). Each problem's time is when it finished, from the call's start."""
    import femtools_amd as fa
    from femtools_amd import problems as PB
    ctx = fa.default_context()
    probs = []
    for w in PB.batch_workloads(count, seed=seed, sizes=tuple(range(32, 257, 16)), log_eta=(-2.0, 4.0)):
        mesh = fa.t7_box_mesh(w.nx, w.ny, ctx=ctx)
        ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
        (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
        dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
        dr.set_phases("v", ph)
        dr.set_phases("P", ph)
        v0 = PB.pure_shear_field(mesh.coords)
        dr.v.x.copyto(v0[0])
        dr.v.y.copyto(v0[1])
        bcs = [fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)]
        kw = dict(dt=w.dt, tolerance=1e-8, check_interval=50, max_iterations=max_iterations, iterMax=max_iterations,
                  max_ph_iterations=w.max_ph_iterations, throw_on_failure=False, precision="fp32")
        probs.append((w, dr, mesh, bcs, kw))
    from femtools_amd import autotune as AT
    tuned = AT.tune_batch([(dr, mesh, bcs, kw) for w, dr, mesh, bcs, kw in probs], iterations=300,
                          workload=f"{count} batched viscous T7 problems, n in 32..256", graphs=GRAPHS)
    log(f"batch tuning: {tuned['tuning']}")
    ctx.sync()
    t0 = time.perf_counter()
    res = fa.solve_batch([(dr, mesh, bcs, kw) for w, dr, mesh, bcs, kw in probs], raise_errors=False)
    ctx.sync()
    wall = time.perf_counter() - t0
    dof_it = sum(w.dofs() * max(st["iter"], 0) for (w, *_r), st in zip(probs, res))
    per = [{"n": w.nx, "inclusions": len(w.circles), "iterations": st["iter"], "converged": bool(st["converged"]),
            "finished_s": round(st["seconds"], 3), "status": st["status"]} for (w, *_r), st in zip(probs, res)]
    return {"workload": f"{count} independent T7/P1-disc viscous problems, n in 32..256 (steps of 16), 1-10 "
                        f"inclusions, log10 eta in U[-2,4], FP32 bulk, tol 1e-8, one batched call",
            "seconds": round(wall, 3), "solves_per_s": round(count / wall, 3),
            "converged": sum(p["converged"] for p in per), "gdof_iter_per_s": round(dof_it / wall / 1e9, 3),
            "problems": per, "tuning": tuned["tuning"], "tuning_measurements": tuned["measurements"]}


class Series:
    """The 512^2 VEP time series: state, boundary conditions and the two solvers."""

    def __init__(self, n, eta_reg):
        import femtools_amd as fa
        from femtools_amd import problems as PB
        self.fa, self.PB = fa, PB
        self.w = w = PB.vep_inclusions(n, eta_reg=eta_reg)
        self.ctx = fa.default_context()
        self.mesh = mesh = fa.t7_box_mesh(w.nx, w.ny)
        self.ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
        (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
        self.bcs = (fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy))
        self.dr = dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K,
                                                           g=w.g), CFL_v=w.CFL_v)
        self.v_bg = PB.pure_shear_field(mesh.coords)
        dr.v.x.copyto(self.v_bg[0])
        dr.v.y.copyto(self.v_bg[1])
        dr.set_phases("v", self.ph)
        dr.set_phases("P", self.ph)
        self.dp = fa.DruckerPrager(w.plastic["phi"], w.plastic["psi"], w.plastic["C"], w.plastic["eta_reg"],
                                   w.plastic["Kb"])
        self.fields = (dr.v.x, dr.v.y, dr.P, dr.tau_old.xx, dr.tau_old.yy, dr.tau_old.xy)

    def snapshot(self):
        return [f.numpy() for f in self.fields]

    def restore(self, snap):
        for f, a in zip(self.fields, snap):
            f.copyto(a)

    def solve(self, prec, timed, tol, max_it, guess=None):
        """One physical step from the current state; guess (vx, vy, P) replaces
        the initial velocity/pressure (the warm start). Returns the stats, with
        'error' set if the solve stopped with an error."""
        fa, dr = self.fa, self.dr
        dr.P0.copyto(dr.P.numpy())
        if guess is not None:
            dr.v.x.copyto(guess[0])
            dr.v.y.copyto(guess[1])
            dr.P.copyto(guess[2])
        kw = self.w.solve_kwargs()
        kw["tolerance"] = tol
        self.ctx.sync()
        t0 = time.perf_counter()
        try:
            st = fa.solve(dr, self.mesh, *self.bcs, dt=self.w.dt, plastic=self.dp, max_iterations=max_it,
                          iterMax=max_it, throw_on_failure=False, time_kernels=timed, precision=prec, **kw)
        except fa.FemError as e:
            st = {"iter": 0, "itPH": 0, "converged": False, "err_abs": float("nan"), "error": str(e),
                  "kernel_ms_element": 0.0, "kernel_ms_gather": 0.0, "timed_launches": 0,
                  "kernel_ms_element_rm": 0.0, "kernel_ms_gather_rm": 0.0, "timed_launches_rm": 0}
        self.ctx.sync()
        st["wall"] = time.perf_counter() - t0
        return st

    def advance(self):
        """Hand the converged step over: current stress -> stress history."""
        fa, dr = self.fa, self.dr
        fa.update_stokes_current_stress(dr, self.mesh, self.w.dt, plastic=self.dp)
        for c in ("xx", "yy", "xy"):
            fa.copyto(getattr(dr.tau_old, c), getattr(dr.tau, c))


def learned_guess(series, model):
    """The learned warm start of the first step: the network's velocity
    perturbation and dynamic pressure for the step's effective (visco-elastic)
    viscosity field, resampled to the mesh."""
    from femtools_amd import warmstart as WS
    w = series.w
    eta_ve = [1.0 / (1.0 / e + 1.0 / (g * w.dt)) for e, g in zip(w.eta, w.G)]
    return WS.predict_on_mesh(model, series.mesh, series.ph, eta_ve, series.v_bg, edot=1.0,
                              eta_scale=eta_ve[0], rho=w.rho0, g=w.g)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=SERIES_STEPS, help="timed physical time steps (<= 10)")
    ap.add_argument("--warmup", type=int, default=1, help="bounded warm-up runs (300 iterations per solver)")
    ap.add_argument("--n", type=int, default=512, help="cells per side")
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--max-iterations", type=int, default=600_000)
    ap.add_argument("--fast-precision", default="fp32-tangent", choices=("fp32", "fp32-tangent"),
                    help="the fast form's solve precision: FP32 corrections with two FP64 returns per point near "
                         "yield, or with the return linearised about the FP64 state at every fold")
    ap.add_argument("--eta-reg", type=float, default=None, help="Perzyna viscosity (default problems.VEP_ETA_REG)")
    ap.add_argument("--deadline-s", type=float, default=470.0,
                    help="no new series step starts when it could end after this many seconds of the run")
    ap.add_argument("--end-s", type=float, default=560.0, help="the measurements after the series fit in this")
    ap.add_argument("--hex27", type=int, default=64, help="cells per side of the 3-D time-to-tolerance pair (0: skip)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every inner iteration from the host (for tracing tools that do not follow "
                         "graph launches); the autotuner then tries no graph replay")
    ap.add_argument("--batch", type=int, default=16, help="problems of the batched measurement (0: skip)")
    ap.add_argument("--no-learned", action="store_true", help="fast form without the learned warm start")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=8.0)
    ap.add_argument("--extras", type=int, default=1, help="measure the 3-D and 64^2 rows when time is left")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch choices instead of the autotuner, e.g. tiles=1,tile_rows=8,graphs=0 (profiling "
                         "runs: counter collection slows the kernels it samples and would mislead the autotuner)")
    ap.add_argument("--advance", type=int, default=0,
                    help="untimed FP64 series steps before the timed ones (to profile the yielding steps)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a single solve would shard by element "
              "strips with an interface-node exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import femtools_amd as fa
    from femtools_amd import problems as PB

    global GRAPHS
    GRAPHS = (0,) if args.no_graphs else (1, 0)
    if args.no_graphs:
        fa.default_context().set_tuning(graphs=0)
    eta_reg = PB.VEP_ETA_REG if args.eta_reg is None else args.eta_reg
    args.advance = max(0, min(args.advance, SERIES_STEPS - 1))
    steps = max(1, min(args.steps, SERIES_STEPS - args.advance))
    if steps != args.steps:
        log(f"--steps {args.steps}: the series has {SERIES_STEPS} physical steps; timing {steps}")
    S = Series(args.n, eta_reg)
    # launch choices (tile shape, graph replay) measured on the series' start state
    from femtools_amd import autotune as AT
    kw_t = S.w.solve_kwargs()
    kw_t.pop("tolerance", None)
    if args.tuning:
        fixed = {k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))}
        S.ctx.set_tuning(**fixed)
        tuned = {"tuning": S.ctx.tuning, "measurements": "fixed by --tuning"}
    else:
        tuned = AT.tune_solve(S.dr, S.mesh, S.bcs, iterations=1500, workload=f"{args.n}x{args.n} VEP step 0",
                              graphs=GRAPHS, dt=S.w.dt, plastic=S.dp, precision=args.fast_precision, **kw_t)
    log(f"tuning: {tuned['tuning']}")
    w, mesh = S.w, S.mesh
    nv, ne, npp = mesh.nnodes, mesh.nels, mesh.nnodesP
    log(f"mesh {w.nx}x{w.ny}: {ne} elements, {w.dofs()} DoFs, eta_reg {eta_reg}")

    model = None
    if not args.no_learned:
        try:
            from femtools_amd import warmstart as WS
            model = WS.load_default()
        except (ImportError, FileNotFoundError, RuntimeError) as e:
            log(f"learned warm start unavailable ({e}); the fast form starts from the pure-shear field")

    # bounded warm-up: a few hundred iterations of each solver from the start
    # state (kernels loaded, graphs instantiated), discarded
    s0 = S.snapshot()
    for i in range(args.warmup):
        for prec in ("fp64", args.fast_precision):
            S.restore(s0)
            st = S.solve(prec, False, args.tolerance, 300)
            log(f"warmup {prec}: {st['iter']} iterations {st['wall']:.2f}s")
        if model is not None:
            learned_guess(S, model)
    S.restore(s0)
    for i in range(args.advance):
        st = S.solve("fp64", False, args.tolerance, 600_000)
        log(f"advance: step {i} fp64 iter {st['iter']} converged {st['converged']} {st['wall']:.2f}s (untimed)")
        S.advance()

    stats = {"fp64": [], "fp32": []}
    diffs, errors, guess_s = [], [], []
    # a step costs at most max_iterations of each solver (~160 us per FP64 +
    # FP32 iteration pair at 512^2); no step starts that could end past the deadline
    worst = args.max_iterations * 160e-6 * (w.nx * w.ny) / 512 ** 2
    for i in range(args.advance, args.advance + steps):
        if time.perf_counter() - T_START + worst > args.deadline_s:
            log(f"deadline {args.deadline_s:.0f}s: the series stops after {i} steps")
            break
        start = S.snapshot()
        sols = {}
        for prec in ("fp64", "fp32"):
            S.restore(start)
            guess = None
            tg = 0.0
            if prec == "fp32" and i == 0 and model is not None:
                t0 = time.perf_counter()
                guess = learned_guess(S, model)
                tg = time.perf_counter() - t0
                guess_s.append(tg)
            st = S.solve(args.fast_precision if prec == "fp32" else prec, True, args.tolerance, args.max_iterations,
                         guess)
            if prec == "fp32" and not st["converged"] and args.fast_precision != "fp32":
                # the fast form's fallback: the step again from its start in
                # precision 1 (two FP64 returns per point near yield); both
                # attempts count in its time to tolerance
                log(f"step {i} {args.fast_precision}: {st.get('error', 'not converged')} after "
                    f"{st['wall']:.2f}s; again in fp32")
                S.restore(start)
                st2 = S.solve("fp32", True, args.tolerance, args.max_iterations, guess)
                st2["wall"] += st["wall"]
                st2["iter"] += st["iter"]
                st2["fallback"] = st.get("error", "not converged")
                st = st2
            st["wall"] += tg          # the warm start's inference counts in the time to tolerance
            stats[prec].append(st)
            sols[prec] = [f.numpy() for f in S.fields[:3]]
            log(f"step {i} {prec}: iter {st['iter']} itPH {st['itPH']} converged {st['converged']} "
                f"err_abs {st['err_abs']:.2e} {st['wall']:.2f}s" + (f" ERROR {st['error']}" if "error" in st else ""))
        if "error" not in stats["fp64"][-1] and "error" not in stats["fp32"][-1]:
            diffs.append({k: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
                          for k, a, b in zip(("vx", "vy", "P"), sols["fp32"], sols["fp64"])})
        ok = [p for p in ("fp64", "fp32") if stats[p][-1]["converged"]]
        if not ok:
            errors.append({"step": i, "fp64": stats["fp64"][-1].get("error", "not converged"),
                           "fp32": stats["fp32"][-1].get("error", "not converged")})
            log(f"step {i}: neither solver converged; the series ends here")
            break
        # the series continues from the FP64 solution (the reference's trajectory)
        for f, a in zip(S.fields[:3], sols[ok[0]]):
            f.copyto(a)
        S.advance()

    done = len(stats["fp32"])
    fast, ref = stats["fp32"], stats["fp64"]
    elapsed = sum(s["wall"] for s in fast)
    iters = sum(s["iter"] for s in fast)
    value = w.dofs() * iters / elapsed / 1e9 if elapsed > 0 else 0.0
    t64 = sum(s["wall"] for s in ref)
    # the time-to-tolerance ratio over the steps both forms finished
    both = [k for k in range(done) if fast[k]["converged"] and ref[k]["converged"]]
    t_fast_both = sum(fast[k]["wall"] for k in both)
    t64_both = sum(ref[k]["wall"] for k in both)

    # roofline on the kernel that takes the largest share of the fast form's
    # timed device time: the return-map correction operator (yielding steps)
    # or the visco-elastic tile kernel
    ms_rm = sum(s["kernel_ms_element_rm"] for s in fast)
    n_rm = sum(s["timed_launches_rm"] for s in fast)
    ms_all = sum(s["kernel_ms_element"] for s in fast)
    n_all = sum(s["timed_launches"] for s in fast)
    ms_ve, n_ve = ms_all - ms_rm, n_all - n_rm
    tot_rm = ms_rm / max(n_rm, 1) * sum(s["iter"] for s in fast) * (n_rm / max(n_all, 1))
    tot_ve = ms_ve / max(n_ve, 1) * sum(s["iter"] for s in fast) * (n_ve / max(n_all, 1))
    nt6 = (2 * w.nx + 1) * (2 * w.ny + 1)
    if n_rm > 0 and tot_rm >= tot_ve and args.fast_precision == "fp32-tangent":
        kname = ("k_t7_element<float,EF_DELTA|EF_TANGENT tile> (FP32 correction operator with the linearised "
                 "return, 16x8-cell tiles; bytes: the tile kernel's, the tangent reads of near-yield elements "
                 "not counted)")
        kbytes = algorithmic_bytes_fp32_tile(w, nv, ne, npp, nt6)
        avg_s = ms_rm / n_rm * 1e-3
        traffic = read_traffic("k_t7_element<float, 1006, 256>")
    elif n_rm > 0 and tot_rm >= tot_ve:
        kname = "k_t7_element<float,EF_DELTA tile> (FP32 return-map correction operator, 16x8-cell tiles)"
        kbytes = algorithmic_bytes_fp32_rm(w, nv, ne, npp)
        avg_s = ms_rm / n_rm * 1e-3
        traffic = read_traffic("k_t7_element<float, 495, 256>")
    else:
        kname = "k_t7_element<float,EF_TILE> (FP32 visco-elastic correction operator, 16x8-cell tiles)"
        kbytes = algorithmic_bytes_fp32_tile(w, nv, ne, npp, nt6)
        avg_s = ms_ve / max(n_ve, 1) * 1e-3
        traffic = read_traffic("k_t7_element<float, 364, 256>")
    achieved = kbytes / avg_s / 1e9 if avg_s > 0 else None
    tun = S.ctx.tuning
    tile64 = bool(tun["tiles"]) and tun["tile_rows"] <= 8 and w.nx % 16 == 0 and w.ny % tun["tile_rows"] == 0
    if tile64:
        e64 = algorithmic_bytes_fp64_tile(w, nv, ne, npp, nt6)
        name64 = "k_t7_element<double,EF_TILE> (fused FP64 element residual + bubble update, 16-cell-wide tiles)"
        traffic64 = read_traffic("k_t7_element<double, 303")
    else:
        e64, _ = algorithmic_bytes(w, nv, ne, npp)
        name64 = "k_t7_element<double> (fused FP64 element residual + bubble update, element-slot gather)"
        traffic64 = None
    n64 = max(1, sum(s["timed_launches"] for s in ref))
    avg64 = sum(s["kernel_ms_element"] for s in ref) / n64 * 1e-3
    out = {
        "metric": "time-to-1e-8 residual (FP32+learned vs plain FP64, same GPU); GDOF.iter/s",
        "value": round(value, 4),
        "unit": "GDOF.iter/s",
        "n_gpus": 1,
        "steps": done,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / max(done, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner iterations, f64 state/residuals/norms/stress history",
        "data": "synthetic (seeded random inclusions, generated in-process)",
        "config": {"workload": w.name + f", dt={w.dt}, DP C=1 phi=30deg psi=0 eta_reg={eta_reg:g}, "
                                        f"K=etab=Kb=100, pure shear, {SERIES_STEPS}-step series, tol {args.tolerance:g}",
                   "cells": [w.nx, w.ny], "elements": ne, "velocity_nodes": nv, "pressure_dofs": npp,
                   "dofs": w.dofs(), "steps_requested": args.steps, "parallelism": "single GPU",
                   "warm_start": ("learned (CNN) at step 0, previous step after" if model is not None
                                  else "pure-shear field at step 0, previous step after")},
        "time_to_tolerance_s": {"fp32_learned": [round(s["wall"], 4) for s in fast],
                                "fp64": [round(s["wall"], 4) for s in ref]},
        "iterations": {"fp32_learned": [s["iter"] for s in fast], "fp64": [s["iter"] for s in ref]},
        "converged": {"fp32_learned": [bool(s["converged"]) for s in fast], "fp64": [bool(s["converged"]) for s in ref]},
        "err_abs": {"fp32_learned": [s["err_abs"] for s in fast], "fp64": [s["err_abs"] for s in ref]},
        "speedup_vs_fp64": round(t64_both / t_fast_both, 3) if t_fast_both > 0 else None,
        "speedup_steps": both,
        "fallbacks": {k: fast[k]["fallback"] for k in range(done) if "fallback" in fast[k]},
        "gdof_iter_per_s_fp64": round(w.dofs() * sum(s["iter"] for s in ref) / t64 / 1e9, 4) if t64 > 0 else None,
        "fp32_vs_fp64_max_rel_diff": diffs,
        "fast_precision": args.fast_precision,
        "tuning": tuned["tuning"],
        "tuning_measurements": tuned["measurements"],
        "warm_start_inference_s": [round(t, 4) for t in guess_s],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic, "kernel": kname, "algorithmic_bytes_per_launch": kbytes,
                     "avg_launch_us": round(avg_s * 1e6, 2),
                     "time_share": {"return_map": round(tot_rm / max(tot_rm + tot_ve, 1e-30), 3),
                                    "visco_elastic": round(tot_ve / max(tot_rm + tot_ve, 1e-30), 3)}},
        "fp64_element_roofline": {"kernel": name64,
                                  "algorithmic_bytes_per_launch": e64, "avg_launch_us": round(avg64 * 1e6, 2),
                                  "achieved_GBs": round(e64 / avg64 / 1e9, 1) if avg64 > 0 else None,
                                  "frac": round(e64 / avg64 / 1e9 / HBM_PEAK_GBS, 4) if avg64 > 0 else None,
                                  "traffic": traffic64},
    }
    if errors:
        out["errors"] = errors
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    log("cpu baseline done")

    def left():
        return args.end_s - (time.perf_counter() - T_START)

    if args.extras and left() > 60:
        try:
            out["viscous64"] = viscous64_gpu()
        except Exception as e:   # reported, never required
            out["extras_error"] = repr(e)
    if args.hex27 > 0 and left() > 120:
        try:
            out["hex27"] = hex27_measure(args.hex27)
        except Exception as e:   # reported, never required
            out["hex27"] = {"error": repr(e)}
    if args.extras and model is not None and left() > 60:
        from femtools_amd import warmstart as WS
        try:
            rows = WS.evaluate_gain(model, 256, [1, 2], log=log, tolerance=1e-8)
            out["warm_start_gain"] = {"n256_steps01": rows, "summary": WS.summarize_gain(rows)}
        except Exception as e:
            out["warm_start_gain"] = {"error": repr(e)}
    if args.batch > 0 and left() > 60:
        try:
            out["batched"] = batch_measure(args.batch)
        except Exception as e:
            out["batched"] = {"error": repr(e)}
    log("done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
