#!/usr/bin/env python
"""Benchmark of the MI355X DR Stokes hot path.

Workload: 2-D visco-elasto-plastic Stokes, 512x512 cells -> 524,288 T7/P1-disc
elements, 20 seeded random circular inclusions (r in U[0.01,0.05],
log10 eta in U[-2,4]), G=1, dt=0.1, Drucker-Prager C=1, phi=30 deg, psi=0,
K = etab = Kb = 100, densities 1 / 1.1, g=(0,-1), pure shear; solver keywords
of the reference's VEP examples (pressure_factor 20, rel_drop0 0.75,
check_interval 100, CFL_v 0.99); each physical time step solved to residual
1e-8, stress history handed over between steps.

A "step" is one physical time step. Each timed step starts from the same state
(velocity, pressure, stress history) for both solvers: the plain FP64 solve
(the reference's iteration) and the mixed-precision solve (FP64 state, outer
residuals and convergence test, FP32 inner iterations), each timed to the same
tolerance 1e-8 with its residual re-checked in FP64; the series then continues
from the FP64 solution. Warmup solves repeat the first step and are discarded.

value = GDOF.iter/s = (2 N_v + N_p) x DR iterations / seconds / 1e9 of the
mixed-precision solve over the timed steps, inputs resident in HBM;
time_to_tolerance_s and speedup_vs_fp64 give the time-to-1e-8 comparison.

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
    """Compulsory traffic per element-kernel launch and per DR iteration of the
    FP64 path (each array counted once; geometry as (D^2+1) words per element)."""
    D, NV, nq = 2, 7, 7
    elastic = any(math.isfinite(g) for g in w.G)
    p_old = any(math.isfinite(b) for b in w.etab)
    tau = 3 * nq * ne if elastic else 0
    elem = wf * (D * nv + npp + (npp if p_old else 0) + npp + tau + (D * D + 1) * ne + D * nv) + wi * (NV * ne + ne)
    it = wf * (2 * D * nv + 2 * D * nv + D * nv + npp + (npp if p_old else 0) + npp + tau + (D * D + 1) * ne) \
        + wi * (NV * ne + ne)
    return elem, it


def algorithmic_bytes_fp32(w, nv, ne, npp):
    """Same for the FP32 inner iteration of the mixed-precision solve on a
    structured mesh (connectivity and geometry from the element index) with
    the return-map correction operator: the element kernel reads the
    correction and the FP32 state copy (velocity, pressure), the scale
    gamma_P/M_P, the right-hand side R(x) and the stress history, and writes
    the residual; the iteration adds the rate and preconditioner passes."""
    D, nq = 2, 7
    elastic = any(math.isfinite(g) for g in w.G)
    plastic = w.plastic is not None
    tau = 3 * nq * ne if (elastic and plastic) else 0
    xcopy = (D * nv + npp) if plastic else 0
    elem = 4 * (D * nv + npp + xcopy + npp + npp + D * nv + tau + D * nv) + 4 * ne
    it = 4 * (2 * D * nv + 2 * D * nv + D * nv + D * nv + npp + xcopy + npp + npp + tau) + 4 * ne
    return elem, it


def algorithmic_bytes_fp32_tile(w, nv, ne, npp, nt6):
    """Compulsory traffic of one launch of the FP32 visco-elastic correction
    kernel in 16 x 8-cell tiles (structured mesh): it reads the correction's
    velocity (every node) and pressure, the scale gamma_P/M_P, the pressure
    right-hand side and the cell phase, writes the summed residual of every
    grid node, and updates the element-private bubble nodes (rate read and
    written, preconditioner and velocity right-hand side read, velocity
    written, Dirichlet flag). Elements of structured meshes carry no
    connectivity or geometry; the visco-elastic operator reads no stress
    history."""
    D = 2
    return 4 * (D * nv + 3 * npp + ne + D * nt6 + D * ne * 5) + ne


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


# FP64 flops the factorised Hex27 evaluation executes per element in the
# viscous DR-step kernel: 27 active lanes x 1234 (velocity gradient at the
# lane's point 480, continuity / stress / fluxes ~106, nodal residual 648)
HEX27_FLOPS_PER_ELEMENT = 27 * 1234
FP64_PEAK_TFLOPS = 78.6


def hex27_measure(n, iters):
    """The 3-D path on its own configuration: n^3 Hex27 cells, 50 seeded
    spheres, viscous, pure shear; iters DR iterations from the far-field
    guess (not to convergence) and the element kernel alone timed with HIP
    events."""
    import femtools_amd as fa
    from femtools_amd import problems as PB
    w = PB.hex27_inclusions(n)
    mesh = fa.hex27_box_mesh(n, n, n)
    ph = PB.hex27_sphere_phases(mesh.coords, mesh.el2n, w.spheres)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    bcs = [fa.DirichletBoundaryCondition(d, v) for d, v in PB.pure_shear_bc3(mesh.coords)]
    v0 = PB.pure_shear_field3(mesh.coords)
    fa.assemble_viscosity_weighted_pressure_scaling(dr, w.pressure_factor, w.dt)
    el_ms = fa.time_element_kernel(dr, w.dt, n=20)
    for f, v in zip(dr.v, v0):
        f.copyto(v)
    dr.P.fill(0.0)
    mesh.ctx.sync()
    t0 = time.perf_counter()
    st = fa.solve(dr, mesh, *bcs, dt=w.dt, max_iterations=iters, iterMax=iters, throw_on_failure=False,
                  **w.solve_kwargs())
    mesh.ctx.sync()
    wall = time.perf_counter() - t0
    flops = HEX27_FLOPS_PER_ELEMENT * mesh.nels
    tf = flops / (el_ms * 1e-3) / 1e12
    return {"workload": w.name + f", {iters} DR iterations from the far-field guess",
            "elements": mesh.nels, "velocity_nodes": mesh.nnodes, "pressure_dofs": mesh.nnodesP, "dofs": w.dofs(),
            "iterations": st["iter"], "seconds": round(wall, 4), "ms_per_iter": round(wall / st["iter"] * 1e3, 4),
            "gdof_iter_per_s": round(w.dofs() * st["iter"] / wall / 1e9, 4),
            "roofline": {"bound": "fp64 alu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tf / FP64_PEAK_TFLOPS, 4), "kernel": "k_hex27_element (viscous DR step)",
                         "flops_per_launch": flops, "avg_launch_us": round(el_ms * 1e3, 2)}}


def _free_slip3(coords):
    import numpy as np
    out = []
    for c in range(3):
        d = np.where((np.abs(coords[:, c]) < 1e-12) | (np.abs(coords[:, c] - 1) < 1e-12))[0].astype(np.int32)
        out.append((d, np.zeros(len(d))))
    return out


def _block_phases3(coords, el2n):
    import numpy as np
    cen = coords[el2n].mean(axis=1)
    return np.where(np.all(np.abs(cen - 0.5) < 0.26, axis=1), 2, 1).astype(np.int32)


# FP64 flops per element of one CellPressureStokesDR iteration (pressure pass:
# 27 points x ~500 for the gradient and the mode sums; momentum pass: 27
# lanes x ~1160 for the gradient, stress and nodal residual)
CP3_FLOPS_PER_ELEMENT = 27 * (500 + 1160)


def cell_pressure_measure(n, iters):
    """CellPressureStokesDR on n^3 Hex27 cells: a stiff, dense block (η 50,
    ρ 2) in a free-slip unit box under gravity; iters relaxation iterations
    (tolerance 0: the loop runs them all), checks every 100."""
    import femtools_amd as fa
    mesh = fa.hex27_box_mesh(n, n, n)
    ph = _block_phases3(mesh.coords, mesh.el2n)
    dr = fa.CellPressureStokesDR(mesh, fa.StokesMaterial(eta=(1.0, 50.0), rho0=(1.0, 2.0), g=(0.0, 0.0, -1.0)),
                                 phases=ph)
    bcs = [fa.DirichletBoundaryCondition(d, v) for d, v in _free_slip3(mesh.coords)]
    fa.solve(dr, mesh, *bcs, max_iterations=20, check_interval=20, tolerance=0.0, throw_on_failure=False)
    for f in dr.v:
        f.fill(0.0)
    dr.P.fill(0.0)
    mesh.ctx.sync()
    t0 = time.perf_counter()
    st = fa.solve(dr, mesh, *bcs, max_iterations=iters, check_interval=100, tolerance=0.0, throw_on_failure=False)
    mesh.ctx.sync()
    wall = time.perf_counter() - t0
    dofs = 3 * mesh.nnodes + 4 * mesh.nels
    tf = CP3_FLOPS_PER_ELEMENT * mesh.nels * st["iterations"] / wall / 1e12
    return {"workload": f"CellPressureStokesDR, {n}^3 Hex27 cells, sinking block (eta 50, rho 2), free slip, "
                        f"{iters} Arrow-Hurwicz/Jacobi iterations",
            "elements": mesh.nels, "dofs": dofs, "iterations": st["iterations"], "seconds": round(wall, 4),
            "ms_per_iter": round(wall / st["iterations"] * 1e3, 4),
            "gdof_iter_per_s": round(dofs * st["iterations"] / wall / 1e9, 4), "residual": st["residual"],
            "fp64_tflops_whole_iteration": round(tf, 2), "fp64_peak_tflops": FP64_PEAK_TFLOPS}


def _cap_setup(fa, n, cap):
    """2-D extension problem with a weak inclusion: the tensile cap (or plain
    Drucker-Prager) under pure-shear boundary velocities."""
    import math
    import numpy as np
    mesh = fa.t7_box_mesh(n, n)
    coords, el2n = mesh.coords, mesh.el2n
    x, y = coords[:, 0], coords[:, 1]
    gam = mesh.boundary_nodes
    nx = gam[(np.abs(x[gam]) < 1e-12) | (np.abs(x[gam] - 1) < 1e-12)]
    ny = gam[(np.abs(y[gam]) < 1e-12) | (np.abs(y[gam] - 1) < 1e-12)]
    cen = coords[el2n[:, :3]].mean(axis=1)
    ph = np.where(((cen - 0.5) ** 2).sum(axis=1) < 0.15 ** 2, 2, 1).astype(np.int32)
    mat = fa.StokesMaterial(eta=(1.0, 0.1), etab=(1.0, 1.0), G=(1.0, 1.0), K=(4.0, 4.0))
    dr = fa.StokesDR(mesh, mat)
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    dr.v.x.copyto(x - 0.5)
    dr.v.y.copyto(-(y - 0.5))
    bcs = (fa.DirichletBoundaryCondition(nx, x[nx] - 0.5), fa.DirichletBoundaryCondition(ny, -(y[ny] - 0.5)))
    ang = (math.radians(30),) * 2
    if cap:
        plastic = fa.DruckerPragerCap(ang, (math.radians(5),) * 2, (1.0, 0.5), (-0.3, -0.15), (0.1, 0.1), (4.0, 4.0))
    else:
        plastic = fa.DruckerPrager(ang, (math.radians(5),) * 2, (1.0, 0.5), (0.1, 0.1), (4.0, 4.0))
    return mesh, dr, bcs, plastic


def cap_measure(n, iters):
    """The tensile cap against plain Drucker-Prager on the same 2-D extension
    problem (n^2 T7 cells, FP64, Jacobian re-assembled at every check):
    iters DR iterations each from the same start."""
    import femtools_amd as fa
    out = {"workload": f"{n}x{n} T7 pure shear with a weak inclusion, VE + yield, {iters} FP64 DR iterations, "
                       "check interval 50"}
    for name, cap in (("cap", True), ("drucker_prager", False)):
        mesh, dr, bcs, plastic = _cap_setup(fa, n, cap)
        kw = dict(dt=0.5, plastic=plastic, pressure_factor=1.0, tolerance=1e-30, check_interval=50,
                  max_iterations=iters, iterMax=iters, throw_on_failure=False, finite_K=True)
        st = fa.solve(dr, mesh, *bcs, **kw)
        out[name] = {"iterations": st["iter"], "seconds": round(st["seconds"], 4),
                     "us_per_iter": round(st["seconds"] / max(1, st["iter"]) * 1e6, 2),
                     "jacobian_assemblies": st["jacobian_assemblies"]}
    out["cap_over_drucker_prager"] = round(out["cap"]["us_per_iter"] / out["drucker_prager"]["us_per_iter"], 3)
    return out


def _buoyancy_setup(fa, n, coupled):
    import math
    import numpy as np
    mesh = fa.t7_box_mesh(n, n)
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    gam = mesh.boundary_nodes
    zero = np.zeros(len(gam))
    mat = fa.StokesMaterial(eta=(1.0,), etab=(1.0,), G=(math.inf,), alpha=(0.05,), rho0=(1.0,), K=(math.inf,),
                            g=(0.0, -10.0))
    dr = fa.StokesDR(mesh, mat)
    th = None
    T0 = 0.2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    if coupled:
        th = fa.ThermalDiffusionDR(mesh, fa.ThermalMaterial(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,),
                                                             K=(math.inf,)), eps=1e-9)
        th.source.copyto(50.0 * np.exp(-((x - 0.4) ** 2 + (y - 0.5) ** 2) / 0.02))
        th.T0.copyto(T0)
        th.T.copyto(T0)
    bcv = tuple(fa.DirichletBoundaryCondition(gam, zero) for _ in range(2))
    return mesh, dr, th, fa.DirichletBoundaryCondition(gam, zero), bcv


def coupled_measure(n, iters):
    """solve_coupled! against the Stokes solve alone on the same 2-D buoyancy
    problem (n^2 T7 cells, FP64): iters inner iterations each (one thermal DR
    iteration and the temperature transfer per Stokes iteration)."""
    import femtools_amd as fa
    out = {"workload": f"{n}x{n} T7 hot blob under gravity (alpha 0.05), {iters} FP64 inner iterations, "
                       "check interval 50"}
    kw = dict(dt=0.01, pressure_factor=1.0, tolerance=1e-30, check_interval=50, max_iterations=iters, iterMax=iters,
              throw_on_failure=False)
    mesh, dr, th, bcT, bcv = _buoyancy_setup(fa, n, True)
    t0 = time.perf_counter()
    st = fa.solve_coupled(th, dr, mesh, mesh, bcT, *bcv, Tref=0.0, **kw)
    wall = time.perf_counter() - t0
    out["coupled"] = {"iterations": st["iter"], "thermal_iterations": st["thermal_iterations"],
                      "seconds": round(wall, 4), "us_per_iter": round(wall / max(1, st["iter"]) * 1e6, 2)}
    mesh, dr, _, _, bcv = _buoyancy_setup(fa, n, False)
    t0 = time.perf_counter()
    st = fa.solve(dr, mesh, *bcv, **kw)
    wall = time.perf_counter() - t0
    out["stokes_alone"] = {"iterations": st["iter"], "seconds": round(wall, 4),
                           "us_per_iter": round(wall / max(1, st["iter"]) * 1e6, 2)}
    out["coupled_over_stokes"] = round(out["coupled"]["us_per_iter"] / out["stokes_alone"]["us_per_iter"], 3)
    return out


def cpu_rows(budget_s=4.0):
    """The oracle on small samples of the rows measured beside the headline
    (cell-pressure 3-D, tensile cap, coupled step): per-iteration times."""
    import math
    import numpy as np
    import pyoracle as O
    rows = {}
    # CellPressureStokesDR, 12^3 cells
    n = 12
    coords, el2n = O.hex27_mesh(n, n, n)
    pr = O.Problem3(coords, el2n, O.material(eta=(1.0, 50.0), rho0=(1.0, 2.0)), g=(0.0, 0.0, -1.0),
                    phases=_block_phases3(coords, el2n), bc=tuple(_free_slip3(coords)))
    t0 = time.perf_counter()
    st = pr.cp3_solve(tolerance=0.0, max_iterations=20, check_interval=20)
    t = time.perf_counter() - t0
    rows["cell_pressure"] = {"sample": f"{n}^3 cells, {st['iterations']} iterations",
                             "us_per_element_iter": round(t / st["iterations"] / len(el2n) * 1e6, 4)}
    # tensile cap, 64^2
    n = 64
    coords, el2n = O.t7_mesh(n, n)
    x, y = coords[:, 0], coords[:, 1]
    gam = O.box_boundary_nodes(coords)
    nx = gam[(np.abs(x[gam]) < 1e-12) | (np.abs(x[gam] - 1) < 1e-12)]
    ny = gam[(np.abs(y[gam]) < 1e-12) | (np.abs(y[gam] - 1) < 1e-12)]
    cen = coords[el2n[:, :3]].mean(axis=1)
    ph = np.where(((cen - 0.5) ** 2).sum(axis=1) < 0.15 ** 2, 2, 1).astype(np.int32)
    cap = O.drucker_prager_cap((math.radians(30),) * 2, (math.radians(5),) * 2, (1.0, 0.5), (-0.3, -0.15),
                               (0.1, 0.1), (4.0, 4.0))
    ne = len(el2n)
    pr = O.Problem(coords, el2n, O.material(eta=(1.0, 0.1), etab=(1.0, 1.0), G=(1.0, 1.0), K=(4.0, 4.0)),
                   phases_v=ph, phases_P=ph, tau_old=tuple(np.zeros((ne, 7)) for _ in range(3)), cap=cap,
                   bc=((nx, x[nx] - 0.5), (ny, -(y[ny] - 0.5))))
    pr.vx[:], pr.vy[:] = x - 0.5, -(y - 0.5)
    k = 100
    t0 = time.perf_counter()
    st = pr.solve(0.5, pressure_factor=1.0, tolerance=1e-30, check_interval=50, max_iterations=k, iter_max=k,
                  finite_K=True)
    t = time.perf_counter() - t0
    rows["cap"] = {"sample": f"{n}x{n} cells, {st['iter']} iterations",
                   "us_per_element_iter": round(t / max(1, st["iter"]) / ne * 1e6, 4)}
    return rows


def batch_measure(count, nthreads, precision):
    """Independent small solves one at a time and then all at once through
    fem_stokes_solve_batch (one context and stream per problem), each from
    the same initial guess to residual 1e-8."""
    import femtools_amd as fa
    from femtools_amd import problems as PB
    probs = []
    for w in PB.batch_workloads(count):
        ctx = fa.Context(0)
        mesh = fa.t7_box_mesh(w.nx, w.ny, ctx=ctx)
        ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
        (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
        dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g))
        dr.set_phases("v", ph)
        dr.set_phases("P", ph)
        bcs = [fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)]
        kw = dict(dt=w.dt, tolerance=1e-8, check_interval=50, max_iterations=400_000, iterMax=400_000,
                  throw_on_failure=False, precision=precision)
        probs.append((w, dr, mesh, bcs, kw, PB.pure_shear_field(mesh.coords)))

    def reset():
        for w, dr, mesh, bcs, kw, v0 in probs:
            dr.v.x.copyto(v0[0])
            dr.v.y.copyto(v0[1])
            dr.P.fill(0.0)
            mesh.ctx.sync()

    def one(dr, mesh, bcs, kw):
        try:
            return fa.solve(dr, mesh, *bcs, **kw)
        except fa.FemError as e:   # a diverging problem counts as a finished, unconverged solve
            return {"iter": -1, "converged": False, "error": str(e)}

    reset()
    t0 = time.perf_counter()
    seq = [one(dr, mesh, bcs, kw) for w, dr, mesh, bcs, kw, v0 in probs]
    t_seq = time.perf_counter() - t0
    reset()
    t0 = time.perf_counter()
    bat = fa.solve_batch([(dr, mesh, bcs, kw) for w, dr, mesh, bcs, kw, v0 in probs], nthreads=nthreads,
                         raise_errors=False)
    t_bat = time.perf_counter() - t0
    dof_it = sum(w.dofs() * max(st["iter"], 0) for (w, *_r), st in zip(probs, bat))
    return {"workload": f"{count} independent T7/P1-disc viscous problems, n in 32..128, 1-10 inclusions, "
                        f"log10 eta in U[-2,4], tol 1e-8, {precision}",
            "sizes": [w.nx for w, *_r in probs], "nthreads": nthreads,
            "converged": sum(bool(st["converged"]) for st in bat),
            "failed": sum(st["status"] != 0 for st in bat),
            "same_iterations_as_sequential": all(a["iter"] == b["iter"] for a, b in zip(seq, bat)
                                                 if b["status"] == 0 and "error" not in a),
            "sequential_s": round(t_seq, 3), "batched_s": round(t_bat, 3),
            "solves_per_s": {"sequential": round(count / t_seq, 3), "batched": round(count / t_bat, 3)},
            "batched_gdof_iter_per_s": round(dof_it / t_bat / 1e9, 3),
            "speedup": round(t_seq / t_bat, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed physical time steps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves of the first step")
    ap.add_argument("--n", type=int, default=512, help="cells per side")
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--max-iterations", type=int, default=400_000)
    ap.add_argument("--precision", choices=("both", "fp64", "fp32"), default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--hex27-n", type=int, default=64, help="cells per side of the 3-D measurement (0: skip)")
    ap.add_argument("--hex27-iters", type=int, default=2000)
    ap.add_argument("--rows-n", type=int, default=256, help="cells per side of the 2-D cap / coupled rows (0: skip)")
    ap.add_argument("--rows-iters", type=int, default=2000)
    ap.add_argument("--cell-n", type=int, default=64, help="cells per side of the CellPressureStokesDR row (0: skip)")
    ap.add_argument("--cell-iters", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=0, help="independent small solves to run batched (0: skip)")
    ap.add_argument("--batch-threads", type=int, default=8)
    ap.add_argument("--batch-precision", choices=("fp64", "fp32"), default="fp32")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a single solve would shard by element "
              "strips with an interface-node exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import femtools_amd as fa
    from femtools_amd import problems as PB

    w = PB.vep_inclusions(args.n, nsteps=args.steps)
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
    fields = (dr.v.x, dr.v.y, dr.P, dr.tau_old.xx, dr.tau_old.yy, dr.tau_old.xy)

    def snapshot():
        return [f.numpy() for f in fields]

    def restore(snap):
        for f, a in zip(fields, snap):
            f.copyto(a)

    def solve(prec, timed):
        dr.P0.copyto(dr.P.numpy())
        kw = w.solve_kwargs()
        kw["tolerance"] = args.tolerance
        ctx.sync()
        t0 = time.perf_counter()
        st = fa.solve(dr, mesh, bcx, bcy, dt=w.dt, plastic=dp, max_iterations=args.max_iterations,
                      iterMax=args.max_iterations, throw_on_failure=False, time_kernels=timed, precision=prec, **kw)
        ctx.sync()
        st["wall"] = time.perf_counter() - t0
        return st

    precs = ("fp64", "fp32") if args.precision == "both" else (args.precision,)
    s0 = snapshot()
    for i in range(args.warmup):
        for prec in precs:
            restore(s0)
            st = solve(prec, False)
            print(f"# warmup {prec}: iter {st['iter']} converged {st['converged']} {st['wall']:.2f}s",
                  file=sys.stderr, flush=True)
    restore(s0)
    stats = {p: [] for p in precs}
    diffs = []
    for i in range(args.steps):
        start = snapshot()
        sols = {}
        for prec in precs:
            restore(start)
            st = solve(prec, True)
            stats[prec].append(st)
            sols[prec] = [f.numpy() for f in fields[:3]]
            print(f"# step {i} {prec}: iter {st['iter']} itPH {st['itPH']} converged {st['converged']} "
                  f"err_abs {st['err_abs']:.2e} {st['wall']:.2f}s", file=sys.stderr, flush=True)
        if len(precs) == 2:
            diffs.append({k: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
                          for k, a, b in zip(("vx", "vy", "P"), sols["fp32"], sols["fp64"])})
        # the series continues from the FP64 solution (the reference's trajectory)
        for f, a in zip(fields[:3], sols[precs[0]]):
            f.copyto(a)
        fa.update_stokes_current_stress(dr, mesh, w.dt, plastic=dp)
        for c in ("xx", "yy", "xy"):
            getattr(dr.tau_old, c).copyto(getattr(dr.tau, c).numpy())

    head = "fp32" if "fp32" in precs else "fp64"
    hs = stats[head]
    elapsed = sum(s["wall"] for s in hs)
    iters = sum(s["iter"] for s in hs)
    value = w.dofs() * iters / elapsed / 1e9
    ms_el = sum(s["kernel_ms_element"] for s in hs)
    ms_ga = sum(s["kernel_ms_gather"] for s in hs)
    nl = max(1, sum(s["timed_launches"] for s in hs))
    if head == "fp32":
        # the dominant kernel of step 0: the visco-elastic correction operator
        # (step 0 has no yielding interval; yielding steps run the return-map
        # variant, reported by its time only since its traffic depends on how
        # many elements are near yield)
        nt6 = (2 * w.nx + 1) * (2 * w.ny + 1)
        elem_bytes = algorithmic_bytes_fp32_tile(w, nv, ne, npp, nt6)
        _, iter_bytes = algorithmic_bytes_fp32(w, nv, ne, npp)
        kname = "k_t7_element<float,364,256> (FP32 visco-elastic correction operator, 16x8-cell tiles), step 0"
        s0 = hs[0]
        n0 = max(1, s0["timed_launches"])
        avg_el_s = s0["kernel_ms_element"] / n0 * 1e-3
        avg_it_s = (s0["kernel_ms_element"] + s0["kernel_ms_gather"]) / n0 * 1e-3
        traffic = read_traffic("k_t7_element<float,364")
    else:
        elem_bytes, iter_bytes = algorithmic_bytes(w, nv, ne, npp)
        kname = "k_t7_element<double> (fused element residual + bubble update)"
        avg_el_s = ms_el / nl * 1e-3
        avg_it_s = (ms_el + ms_ga) / nl * 1e-3
        traffic = read_traffic("k_t7_element<double>")
    achieved = elem_bytes / avg_el_s / 1e9 if avg_el_s > 0 else None
    out = {
        "metric": "time-to-1e-8 residual (FP32+learned vs plain FP64, same GPU); GDOF.iter/s",
        "value": round(value, 4),
        "unit": "GDOF.iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner iterations, f64 state/residuals/norms" if head == "fp32" else "f64",
        "data": "synthetic (seeded random inclusions, generated in-process)",
        "config": {"workload": w.name + f", dt={w.dt}, DP C=1 phi=30deg psi=0 eta_reg=8e-3, K=etab=Kb=100, "
                                        f"pure shear, tol {args.tolerance:g}",
                   "cells": [w.nx, w.ny], "elements": ne, "velocity_nodes": nv, "pressure_dofs": npp,
                   "dofs": w.dofs(), "time_steps_timed": args.steps, "parallelism": "single GPU",
                   "warm_start": "previous time step (no learned warm start in this build)"},
        "time_to_tolerance_s": {p: [round(s["wall"], 4) for s in stats[p]] for p in precs},
        "iterations": {p: [s["iter"] for s in stats[p]] for p in precs},
        "converged": {p: [bool(s["converged"]) for s in stats[p]] for p in precs},
        "err_abs": {p: [s["err_abs"] for s in stats[p]] for p in precs},
        "gdof_iter_per_s": {p: round(w.dofs() * sum(s["iter"] for s in stats[p]) /
                                     sum(s["wall"] for s in stats[p]) / 1e9, 4) for p in precs},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic, "kernel": kname,
                     "algorithmic_bytes_per_launch": elem_bytes,
                     "avg_launch_us": round(avg_el_s * 1e6, 2)},
        "iteration_roofline": {"algorithmic_bytes": iter_bytes, "avg_us": round(avg_it_s * 1e6, 2),
                               "achieved_GBs": round(iter_bytes / avg_it_s / 1e9, 1) if avg_it_s > 0 else None,
                               "avg_gather_us": round(ms_ga / nl * 1e3, 2)},
    }
    if head == "fp32" and len(hs) > 1:
        rest = hs[1:]
        nr = max(1, sum(s["timed_launches"] for s in rest))
        out["return_map_steps"] = {"avg_element_us": round(sum(s["kernel_ms_element"] for s in rest) / nr * 1e3, 2),
                                   "avg_update_us": round(sum(s["kernel_ms_gather"] for s in rest) / nr * 1e3, 2),
                                   "us_per_iteration_wall": round(sum(s["wall"] for s in rest) /
                                                                  max(1, sum(s["iter"] for s in rest)) * 1e6, 2)}
    if len(precs) == 2:
        t64 = sum(s["wall"] for s in stats["fp64"])
        out["speedup_vs_fp64"] = round(t64 / elapsed, 3)
        out["fp32_vs_fp64_max_rel_diff"] = diffs
    if args.hex27_n > 0:
        out["hex27"] = hex27_measure(args.hex27_n, args.hex27_iters)
    if args.cell_n > 0:
        out["cell_pressure"] = cell_pressure_measure(args.cell_n, args.cell_iters)
    if args.rows_n > 0:
        out["cap"] = cap_measure(args.rows_n, args.rows_iters)
        out["coupled"] = coupled_measure(args.rows_n, args.rows_iters)
    if args.batch > 0:
        out["batched"] = batch_measure(args.batch, args.batch_threads, args.batch_precision)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
            out["cpu_baseline"]["rows"] = cpu_rows()
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
