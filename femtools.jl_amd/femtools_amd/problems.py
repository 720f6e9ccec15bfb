"""Synthetic inputs of the DR Stokes workloads, as plain numpy arrays so the
same problem can be handed to the GPU library and to a CPU checker.

* viscous inclusion: 2-D incompressible Stokes on the unit box, T7/P1-disc,
  matrix η=1 with a centred circular inclusion (r=0.1, η=1e3), pure-shear
  velocity boundary conditions (ε̇ = 1), no gravity, G=K=ηb=∞ -- the setting
  of the pure-shear triangle example
  (examples/miniapps/stokes/stokes_2D_pure_shear_triangle/...jl:40-154).
* visco-elasto-plastic inclusions: seeded random circular inclusions
  (r ∈ U[0.01, 0.05], log10 η ∈ U[-2, 4], later inclusion wins), shear
  modulus G=1 and Δt = 0.1 η_matrix/G, Drucker-Prager C=1, ϕ=30°, Ψ=0,
  densities 1 (matrix) / 1.1 (inclusions), g=(0,-1), pure shear. The bulk
  modulus, pressure storage modulus and Kb are 100 and the solver keywords
  are those of the reference's VEP examples (γfact=20, rel_drop0=0.75,
  check_interval=100, CFL_v=0.99; examples/miniapps/stokes/vevp/
  stokes_2D_shear_bands_triangle.jl:48-79, 199-214); the initial velocity is
  the background pure-shear field, as there (:189-195).

Phases are per element (1 x nels layout), taken at the element centroid.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def element_centroids(coords, el2n):
    return coords[el2n[:, :3]].mean(axis=1)


def inclusion_phases(coords, el2n, circles):
    """1-based cell phases: 1 = matrix, k+1 = k-th circle (later circles win)."""
    c = element_centroids(coords, el2n)
    ph = np.ones(len(el2n), np.int32)
    for k, (cx, cy, r) in enumerate(circles):
        inside = (c[:, 0] - cx) ** 2 + (c[:, 1] - cy) ** 2 <= r * r
        ph[inside] = k + 2
    return ph


def box_boundary_nodes(coords, box):
    x0, x1, y0, y1 = box
    tol = 32 * np.finfo(float).eps * max(1.0, *(abs(v) for v in box))
    x, y = coords[:, 0], coords[:, 1]
    return np.nonzero((np.abs(x - x0) <= tol) | (np.abs(x - x1) <= tol) |
                      (np.abs(y - y0) <= tol) | (np.abs(y - y1) <= tol))[0].astype(np.int32)


def pure_shear_bc(coords, box=(0.0, 1.0, 0.0, 1.0), edot=1.0):
    """vx = ε̇(x - xc) on the x-walls, vy = -ε̇(y - yc) on the y-walls
    (stokes_2D_pure_shear_triangle.jl:145-154). Returns ((dofs_x, vals_x), (dofs_y, vals_y))."""
    x0, x1, y0, y1 = box
    gam = box_boundary_nodes(coords, box)
    tol = max(abs(x1 - x0), abs(y1 - y0)) * np.finfo(float).eps * 32
    x, y = coords[gam, 0], coords[gam, 1]
    nx = gam[(np.abs(x - x0) <= tol) | (np.abs(x - x1) <= tol)]
    ny = gam[(np.abs(y - y0) <= tol) | (np.abs(y - y1) <= tol)]
    xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    return (nx, edot * (coords[nx, 0] - xm)), (ny, -edot * (coords[ny, 1] - ym))


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    eta: tuple
    etab: tuple
    G: tuple
    rho0: tuple
    K: tuple
    g: tuple
    dt: float
    circles: list
    plastic: dict | None = None
    tolerance: float = 1e-8
    check_interval: int = 50
    pressure_factor: float = 50.0
    rel_drop0: float = 1e-2
    CFL_v: float = 0.98
    nsteps: int = 1
    seed_pure_shear: bool = False
    # outer (pressure) iterations: warm-started and high-contrast solves take
    # more than the solver's default cap of 1000 to reach 1e-8
    max_ph_iterations: int = 20_000
    extra: dict = field(default_factory=dict)

    def solve_kwargs(self):
        return dict(tolerance=self.tolerance, check_interval=self.check_interval,
                    pressure_factor=self.pressure_factor, rel_drop0=self.rel_drop0,
                    max_ph_iterations=self.max_ph_iterations)

    @property
    def nphases(self):
        return len(self.eta)

    def sizes(self):
        nn = (2 * self.nx + 1) * (2 * self.ny + 1) + 2 * self.nx * self.ny
        ne = 2 * self.nx * self.ny
        return nn, ne, 3 * ne

    def dofs(self):
        nn, ne, npp = self.sizes()
        return 2 * nn + npp


def viscous_inclusion(n=64) -> Workload:
    """Plain FP64 viscous inclusion workload on an n x n cell grid."""
    inf = math.inf
    return Workload(name=f"T7/P1-disc {n}x{n} viscous inclusion", nx=n, ny=n, eta=(1.0, 1e3), etab=(inf, inf),
                    G=(inf, inf), rho0=(1.0, 1.0), K=(inf, inf), g=(0.0, 0.0), dt=1.0,
                    circles=[(0.5, 0.5, 0.1)])


def pure_shear_field(coords, box=(0.0, 1.0, 0.0, 1.0), edot=1.0):
    """Background pure-shear velocity (vx, vy) = ε̇ (x - xc, -(y - yc))."""
    xm, ym = 0.5 * (box[0] + box[1]), 0.5 * (box[2] + box[3])
    return edot * (coords[:, 0] - xm), -edot * (coords[:, 1] - ym)


# Perzyna viscosity of the VEP series. With 8e-3 the reference's algorithm (the
# oracle, 64^2) blows up at the first check intervals of the fifth step, once
# yielding spreads: the absolute residual grows from 5e-6 to 2e3 within 100
# iterations (a genuine divergence of the DR iteration, not a relative-error
# artefact); 1e-2 fails the same way. The reference warns that values well
# below its default 1e-2 diverge the DR iteration on fine meshes
# (stokes_2D_shear_bands_triangle.jl:42-44). At 512^2 the FP64 series with
# 5e-2 stalls in the third step (the pressure residual err_P stays at 2.9e-8
# after 400k iterations while err_v is 1e-11), with 0.1 from the fourth on;
# 0.2 carries nine of the ten steps to 1e-8 (the ninth ends at 1.2e-8).
VEP_ETA_REG = 0.2


def vep_inclusions(n=512, ninc=20, seed=0, nsteps=10, eta_reg=VEP_ETA_REG, Kb=100.0, K=100.0, C=1.0) -> Workload:
    """Visco-elasto-plastic multi-inclusion workload on an n x n cell grid."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, 1, ninc)
    cy = rng.uniform(0, 1, ninc)
    r = rng.uniform(0.01, 0.05, ninc)
    leta = rng.uniform(-2, 4, ninc)
    inf = math.inf
    nph = ninc + 1
    eta = (1.0,) + tuple(10.0 ** leta)
    G = 1.0
    return Workload(name=f"T7/P1-disc {n}x{n} VEP {ninc} inclusions", nx=n, ny=n, eta=eta, etab=(K,) * nph,
                    G=(G,) * nph, rho0=(1.0,) + (1.1,) * ninc, K=(K,) * nph, g=(0.0, -1.0), dt=0.1 * 1.0 / G,
                    circles=list(zip(cx, cy, r)),
                    plastic=dict(phi=(math.radians(30),) * nph, psi=(0.0,) * nph, C=(C,) * nph,
                                 eta_reg=(eta_reg,) * nph, Kb=(Kb,) * nph),
                    pressure_factor=20.0, rel_drop0=0.75, check_interval=100, CFL_v=0.99,
                    nsteps=nsteps, seed_pure_shear=True, extra=dict(seed=seed))


# ---------------------------------------------------------------------------
# 3-D: Hex27 velocity / P1-disc (four modes) pressure on the unit cube
# ---------------------------------------------------------------------------
@dataclass
class Workload3:
    name: str
    n: int
    eta: tuple
    etab: tuple
    G: tuple
    rho0: tuple
    K: tuple
    g: tuple
    dt: float
    spheres: list
    plastic: dict | None = None
    tolerance: float = 1e-7
    check_interval: int = 100
    pressure_factor: float = 50.0
    rel_drop0: float = 1e-2
    # outer (pressure) iterations: warm-started and high-contrast solves take
    # more than the solver's default cap of 1000 to reach 1e-8
    max_ph_iterations: int = 20_000
    extra: dict = field(default_factory=dict)

    def solve_kwargs(self):
        return dict(tolerance=self.tolerance, check_interval=self.check_interval,
                    pressure_factor=self.pressure_factor, rel_drop0=self.rel_drop0,
                    max_ph_iterations=self.max_ph_iterations)

    def sizes(self):
        s = 2 * self.n + 1
        ne = self.n ** 3
        return s ** 3, ne, 4 * ne

    def dofs(self):
        nn, ne, npp = self.sizes()
        return 3 * nn + npp


def hex27_sphere_phases(coords, el2n, spheres):
    """1-based cell phases from the cell centre (local node 27): 1 = matrix,
    k+1 = k-th sphere (later spheres win)."""
    c = coords[el2n[:, 26]]
    ph = np.ones(len(el2n), np.int32)
    for k, (cx, cy, cz, r) in enumerate(spheres):
        inside = (c[:, 0] - cx) ** 2 + (c[:, 1] - cy) ** 2 + (c[:, 2] - cz) ** 2 <= r * r
        ph[inside] = k + 2
    return ph


def pure_shear_bc3(coords, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), edot=1.0):
    """Every wall node carries the far-field pure shear (ε̇(x - xc), 0,
    -ε̇(z - zc)) in all three components
    (stokes_3D_viscous_inclusion.jl:237-243). Returns three (dofs, vals)."""
    on = np.zeros(len(coords), bool)
    tol = 32 * np.finfo(float).eps * max(1.0, *(abs(v) for v in box))
    for d in range(3):
        on |= np.abs(coords[:, d] - box[2 * d]) <= tol
        on |= np.abs(coords[:, d] - box[2 * d + 1]) <= tol
    w = np.nonzero(on)[0].astype(np.int32)
    v = pure_shear_field3(coords[w], box, edot)
    return tuple((w, v[c]) for c in range(3))


def pure_shear_field3(coords, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), edot=1.0):
    xm, zm = 0.5 * (box[0] + box[1]), 0.5 * (box[4] + box[5])
    return edot * (coords[:, 0] - xm), np.zeros(len(coords)), -edot * (coords[:, 2] - zm)


def hex27_inclusions(n=64, nsph=50, seed=0, plastic=False, eta_reg=8e-3, Kb=100.0, C=1.0,
                     elastic=True) -> Workload3:
    """Viscous (or Drucker-Prager) multi-sphere workload on an n^3 Hex27 grid:
    spheres r ∈ U[0.02, 0.08], log10 η ∈ U[-1, 3], matrix η = 1, pure shear,
    no gravity, incompressible (ηb = K = ∞); with plasticity C = 1, ϕ = 30°,
    Ψ = 0, finite Kb, and G = 1, Δt = 0.1 unless elastic is False (viscoplastic,
    G = ∞)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (nsph, 3))
    r = rng.uniform(0.02, 0.08, nsph)
    leta = rng.uniform(-1, 3, nsph)
    inf = math.inf
    nph = nsph + 1
    eta = (1.0,) + tuple(10.0 ** leta)
    ve = plastic and elastic
    G = (1.0,) * nph if ve else (inf,) * nph
    kind = ("VEP" if ve else "VP") if plastic else "viscous"
    return Workload3(name=f"Hex27/P1-disc {n}^3 {kind} {nsph} spheres", n=n, eta=eta,
                     etab=(inf,) * nph, G=G, rho0=(0.0,) * nph, K=(inf,) * nph, g=(0.0, 0.0, 0.0),
                     dt=0.1 if ve else 1.0, spheres=[(*c[k], r[k]) for k in range(nsph)],
                     plastic=dict(phi=(math.radians(30),) * nph, psi=(0.0,) * nph, C=(C,) * nph,
                                  eta_reg=(eta_reg,) * nph, Kb=(Kb,) * nph) if plastic else None,
                     extra=dict(seed=seed))


def batch_workloads(count=32, seed=0, sizes=(32, 48, 64, 80, 96, 112, 128), log_eta=(-2.0, 4.0)):
    """Many small independent viscous problems: n x n T7 cells with n drawn
    from sizes, 1-10 circular inclusions (r ∈ U[0.02, 0.1], log10 η ∈
    U[log_eta], later inclusion wins), pure shear, incompressible."""
    rng = np.random.default_rng(seed)
    out = []
    inf = math.inf
    for k in range(count):
        n = int(rng.choice(sizes))
        ninc = int(rng.integers(1, 11))
        circles = list(zip(rng.uniform(0.1, 0.9, ninc), rng.uniform(0.1, 0.9, ninc), rng.uniform(0.02, 0.1, ninc)))
        eta = (1.0,) + tuple(10.0 ** rng.uniform(log_eta[0], log_eta[1], ninc))
        nph = ninc + 1
        out.append(Workload(name=f"T7/P1-disc {n}x{n} viscous {ninc} inclusions", nx=n, ny=n, eta=eta,
                            etab=(inf,) * nph, G=(inf,) * nph, rho0=(1.0,) * nph, K=(inf,) * nph, g=(0.0, 0.0),
                            dt=1.0, circles=circles, extra=dict(seed=seed, index=k)))
    return out



def q9_thermal_step(coords, dt=1e-2):
    """One implicit heat-diffusion step on a Q9 mesh of the unit square with
    node coordinates coords (the Q9 thermal step of the reference's
    performance notes is 48 x 48 cells): k = ρ0 = Cp = 1, no expansion, a
    Gaussian temperature anomaly T0 = exp(-r²/0.01) about the centre with a
    heat source of the same shape (amplitude 10), T = 0 held on the whole
    boundary, the start T = T0, and the solver defaults (tolerance 1e-6,
    10 000 iterations, checks every 100)."""
    x, y = coords[:, 0], coords[:, 1]
    g = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    on = (x == 0.0) | (x == 1.0) | (y == 0.0) | (y == 1.0)
    bc = np.nonzero(on)[0].astype(np.int32)
    return dict(material=dict(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,)),
                T=g.copy(), T0=g, source=10.0 * g, bc_dofs=bc, bc_vals=np.zeros(len(bc)),
                solve=dict(dt=dt, tolerance=1e-6, max_iterations=10_000, check_interval=100, Tref=273.0))
