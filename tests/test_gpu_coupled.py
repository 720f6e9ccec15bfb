"""GPU: the coupled thermal--Stokes step (solve_coupled!, src/stokes/solvers/
DR.jl:44-165, 514-661) and the shear dissipation (assemble_shear_heating!,
src/stokes/assemblers/shear_heating.jl) through the C ABI, against the CPU
oracle (identical iteration counts, fields to 1e-10, Φ to 1e-12) and the
reference's own test cases (test/test_coupled_solver.jl)."""
import math
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

fa = pytest.importorskip("femtools_amd")

from test_gpu_parity import make_pair  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


UNIT_T = dict(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(0.0,), K=(math.inf,))


class Coupled:
    """The same coupled problem on the device and in the oracle."""

    def __init__(self, n, stokes_mat, thermal_mat, bc_v, bc_T, eps_T, eps_v=1e-6, structured=True, init=None):
        coords, el2n = O.t7_mesh(n, n)
        self.coords, self.el2n = coords, el2n
        self.mesh = fa.t7_box_mesh(n, n) if structured else fa.t7_mesh_from_arrays(coords, el2n)
        assert np.array_equal(self.mesh.el2n, el2n)
        self.th = fa.ThermalDiffusionDR(self.mesh, fa.ThermalMaterial(**thermal_mat), eps=eps_T)
        self.dr = fa.StokesDR(self.mesh, fa.StokesMaterial(**stokes_mat), eps=eps_v)
        self.th_o = O.ThermalProblem(coords, el2n, O.thermal_material(**thermal_mat), bc=bc_T)
        self.pr = O.Problem(coords, el2n, O.material(**stokes_mat), bc=bc_v)
        self.bc_T = fa.DirichletBoundaryCondition(*bc_T)
        self.bc_v = tuple(fa.DirichletBoundaryCondition(*b) for b in bc_v)
        self.eps_T = eps_T
        for name, v in (init or {}).items():
            if name in ("vx", "vy"):
                getattr(self.pr, name)[:] = v
                getattr(self.dr.v, name[1]).copyto(v)
            else:
                getattr(self.th_o, name)[:] = v
                getattr(self.th, name).copyto(v)

    def solve_gpu(self, dt, **kw):
        return fa.solve_coupled(self.th, self.dr, self.mesh, self.mesh, self.bc_T, *self.bc_v, dt=dt, **kw)

    def solve_oracle(self, dt, shear_heating=False, **kw):
        kw = dict(kw)
        if "iterMax" in kw:
            kw["iter_max"] = kw.pop("iterMax")
        return O.coupled_solve(self.th_o, self.pr, dt, tolerance_T=self.eps_T, shear_heating=shear_heating, **kw)


def test_gpu_coupled_reference_case():
    # test/test_coupled_solver.jl:8-56
    say("[gpu] coupled reference case (1 x 1 cell)")
    coords, el2n = O.t7_mesh(1, 1)
    gam = O.box_boundary_nodes(coords)
    zero = np.zeros(len(gam))
    cp = Coupled(1, dict(eta=(1.0,), etab=(1.0,), G=(math.inf,), alpha=(0.0,)), UNIT_T, ((gam, zero), (gam, zero)),
                 (gam, zero), 1e-8, eps_v=1e-8, init={"source": np.ones(len(coords))})
    kw = dict(pressure_factor=1.0, check_interval=1, iterMax=2000, max_iterations=2000, max_ph_iterations=5,
              tolerance=1e-8)
    st = cp.solve_gpu(1.0, **kw)
    say(f"  gpu: iter {st['iter']} itPH {st['itPH']} err_T {st['err_T']:.3e} thermal_iterations "
        f"{st['thermal_iterations']}")
    assert st["converged"]
    assert st["err_T"] < cp.th.eps
    assert st["thermal_iterations"] == st["iter"]
    T = cp.th.T.numpy()
    # the standalone thermal solve of the same problem
    ref = fa.ThermalDiffusionDR(cp.mesh, fa.ThermalMaterial(**UNIT_T), eps=1e-8)
    ref.source.fill(1.0)
    fa.solve(ref, cp.mesh, cp.bc_T, dt=1.0, check_interval=1, max_iterations=2000)
    assert np.allclose(T, ref.T.numpy(), rtol=1e-12, atol=0)
    assert np.array_equal(cp.dr.T.numpy()[cp.mesh.DoFsP], T[cp.mesh.el2nP])
    so = cp.solve_oracle(1.0, **kw)
    assert (so["iter"], so["itPH"], so["thermal_iterations"]) == (st["iter"], st["itPH"], st["thermal_iterations"])
    assert rel(T, cp.th_o.T) < 1e-10


def _buoyancy(n, structured, alpha_T=0.0):
    coords, _ = O.t7_mesh(n, n)
    gam = O.box_boundary_nodes(coords)
    zero = np.zeros(len(gam))
    x, y = coords[:, 0], coords[:, 1]
    T0 = 0.2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    init = {"source": 50.0 * np.exp(-((x - 0.4) ** 2 + (y - 0.5) ** 2) / 0.02), "T0": T0, "T": T0}
    tmat = dict(UNIT_T, alpha=(alpha_T,))
    smat = dict(eta=(1.0,), etab=(1.0,), G=(math.inf,), alpha=(0.05,), rho0=(1.0,), K=(math.inf,), g=(0.0, -10.0))
    return Coupled(n, smat, tmat, ((gam, zero), (gam, zero)), (gam, zero), 1e-9, structured=structured, init=init)


@pytest.mark.parametrize("n,structured,ncheck", [(8, True, 10), (6, False, 7), (16, True, 25)])
def test_gpu_coupled_buoyancy_parity(n, structured, ncheck):
    # a hot blob under gravity: the temperature drives the density and the
    # thermal-expansion term of the continuity residual
    say(f"[gpu] coupled buoyancy parity n={n} structured={structured} ncheck={ncheck}")
    cp = _buoyancy(n, structured, alpha_T=1e-3)
    kw = dict(pressure_factor=1.0, check_interval=ncheck, tolerance=1e-9, iterMax=20000, max_iterations=40000,
              Tref=0.0)
    st = cp.solve_gpu(0.01, **kw)
    so = cp.solve_oracle(0.01, **{k: v for k, v in kw.items() if k != "Tref"}, Tref=0.0)
    say(f"  gpu: iter {st['iter']} itPH {st['itPH']} err {st['err']:.3e} err_T {st['err_T']:.3e} | oracle: iter "
        f"{so['iter']} itPH {so['itPH']} err_T {so['err_T']:.3e}")
    assert st["converged"] and so["converged"]
    assert (st["iter"], st["itPH"], st["thermal_iterations"]) == (so["iter"], so["itPH"], so["thermal_iterations"])
    errs = dict(vx=rel(cp.dr.v.x.numpy(), cp.pr.vx), vy=rel(cp.dr.v.y.numpy(), cp.pr.vy),
                P=rel(cp.dr.P.numpy(), cp.pr.P), T=rel(cp.th.T.numpy(), cp.th_o.T))
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-10
    T = cp.th.T.numpy()
    assert np.array_equal(cp.dr.T.numpy()[cp.mesh.DoFsP], T[cp.mesh.el2nP])
    T0 = cp.th.T0.numpy()
    assert np.array_equal(cp.dr.T0.numpy()[cp.mesh.DoFsP], T0[cp.mesh.el2nP])


@pytest.mark.parametrize("case", ["viscous", "elastic", "plastic"])
def test_gpu_shear_heating_parity(case):
    say(f"[gpu] shear heating parity case={case}")
    rng = np.random.default_rng(4)
    nx, ny = 6, 5
    ne = 2 * nx * ny
    eta = (1.0, 7.0, 0.3)
    phases = rng.integers(1, 4, ne).astype(np.int32)
    G = None if case == "viscous" else (0.7, 3.0, 1.5)
    tau_old = tuple(0.2 * rng.normal(size=(ne, 7)) for _ in range(3)) if G else None
    dp = dict(phi=(math.radians(30),) * 3, psi=(math.radians(5),) * 3, C_=(0.3, 0.5, 0.2), eta_reg=(0.01,) * 3,
              Kb=(2.0,) * 3) if case == "plastic" else None
    pr, dr, mesh, dp_o, dp_g = make_pair(nx, ny, eta, G=G, dp=dp, phases=phases, tau_old=tau_old)
    dt = 0.4
    got = fa.assemble_shear_heating(dr, mesh, dt, plastic=dp_g)
    ref = pr.shear_heating(dt)
    e = rel(got, ref)
    say(f"  Φ max rel err {e:.2e} (|Φ| max {np.abs(ref).max():.3e})")
    assert got.shape == (ne, 7)
    assert e < 1e-12


def _simple_shear(shear):
    # test/test_coupled_solver.jl:59-106
    dt, eta, G, gd = 0.5, 2.0, 1.0, 1.0
    eve = 1 / (1 / eta + 1 / (G * dt))
    Texp = dt * (eve * gd) ** 2 / eta
    coords, _ = O.t7_mesh(2, 2)
    gam = O.box_boundary_nodes(coords)
    bc_v = ((gam, gd * coords[gam, 1]), (gam, np.zeros(len(gam))))
    bc_T = (gam, np.full(len(gam), Texp if shear else 0.0))
    vx = gd * coords[:, 1] + 0.1 * np.sin(np.pi * coords[:, 0]) * np.sin(np.pi * coords[:, 1])
    cp = Coupled(2, dict(eta=(eta,), etab=(1.0,), G=(G,), alpha=(0.0,)), UNIT_T, bc_v, bc_T, 1e-10, eps_v=1e-10,
                 init={"vx": vx})
    kw = dict(pressure_factor=1.0, check_interval=10, iterMax=20_000, max_iterations=20_000, tolerance=1e-10)
    return cp, kw, dt, Texp


@pytest.mark.parametrize("shear", [True, False])
def test_gpu_coupled_shear_heating_closed_form(shear):
    say(f"[gpu] coupled shear heating shear={shear}")
    cp, kw, dt, Texp = _simple_shear(shear)
    st = cp.solve_gpu(dt, shear_heating=shear, **kw)
    so = cp.solve_oracle(dt, shear_heating=shear, **kw)
    T = cp.th.T.numpy()
    say(f"  gpu: iter {st['iter']} err_T {st['err_T']:.3e}; max |T - T_expected| "
        f"{np.abs(T - (Texp if shear else 0.0)).max():.2e}")
    assert st["converged"]
    if shear:
        assert np.all(np.abs(T - Texp) < 1e-8)
    else:
        assert np.all(np.abs(T) < 1e-12)
    assert np.all(cp.th.source.numpy() == 0.0)
    assert (st["iter"], st["itPH"], st["thermal_iterations"]) == (so["iter"], so["itPH"], so["thermal_iterations"])
    assert rel(cp.dr.v.x.numpy(), cp.pr.vx) < 1e-10
    if shear:
        assert rel(T, cp.th_o.T) < 1e-10


def test_gpu_coupled_argument_checks():
    say("[gpu] coupled argument checks")
    coords, _ = O.t7_mesh(2, 2)
    gam = O.box_boundary_nodes(coords)
    zero = np.zeros(len(gam))
    cp = Coupled(2, dict(eta=(1.0,)), UNIT_T, ((gam, zero), (gam, zero)), (gam, zero), 1e-8)
    coarse = fa.t7_box_mesh(1, 1)
    th1 = fa.ThermalDiffusionDR(coarse, fa.ThermalMaterial(**UNIT_T))
    with pytest.raises(fa.DimensionMismatch, match="must match the Stokes velocity-node layout"):
        fa.solve_coupled(th1, cp.dr, coarse, cp.mesh, cp.bc_T, *cp.bc_v, dt=1.0)
    with pytest.raises(ValueError, match="FP64"):
        cp.solve_gpu(1.0, precision="fp32")
    other = fa.Context()
    m2 = fa.t7_box_mesh(2, 2, ctx=other)
    th2 = fa.ThermalDiffusionDR(m2, fa.ThermalMaterial(**UNIT_T))
    with pytest.raises(ValueError, match="share one Context"):
        fa.solve_coupled(th2, cp.dr, m2, cp.mesh, cp.bc_T, *cp.bc_v, dt=1.0)
    # a non-finite thermal residual is reported with the reference's message:
    # T0 enters the residual but not its Jacobian, here at a bubble node (the
    # Stokes pressure DoFs see only vertex temperatures)
    T0 = np.zeros(len(coords))
    T0[cp.el2n[0, 6]] = np.nan
    cp.th.T0.copyto(T0)
    cp.th_o.T0[:] = T0
    with pytest.raises(fa.FemError, match="non-finite thermal residual at coupled iteration 1"):
        cp.solve_gpu(1.0)
    with pytest.raises(RuntimeError, match="non-finite thermal residual at coupled iteration 1"):
        cp.solve_oracle(1.0)
    # a non-finite temperature reaches the Jacobian (ρ(T) is NaN even for α = 0)
    T = np.zeros(len(coords))
    T[cp.el2n[0, 6]] = np.nan
    cp.th.T0.fill(0.0)
    cp.th.T.copyto(T)
    with pytest.raises(fa.FemError, match="thermal diffusion preconditioner produced invalid λmax"):
        cp.solve_gpu(1.0)
