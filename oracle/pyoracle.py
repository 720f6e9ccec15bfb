"""ctypes binding of the CPU oracle (liboracle_femdr.so).

TEST INFRASTRUCTURE ONLY. Imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, as the checker; the product (libfemdr.so and
the ``femtools_amd`` package) never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_femdr.so")
MAXPH = 64

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


D = C.c_double
I32 = C.c_int32
I64 = C.c_int64
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)


class Material(C.Structure):
    _fields_ = [("nphases", I32), ("eta", D * MAXPH), ("etab", D * MAXPH), ("G", D * MAXPH),
                ("alpha", D * MAXPH), ("rho0", D * MAXPH), ("K", D * MAXPH), ("g", D * 2), ("Tref", D)]


class DP(C.Structure):
    _fields_ = [("nphases", I32), ("cos_phi", D * MAXPH), ("sin_phi", D * MAXPH), ("sin_psi", D * MAXPH),
                ("C", D * MAXPH), ("eta_reg", D * MAXPH), ("Kb", D * MAXPH)]


class Cap(C.Structure):
    _fields_ = [("nphases", I32), ("cos_phi", D * MAXPH), ("sin_phi", D * MAXPH), ("sin_psi", D * MAXPH),
                ("C", D * MAXPH), ("pT", D * MAXPH), ("eta_reg", D * MAXPH), ("Kb", D * MAXPH),
                ("C_min", D * MAXPH), ("H_C", D * MAXPH)]


class Stokes(C.Structure):
    _fields_ = [("nnodes", I64), ("nels", I64), ("nnodesP", I64),
                ("coords", PD), ("el2n", PI), ("invJ", PD), ("dOmega", PD), ("dOmegaP", PD),
                ("phases_v", PI), ("phase_mode_v", I32), ("phases_P", PI), ("phase_mode_P", I32),
                ("mat", Material), ("dp", C.POINTER(DP)),
                ("vx", PD), ("vy", PD), ("rate_x", PD), ("rate_y", PD), ("Rvx", PD), ("Rvy", PD),
                ("Rv0x", PD), ("Rv0y", PD), ("dRdvx", PD), ("dRdvy", PD), ("PCx", PD), ("PCy", PD),
                ("P", PD), ("P0", PD), ("gammaP", PD), ("T", PD), ("T0", PD), ("Q", PD), ("RP", PD),
                ("MP", PD), ("Pnum", PD), ("Pf", PD), ("tau_old", PD * 3), ("tau_store", PD * 3),
                ("nbc", I64 * 2), ("bc_dofs", PI * 2), ("bc_vals", PD * 2), ("cap", C.POINTER(Cap)), ("gamma_hist", PD), ("theta_hist", PD), ("P_store", PD),
                ("body_force", PD * 2), ("eta_ip", PD), ("eta_scale_ip", PD), ("Pold_ip", PD),
                ("history", PD), ("history_cap", I64), ("history_len", I64)]


class Params(C.Structure):
    _fields_ = [("dt", D), ("pressure_factor", D), ("tolerance", D), ("max_iterations", I64),
                ("iter_max", I64), ("check_interval", I64), ("max_ph_iterations", I64), ("rel_drop0", D),
                ("CFL_v", D), ("c_fact", D), ("finite_K", I32), ("freeze_jacobian", I32),
                ("skip_scaling", I32), ("verbose", I32), ("measure_lmax", I32), ("lmax_power_iterations", I32),
                ("lmax_power_rtol", D), ("lmax_safety", D)]


class Stats(C.Structure):
    _fields_ = [("itPH", I64), ("iter", I64), ("err", D), ("err_abs", D), ("err_rel", D), ("err_v", D),
                ("err_P", D), ("converged", I32), ("reached_total_iter", I32), ("lambda_max", D),
                ("lambda_max_gershgorin", D), ("jacobian_assemblies", I64), ("status", I32),
                ("lambda_max_iterations", I64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ThermalMaterialC(C.Structure):
    _fields_ = [("nphases", I32), ("k", D * MAXPH), ("Cp", D * MAXPH), ("rho0", D * MAXPH), ("alpha", D * MAXPH),
                ("K", D * MAXPH)]


class ThermalC(C.Structure):
    _fields_ = [("nnodes", I64), ("nels", I64), ("el2n", PI), ("invJ", PD), ("dOmega", PD), ("phases", PI),
                ("mat", ThermalMaterialC), ("R", PD), ("R0", PD), ("dRdT", PD), ("PC", PD), ("T", PD), ("T0", PD),
                ("rate", PD), ("P", PD), ("source", PD), ("nbc", I64), ("bc_dofs", PI), ("bc_vals", PD),
                ("CFL", D), ("c_fact", D), ("source_ip", PD), ("dim", I32), ("nloc", I32)]


class CoupledParamsC(C.Structure):
    _fields_ = [("Tref", D), ("tolerance_T", D), ("shear_heating", I32)]


class CoupledStatsC(C.Structure):
    _fields_ = [("err_T", D), ("thermal_iterations", I64), ("thermal_converged", I32)]


class ThermalParamsC(C.Structure):
    _fields_ = [("dt", D), ("tolerance", D), ("Tref", D), ("max_iterations", I64), ("check_interval", I64),
                ("verbose", I32)]


class ThermalStatsC(C.Structure):
    _fields_ = [("converged", I32), ("iterations", I64), ("residual", D), ("status", I32)]


class AdjointParamsC(C.Structure):
    _fields_ = [("dt", D), ("tolerance", D), ("rel_drop", D), ("CFL_v", D), ("c_fact", D), ("max_iterations", I64),
                ("iter_max", I64), ("check_interval", I64), ("max_ph_iterations", I64), ("measure_lmax", I32),
                ("verbose", I32)]


class AdjointStatsC(C.Structure):
    _fields_ = [("itPH", I64), ("iter", I64), ("lmax_iterations", I64), ("err", D), ("err_v", D), ("err_P", D),
                ("lmax", D), ("lmax_gershgorin", D), ("converged", I32), ("status", I32)]


class CP3Params(C.Structure):
    _fields_ = [("tolerance", D), ("max_iterations", I64), ("check_interval", I64), ("velocity_step", D),
                ("pressure_step", D), ("verbose", I32)]


class CP3Stats(C.Structure):
    _fields_ = [("converged", I32), ("reached_total_iter", I32), ("iterations", I64), ("residual", D),
                ("err_v", D), ("err_P", D), ("status", I32)]


class Stokes3(C.Structure):
    _fields_ = [("nnodes", I64), ("nels", I64), ("nnodesP", I64), ("coords", PD), ("el2n", PI), ("invJ", PD),
                ("dOmega", PD), ("phases", PI), ("mat", Material), ("g", D * 3), ("dp", C.POINTER(DP)),
                ("v", PD * 3), ("rate", PD * 3), ("Rv", PD * 3), ("Rv0", PD * 3), ("dRdv", PD * 3), ("PC", PD * 3),
                ("P", PD), ("P0", PD), ("gammaP", PD), ("RP", PD), ("MP", PD), ("Pnum", PD),
                ("tau_old", PD * 6), ("tau_store", PD * 6), ("nbc", I64 * 3), ("bc_dofs", PI * 3),
                ("bc_vals", PD * 3), ("body_force", PD * 3), ("T", PD), ("T0", PD)]


def _declare(L):
    L.orc_stokes_lmin.argtypes = [D, PD, PD, PD, PD, I64]
    L.orc_stokes_lmin.restype = D
    L.orc_stokes_cheb.argtypes = [D, D, D, PD, PD]
    L.orc_hex27_basis.argtypes = [D, D, D, PD, PD]
    L.orc_hex27_rule.argtypes = [PD, PD, PD, PD]
    L.orc_hex27_tables.argtypes = [PD, PD, PD, PD]
    L.orc_hex27_mesh_sizes.argtypes = [I32, I32, I32, C.POINTER(I64), C.POINTER(I64)]
    L.orc_hex27_mesh.argtypes = [I32, I32, I32, D, D, D, D, D, D, PD, PI]
    L.orc_hex27_geometry.argtypes = [I64, PD, PI, PD, PD]
    L.orc_adjoint_blocks.argtypes = [C.POINTER(Stokes), D, PD, PD, PD, PD, PD, PD]
    L.orc_adjoint_apply.argtypes = [C.POINTER(Stokes), PD, PD, PD, PD, PD, I32, I32, PD, PD, PD, PD, PD, PD]
    L.orc_adjoint_solve.argtypes = [C.POINTER(Stokes), C.POINTER(AdjointParamsC), PD, PD, PD, PD, PD,
                                    C.POINTER(AdjointStatsC)]
    L.orc_q9_basis.argtypes = [D, D, PD, PD]
    L.orc_q9_rule.argtypes = [PD, PD, PD]
    L.orc_q9_mesh_sizes.argtypes = [I32, I32, C.POINTER(I64), C.POINTER(I64)]
    L.orc_q9_mesh.argtypes = [I32, I32, D, D, D, D, PD, PI]
    L.orc_q9_geometry.argtypes = [I64, PD, PI, PD, PD]
    L.orc_momentum_element3.argtypes = [I32, I32, I32, PD, PD, PD, PD, PD, PD, PD, PD, PD, PI, C.POINTER(Material),
                                        PD, C.POINTER(DP), D, PD, PD, PD, PD, PD]
    for f in ("orc_pressure_scaling3", "orc_augmented_jacobian3"):
        getattr(L, f).argtypes = [C.POINTER(Stokes3), D, D, I32] if f == "orc_pressure_scaling3" \
            else [C.POINTER(Stokes3), D, I32]
    L.orc_pressure_residual3.argtypes = [C.POINTER(Stokes3), D, I32]
    L.orc_momentum_residual3.argtypes = [C.POINTER(Stokes3), D, PD]
    L.orc_stokes_solve3.argtypes = [C.POINTER(Stokes3), C.POINTER(Params), C.POINTER(Stats)]
    L.orc_set_pressure_source.argtypes = [C.POINTER(Stokes), PD, D, I32]
    L.orc_set_pressure_source.restype = C.c_int
    L.orc_stokes_solve3.restype = C.c_int
    L.orc_update_current_stress3.argtypes = [C.POINTER(Stokes3), D]
    L.orc_cp3_pressure_residual.argtypes = [C.POINTER(Stokes3), PD]
    L.orc_cp3_momentum_residual.argtypes = [C.POINTER(Stokes3), C.POINTER(PD)]
    L.orc_cp3_preconditioner.argtypes = [C.POINTER(Stokes3), C.POINTER(PD), PD]
    L.orc_cp3_solve.argtypes = [C.POINTER(Stokes3), C.POINTER(CP3Params), C.POINTER(CP3Stats), PD, I64,
                                C.POINTER(I64)]
    L.orc_cp3_solve.restype = C.c_int
    L.orc_thermal_integrate.argtypes = [I32, I32, PD, PD, PD, PD, PD, PD, PI, C.POINTER(ThermalMaterialC), PD, D, D,
                                        PD, PD]
    L.orc_thermal_residual.argtypes = [C.POINTER(ThermalC), D, D, I32]
    L.orc_thermal_solve.argtypes = [C.POINTER(ThermalC), C.POINTER(ThermalParamsC), C.POINTER(ThermalStatsC)]
    L.orc_thermal_solve.restype = C.c_int
    L.orc_shear_heating.argtypes = [C.POINTER(Stokes), D, PD]
    L.orc_cap_geometry.argtypes = [D, D, D, D, PD]
    L.orc_cap_yield_function.argtypes = [D, D, D, D, D, D]
    L.orc_cap_yield_function.restype = D
    L.orc_cap_invariants.argtypes = [D, D, D, D, D, D, PD]
    L.orc_cap_return_map.argtypes = [D, D, D, D, D, D, D, D, D, I32, PD]
    L.orc_cap_return_map.restype = C.c_int
    L.orc_cap_local_update.argtypes = [PD, D, D, D, D, D, D, D, D, PD, PD]
    L.orc_commit_plastic_history.argtypes = [C.POINTER(Stokes), D]
    L.orc_commit_plastic_history.restype = C.c_int
    L.orc_rotate_elastic_stress2.argtypes = [D, PD, D, PD]
    L.orc_rotate_elastic_stress3.argtypes = [PD, PD, D, PD]
    L.orc_rotate_stress.argtypes = [C.POINTER(Stokes), D, C.POINTER(PD), C.POINTER(PD)]
    L.orc_rotate_stress3.argtypes = [C.POINTER(Stokes3), D, C.POINTER(PD), C.POINTER(PD)]
    L.orc_coupled_solve.argtypes = [C.POINTER(ThermalC), C.POINTER(Stokes), C.POINTER(Params),
                                    C.POINTER(CoupledParamsC), C.POINTER(Stats), C.POINTER(CoupledStatsC)]
    L.orc_coupled_solve.restype = C.c_int
    L.orc_shear_heating3.argtypes = [C.POINTER(Stokes3), D, PD]
    L.orc_shear_heating3.restype = None
    L.orc_coupled_solve3.argtypes = [C.POINTER(ThermalC), C.POINTER(Stokes3), C.POINTER(Params),
                                     C.POINTER(CoupledParamsC), C.POINTER(Stats), C.POINTER(CoupledStatsC)]
    L.orc_coupled_solve3.restype = C.c_int
    L.orc_last_error.restype = C.c_char_p
    L.orc_t7_basis.argtypes = [D, D, PD, PD]
    L.orc_t7_rule.argtypes = [PD, PD, PD]
    L.orc_t7_tables.argtypes = [PD, PD, PD, PD]
    L.orc_t7_mesh_sizes.argtypes = [I32, I32, C.POINTER(I64), C.POINTER(I64)]
    L.orc_t7_mesh.argtypes = [I32, I32, D, D, D, D, PD, PI]
    L.orc_box_boundary_nodes.argtypes = [I64, PD, D, D, D, D, PI]
    L.orc_box_boundary_nodes.restype = I64
    L.orc_t7_geometry.argtypes = [I64, PD, PI, PD, PD, PD]
    L.orc_pressure_element.argtypes = [I32, I32, I32, PD, PD, PD, PD, PD, PD, PD, PD, PD, PD, PD, PI, PD, PD, D, PD]
    L.orc_momentum_element.argtypes = [I32, I32, I32, PD, PD, PD, PD, PD, PD, PD, PD, PD, PI,
                                       C.POINTER(Material), PD, C.POINTER(DP), D, PD, PD, PD, PD]
    L.orc_stress_at_ip.argtypes = [PD, PD, PD, I32, PI, C.POINTER(Material), C.POINTER(DP), D, PD, D, PD]
    L.orc_stress_at_ip.restype = D
    L.orc_pressure_scaling.argtypes = [C.POINTER(Stokes), D, D, I32]
    L.orc_pressure_residual.argtypes = [C.POINTER(Stokes), D, I32, PD]
    L.orc_momentum_residual.argtypes = [C.POINTER(Stokes), D, PD]
    L.orc_augmented_jacobian.argtypes = [C.POINTER(Stokes), D, I32]
    L.orc_dr_update.argtypes = [C.POINTER(Stokes), PD, PD]
    L.orc_stokes_solve.argtypes = [C.POINTER(Stokes), C.POINTER(Params), C.POINTER(Stats)]
    L.orc_stokes_solve.restype = C.c_int
    L.orc_update_current_stress.argtypes = [C.POINTER(Stokes), D]
    L.orc_dr_iterations.argtypes = [C.POINTER(Stokes), D, I32, PD, PD, I64]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_num_threads.restype = C.c_int


def ptr(a, kind=PD):
    if a is None:
        return C.cast(None, kind)
    return a.ctypes.data_as(kind)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def stokes_lmin(step, rate, R, R0, PC):
    """_stokes_λmin (DR.jl:32-37) over one velocity component."""
    a = [f64(x) for x in (rate, R, R0, PC)]
    return lib().orc_stokes_lmin(float(step), *(ptr(x) for x in a), len(a[0]))


def stokes_cheb(dtau, lmin, c_fact):
    """_stokes_cheb (DR.jl:39-42): (α, β)."""
    a, b = C.c_double(), C.c_double()
    lib().orc_stokes_cheb(float(dtau), float(lmin), float(c_fact), C.byref(a), C.byref(b))
    return a.value, b.value


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------------------
def t7_basis(xi, eta):
    N = np.zeros(7)
    dN = np.zeros((7, 2))
    lib().orc_t7_basis(xi, eta, ptr(N), ptr(dN))
    return N, dN


def t7_rule():
    xi, eta, w = np.zeros(7), np.zeros(7), np.zeros(7)
    lib().orc_t7_rule(ptr(xi), ptr(eta), ptr(w))
    return xi, eta, w


def t7_tables():
    Nq, NqP, dNdxi, w = np.zeros((7, 7)), np.zeros((7, 3)), np.zeros((7, 7, 2)), np.zeros(7)
    lib().orc_t7_tables(ptr(Nq), ptr(NqP), ptr(dNdxi), ptr(w))
    return Nq, NqP, dNdxi, w


def t7_mesh(nx, ny, box=(0.0, 1.0, 0.0, 1.0)):
    nn, ne = I64(), I64()
    lib().orc_t7_mesh_sizes(nx, ny, C.byref(nn), C.byref(ne))
    coords = np.zeros((nn.value, 2))
    el2n = np.zeros((ne.value, 7), dtype=np.int32)
    lib().orc_t7_mesh(nx, ny, *box, ptr(coords), ptr(el2n, PI))
    return coords, el2n


def box_boundary_nodes(coords, box=(0.0, 1.0, 0.0, 1.0)):
    n = lib().orc_box_boundary_nodes(len(coords), ptr(coords), *box, ptr(None, PI))
    out = np.zeros(n, dtype=np.int32)
    lib().orc_box_boundary_nodes(len(coords), ptr(coords), *box, ptr(out, PI))
    return out


def t7_geometry(coords, el2n):
    ne = len(el2n)
    invJ = np.zeros((ne, 7, 4))
    dO = np.zeros((ne, 7))
    dOP = np.zeros((ne, 7))
    lib().orc_t7_geometry(ne, ptr(coords), ptr(el2n, PI), ptr(invJ), ptr(dO), ptr(dOP))
    return invJ, dO, dOP


def material(eta=(1.0,), etab=(1.0,), G=(math.inf,), alpha=(0.0,), rho0=(1.0,), K=(math.inf,),
             g=(0.0, 0.0), Tref=0.0):
    """StokesMaterial defaults (src/stokes/types/stokes_types.jl:45-60)."""
    props = [eta, etab, G, alpha, rho0, K]
    n = max(len(p) if isinstance(p, (tuple, list, np.ndarray)) else 1 for p in props)
    m = Material()
    m.nphases = n
    for name, val in zip(("eta", "etab", "G", "alpha", "rho0", "K"), props):
        vals = list(val) if isinstance(val, (tuple, list, np.ndarray)) else [val] * n
        if len(vals) == 1:
            vals = vals * n
        arr = getattr(m, name)
        for i, x in enumerate(vals):
            arr[i] = float(x)
    m.g[0], m.g[1] = float(g[0]), float(g[1])
    m.Tref = float(Tref)
    return m


def drucker_prager(phi, psi, C_, eta_reg, Kb):
    """DruckerPrager(ϕ, Ψ, C, η_reg, Kb) (stokes_types.jl:411-423); angles in radians."""
    n = len(phi)
    dp = DP()
    dp.nphases = n
    for i in range(n):
        dp.cos_phi[i] = math.cos(phi[i])
        dp.sin_phi[i] = math.sin(phi[i])
        dp.sin_psi[i] = math.sin(psi[i])
        dp.C[i] = C_[i]
        dp.eta_reg[i] = eta_reg[i]
        dp.Kb[i] = Kb[i]
    return dp


@dataclass
class Problem:
    """Host arrays of one T7/P1-disc Stokes problem, mirroring StokesDR fields."""
    coords: np.ndarray
    el2n: np.ndarray
    mat: Material
    dp: DP | None = None
    phases_v: np.ndarray | None = None
    phase_mode_v: int = 0
    phases_P: np.ndarray | None = None
    phase_mode_P: int = 0
    tau_old: tuple | None = None
    bc: tuple = ((np.zeros(0, np.int32), np.zeros(0)), (np.zeros(0, np.int32), np.zeros(0)))
    arrays: dict = field(default_factory=dict)
    cap: Cap | None = None
    history: bool = False
    # array keywords of solve! (DR.jl:183-186, 217-226, 302-337): nodal loads
    # (bf_x, bf_y), viscosity at the points [nels][7], viscosity of the pressure
    # scale [nels][7], previous-step pressure at the points [nels][7]
    body_force: tuple | None = None
    eta_ip: np.ndarray | None = None
    eta_scale_ip: np.ndarray | None = None
    Pold_ip: np.ndarray | None = None

    def __post_init__(self):
        self.coords = f64(self.coords)
        self.el2n = i32(self.el2n)
        nn, ne = len(self.coords), len(self.el2n)
        npp = 3 * ne
        self.nnodes, self.nels, self.nnodesP = nn, ne, npp
        self.invJ, self.dOmega, self.dOmegaP = t7_geometry(self.coords, self.el2n)
        if self.phases_v is None:
            self.phases_v, self.phase_mode_v = np.ones(ne, np.int32), 0
        if self.phases_P is None:
            self.phases_P, self.phase_mode_P = np.ones(ne, np.int32), 0
        self.phases_v = i32(self.phases_v)
        self.phases_P = i32(self.phases_P)
        for k in ("vx", "vy", "rate_x", "rate_y", "Rvx", "Rvy", "Rv0x", "Rv0y", "dRdvx", "dRdvy", "PCx", "PCy"):
            self.arrays[k] = np.zeros(nn)
        for k in ("P", "P0", "gammaP", "T", "T0", "Q", "RP", "MP", "Pnum"):
            self.arrays[k] = np.zeros(npp)
        self.tau_store = [np.zeros((ne, 7)) for _ in range(3)]
        self.P_store = np.zeros((ne, 7))
        self.gamma_hist = np.zeros((ne, 7)) if self.history else None
        self.theta_hist = np.zeros((ne, 7)) if self.history else None
        if self.tau_old is not None:
            self.tau_old = tuple(f64(t) for t in self.tau_old)
        self.bc = tuple((i32(d), f64(v)) for d, v in self.bc)
        self._dp_ref = self.dp

    def __getattr__(self, k):
        arrays = self.__dict__.get("arrays", {})
        if k in arrays:
            return arrays[k]
        raise AttributeError(k)

    def struct(self):
        s = Stokes()
        s.nnodes, s.nels, s.nnodesP = self.nnodes, self.nels, self.nnodesP
        s.coords, s.el2n = ptr(self.coords), ptr(self.el2n, PI)
        s.invJ, s.dOmega, s.dOmegaP = ptr(self.invJ), ptr(self.dOmega), ptr(self.dOmegaP)
        s.phases_v, s.phase_mode_v = ptr(self.phases_v, PI), self.phase_mode_v
        s.phases_P, s.phase_mode_P = ptr(self.phases_P, PI), self.phase_mode_P
        s.mat = self.mat
        s.dp = C.pointer(self.dp) if self.dp is not None else C.cast(None, C.POINTER(DP))
        for k, a in self.arrays.items():
            setattr(s, k, ptr(a))
        s.Pf = ptr(None)
        for c in range(3):
            s.tau_old[c] = ptr(self.tau_old[c]) if self.tau_old is not None else ptr(None)
            s.tau_store[c] = ptr(self.tau_store[c])
        for c in range(2):
            d, v = self.bc[c]
            s.nbc[c] = len(d)
            s.bc_dofs[c] = ptr(d, PI)
            s.bc_vals[c] = ptr(v)
        s.cap = C.pointer(self.cap) if self.cap is not None else C.cast(None, C.POINTER(Cap))
        s.gamma_hist = ptr(self.gamma_hist) if self.gamma_hist is not None else ptr(None)
        s.theta_hist = ptr(self.theta_hist) if self.theta_hist is not None else ptr(None)
        s.P_store = ptr(self.P_store)
        self._kw = [None if a is None else f64(a) for a in
                    (*(self.body_force or (None, None)), self.eta_ip, self.eta_scale_ip, self.Pold_ip)]
        bfx, bfy, eip, esc, pold = self._kw
        s.body_force[0], s.body_force[1] = ptr(bfx) if bfx is not None else ptr(None), \
            ptr(bfy) if bfy is not None else ptr(None)
        s.eta_ip = ptr(eip) if eip is not None else ptr(None)
        s.eta_scale_ip = ptr(esc) if esc is not None else ptr(None)
        s.Pold_ip = ptr(pold) if pold is not None else ptr(None)
        s.history, s.history_cap, s.history_len = ptr(None), 0, 0
        self._struct = s
        return s

    def set_pressure_source(self, Qq, Q2D=None):
        """_set_pressure_source! (pressure_residual.jl:32-77) into Q."""
        q = f64(Qq)
        rc = lib().orc_set_pressure_source(C.byref(self.struct()), ptr(q), 0.0 if Q2D is None else float(Q2D),
                                           int(Q2D is not None))
        if rc != 0:
            raise ValueError(lib().orc_last_error().decode())

    def commit_plastic_history(self, dt):
        """commit_stokes_plastic_history! (DR.jl:899-958)."""
        if lib().orc_commit_plastic_history(C.byref(self.struct()), dt) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    # assemblers -------------------------------------------------------------
    def pressure_scaling(self, gamma_fact, dt, finite_K=False):
        lib().orc_pressure_scaling(C.byref(self.struct()), gamma_fact, dt, int(finite_K))

    def pressure_residual(self, dt, finite_K=False):
        lib().orc_pressure_residual(C.byref(self.struct()), dt, int(finite_K), ptr(None))

    def momentum_residual(self, dt, Pnum=None):
        lib().orc_momentum_residual(C.byref(self.struct()), dt, ptr(f64(Pnum)) if Pnum is not None else ptr(None))

    def augmented_jacobian(self, dt, finite_K=False):
        lib().orc_augmented_jacobian(C.byref(self.struct()), dt, int(finite_K))

    def dr_iterations(self, dt, alpha, beta, k, finite_K=False):
        a = f64(alpha)
        b = f64(beta)
        lib().orc_dr_iterations(C.byref(self.struct()), dt, int(finite_K), ptr(a), ptr(b), int(k))

    def update_current_stress(self, dt):
        lib().orc_update_current_stress(C.byref(self.struct()), dt)
        return self.tau_store

    def rotate_stress(self, dt, tau):
        """rotate_stress! (stress_rotation.jl:28-61): the rotated tau (3 arrays [nels][7])."""
        tau = [f64(t).reshape(self.nels, 7) for t in tau]
        out = [np.zeros((self.nels, 7)) for _ in range(3)]
        lib().orc_rotate_stress(C.byref(self.struct()), dt, (PD * 3)(*[ptr(t) for t in tau]),
                                (PD * 3)(*[ptr(o) for o in out]))
        return out

    def shear_heating(self, dt):
        """assemble_shear_heating! (shear_heating.jl:18-92): Φ [nels][7]."""
        phi = np.zeros((self.nels, 7))
        lib().orc_shear_heating(C.byref(self.struct()), dt, ptr(phi))
        return phi

    # the frozen-operator adjoint (DR_adjoint.jl, adjoint_operator.jl) ------------
    def adjoint_flags(self):
        """(symmetric, stores) of the block layout (adjoint_operator.jl:172-182, 455-458)."""
        m = self.mat
        n = m.nphases
        sym = self.dp is None and self.cap is None and all(math.isinf(m.K[i]) for i in range(n))
        stores = any(math.isfinite(m.etab[i]) for i in range(n))
        return sym, stores

    def adjoint_blocks(self, dt):
        """Element blocks A, B, C, D, W ([nels][14][14], [14][3], [3][14], [3][3],
        [14][3]) and the symmetry defects; fills dRdvx/y and PCx/y."""
        ne = self.nels
        A, B, Cm = np.zeros((ne, 14, 14)), np.zeros((ne, 14, 3)), np.zeros((ne, 3, 14))
        Dm, W, df = np.zeros((ne, 3, 3)), np.zeros((ne, 14, 3)), np.zeros(2)
        if lib().orc_adjoint_blocks(C.byref(self.struct()), dt, ptr(A), ptr(B), ptr(Cm), ptr(Dm), ptr(W), ptr(df)):
            raise ValueError(lib().orc_last_error().decode())
        return dict(A=A, B=B, C=Cm, D=Dm, W=W, defects=df)

    def adjoint_apply(self, blocks, lx, ly, lP):
        sym, stores = self.adjoint_flags()
        dvx, dvy, dP = np.zeros(self.nnodes), np.zeros(self.nnodes), np.zeros(self.nnodesP)
        b = {k: f64(v) for k, v in blocks.items()}
        lib().orc_adjoint_apply(C.byref(self.struct()), ptr(b["A"]), ptr(b["B"]), ptr(b["C"]), ptr(b["D"]),
                                ptr(b["W"]), int(sym), int(stores), ptr(f64(lx)), ptr(f64(ly)), ptr(f64(lP)),
                                ptr(dvx), ptr(dvy), ptr(dP))
        return dvx, dvy, dP

    def solve_adjoint(self, dt, objective, lam, tolerance=1e-6, max_iterations=50_000, iter_max=50_000,
                      check_interval=50, max_ph_iterations=100, rel_drop=0.1, CFL_v=0.98, c_fact=0.9,
                      measure_lmax=True, verbose=False, throw_on_failure=True):
        """solve_adjoint! (DR_adjoint.jl:396-427) with operator = :blocks. objective =
        (∂J/∂vx, ∂J/∂vy); lam = [λvx, λvy, λP] (float64, updated in place from
        their values as the initial iterate). Returns the statistics."""
        p = AdjointParamsC(dt=dt, tolerance=tolerance, rel_drop=rel_drop, CFL_v=CFL_v, c_fact=c_fact,
                           max_iterations=max_iterations, iter_max=iter_max, check_interval=check_interval,
                           max_ph_iterations=max_ph_iterations, measure_lmax=int(bool(measure_lmax)),
                           verbose=int(bool(verbose)))
        st = AdjointStatsC()
        ox, oy = f64(objective[0]), f64(objective[1])
        for a in lam:
            assert a.dtype == np.float64 and a.flags.c_contiguous
        rc = lib().orc_adjoint_solve(C.byref(self.struct()), C.byref(p), ptr(ox), ptr(oy), ptr(lam[0]), ptr(lam[1]),
                                     ptr(lam[2]), C.byref(st))
        if rc != 0:
            msg = lib().orc_last_error().decode()
            raise (ValueError(msg) if rc == -1 else RuntimeError(msg))
        out = {k: getattr(st, k) for k, _ in AdjointStatsC._fields_}
        out["converged"] = bool(out["converged"])
        out["iterations"], out["residual"] = out["iter"], out["err"]
        if throw_on_failure and not out["converged"]:
            raise RuntimeError(f"Stokes adjoint solve did not converge after {out['iter']} iterations "
                               f"(residual = {out['err']}, tolerance = {tolerance})")
        return out

    @staticmethod
    def params(dt, pressure_factor=50.0, tolerance=1e-6, max_iterations=50_000, iter_max=50_000,
               check_interval=50, max_ph_iterations=1000, rel_drop0=1e-2, CFL_v=0.98, c_fact=0.9,
               finite_K=False, freeze_jacobian=-1, skip_scaling=False, verbose=False, measure_lmax=False,
               lmax_power_iterations=12, lmax_power_rtol=1e-2, lmax_safety=1.1):
        return Params(dt=dt, pressure_factor=pressure_factor, tolerance=tolerance, max_iterations=max_iterations,
                      iter_max=iter_max, check_interval=check_interval, max_ph_iterations=max_ph_iterations,
                      rel_drop0=rel_drop0, CFL_v=CFL_v, c_fact=c_fact, finite_K=int(finite_K),
                      freeze_jacobian=int(freeze_jacobian), skip_scaling=int(skip_scaling), verbose=int(verbose),
                      measure_lmax=int(bool(measure_lmax)), lmax_power_iterations=int(lmax_power_iterations),
                      lmax_power_rtol=float(lmax_power_rtol), lmax_safety=float(lmax_safety))

    def solve(self, dt, collect_history=False, **kw):
        p = self.params(dt, **kw)
        st = Stats()
        s = self.struct()
        hist = None
        if collect_history:
            cap = 2 * (p.max_iterations // max(p.check_interval, 1) + p.max_ph_iterations) + 4
            hist = np.zeros((min(cap, 1_000_000), 6))
            s.history, s.history_cap = ptr(hist), len(hist)
        rc = lib().orc_stokes_solve(C.byref(s), C.byref(p), C.byref(st))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        out = st.as_dict()
        if collect_history:
            out["history"] = hist[:s.history_len].copy()
        return out


def drucker_prager_cap(phi, psi, C_, pT, eta_reg, Kb, C_min=None, H_C=None):
    """DruckerPragerCap(ϕ, Ψ, C, pT, η_reg, Kb; C_min=C, H_C=0)
    (stokes_types.jl:492-544); angles in radians, no validation."""
    n = len(phi)
    cp = Cap()
    cp.nphases = n
    C_min = C_ if C_min is None else C_min
    H_C = (0.0,) * n if H_C is None else H_C
    for i in range(n):
        cp.cos_phi[i], cp.sin_phi[i], cp.sin_psi[i] = math.cos(phi[i]), math.sin(phi[i]), math.sin(psi[i])
        cp.C[i], cp.pT[i], cp.eta_reg[i], cp.Kb[i] = C_[i], pT[i], eta_reg[i], Kb[i]
        cp.C_min[i], cp.H_C[i] = C_min[i], H_C[i]
    return cp


def cap_geometry(k, kq, c, pT):
    out = np.zeros(7)
    lib().orc_cap_geometry(k, kq, c, pT, ptr(out))
    return dict(zip(("a", "b", "p_y", "R_y", "p_d", "tau_d", "p_q"), out))


def cap_yield_function(tII, P, k, kq, c, pT):
    return lib().orc_cap_yield_function(tII, P, k, kq, c, pT)


def cap_invariants(s, p, k, kq, c, pT):
    out = np.zeros(3)
    lib().orc_cap_invariants(s, p, k, kq, c, pT, ptr(out))
    return tuple(out)


def cap_return_map(s_trial, p_trial, eta_ve, KDt, k, kq, c, pT, eta_reg, maxiter=40):
    out = np.zeros(5)
    conv = lib().orc_cap_return_map(s_trial, p_trial, eta_ve, KDt, k, kq, c, pT, eta_reg, maxiter, ptr(out))
    return dict(tII=out[0], P=out[1], lam=out[2], At=out[3], Ap=out[4], converged=bool(conv))


def cap_local_update(tau_trial, P_trial, eta_ve, KDt, k, kq, c, pT, eta_reg):
    tau = np.zeros(3)
    out = np.zeros(3)
    lib().orc_cap_local_update(ptr(f64(tau_trial)), P_trial, eta_ve, KDt, k, kq, c, pT, eta_reg, ptr(tau), ptr(out))
    return dict(tau=tau, P=out[0], gdot=out[1], thdot=out[2])


def pure_shear_bc(coords, box=(0.0, 1.0, 0.0, 1.0), edot=1.0):
    """Pure-shear velocity BCs of the pure-shear triangle example
    (examples/miniapps/stokes/stokes_2D_pure_shear_triangle/...jl:145-154)."""
    x0, x1, y0, y1 = box
    gam = box_boundary_nodes(coords, box)
    tol = max(abs(x1 - x0), abs(y1 - y0)) * np.finfo(float).eps * 32
    x, y = coords[gam, 0], coords[gam, 1]
    nx = gam[(np.abs(x - x0) <= tol) | (np.abs(x - x1) <= tol)]
    ny = gam[(np.abs(y - y0) <= tol) | (np.abs(y - y1) <= tol)]
    xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    vx = edot * (coords[nx, 0] - xm)
    vy = -edot * (coords[ny, 1] - ym)
    return (nx, vx), (ny, vy)


# ---------------------------------------------------------------------------
# thermal diffusion
def thermal_material(k=(1.0,), Cp=(1.0,), rho0=(1.0,), alpha=(1.0,), K=(1.0,)):
    """ThermalMaterial (heat_diffusion_types.jl:15-41); defaults are the reference's unit values."""
    m = ThermalMaterialC()
    n = len(k)
    m.nphases = n
    for i in range(n):
        m.k[i], m.Cp[i], m.rho0[i], m.alpha[i], m.K[i] = k[i], Cp[i], rho0[i], alpha[i], K[i]
    return m


def thermal_integrate(N, dNdx, dOm, T, T0, s, phases, mat, P, dt, Tref, jacobian=False):
    """integrate_residual (heat residual.jl:58-87) on caller tables: N [nq][nn],
    dNdx [nq][nn][2], dOm [nq]. Returns Re (and J [nn][nn])."""
    N = f64(N)
    nq, nn = N.shape
    Re = np.zeros(nn)
    J = np.zeros((nn, nn)) if jacobian else None
    lib().orc_thermal_integrate(nn, nq, ptr(N), ptr(f64(dNdx)), ptr(f64(dOm)), ptr(f64(T)), ptr(f64(T0)),
                                ptr(f64(s)), ptr(i32(phases), PI), C.byref(mat), ptr(f64(P)), dt, Tref, ptr(Re),
                                ptr(J))
    return (Re, J) if jacobian else Re


class ThermalProblem:
    """Host arrays of a ThermalDiffusionDR on a T7 mesh (heat_diffusion_types.jl:90-153)."""

    def __init__(self, coords, el2n, mat, phases=None, bc=(np.zeros(0, np.int32), np.zeros(0)), CFL=0.98,
                 c_fact=0.9):
        self.coords = f64(coords)
        self.el2n = i32(el2n)
        nn, ne = len(self.coords), len(self.el2n)
        self.nnodes, self.nels = nn, ne
        self.invJ, self.dOmega, _ = t7_geometry(self.coords, self.el2n)
        self.phases = i32(np.ones(nn) if phases is None else phases)
        self.mat = mat
        for k in ("R", "R0", "dRdT", "PC", "T", "T0", "rate", "P", "source"):
            setattr(self, k, np.zeros(nn))
        self.bc = (i32(bc[0]), f64(bc[1]))
        self.CFL, self.c_fact = CFL, c_fact

    def struct(self):
        t = ThermalC()
        t.nnodes, t.nels = self.nnodes, self.nels
        t.el2n, t.invJ, t.dOmega, t.phases = ptr(self.el2n, PI), ptr(self.invJ), ptr(self.dOmega), ptr(self.phases, PI)
        t.mat = self.mat
        for k in ("R", "R0", "dRdT", "PC", "T", "T0", "rate", "P", "source"):
            setattr(t, k, ptr(getattr(self, k)))
        t.nbc = len(self.bc[0])
        t.bc_dofs, t.bc_vals = ptr(self.bc[0], PI), ptr(self.bc[1])
        t.CFL, t.c_fact = self.CFL, self.c_fact
        t.source_ip = ptr(getattr(self, "source_ip", None)) if getattr(self, "source_ip", None) is not None \
            else ptr(None)
        t.dim = getattr(self, "dim", 2)
        t.nloc = getattr(self, "nloc", 0)
        self._struct = t
        return t

    def residual(self, dt, Tref=273.0, jacobian=False):
        lib().orc_thermal_residual(C.byref(self.struct()), dt, Tref, int(jacobian))

    def solve(self, dt, tolerance=1e-6, max_iterations=10_000, check_interval=100, Tref=273.0, verbose=False):
        p = ThermalParamsC(dt=dt, tolerance=tolerance, Tref=Tref, max_iterations=max_iterations,
                           check_interval=check_interval, verbose=int(verbose))
        st = ThermalStatsC()
        rc = lib().orc_thermal_solve(C.byref(self.struct()), C.byref(p), C.byref(st))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return {"converged": bool(st.converged), "iterations": st.iterations, "residual": st.residual}


class ThermalProblem3(ThermalProblem):
    """A ThermalDiffusionDR on a Hex27 mesh (27 nodes, 3x3x3 Gauss points)."""

    def __init__(self, coords, el2n, mat, phases=None, bc=(np.zeros(0, np.int32), np.zeros(0)), CFL=0.98,
                 c_fact=0.9):
        self.coords = f64(coords)
        self.el2n = i32(el2n)
        nn, ne = len(self.coords), len(self.el2n)
        self.nnodes, self.nels = nn, ne
        self.invJ, self.dOmega = hex27_geometry(self.coords, self.el2n)
        self.phases = i32(np.ones(nn) if phases is None else phases)
        self.mat = mat
        for k in ("R", "R0", "dRdT", "PC", "T", "T0", "rate", "P", "source"):
            setattr(self, k, np.zeros(nn))
        self.bc = (i32(bc[0]), f64(bc[1]))
        self.CFL, self.c_fact = CFL, c_fact
        self.dim = 3
        self.source_ip = None


class ThermalProblemQ9(ThermalProblem):
    """A ThermalDiffusionDR on a Q9 mesh (9 nodes, 3x3 Gauss points)."""

    def __init__(self, coords, el2n, mat, phases=None, bc=(np.zeros(0, np.int32), np.zeros(0)), CFL=0.98,
                 c_fact=0.9):
        self.coords = f64(coords)
        self.el2n = i32(el2n)
        nn, ne = len(self.coords), len(self.el2n)
        self.nnodes, self.nels = nn, ne
        self.invJ, self.dOmega = q9_geometry(self.coords, self.el2n)
        self.phases = i32(np.ones(nn) if phases is None else phases)
        self.mat = mat
        for k in ("R", "R0", "dRdT", "PC", "T", "T0", "rate", "P", "source"):
            setattr(self, k, np.zeros(nn))
        self.bc = (i32(bc[0]), f64(bc[1]))
        self.CFL, self.c_fact = CFL, c_fact
        self.dim, self.nloc = 2, 9
        self.source_ip = None


def q9_basis(xi, eta):
    N, dN = np.zeros(9), np.zeros((9, 2))
    lib().orc_q9_basis(float(xi), float(eta), ptr(N), ptr(dN))
    return N, dN


def q9_rule():
    xi, eta, w = np.zeros(9), np.zeros(9), np.zeros(9)
    lib().orc_q9_rule(ptr(xi), ptr(eta), ptr(w))
    return xi, eta, w


def q9_mesh(nx, ny, box=(0.0, 1.0, 0.0, 1.0)):
    nn, ne = I64(), I64()
    lib().orc_q9_mesh_sizes(nx, ny, C.byref(nn), C.byref(ne))
    coords = np.zeros((nn.value, 2))
    el2n = np.zeros((ne.value, 9), np.int32)
    lib().orc_q9_mesh(nx, ny, *box, ptr(coords), ptr(el2n, PI))
    return coords, el2n


def q9_geometry(coords, el2n):
    ne = len(el2n)
    invJ = np.zeros((ne, 9, 4))
    dOm = np.zeros((ne, 9))
    lib().orc_q9_geometry(ne, ptr(f64(coords)), ptr(i32(el2n), PI), ptr(invJ), ptr(dOm))
    return invJ, dOm


def rotate_elastic_stress(omega, tau, dt):
    """GeoParams.rotate_elastic_stress as restated: 2-D (scalar ω, τ = xx, yy,
    xy) or 3-D (ω = curl v, τ in assembler order xx, yy, zz, xy, xz, yz)."""
    tau = f64(tau)
    out = np.zeros(len(tau))
    if np.ndim(omega) == 0:
        lib().orc_rotate_elastic_stress2(float(omega), ptr(tau), dt, ptr(out))
    else:
        lib().orc_rotate_elastic_stress3(ptr(f64(omega)), ptr(tau), dt, ptr(out))
    return out


def coupled_solve(thermal: ThermalProblem, stokes: Problem, dt, Tref=273.0, tolerance_T=1e-6, shear_heating=False,
                  **kw):
    """solve_coupled! (src/stokes/solvers/DR.jl:81-131): one thermal DR
    iteration per Stokes inner iteration. Keywords as Problem.solve; returns
    the Stokes stats with err_T, thermal_iterations and thermal_converged."""
    p = Problem.params(dt, **kw)
    cp = CoupledParamsC(Tref=Tref, tolerance_T=tolerance_T, shear_heating=int(bool(shear_heating)))
    st, cst = Stats(), CoupledStatsC()
    ts, ss = thermal.struct(), stokes.struct()
    rc = lib().orc_coupled_solve(C.byref(ts), C.byref(ss), C.byref(p), C.byref(cp), C.byref(st), C.byref(cst))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    out = st.as_dict()
    out.update(err_T=cst.err_T, thermal_iterations=cst.thermal_iterations,
               thermal_converged=bool(cst.thermal_converged))
    return out


def coupled_solve3(thermal: "ThermalProblem3", stokes: "Problem3", dt, Tref=273.0, tolerance_T=1e-6,
                   shear_heating=False, **kw):
    """solve_coupled! on a Hex27 state (DR.jl:81-131): as coupled_solve; the
    Stokes problem's T and T0 arrays receive the transferred temperature."""
    if stokes.T is None:
        stokes.T, stokes.T0 = np.zeros(4 * stokes.nels), np.zeros(4 * stokes.nels)
    p = Problem.params(dt, **kw)
    cp = CoupledParamsC(Tref=Tref, tolerance_T=tolerance_T, shear_heating=int(bool(shear_heating)))
    st, cst = Stats(), CoupledStatsC()
    ts, ss = thermal.struct(), stokes.struct()
    rc = lib().orc_coupled_solve3(C.byref(ts), C.byref(ss), C.byref(p), C.byref(cp), C.byref(st), C.byref(cst))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    out = st.as_dict()
    out.update(err_T=cst.err_T, thermal_iterations=cst.thermal_iterations,
               thermal_converged=bool(cst.thermal_converged))
    return out


# ---------------------------------------------------------------------------
# 3-D: Hex27 / P1-disc (4 modes)
def hex27_basis(xi, eta, zeta):
    N = np.zeros(27)
    dN = np.zeros((27, 3))
    lib().orc_hex27_basis(xi, eta, zeta, ptr(N), ptr(dN))
    return N, dN


def hex27_rule():
    xi, eta, zeta, w = (np.zeros(27) for _ in range(4))
    lib().orc_hex27_rule(ptr(xi), ptr(eta), ptr(zeta), ptr(w))
    return xi, eta, zeta, w


def hex27_tables():
    Nq, NqP, dN, w = np.zeros((27, 27)), np.zeros((27, 4)), np.zeros((27, 27, 3)), np.zeros(27)
    lib().orc_hex27_tables(ptr(Nq), ptr(NqP), ptr(dN), ptr(w))
    return Nq, NqP, dN, w


def hex27_mesh(nx, ny, nz, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    nn, ne = I64(), I64()
    lib().orc_hex27_mesh_sizes(nx, ny, nz, C.byref(nn), C.byref(ne))
    coords = np.zeros((nn.value, 3))
    el2n = np.zeros((ne.value, 27), np.int32)
    lib().orc_hex27_mesh(nx, ny, nz, *box, ptr(coords), ptr(el2n, PI))
    return coords, el2n


def hex27_geometry(coords, el2n):
    ne = len(el2n)
    invJ = np.zeros((ne, 27, 9))
    dOm = np.zeros((ne, 27))
    lib().orc_hex27_geometry(ne, ptr(f64(coords)), ptr(i32(el2n), PI), ptr(invJ), ptr(dOm))
    return invJ, dOm


def box_boundary_nodes3(coords, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), tol=1e-12):
    on = np.zeros(len(coords), bool)
    for d in range(3):
        on |= np.abs(coords[:, d] - box[2 * d]) <= tol
        on |= np.abs(coords[:, d] - box[2 * d + 1]) <= tol
    return np.nonzero(on)[0].astype(np.int32)


class Problem3:
    """Host arrays of one 3-D Hex27/P1-disc Stokes problem (cell phases)."""

    def __init__(self, coords, el2n, mat, g=(0.0, 0.0, 0.0), dp=None, phases=None, tau_old=None, bc=None):
        self.coords, self.el2n = f64(coords), i32(el2n)
        nn, ne = len(self.coords), len(self.el2n)
        self.nnodes, self.nels, self.nnodesP = nn, ne, 4 * ne
        self.invJ, self.dOmega = hex27_geometry(self.coords, self.el2n)
        self.phases = i32(np.ones(ne) if phases is None else phases)
        self.mat, self.g, self.dp = mat, tuple(g), dp
        for k in ("v", "rate", "Rv", "Rv0", "dRdv", "PC"):
            setattr(self, k, [np.zeros(nn) for _ in range(3)])
        for k in ("P", "P0", "gammaP", "RP", "MP", "Pnum"):
            setattr(self, k, np.zeros(4 * ne))
        self.tau_old = None if tau_old is None else [f64(t) for t in tau_old]
        self.tau_store = [np.zeros((ne, 27)) for _ in range(6)]
        e = (np.zeros(0, np.int32), np.zeros(0))
        self.bc = tuple((i32(d), f64(v)) for d, v in (bc if bc is not None else (e, e, e)))
        self.body_force = None   # nullable: 3 nodal loads of the cell-pressure relaxation
        self.T = None            # nullable (with T0): temperature at the pressure DoFs [4 nels]
        self.T0 = None

    def struct(self):
        s = Stokes3()
        s.nnodes, s.nels, s.nnodesP = self.nnodes, self.nels, self.nnodesP
        s.coords, s.el2n = ptr(self.coords), ptr(self.el2n, PI)
        s.invJ, s.dOmega, s.phases = ptr(self.invJ), ptr(self.dOmega), ptr(self.phases, PI)
        s.mat = self.mat
        for i in range(3):
            s.g[i] = self.g[i]
        s.dp = C.pointer(self.dp) if self.dp is not None else C.cast(None, C.POINTER(DP))
        for k in ("v", "rate", "Rv", "Rv0", "dRdv", "PC"):
            arr = getattr(s, k)
            for c in range(3):
                arr[c] = ptr(getattr(self, k)[c])
        for k in ("P", "P0", "gammaP", "RP", "MP", "Pnum"):
            setattr(s, k, ptr(getattr(self, k)))
        for k in range(6):
            s.tau_old[k] = ptr(self.tau_old[k]) if self.tau_old is not None else ptr(None)
            s.tau_store[k] = ptr(self.tau_store[k])
        for c in range(3):
            s.nbc[c] = len(self.bc[c][0])
            s.bc_dofs[c], s.bc_vals[c] = ptr(self.bc[c][0], PI), ptr(self.bc[c][1])
        self._bf = None if self.body_force is None else [f64(b) for b in self.body_force]
        for c in range(3):
            s.body_force[c] = ptr(self._bf[c]) if self._bf is not None else ptr(None)
        s.T = ptr(self.T) if self.T is not None else ptr(None)
        s.T0 = ptr(self.T0) if self.T is not None else ptr(None)
        self._struct = s
        return s

    def shear_heating(self, dt):
        """assemble_shear_heating! (shear_heating.jl:18-103), 3-D: (nels, 27)."""
        phi = np.zeros((self.nels, 27))
        lib().orc_shear_heating3(C.byref(self.struct()), dt, ptr(phi))
        return phi

    def cp3_solve_adjoint(self, objective, lam_v=None, lam_P=None, **kw):
        """solve_adjoint!(dr::CellPressureStokesDR, mesh, bc_v; objective_v, λv, λP)
        (DR_adjoint.jl:446-480): the relaxation on (λv, λP) with the objective as
        momentum load, zero density and gravity, homogeneous conditions on the
        constrained nodes; v, P and g of this problem are left as they were.
        Returns (λv, λP, stats)."""
        lam_v = [np.zeros(self.nnodes) if lam_v is None else f64(lam_v[c]).copy() for c in range(3)]
        lam_P = np.zeros(4 * self.nels) if lam_P is None else f64(lam_P).reshape(-1).copy()
        saved = (self.v, self.P, self.g, self.bc, self.body_force)
        try:
            self.v, self.P, self.g = lam_v, lam_P, (0.0, 0.0, 0.0)
            self.bc = tuple((d, np.zeros(len(d))) for d, _ in self.bc)
            self.body_force = [f64(o) for o in objective]
            st = self.cp3_solve(**kw)
        finally:
            self.v, self.P, self.g, self.bc, self.body_force = saved
        return lam_v, lam_P, st

    def cp3_material_gradient(self, v_fwd, lam_v):
        """stokes_material_gradient_3d (DR_adjoint.jl:483-531): per phase
        -λ·R(v=0, P=0, ρ=e_p) and -λ·R(v_fwd, P=0, η=e_p, ρ=0, g=0)."""
        n = self.mat.nphases
        saved = (self.v, self.P, self.g, Material.from_buffer_copy(self.mat))
        dens, visc = [], []
        try:
            self.P = np.zeros(4 * self.nels)
            for which in ("rho", "eta"):
                for p in range(n):
                    m = Material.from_buffer_copy(saved[3])
                    for q in range(n):
                        if which == "rho":
                            m.rho0[q] = 1.0 if q == p else 0.0
                        else:
                            m.eta[q] = 1.0 if q == p else 0.0
                            m.rho0[q] = 0.0
                    self.mat = m
                    if which == "rho":
                        self.v, self.g = [np.zeros(self.nnodes) for _ in range(3)], saved[2]
                    else:
                        self.v, self.g = [f64(x) for x in v_fwd], (0.0, 0.0, 0.0)
                    R = self.cp3_momentum_residual()
                    val = -sum(float(np.dot(lam_v[c], R[c])) for c in range(3))
                    (dens if which == "rho" else visc).append(val)
        finally:
            self.v, self.P, self.g, self.mat = saved
        return dens, visc

    def pressure_scaling(self, gamma_fact, dt, finite_K=False):
        lib().orc_pressure_scaling3(C.byref(self.struct()), gamma_fact, dt, int(finite_K))

    def pressure_residual(self, dt, finite_K=False):
        lib().orc_pressure_residual3(C.byref(self.struct()), dt, int(finite_K))

    def momentum_residual(self, dt, Pnum=None):
        lib().orc_momentum_residual3(C.byref(self.struct()), dt, ptr(f64(Pnum)) if Pnum is not None else ptr(None))

    def augmented_jacobian(self, dt, finite_K=False):
        lib().orc_augmented_jacobian3(C.byref(self.struct()), dt, int(finite_K))

    def update_current_stress(self, dt):
        lib().orc_update_current_stress3(C.byref(self.struct()), dt)
        return self.tau_store

    # CellPressureStokesDR (DR.jl:960-1057): P holds the modes (1, ξ, η, ζ)
    def cp3_pressure_residual(self):
        RP = np.zeros(4 * self.nels)
        lib().orc_cp3_pressure_residual(C.byref(self.struct()), ptr(RP))
        return RP

    def cp3_momentum_residual(self):
        R = [np.zeros(self.nnodes) for _ in range(3)]
        lib().orc_cp3_momentum_residual(C.byref(self.struct()), (PD * 3)(*[ptr(r) for r in R]))
        return R

    def cp3_preconditioner(self):
        dg = [np.zeros(self.nnodes) for _ in range(3)]
        M = np.zeros(4 * self.nels)
        lib().orc_cp3_preconditioner(C.byref(self.struct()), (PD * 3)(*[ptr(d) for d in dg]), ptr(M))
        return dg, M

    def cp3_solve(self, tolerance=1e-5, max_iterations=3000, check_interval=100, velocity_step=0.6,
                  pressure_step=0.2, verbose=False, history_cap=0):
        prm = CP3Params(tolerance=tolerance, max_iterations=max_iterations, check_interval=check_interval,
                        velocity_step=velocity_step, pressure_step=pressure_step, verbose=int(verbose))
        st = CP3Stats()
        hist = np.zeros((max(history_cap, 1), 3))
        nh = I64(0)
        rc = lib().orc_cp3_solve(C.byref(self.struct()), C.byref(prm), C.byref(st),
                                 ptr(hist) if history_cap else ptr(None), int(history_cap), C.byref(nh))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        out = {k: getattr(st, k) for k, _ in CP3Stats._fields_}
        out["history"] = hist[:nh.value].copy()
        return out

    def rotate_stress(self, dt, tau):
        """rotate_stress! (stress_rotation.jl:28-61), 3-D: 6 arrays [nels][27] in assembler order."""
        tau = [f64(t).reshape(self.nels, 27) for t in tau]
        out = [np.zeros((self.nels, 27)) for _ in range(6)]
        lib().orc_rotate_stress3(C.byref(self.struct()), dt, (PD * 6)(*[ptr(t) for t in tau]),
                                 (PD * 6)(*[ptr(o) for o in out]))
        return out

    def solve(self, dt, pressure_factor=50.0, tolerance=1e-6, max_iterations=50_000, iter_max=50_000,
              check_interval=50, max_ph_iterations=1000, rel_drop0=1e-2, CFL_v=0.98, c_fact=0.9,
              finite_K=False, freeze_jacobian=-1, skip_scaling=False, verbose=False):
        p = Params(dt=dt, pressure_factor=pressure_factor, tolerance=tolerance, max_iterations=max_iterations,
                   iter_max=iter_max, check_interval=check_interval, max_ph_iterations=max_ph_iterations,
                   rel_drop0=rel_drop0, CFL_v=CFL_v, c_fact=c_fact, finite_K=int(finite_K),
                   freeze_jacobian=int(freeze_jacobian), skip_scaling=int(skip_scaling), verbose=int(verbose))
        st = Stats()
        rc = lib().orc_stokes_solve3(C.byref(self.struct()), C.byref(p), C.byref(st))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return st.as_dict()


def postprocess2(coords, el2n, vx, vy, tau):
    """compute_strain_rate_stress_postprocess (postprocess.jl:17-100): per
    element the dΩ-weighted averages of (ε̇xx, ε̇yy, ε̇zz = 0, ε̇xy, ε̇II of the
    deviatoric rate, τxx, τyy, τzz = −(τxx + τyy), τxy, τII) from the nodal
    velocity and the point stresses tau = (τxx, τyy, τxy), each [nels][7].
    Returns a dict keyed exx, eyy, ezz, exy, eII, txx, tyy, tzz, txy, tII."""
    _, _, dNdxi, _ = t7_tables()
    invJ, dO, _ = t7_geometry(coords, el2n)
    ux, uy = np.asarray(vx)[el2n], np.asarray(vy)[el2n]           # [ne][7]
    # dN/dx = dN/dξ J⁻¹ per point: [ne][q][a][x]
    dNdx = np.einsum("qak,eqkx->eqax", dNdxi, invJ.reshape(-1, 7, 2, 2))
    gx = np.einsum("eqax,ea->eqx", dNdx, ux)
    gy = np.einsum("eqax,ea->eqx", dNdx, uy)
    exx, eyy = gx[..., 0], gy[..., 1]
    exy = (gx[..., 1] + gy[..., 0]) / 2
    tr = (exx + eyy) / 3
    d0, d1, d2 = exx - tr, eyy - tr, -tr
    txx, tyy, txy = (np.asarray(t).reshape(-1, 7) for t in tau)
    tzz = -(txx + tyy)
    fields = dict(exx=exx, eyy=eyy, ezz=np.zeros_like(exx), exy=exy,
                  eII=np.sqrt((d0 ** 2 + d1 ** 2 + d2 ** 2) / 2 + exy ** 2),
                  txx=txx, tyy=tyy, tzz=tzz, txy=txy, tII=np.sqrt((txx ** 2 + tyy ** 2 + tzz ** 2) / 2 + txy ** 2))
    vol = dO.sum(axis=1)
    return {k: (v * dO).sum(axis=1) / vol for k, v in fields.items()}


def postprocess3(coords, el2n, v, tau):
    """compute_strain_rate_stress_postprocess, 3-D (postprocess.jl:169-212):
    per cell the dΩ-weighted averages of ε̇ (xx, yy, zz, xy, xz, yz), ε̇II of the
    deviatoric rate, the six point stresses tau ([nels][27] each) and τII; the
    invariants carry the eps² floor of second_invariant (rheology.jl:395-399)."""
    _, _, dNdxi, _ = hex27_tables()
    invJ, dO = hex27_geometry(coords, el2n)
    ne = len(el2n)
    dNdx = np.einsum("qak,eqkx->eqax", dNdxi, invJ.reshape(ne, 27, 3, 3))
    G = [np.einsum("eqax,ea->eqx", dNdx, np.asarray(v[c])[el2n]) for c in range(3)]   # G[c][e,q,x]
    e = dict(exx=G[0][..., 0], eyy=G[1][..., 1], ezz=G[2][..., 2], exy=(G[0][..., 1] + G[1][..., 0]) / 2,
             exz=(G[0][..., 2] + G[2][..., 0]) / 2, eyz=(G[1][..., 2] + G[2][..., 1]) / 2)
    tr = (e["exx"] + e["eyy"] + e["ezz"]) / 3
    eps2 = np.finfo(float).eps ** 2
    d = [e["exx"] - tr, e["eyy"] - tr, e["ezz"] - tr]
    e["eII"] = np.sqrt((d[0] ** 2 + d[1] ** 2 + d[2] ** 2) / 2 + e["exy"] ** 2 + e["exz"] ** 2 + e["eyz"] ** 2 + eps2)
    t = [np.asarray(x).reshape(ne, 27) for x in tau]
    for k, x in zip(("txx", "tyy", "tzz", "txy", "txz", "tyz"), t):
        e[k] = x
    e["tII"] = np.sqrt((t[0] ** 2 + t[1] ** 2 + t[2] ** 2) / 2 + t[3] ** 2 + t[4] ** 2 + t[5] ** 2 + eps2)
    vol = dO.sum(axis=1)
    return {k: (x * dO).sum(axis=1) / vol for k, x in e.items()}
