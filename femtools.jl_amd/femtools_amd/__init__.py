"""femtools_amd -- host-side mirror of FEMTools.jl's DR Stokes API over the
C ABI of libfemdr.so (include/femdr.h), for MI355X (gfx950).

Names follow the reference: ``StokesMaterial``, ``DruckerPrager``,
``StokesDR`` with struct-of-arrays fields ``dr.v.x``, ``dr.tau.xy`` ...,
``DirichletBoundaryCondition``, ``solve`` (the reference's ``solve!``, which
mutates ``dr`` in place) and ``update_stokes_current_stress``.

Indices are 0-based on the Python side (node ids, DoFs); the ABI takes
Julia's 1-based ids and the conversion happens here. Integration-point
arrays (Julia ``nq x nels``) appear as numpy ``(nels, nq)``.

There is no CPU fallback: every operation runs in libfemdr.so on the GPU,
and importing the package fails loudly if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

__all__ = [
    "FemError", "DimensionMismatch", "Context", "DeviceArray", "VectorField2D", "SymmetricTensor2D",
    "StokesMaterial", "DruckerPrager", "MixedMesh", "t7_box_mesh", "t7_mesh_from_arrays", "hex27_mesh_from_arrays",
    "DirichletBoundaryCondition", "StokesDR", "solve", "update_stokes_current_stress",
    "assemble_viscosity_weighted_pressure_scaling", "lib_path", "default_context", "solve_batch",
    "hex27_box_mesh", "q9_box_mesh", "VectorField3D", "SymmetricTensor3D", "ThermalMaterial", "ThermalDiffusionDR",
    "solve_thermal", "thermal_residual", "residual", "jacobian_diagnostics", "time_element_kernel", "time_correction_kernel",
    "solve_coupled", "assemble_shear_heating", "rotate_stress", "copyto", "DruckerPragerCap", "set_pressure_source",
    "solve_adjoint", "solve_stokes_adjoint", "adjoint_blocks", "stokes_material_gradient_3d", "compute_strain_rate_stress_postprocess", "write_vtk",
    "write_stokes_vtk", "VtkMesh",
    "commit_stokes_plastic_history", "PlasticHistory", "CellPressureStokesDR",
    "assemble_stokes_momentum_residual_3d", "assemble_stokes_pressure_residual_3d", "stokes_preconditioner_3d",
    "norms", "device_array", "dr_velocity_update", "dp_return_map_batch",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
# FEMDR_LIBRARY selects another build of the same library (e.g. a variant
# compiled with different launch parameters); default: the in-tree build.
_LIB_PATH = os.environ.get("FEMDR_LIBRARY") or os.path.join(_HERE, "libfemdr.so")
MAXPH = 64

FEM_OK, FEM_EARG, FEM_EDIM, FEM_ENOCONV, FEM_ENAN, FEM_EKABOOM, FEM_EPRECOND, FEM_EHIP, FEM_ENOMEM = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)


def lib_path() -> str:
    return _LIB_PATH


if not os.path.exists(_LIB_PATH):
    raise ImportError(f"libfemdr.so is not built ({_LIB_PATH}); run __graft_entry__.build() "
                      "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
_lib = C.CDLL(_LIB_PATH)

D, I32, I64 = C.c_double, C.c_int32, C.c_int64
VP = C.c_void_p


class FemError(RuntimeError):
    """Error raised by the solver (Julia ``ErrorException`` counterpart)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class DimensionMismatch(ValueError):
    pass


class _Material(C.Structure):
    _fields_ = [("nphases", I32), ("ndim", I32), ("eta", D * MAXPH), ("etab", D * MAXPH), ("G", D * MAXPH),
                ("alpha", D * MAXPH), ("rho0", D * MAXPH), ("K", D * MAXPH), ("g", D * 3), ("Tref", D)]


class _DP(C.Structure):
    _fields_ = [("nphases", I32), ("cos_phi", D * MAXPH), ("sin_phi", D * MAXPH), ("sin_psi", D * MAXPH),
                ("C", D * MAXPH), ("eta_reg", D * MAXPH), ("Kb", D * MAXPH)]


class _Cap(C.Structure):
    _fields_ = [("nphases", I32), ("cos_phi", D * MAXPH), ("sin_phi", D * MAXPH), ("sin_psi", D * MAXPH),
                ("C", D * MAXPH), ("pT", D * MAXPH), ("eta_reg", D * MAXPH), ("Kb", D * MAXPH),
                ("C_min", D * MAXPH), ("H_C", D * MAXPH)]


class _MeshInfo(C.Structure):
    _fields_ = [("dim", I32), ("nv_local", I32), ("np_local", I32), ("nq", I32), ("nnodes", I64), ("nels", I64),
                ("nnodesP", I64), ("n_shared", I64), ("n_boundary", I64)]


class _StateOpts(C.Structure):
    _fields_ = [("CFL_v", D), ("CFL_P", D), ("c_fact", D), ("eps", D)]


class _BC(C.Structure):
    _fields_ = [("n", I64), ("host_dofs_1based", C.POINTER(I32)), ("host_vals", C.POINTER(D))]


class _SolveParams(C.Structure):
    _fields_ = [("dt", D), ("pressure_factor", D), ("tolerance", D), ("max_iterations", I64), ("iter_max", I64),
                ("check_interval", I64), ("max_ph_iterations", I64), ("rel_drop0", D), ("finite_K", I32),
                ("freeze_jacobian", I32), ("throw_on_failure", I32), ("use_tau_old", I32), ("skip_scaling", I32),
                ("verbose", I32), ("precision", I32), ("time_kernels", I32), ("measure_lmax", I32),
                ("lmax_power_iterations", I32), ("lmax_power_rtol", D), ("lmax_safety", D)]


class _SolveInputs(C.Structure):
    _fields_ = [("body_force", VP * 3), ("viscosity", VP), ("scaling_viscosity", VP), ("P_old", VP),
                ("P_old_at_points", I32), ("history", C.POINTER(D)), ("history_cap", I64),
                ("history_len", C.POINTER(I64))]


class _SolveStats(C.Structure):
    _fields_ = [("itPH", I64), ("iter", I64), ("err", D), ("err_abs", D), ("err_rel", D), ("err_v", D),
                ("err_P", D), ("converged", I32), ("reached_total_iter", I32), ("lambda_max", D),
                ("lambda_max_gershgorin", D), ("lambda_max_iterations", I64), ("jacobian_assemblies", I64),
                ("seconds", D), ("kernel_ms_element", D), ("kernel_ms_gather", D), ("element_launches", I64),
                ("timed_launches", I64), ("kernel_ms_element_rm", D), ("kernel_ms_gather_rm", D),
                ("timed_launches_rm", I64)]


class _ThermalMaterial(C.Structure):
    _fields_ = [("nphases", I32), ("k", D * MAXPH), ("Cp", D * MAXPH), ("rho0", D * MAXPH), ("alpha", D * MAXPH),
                ("K", D * MAXPH)]


class _ThermalOpts(C.Structure):
    _fields_ = [("CFL", D), ("c_fact", D), ("eps", D)]


class _ThermalParams(C.Structure):
    _fields_ = [("dt", D), ("tolerance", D), ("max_iterations", I64), ("check_interval", I64), ("Tref", D),
                ("throw_on_failure", I32), ("verbose", I32)]


class _ThermalStats(C.Structure):
    _fields_ = [("converged", I32), ("iterations", I64), ("residual", D), ("seconds", D)]


class _CellParams(C.Structure):
    _fields_ = [("tolerance", D), ("max_iterations", I64), ("check_interval", I64), ("velocity_step", D),
                ("pressure_step", D), ("verbose", I32), ("throw_on_failure", I32)]


class _AdjointParams(C.Structure):
    _fields_ = [("dt", D), ("tolerance", D), ("max_iterations", I64), ("check_interval", I64), ("iter_max", I64),
                ("max_ph_iterations", I64), ("rel_drop", D), ("measure_lmax", I32), ("throw_on_failure", I32),
                ("verbose", I32)]


class _AdjointStats(C.Structure):
    _fields_ = [("converged", I32), ("itPH", I64), ("iter", I64), ("lmax_iterations", I64), ("err", D),
                ("err_v", D), ("err_P", D), ("lmax", D), ("lmax_gershgorin", D), ("seconds", D)]


class _CellStats(C.Structure):
    _fields_ = [("converged", I32), ("reached_total_iter", I32), ("iterations", I64), ("residual", D),
                ("err_v", D), ("err_P", D), ("seconds", D)]


class _CoupledParams(C.Structure):
    _fields_ = [("Tref", D), ("shear_heating", I32)]


class _CoupledStats(C.Structure):
    _fields_ = [("err_T", D), ("thermal_iterations", I64), ("thermal_converged", I32)]


def _sig(name, restype, *argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_P = C.POINTER
_sig("fem_ctx_create", C.c_int, C.c_int, _P(VP))
_sig("fem_ctx_destroy", C.c_int, VP)
_sig("fem_last_error", C.c_int, VP, C.c_char_p, C.c_size_t)
_sig("fem_sync", C.c_int, VP)
_sig("fem_buf_alloc", C.c_int, VP, C.c_size_t, _P(VP))
_sig("fem_buf_free", C.c_int, VP, VP)
_sig("fem_copy_h2d", C.c_int, VP, VP, VP, C.c_size_t)
_sig("fem_copy_d2h", C.c_int, VP, VP, VP, C.c_size_t)
_sig("fem_copy_d2d", C.c_int, VP, VP, VP, C.c_size_t)
_sig("fem_fill_f64", C.c_int, VP, VP, D, I64)
_sig("fem_fill_f32", C.c_int, VP, VP, C.c_float, I64)
_sig("fem_fill_i32", C.c_int, VP, VP, I32, I64)
_sig("fem_mesh_t7_create", C.c_int, VP, I32, I32, D, D, D, D, _P(VP))
_sig("fem_mesh_t7_upload", C.c_int, VP, _P(D), I64, _P(I32), I64, _P(VP))
_sig("fem_mesh_hex27_create", C.c_int, VP, I32, I32, I32, D, D, D, D, D, D, _P(VP))
_sig("fem_mesh_hex27_upload", C.c_int, VP, _P(D), I64, _P(I32), I64, _P(VP))
_sig("fem_mesh_q9_create", C.c_int, VP, I32, I32, D, D, D, D, _P(VP))
_sig("fem_mesh_get_info", C.c_int, VP, _P(_MeshInfo))
_sig("fem_mesh_get_coords", C.c_int, VP, _P(D))
_sig("fem_mesh_get_el2n", C.c_int, VP, _P(I32))
_sig("fem_mesh_get_boundary_nodes", C.c_int, VP, _P(I32))
_sig("fem_mesh_destroy", C.c_int, VP)
_sig("fem_stokes_create", C.c_int, VP, VP, _P(_Material), _P(_StateOpts), _P(VP))
_sig("fem_stokes_destroy", C.c_int, VP)
_sig("fem_stokes_field", C.c_int, VP, I32, _P(VP), _P(I64), _P(I32))
_sig("fem_stokes_set_phases", C.c_int, VP, I32, I32, _P(I32), I64)
_sig("fem_stokes_solve", C.c_int, VP, _P(_BC), _P(_DP), _P(_SolveParams), _P(_SolveStats))
_sig("fem_stokes_solve_batch", C.c_int, I32, _P(VP), _P(_P(_BC)), _P(_P(_DP)), _P(_SolveParams), _P(_SolveStats),
     _P(I32), I32)
_sig("fem_stokes_pressure_scaling", C.c_int, VP, D, D, I32)
_sig("fem_stokes_update_current_stress", C.c_int, VP, _P(_DP), D)
_sig("fem_stokes_residual", C.c_int, VP, _P(_DP), D, I32, I32)
_sig("fem_stokes_jacobian_diagnostics", C.c_int, VP, _P(_DP), D, I32)
_sig("fem_stokes_time_element_kernel", C.c_int, VP, _P(_DP), D, I32, _P(D))
_sig("fem_stokes_time_correction_kernel", C.c_int, VP, _P(_DP), D, I32, _P(D))
_sig("fem_ctx_set_tuning", C.c_int, VP, VP)
_sig("fem_ctx_get_tuning", C.c_int, VP, VP)
_sig("fem_thermal_create", C.c_int, VP, VP, _P(_ThermalMaterial), _P(_ThermalOpts), _P(VP))
_sig("fem_thermal_destroy", C.c_int, VP)
_sig("fem_thermal_field", C.c_int, VP, I32, _P(VP), _P(I64), _P(I32))
_sig("fem_thermal_solve", C.c_int, VP, _P(_BC), _P(_ThermalParams), _P(_ThermalStats))
_sig("fem_thermal_residual", C.c_int, VP, D, D, I32)
_sig("fem_stokes_shear_heating", C.c_int, VP, _P(_DP), D, VP)
_sig("fem_stokes_rotate_stress", C.c_int, VP, D)
_sig("fem_stokes_solve_cap", C.c_int, VP, _P(_BC), _P(_Cap), _P(_SolveParams), _P(_SolveStats))
_sig("fem_cell_stokes_preconditioner", C.c_int, VP)
_sig("fem_cell_stokes_pressure_residual", C.c_int, VP)
_sig("fem_cell_stokes_momentum_residual", C.c_int, VP)
_sig("fem_cell_stokes_solve", C.c_int, VP, _P(_BC), _P(_CellParams), _P(_CellStats), _P(D), I64, _P(I64))
_sig("fem_stokes_residual_cap", C.c_int, VP, _P(_Cap), D, I32, I32)
_sig("fem_stokes_jacobian_diagnostics_cap", C.c_int, VP, _P(_Cap), D, I32)
_sig("fem_stokes_update_current_stress_cap", C.c_int, VP, _P(_Cap), D)
_sig("fem_stokes_commit_plastic_history", C.c_int, VP, _P(_Cap), D)
_sig("fem_stokes_set_solve_inputs", C.c_int, VP, _P(_SolveInputs))
_sig("fem_stokes_postprocess", C.c_int, VP, VP)
_sig("fem_norms", C.c_int, VP, VP, VP, I64, VP, VP, I64, _P(D))
_sig("fem_norms_noalloc", C.c_int, VP, VP, VP, I64, VP, VP, I64, VP)
_sig("fem_dr_velocity_update", C.c_int, VP, I64, VP, VP, VP, VP, D, D)


class _DPPoints(C.Structure):
    _fields_ = [(k, VP) for k in ("exx", "eyy", "exy", "txx_old", "tyy_old", "txy_old", "P", "eta", "G", "C", "phi",
                                  "psi", "eta_reg", "Kb", "txx", "tyy", "txy", "yields")]


_sig("fem_dp_return_map_batch", C.c_int, VP, I64, D, _P(_DPPoints))
_sig("fem_cnn_create", C.c_int, VP, I32, I32, I32, _P(VP))
_sig("fem_cnn_destroy", C.c_int, VP)
_sig("fem_cnn_parameter_count", C.c_int, _P(I64), _P(I64))
_sig("fem_cnn_set_weights", C.c_int, VP, VP, VP)
_sig("fem_cnn_get_weights", C.c_int, VP, VP, VP)
_sig("fem_cnn_get_grads", C.c_int, VP, VP, VP)
_sig("fem_cnn_forward", C.c_int, VP, VP, I32, VP)
_sig("fem_cnn_train_step", C.c_int, VP, VP, VP, I32, D, I32, _P(D))
_sig("fem_cnn_gemm_bf16", C.c_int, VP, I64, I64, I64, VP, VP, VP)
_sig("fem_cell_stokes_solve_adjoint", C.c_int, VP, _P(_BC), _P(VP), _P(VP), VP, _P(_CellParams), _P(_CellStats),
     _P(D), I64, _P(I64))
_sig("fem_cell_stokes_material_gradient", C.c_int, VP, _P(VP), _P(VP), _P(D), _P(D))
_sig("fem_stokes_solve_adjoint", C.c_int, VP, _P(_BC), _P(_DP), VP, VP, VP, VP, VP, _P(_AdjointParams),
     _P(_AdjointStats))
_sig("fem_stokes_adjoint_blocks", C.c_int, VP, _P(_DP), D, _P(D), _P(D), _P(D), _P(D), _P(D))
_sig("fem_stokes_set_pressure_source", C.c_int, VP, VP, D, I32)
_sig("fem_coupled_solve", C.c_int, VP, VP, _P(_BC), _P(_BC), _P(_DP), _P(_SolveParams), _P(_CoupledParams),
     _P(_SolveStats), _P(_CoupledStats))

EXPORTED = [
    "fem_ctx_create", "fem_ctx_destroy", "fem_last_error", "fem_sync", "fem_buf_alloc", "fem_buf_free",
    "fem_copy_h2d", "fem_copy_d2h", "fem_copy_d2d", "fem_fill_f64", "fem_fill_f32", "fem_fill_i32",
    "fem_mesh_t7_create", "fem_mesh_t7_upload", "fem_mesh_hex27_create", "fem_mesh_hex27_upload", "fem_mesh_q9_create", "fem_mesh_get_info", "fem_mesh_get_coords", "fem_mesh_get_el2n",
    "fem_mesh_get_boundary_nodes", "fem_mesh_destroy", "fem_stokes_create", "fem_stokes_destroy",
    "fem_stokes_field", "fem_stokes_set_phases", "fem_stokes_solve", "fem_stokes_solve_batch",
    "fem_stokes_pressure_scaling",
    "fem_stokes_update_current_stress", "fem_stokes_residual", "fem_stokes_jacobian_diagnostics",
    "fem_stokes_solve_adjoint", "fem_stokes_adjoint_blocks",
    "fem_stokes_time_element_kernel", "fem_stokes_time_correction_kernel", "fem_ctx_set_tuning", "fem_ctx_get_tuning", "fem_thermal_create", "fem_thermal_destroy", "fem_thermal_field",
    "fem_thermal_solve", "fem_thermal_residual", "fem_stokes_shear_heating", "fem_coupled_solve",
    "fem_stokes_rotate_stress", "fem_stokes_solve_cap", "fem_stokes_residual_cap",
    "fem_stokes_jacobian_diagnostics_cap", "fem_stokes_update_current_stress_cap",
    "fem_stokes_commit_plastic_history", "fem_cell_stokes_preconditioner", "fem_cell_stokes_pressure_residual",
    "fem_cell_stokes_momentum_residual", "fem_cell_stokes_solve", "fem_stokes_set_solve_inputs",
    "fem_stokes_set_pressure_source", "fem_cell_stokes_solve_adjoint", "fem_cell_stokes_material_gradient",
    "fem_stokes_postprocess", "fem_norms", "fem_cnn_create", "fem_cnn_destroy", "fem_cnn_parameter_count",
    "fem_cnn_set_weights", "fem_cnn_get_weights", "fem_cnn_get_grads", "fem_cnn_forward", "fem_cnn_train_step",
    "fem_cnn_gemm_bf16", "fem_norms_noalloc", "fem_dr_velocity_update", "fem_dp_return_map_batch",
]


class _Tuning(C.Structure):
    _fields_ = [("tiles", I32), ("tile_rows", I32), ("graphs", I32), ("fold_interval", I32)]


class Context:
    """One device + one HIP stream (fem_ctx)."""

    def __init__(self, device: int = 0):
        h = VP()
        rc = _lib.fem_ctx_create(device, C.byref(h))
        if rc != FEM_OK:
            raise FemError(rc, f"fem_ctx_create(device={device}) failed with code {rc} (no HIP device?)")
        self.handle = h
        self.device = device

    def error(self) -> str:
        buf = C.create_string_buffer(2048)
        _lib.fem_last_error(self.handle, buf, len(buf))
        return buf.value.decode(errors="replace")

    def check(self, rc: int):
        if rc == FEM_OK:
            return
        msg = self.error()
        if rc == FEM_EARG:
            raise ValueError(msg)
        if rc == FEM_EDIM:
            raise DimensionMismatch(msg)
        raise FemError(rc, msg)

    def sync(self):
        self.check(_lib.fem_sync(self.handle))

    @property
    def tuning(self) -> dict:
        """The context's launch choices (fem_tuning): tiles, tile_rows, graphs,
        fold_interval."""
        t = _Tuning()
        self.check(_lib.fem_ctx_get_tuning(self.handle, C.byref(t)))
        return {k: int(getattr(t, k)) for k, _ in _Tuning._fields_}

    def set_tuning(self, **kw):
        """Changes some of the launch choices, e.g. set_tuning(tile_rows=16,
        graphs=0); unknown names raise ValueError."""
        cur = self.tuning
        for k, v in kw.items():
            if k not in cur:
                raise ValueError(f"unknown tuning parameter {k!r} (known: {sorted(cur)})")
            cur[k] = int(v)
        self.check(_lib.fem_ctx_set_tuning(self.handle, C.byref(_Tuning(**cur))))


_default_ctx = None


def default_context() -> Context:
    """The process's context on device FEMDR_DEVICE (default 0). If
    FEMDR_TUNING names a JSON file written by femtools_amd.autotune, its
    "tuning" object is applied."""
    global _default_ctx
    if _default_ctx is None:
        ctx = Context(int(os.environ.get("FEMDR_DEVICE", "0")))
        path = os.environ.get("FEMDR_TUNING")
        if path:
            import json
            with open(path) as f:
                ctx.set_tuning(**json.load(f)["tuning"])
        _default_ctx = ctx
    return _default_ctx


class DeviceArray:
    """A device array owned by a StokesDR field (or by fem_buf_alloc)."""

    def __init__(self, ctx: Context, ptr: int, shape, dtype=np.float64, owner=None):
        self.ctx = ctx
        self.ptr = int(ptr)
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self._owner = owner

    @property
    def size(self):
        return int(np.prod(self.shape))

    def __len__(self):
        return self.shape[0]

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.size:
            self.ctx.check(_lib.fem_copy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    def copyto(self, src):
        """copyto!(dest, src); a DeviceArray source is copied on the device."""
        if isinstance(src, DeviceArray):
            if src.size != self.size or src.dtype != self.dtype:
                raise DimensionMismatch(f"copyto! from {src} into {self}")
            if src.ctx is not self.ctx:
                raise ValueError("copyto! between device arrays of different contexts")
            if self.size:
                self.ctx.check(_lib.fem_copy_d2d(self.ctx.handle, self.ptr, src.ptr, self.nbytes))
            return
        a = np.ascontiguousarray(np.asarray(src, dtype=self.dtype).reshape(self.shape))
        if self.size:
            self.ctx.check(_lib.fem_copy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes))

    def fill(self, v):
        """fill!(A, v)"""
        if self.dtype == np.float64:
            self.ctx.check(_lib.fem_fill_f64(self.ctx.handle, self.ptr, float(v), self.size))
        elif self.dtype == np.float32:
            self.ctx.check(_lib.fem_fill_f32(self.ctx.handle, self.ptr, float(v), self.size))
        else:
            self.ctx.check(_lib.fem_fill_i32(self.ctx.handle, self.ptr, int(v), self.size))

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


@dataclass
class VectorField2D:
    """VectorField2D (src/vector_fields/vector_fields.jl:24-32)."""
    x: DeviceArray
    y: DeviceArray

    def __iter__(self):
        return iter((self.x, self.y))

    def __getitem__(self, i):
        return tuple(self)[i]

    def copyto(self, src):
        """copyto!(dst, src), component by component on the device
        (symmetric_tensor.jl:133-150 advances the stress history this way)."""
        mine, theirs = tuple(self), tuple(src)
        if len(mine) != len(theirs):
            raise DimensionMismatch("copyto! between fields of different dimension")
        for d, c in zip(mine, theirs):
            d.copyto(c)


@dataclass
class VectorField3D:
    """VectorField3D (src/vector_fields/vector_fields.jl:34-43)."""
    x: DeviceArray
    y: DeviceArray
    z: DeviceArray

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __getitem__(self, i):
        return tuple(self)[i]

    def copyto(self, src):
        """copyto!(dst, src), component by component on the device
        (symmetric_tensor.jl:133-150 advances the stress history this way)."""
        mine, theirs = tuple(self), tuple(src)
        if len(mine) != len(theirs):
            raise DimensionMismatch("copyto! between fields of different dimension")
        for d, c in zip(mine, theirs):
            d.copyto(c)


@dataclass
class SymmetricTensor3D:
    """SymmetricTensor3D (src/tensor_fields/symmetric_tensor.jl:53-63), components
    xx, yy, zz, xy, xz, yz; the invariant slot is not stored."""
    xx: DeviceArray
    yy: DeviceArray
    zz: DeviceArray
    xy: DeviceArray
    xz: DeviceArray
    yz: DeviceArray

    def __iter__(self):
        return iter((self.xx, self.yy, self.zz, self.xy, self.xz, self.yz))

    def __getitem__(self, i):
        return tuple(self)[i]

    def copyto(self, src):
        """copyto!(dst, src), component by component on the device
        (symmetric_tensor.jl:133-150 advances the stress history this way)."""
        mine, theirs = tuple(self), tuple(src)
        if len(mine) != len(theirs):
            raise DimensionMismatch("copyto! between fields of different dimension")
        for d, c in zip(mine, theirs):
            d.copyto(c)


@dataclass
class PlasticHistory:
    """CapPlasticHistory (stokes_types.jl:83-96): accepted deviatoric (γ) and
    volumetric (θ) plastic strain at the integration points."""
    gamma: DeviceArray
    theta: DeviceArray


@dataclass
class SymmetricTensor2D:
    """SymmetricTensor2D (src/tensor_fields/symmetric_tensor.jl:44-51); the
    invariant slot is not stored (stokes_types.jl:71-79)."""
    xx: DeviceArray
    yy: DeviceArray
    xy: DeviceArray

    def __iter__(self):
        return iter((self.xx, self.yy, self.xy))

    def __getitem__(self, i):
        return tuple(self)[i]

    def copyto(self, src):
        """copyto!(dst, src), component by component on the device
        (symmetric_tensor.jl:133-150 advances the stress history this way)."""
        mine, theirs = tuple(self), tuple(src)
        if len(mine) != len(theirs):
            raise DimensionMismatch("copyto! between fields of different dimension")
        for d, c in zip(mine, theirs):
            d.copyto(c)


def _tup(v, n=None):
    if isinstance(v, (tuple, list, np.ndarray)):
        return tuple(float(x) for x in v)
    return (float(v),) * (n or 1)


class StokesMaterial:
    """StokesMaterial(; η=(1.0,), ηb=(1.0,), G=(Inf,), α=(0.0,), ρ0=(1.0,),
    K=(Inf,), g=(0.0, 0.0), Tref=0.0) (src/stokes/types/stokes_types.jl:19-60).
    The first tuple given fixes the phase count; scalars apply to every phase."""

    def __init__(self, eta=None, etab=None, G=None, alpha=None, rho0=None, K=None, g=None, Tref=None):
        given = [v for v in (eta, etab, G, alpha, rho0, K) if isinstance(v, (tuple, list, np.ndarray))]
        n = len(given[0]) if given else 1
        if n < 1:
            raise ValueError("material properties cannot be empty")
        if n > MAXPH:
            raise ValueError(f"at most {MAXPH} phases are supported")

        def prop(v, default):
            t = _tup(default if v is None else v, n)
            if len(t) != n:
                raise DimensionMismatch("Stokes material property tuples must have the same length")
            return t

        self.eta = prop(eta, 1.0)
        self.etab = prop(etab, 1.0)
        self.G = prop(G, math.inf)
        self.alpha = prop(alpha, 0.0)
        self.rho0 = prop(rho0, 1.0)
        self.K = prop(K, math.inf)
        self.g = tuple(float(x) for x in (g if g is not None else (0.0, 0.0)))
        if len(self.g) not in (2, 3):
            raise ValueError(f"gravity must have 2 or 3 components, got {len(self.g)}")
        self.Tref = float(Tref) if Tref is not None else 0.0
        self.nphases = n

    def _c(self):
        m = _Material()
        m.nphases = self.nphases
        m.ndim = len(self.g)
        for name in ("eta", "etab", "G", "alpha", "rho0", "K"):
            arr = getattr(m, name)
            for i, x in enumerate(getattr(self, name)):
                arr[i] = x
        for i, x in enumerate(self.g):
            m.g[i] = x
        m.Tref = self.Tref
        return m


class DruckerPrager:
    """DruckerPrager(ϕ, Ψ, C, η_reg, Kb) (stokes_types.jl:388-423): friction and
    dilation angles in radians, stored as cos ϕ, sin ϕ, sin Ψ."""

    def __init__(self, phi, psi, C_, eta_reg, Kb):
        phi, psi, C_, eta_reg, Kb = (_tup(x) for x in (phi, psi, C_, eta_reg, Kb))
        n = len(phi)
        if not all(len(t) == n for t in (psi, C_, eta_reg, Kb)):
            raise DimensionMismatch("DruckerPrager parameter tuples must have the same length")
        self.cos_phi = tuple(math.cos(x) for x in phi)
        self.sin_phi = tuple(math.sin(x) for x in phi)
        self.sin_psi = tuple(math.sin(x) for x in psi)
        self.C, self.eta_reg, self.Kb = C_, eta_reg, Kb
        self.nphases = n

    def _c(self):
        d = _DP()
        d.nphases = self.nphases
        for name in ("cos_phi", "sin_phi", "sin_psi", "C", "eta_reg", "Kb"):
            arr = getattr(d, name)
            for i, x in enumerate(getattr(self, name)):
                arr[i] = x
        return d


class DruckerPragerCap:
    """DruckerPragerCap(ϕ, Ψ, C, pT, η_reg, Kb; C_min=C, H_C=0)
    (stokes_types.jl:449-544): Drucker-Prager closed on the tensile side by
    the circular cap of Popov, Berlie & Kaus (2025); angles in radians, pT <= 0
    (compression-positive pressure), optional linear cohesion softening with
    the accepted plastic strain. Validated per phase as the reference does."""

    def __init__(self, phi, psi, C_, pT, eta_reg, Kb, C_min=None, H_C=None):
        phi, psi, C_, pT, eta_reg, Kb = (_tup(x) for x in (phi, psi, C_, pT, eta_reg, Kb))
        n = len(phi)
        C_min = C_ if C_min is None else _tup(C_min)
        H_C = (0.0,) * n if H_C is None else _tup(H_C)
        if not all(len(t) == n for t in (psi, C_, pT, eta_reg, Kb, C_min, H_C)):
            raise DimensionMismatch("DruckerPragerCap parameter tuples must have the same length")
        for i in range(n):
            vals = (phi[i], psi[i], C_[i], pT[i], eta_reg[i], Kb[i], C_min[i], H_C[i])
            if not all(math.isfinite(x) for x in vals):
                raise ValueError(f"phase {i + 1}: tensile-cap parameters must all be finite")
            if pT[i] > 0:
                raise ValueError(f"phase {i + 1}: tensile strength pT must be ≤ 0 with compression-positive "
                                 f"pressure, got {pT[i]}")
            r = C_[i] * math.cos(phi[i]) + math.sin(phi[i]) * pT[i]
            if not r > 0:
                raise ValueError(f"phase {i + 1}: cap radius is not positive; need C·cos(ϕ) + sin(ϕ)·pT > 0, "
                                 f"got {r}. Cohesion must exceed sin(ϕ)·|pT|.")
            if not Kb[i] > 0:
                raise ValueError(f"phase {i + 1}: dilatant plasticity requires a positive finite bulk modulus, "
                                 f"got {Kb[i]}")
            if not eta_reg[i] >= 0:
                raise ValueError(f"phase {i + 1}: η_reg must be ≥ 0, got {eta_reg[i]}")
            if not C_min[i] >= 0:
                raise ValueError(f"phase {i + 1}: C_min must be ≥ 0, got {C_min[i]}")
            if not C_min[i] <= C_[i]:
                raise ValueError(f"phase {i + 1}: C_min must be ≤ C, got C_min={C_min[i]}, C={C_[i]}")
            if not H_C[i] <= 0:
                raise ValueError(f"phase {i + 1}: H_C must be ≤ 0 for softening, got {H_C[i]}")
            if not C_min[i] * math.cos(phi[i]) + math.sin(phi[i]) * pT[i] > 0:
                raise ValueError(f"phase {i + 1}: softened cap radius must remain positive; need "
                                 "C_min·cos(ϕ) + sin(ϕ)·pT > 0")
        self.cos_phi = tuple(math.cos(x) for x in phi)
        self.sin_phi = tuple(math.sin(x) for x in phi)
        self.sin_psi = tuple(math.sin(x) for x in psi)
        self.C, self.pT, self.eta_reg, self.Kb, self.C_min, self.H_C = C_, pT, eta_reg, Kb, C_min, H_C
        self.nphases = n

    def _c(self):
        d = _Cap()
        d.nphases = self.nphases
        for name in ("cos_phi", "sin_phi", "sin_psi", "C", "pT", "eta_reg", "Kb", "C_min", "H_C"):
            arr = getattr(d, name)
            for i, x in enumerate(getattr(self, name)):
                arr[i] = x
        return d


class MixedMesh:
    """MixedMesh of T7 (2-D) or Hex27 (3-D) velocity / P1-disc pressure
    elements (src/mesh/mixed_mesh.jl:173-211, 358-417) resident on the device;
    also the plain Q9 Mesh of thermal states (no pressure space)."""

    def __init__(self, ctx: Context, handle, box=None):
        self.ctx = ctx
        self.handle = handle
        self.nx = self.ny = self.nz = 0   # cells per side of generated box meshes (0: uploaded mesh)
        info = _MeshInfo()
        ctx.check(_lib.fem_mesh_get_info(handle, C.byref(info)))
        self.nnodes, self.nels, self.nnodesP = info.nnodes, info.nels, info.nnodesP
        self.dim, self.nv_local, self.np_local = info.dim, info.nv_local, info.np_local
        self.n_shared = info.n_shared
        self.nq = info.nq
        self.box = box
        coords = np.empty((self.nnodes, self.dim))
        ctx.check(_lib.fem_mesh_get_coords(handle, coords.ctypes.data_as(_P(D))))
        el2n = np.empty((self.nels, self.nv_local), np.int32)
        ctx.check(_lib.fem_mesh_get_el2n(handle, el2n.ctypes.data_as(_P(I32))))
        self.coords = coords
        self.el2n = el2n - 1
        self.el2nP = self.el2n[:, :self.np_local]
        self.DoFsP = np.arange(self.nnodesP, dtype=np.int32).reshape(self.nels, self.np_local)
        nb = info.n_boundary
        b = np.empty(nb, np.int32)
        if nb:
            ctx.check(_lib.fem_mesh_get_boundary_nodes(handle, b.ctypes.data_as(_P(I32))))
        self.boundary_nodes = b - 1

    def __del__(self):
        try:
            _lib.fem_mesh_destroy(self.handle)
        except Exception:
            pass


def t7_box_mesh(nx: int, ny: int, domain=((0.0, 1.0), (0.0, 1.0)), ctx: Context | None = None) -> MixedMesh:
    """MixedMesh(Mesh(backend, (x0..x1)×(y0..y1), T7, (nx, ny)), P1)."""
    ctx = ctx or default_context()
    (x0, x1), (y0, y1) = domain
    h = VP()
    ctx.check(_lib.fem_mesh_t7_create(ctx.handle, int(nx), int(ny), x0, x1, y0, y1, C.byref(h)))
    m = MixedMesh(ctx, h, box=(x0, x1, y0, y1))
    m.nx, m.ny = int(nx), int(ny)
    return m


def hex27_box_mesh(nx: int, ny: int, nz: int, domain=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)),
                   ctx: Context | None = None) -> MixedMesh:
    """MixedMesh(Mesh(backend, (x0..x1)×(y0..y1)×(z0..z1), Hex27, (nx, ny, nz)), P1)."""
    ctx = ctx or default_context()
    (x0, x1), (y0, y1), (z0, z1) = domain
    h = VP()
    ctx.check(_lib.fem_mesh_hex27_create(ctx.handle, int(nx), int(ny), int(nz), x0, x1, y0, y1, z0, z1,
                                         C.byref(h)))
    m = MixedMesh(ctx, h, box=(x0, x1, y0, y1, z0, z1))
    m.nx, m.ny, m.nz = int(nx), int(ny), int(nz)
    return m


def q9_box_mesh(nx: int, ny: int, domain=((0.0, 1.0), (0.0, 1.0)), ctx: Context | None = None) -> MixedMesh:
    """Mesh(backend, (x0..x1)×(y0..y1), ReferenceElement(QuadraticElement{2,9}), (nx, ny))
    (mesh.jl:340-398, connectivity.jl:193-227): a structured Q9 mesh for
    ThermalDiffusionDR states. It has no pressure space (nnodesP = 0) and
    StokesDR rejects it."""
    ctx = ctx or default_context()
    (x0, x1), (y0, y1) = domain
    h = VP()
    ctx.check(_lib.fem_mesh_q9_create(ctx.handle, int(nx), int(ny), x0, x1, y0, y1, C.byref(h)))
    m = MixedMesh(ctx, h, box=(x0, x1, y0, y1))
    m.nx, m.ny = int(nx), int(ny)
    return m


def t7_mesh_from_arrays(coords, el2n, ctx: Context | None = None) -> MixedMesh:
    """MixedMesh from unstructured T7 arrays (0-based el2n, local node 7 = centroid)."""
    ctx = ctx or default_context()
    coords = np.ascontiguousarray(coords, np.float64)
    e1 = np.ascontiguousarray(np.asarray(el2n, np.int32) + 1)
    h = VP()
    ctx.check(_lib.fem_mesh_t7_upload(ctx.handle, coords.ctypes.data_as(_P(D)), len(coords),
                                      e1.ctypes.data_as(_P(I32)), len(e1), C.byref(h)))
    return MixedMesh(ctx, h)


def hex27_mesh_from_arrays(coords, el2n, ctx: Context | None = None) -> MixedMesh:
    """MixedMesh(Mesh(backend, coords, el2n, Hex27), P1) (mesh.jl:147-188) from
    host arrays (0-based el2n in the local order of connectivity.jl:273-322).
    Elements must be axis-aligned boxes with their nodes at the corner, edge,
    face and centre points; their sizes and the node and element order are free.
    Raises ValueError naming the first element that is not."""
    ctx = ctx or default_context()
    coords = np.ascontiguousarray(coords, np.float64)
    e1 = np.ascontiguousarray(np.asarray(el2n, np.int32) + 1)
    if coords.ndim != 2 or coords.shape[1] != 3 or e1.ndim != 2 or e1.shape[1] != 27:
        raise ValueError("hex27_mesh_from_arrays takes coords [nnodes, 3] and el2n [nels, 27]")
    h = VP()
    ctx.check(_lib.fem_mesh_hex27_upload(ctx.handle, coords.ctypes.data_as(_P(D)), len(coords),
                                         e1.ctypes.data_as(_P(I32)), len(e1), C.byref(h)))
    return MixedMesh(ctx, h)


class DirichletBoundaryCondition:
    """DirichletBoundaryCondition(DoFs, vals) (boundary_conditions.jl:30-42), 0-based DoFs."""

    def __init__(self, dofs, vals):
        self.DoFs = np.ascontiguousarray(dofs, np.int32)
        self.vals = np.ascontiguousarray(vals, np.float64)
        if self.DoFs.shape != self.vals.shape:
            raise DimensionMismatch("DoFs and vals must have equal lengths")
        self.zero_vals = np.zeros_like(self.vals)

    def _c(self, keep):
        d1 = np.ascontiguousarray(self.DoFs + 1, np.int32)
        keep.append(d1)
        b = _BC()
        b.n = len(d1)
        b.host_dofs_1based = d1.ctypes.data_as(_P(I32))
        b.host_vals = self.vals.ctypes.data_as(_P(D))
        return b


_FIELDS = ["v.x", "v.y", "dvdtau.x", "dvdtau.y", "Rv.x", "Rv.y", "Rv0.x", "Rv0.y", "dRvdv.x", "dRvdv.y",
           "PC_v.x", "PC_v.y", "tau.xx", "tau.yy", "tau.xy", "tau_old.xx", "tau_old.yy", "tau_old.xy",
           "P", "P0", "gammaP", "T", "T0", "Q", "Pf", "RP", "RP0", "M_P", "Pnum",
           "v.z", "dvdtau.z", "Rv.z", "Rv0.z", "dRvdv.z", "PC_v.z", "tau.zz", "tau.xz", "tau.yz",
           "tau_old.zz", "tau_old.xz", "tau_old.yz", "plastic_history.gamma", "plastic_history.theta", "P_ip"]


class StokesDR:
    """StokesDR(mesh, material; CFL_v=0.98, CFL_P=0.98, c_fact=0.9, ϵ=1e-6)
    (stokes_types.jl:197-298, 359-369). Fields are device arrays:
    ``v``, ``dvdtau`` (∂v∂τ), ``Rv``, ``Rv0``, ``dRvdv`` (∂Rv∂v), ``PC_v``
    (VectorField2D / VectorField3D); ``tau``, ``tau_old`` (SymmetricTensor2D /
    SymmetricTensor3D at integration points, shape (nels, nq)); ``P``, ``P0``,
    ``gammaP``, ``T``, ``T0``, ``Q``, ``Pf``, ``RP``, ``RP0``, ``M_P``, ``Pnum``."""

    def __init__(self, mesh: MixedMesh, material: StokesMaterial, CFL_v=0.98, CFL_P=0.98, c_fact=0.9, eps=1e-6):
        self.ctx = mesh.ctx
        self.mesh = mesh
        self.material = material
        if len(material.g) != mesh.dim:
            raise DimensionMismatch(f"material gravity must have {mesh.dim} components for a {mesh.dim}-D mesh")
        self._mat = material._c()
        opts = _StateOpts(CFL_v, CFL_P, c_fact, eps)
        h = VP()
        self.ctx.check(_lib.fem_stokes_create(self.ctx.handle, mesh.handle, C.byref(self._mat), C.byref(opts),
                                              C.byref(h)))
        self.handle = h
        self.CFL_v, self.CFL_P, self.c_fact, self.eps = CFL_v, CFL_P, c_fact, eps
        arrs = []
        for k, name in enumerate(_FIELDS):
            p, n, dt = VP(), I64(), I32()
            self.ctx.check(_lib.fem_stokes_field(h, k, C.byref(p), C.byref(n), C.byref(dt)))
            ip = name.startswith(("tau", "plastic_history", "P_ip"))
            shape = (mesh.nels, mesh.nq) if ip and n.value else (n.value,)
            arrs.append(DeviceArray(self.ctx, p.value, shape, np.float64, owner=self))
        a = dict(zip(_FIELDS, arrs))
        if mesh.dim == 3:
            for name in ("v", "dvdtau", "Rv", "Rv0", "dRvdv", "PC_v"):
                setattr(self, name, VectorField3D(a[name + ".x"], a[name + ".y"], a[name + ".z"]))
            for name in ("tau", "tau_old"):
                setattr(self, name, SymmetricTensor3D(*(a[f"{name}.{c}"] for c in ("xx", "yy", "zz", "xy", "xz", "yz"))))
        else:
            for name in ("v", "dvdtau", "Rv", "Rv0", "dRvdv", "PC_v"):
                setattr(self, name, VectorField2D(a[name + ".x"], a[name + ".y"]))
            for name in ("tau", "tau_old"):
                setattr(self, name, SymmetricTensor2D(*(a[f"{name}.{c}"] for c in ("xx", "yy", "xy"))))
        for name in ("P", "P0", "gammaP", "T", "T0", "Q", "Pf", "RP", "RP0", "M_P", "Pnum"):
            setattr(self, name, a[name])
        # the tensile cap's integration-point history (zero: no softening) and
        # the pressure of the last stress update (cap-corrected); 2-D states
        self.plastic_history = PlasticHistory(a["plastic_history.gamma"], a["plastic_history.theta"])
        self.P_ip = a["P_ip"]

    def __del__(self):
        try:
            _lib.fem_stokes_destroy(self.handle)
        except Exception:
            pass

    def set_phases(self, which: str, phases):
        """phases_v / phases_P: (nels,) cell phases, (nels, NV|NP) element-local
        phases, or a per-node (per-DoF) vector. 1-based ids as in the reference."""
        ph = np.asarray(phases)
        w = 0 if which == "v" else 1
        nloc = self.mesh.nv_local if w == 0 else self.mesh.np_local
        nnod = self.mesh.nnodes if w == 0 else self.mesh.nnodesP
        if ph.ndim == 2 and ph.shape == (self.mesh.nels, nloc):
            layout = 1
        elif ph.ndim == 2 and ph.shape == (1, self.mesh.nels):
            layout, ph = 0, ph[0]
        elif ph.ndim == 1 and len(ph) == self.mesh.nels and len(ph) != nnod:
            layout = 0
        elif ph.ndim == 1 and len(ph) == nnod:
            layout = 2
        else:
            raise DimensionMismatch(f"phase array of shape {ph.shape} matches no layout")
        a = np.ascontiguousarray(ph, np.int32).ravel()
        self.ctx.check(_lib.fem_stokes_set_phases(self.handle, w, layout, a.ctypes.data_as(_P(I32)), len(a)))


def _dp_ptr(plastic):
    if plastic is None:
        return None, C.cast(None, _P(_DP))
    d = plastic._c()
    return d, C.pointer(d)


def _is_cap(plastic):
    return isinstance(plastic, DruckerPragerCap)


class _Inputs:
    """The array keywords of one solve!/residual call (DR.jl:183-186, 217-226,
    302-337, 386-418), attached with fem_stokes_set_solve_inputs on entry and
    cleared on exit. Host arrays are staged in device buffers freed on exit;
    DeviceArrays are used in place. Shapes follow this package's layouts:
    nodal vectors, point fields (nels, nq), pressure (nnodesP,) or (nels, np)."""

    def __init__(self, dr, *, body_force=None, viscosity=None, scaling_viscosity=None, P_old=None,
                 history_cap=0):
        self.dr, self.ctx = dr, dr.ctx
        self.bufs, self.keep = [], []
        m = dr.mesh
        nq = 7 if m.dim == 2 else 27
        c = _SolveInputs()
        self.history = None
        self.history_len = I64(0)
        if body_force is not None:
            if not isinstance(body_force, (tuple, list)) or len(body_force) != m.dim:
                raise ValueError("body_force must contain one assembled load vector per velocity component")
            for k, f in enumerate(body_force):
                if _size(f) != m.nnodes:
                    raise DimensionMismatch("body_force must match velocity nodes")
                c.body_force[k] = self._dev(f, (m.nnodes,))
        for name in ("viscosity", "scaling_viscosity"):
            eta = locals()[name]
            if eta is None:
                continue
            if isinstance(eta, (tuple, list)):
                raise ValueError(f"{name} takes a viscosity at the points, an (nels, {nq}) array; per-phase "
                                 "viscosities are set on the StokesMaterial")
            if _shape(eta) not in ((m.nels, nq), (m.nels * nq,)):
                raise DimensionMismatch(f"viscosity must be nq × nels = ({nq}, {m.nels}) (here (nels, nq))")
            if _dtype(eta) != np.float64:
                raise ValueError("viscosity must match the state precision")
            setattr(c, name, self._dev(eta, (m.nels, nq)))
        if P_old is not None:
            sh = _shape(P_old)
            npl = m.nnodesP // m.nels
            if len(sh) == 2 and sh != (m.nels, npl):
                if sh != (m.nels, nq):
                    raise DimensionMismatch(f"integration-point P_old must be nq × nels = ({nq}, {m.nels}), got "
                                            f"{sh[::-1]}")
                c.P_old = self._dev(P_old, sh)
                c.P_old_at_points = 1
            elif _size(P_old) == m.nnodesP:
                c.P_old = self._dev(P_old, (m.nnodesP,))
            else:
                raise DimensionMismatch("nodal P_old must match the pressure field shape")
        if history_cap:
            self.history = np.zeros((int(history_cap), 6))
            c.history = self.history.ctypes.data_as(C.POINTER(D))
            c.history_cap = int(history_cap)
            c.history_len = C.pointer(self.history_len)
        self.c = c

    def _dev(self, a, shape):
        if isinstance(a, DeviceArray):
            if a.ctx is not self.ctx:
                raise ValueError("keyword arrays and Stokes state must use the same device context")
            return a.ptr
        h = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))
        buf = VP()
        self.ctx.check(_lib.fem_buf_alloc(self.ctx.handle, max(h.nbytes, 8), C.byref(buf)))
        self.bufs.append(buf)
        if h.nbytes:
            self.ctx.check(_lib.fem_copy_h2d(self.ctx.handle, buf, h.ctypes.data, h.nbytes))
        return buf.value

    def rows(self):
        """collect_history rows as the reference's NamedTuples (DR.jl:558-561, 722-727)."""
        h = self.history[:self.history_len.value]
        nd = self.dr.mesh.dim
        return [dict(iter=int(r[0]), err_v=r[1], err_v_components=tuple(r[2:2 + nd]), err_P=r[5]) for r in h]

    def __enter__(self):
        try:
            self.ctx.check(_lib.fem_stokes_set_solve_inputs(self.dr.handle, C.byref(self.c)))
        except Exception:
            self._free()
            raise
        return self

    def __exit__(self, *exc):
        _lib.fem_stokes_set_solve_inputs(self.dr.handle, None)
        self._free()
        return False

    def _free(self):
        for b in self.bufs:
            _lib.fem_buf_free(self.ctx.handle, b)
        self.bufs = []


def _shape(a):
    return tuple(a.shape) if isinstance(a, DeviceArray) else np.shape(a)


def _size(a):
    return a.size if isinstance(a, DeviceArray) else int(np.size(a))


def _dtype(a):
    return a.dtype if isinstance(a, DeviceArray) else np.asarray(a).dtype


_INPUT_KEYS = ("body_force", "viscosity", "scaling_viscosity", "P_old")


def _pop_inputs(kwargs):
    return {k: kwargs.pop(k) for k in _INPUT_KEYS if k in kwargs and kwargs[k] is not None}


def set_pressure_source(dr: StokesDR, Qq, Q2D=None):
    """_set_pressure_source!(dr.Q, Qq, Q2D, ...) (pressure_residual.jl:32-77),
    as solve_stokes_dyrel! runs it for its Qq/Q2D keywords (DR.jl:451-457):
    dr.Q = Qq scaled so that ∫ Q dΩ = Q2D; without Q2D, Qq is copied."""
    if _size(Qq) != dr.mesh.nnodesP:
        raise DimensionMismatch("Qq must match the pressure field shape")
    with _Inputs(dr) as io:
        ptr = io._dev(Qq, (dr.mesh.nnodesP,))
        dr.ctx.check(_lib.fem_stokes_set_pressure_source(dr.handle, ptr, 0.0 if Q2D is None else float(Q2D),
                                                          int(Q2D is not None)))


def assemble_viscosity_weighted_pressure_scaling(dr: StokesDR, pressure_factor, dt, finite_K=False, eta=None):
    """assemble_viscosity_weighted_pressure_scaling! (pressure_scaling.jl:8-149):
    fills dr.M_P and dr.gammaP; eta (nels, nq) samples the viscosity at the
    points in place of the phase viscosities (pressure_scaling.jl:60, 185)."""
    with _Inputs(dr, scaling_viscosity=eta):
        dr.ctx.check(_lib.fem_stokes_pressure_scaling(dr.handle, float(pressure_factor), float(dt),
                                                      int(finite_K)))


def solve(dr, mesh, *bcs, **kwargs):
    """solve!(dr, mesh, bc...; kwargs...): dispatches on the state type like
    the reference's solve! methods: StokesDR -> solve_stokes (DR.jl:302-337),
    ThermalDiffusionDR -> solve_thermal (heat_diffusion/solvers/DR.jl:50-54),
    CellPressureStokesDR -> solve_cell_pressure (DR.jl:980-996)."""
    if isinstance(dr, ThermalDiffusionDR):
        return solve_thermal(dr, mesh, *bcs, **kwargs)
    if isinstance(dr, CellPressureStokesDR):
        return solve_cell_pressure(dr, mesh, *bcs, **kwargs)
    return solve_stokes(dr, mesh, *bcs, **kwargs)


# fem_solve_params.precision: FP64; FP32 inner iterations on corrections of the
# FP64 state (with a yield model, two FP64 returns per point near yield); the
# same with the return linearised about the FP64 state (2-D Drucker-Prager)
_PRECISIONS = {"fp64": 0, "fp32": 1, "fp32-tangent": 2}


def _stokes_call(dr: StokesDR, mesh: MixedMesh, bcs, keep, *, dt, pressure_factor=50.0, tolerance=1e-6,
                 max_iterations=50_000, check_interval=50, iterMax=50_000, max_ph_iterations=1000, rel_drop0=1e-2,
                 plastic=None, phases_v=None, phases_P=None, finite_K=None, freeze_jacobian=None,
                 throw_on_failure=True, tau_old=True, skip_scaling=False, verbose=False, time_kernels=False,
                 precision="fp64", measure_lmax=False, lmax_power_iterations=12, lmax_power_rtol=1e-2,
                 lmax_safety=1.1):
    """Checks and C arguments of one solve! call: (bc array, plastic pointer, params)."""
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    if phases_v is not None:
        dr.set_phases("v", phases_v)
    if phases_P is not None:
        dr.set_phases("P", phases_P)
    given = [b for b in bcs if b is not None]
    if len(given) != mesh.dim:
        raise DimensionMismatch(f"a {mesh.dim}-D solve takes {mesh.dim} boundary conditions, got {len(given)}")
    arr = (_BC * len(given))(*(b._c(keep) for b in given))
    keep.append(arr)
    if finite_K is None:   # DR.jl:312: finite_K = plastic isa DruckerPragerCap
        finite_K = _is_cap(plastic)
    d, dp = _dp_ptr(plastic)
    keep.append(d)
    p = _SolveParams(dt=float(dt), pressure_factor=float(pressure_factor), tolerance=float(tolerance),
                     max_iterations=int(max_iterations), iter_max=int(iterMax), check_interval=int(check_interval),
                     max_ph_iterations=int(max_ph_iterations), rel_drop0=float(rel_drop0), finite_K=int(finite_K),
                     freeze_jacobian=-1 if freeze_jacobian is None else int(bool(freeze_jacobian)),
                     throw_on_failure=int(bool(throw_on_failure)), use_tau_old=int(bool(tau_old)),
                     skip_scaling=int(bool(skip_scaling)), verbose=int(verbose),
                     precision=_PRECISIONS[precision], time_kernels=int(bool(time_kernels)),
                     measure_lmax=int(bool(measure_lmax)), lmax_power_iterations=int(lmax_power_iterations),
                     lmax_power_rtol=float(lmax_power_rtol), lmax_safety=float(lmax_safety))
    return arr, dp, p


def _stats_dict(st):
    stats = {k: getattr(st, k) for k, _ in _SolveStats._fields_}
    stats["converged"] = bool(stats["converged"])
    stats["reached_total_iter"] = bool(stats["reached_total_iter"])
    stats["iterations"] = stats["iter"]
    stats["residual"] = stats["err_abs"]
    return stats


def solve_stokes(dr: StokesDR, mesh: MixedMesh, bc_vx: DirichletBoundaryCondition, bc_vy: DirichletBoundaryCondition,
                 bc_vz: DirichletBoundaryCondition | None = None, **kwargs):
    """solve!(dr, mesh, bc_vx, bc_vy[, bc_vz]; dt, ...) (src/stokes/solvers/DR.jl:302-337).
    Mutates dr in place and returns the stats dict of DR.jl:783-801. A 3-D
    state takes three boundary conditions and runs in FP64. Keywords: dt,
    pressure_factor=50, tolerance=1e-6, max_iterations=50_000,
    check_interval=50, iterMax=50_000, max_ph_iterations=1000,
    rel_drop0=1e-2, plastic, phases_v, phases_P, finite_K, freeze_jacobian,
    throw_on_failure=True, tau_old=True, skip_scaling, verbose,
    time_kernels, precision.

    precision="fp64" runs the reference's iteration in FP64. precision="fp32"
    keeps the state, the outer residuals, the norms and the convergence test
    in FP64 (same criterion and tolerance) and runs the inner pseudo-transient
    iterations in FP32 on a correction of the linearised visco-elastic
    problem, folded into the FP64 state at every outer iteration.
    precision="fp32-tangent" does the same, with the Drucker-Prager return of
    2-D states linearised about the FP64 state at every fold instead of two
    FP64 returns per point near yield."""
    keep = []
    inputs = _pop_inputs(kwargs)
    collect = bool(kwargs.pop("collect_history", False))
    if kwargs.pop("verbose_inner", False):   # the inner-check trace (DR.jl:729) is verbose level 2
        kwargs["verbose"] = 2
    Qq, Q2D = kwargs.pop("Qq", None), kwargs.pop("Q2D", None)
    if Q2D is not None and Qq is None:
        raise ValueError("Qq is required when Q2D is supplied")
    bcs, dp, p = _stokes_call(dr, mesh, (bc_vx, bc_vy, bc_vz), keep, **kwargs)
    if Qq is not None:
        set_pressure_source(dr, Qq, Q2D)
    hcap = min(2 * (p.max_iterations // max(p.check_interval, 1) + p.max_ph_iterations) + 4, 1_000_000) \
        if collect else 0
    st = _SolveStats()
    with _Inputs(dr, history_cap=hcap, **inputs) as io:
        if _is_cap(kwargs.get("plastic")):
            rc = _lib.fem_stokes_solve_cap(dr.handle, bcs, C.cast(dp, _P(_Cap)), C.byref(p), C.byref(st))
        else:
            rc = _lib.fem_stokes_solve(dr.handle, bcs, dp, C.byref(p), C.byref(st))
        stats = _stats_dict(st)
        stats["history"] = io.rows() if collect else []
    dr.ctx.check(rc)
    return stats


def solve_batch(problems, nthreads: int = 0, raise_errors: bool = True):
    """Independent Stokes solves in one call (fem_stokes_solve_batch):
    ``problems`` is a list of (dr, mesh, bcs, kwargs) with the keywords of
    solve_stokes; every dr shares one Context. Mixed-precision solves
    without a yield model on structured meshes with cell counts multiples of
    16 x 8 run their inner iterations as grouped launches; the rest are
    solved one after the other. nthreads is accepted and ignored. Returns
    the list of stats dicts (with "status"); raises the first failure after
    all ran unless raise_errors is False (then a failed solve shows as a
    nonzero "status" and its message as "error")."""
    n = len(problems)
    keep = []
    states = (VP * n)()
    bcp = (_P(_BC) * n)()
    dps = (_P(_DP) * n)()
    params = (_SolveParams * n)()
    stats = (_SolveStats * n)()
    status = (I32 * n)()
    for i, (dr, mesh, bcs, kw) in enumerate(problems):
        if _is_cap(kw.get("plastic")):
            raise ValueError("batched solves take DruckerPrager or no yield model, not the tensile cap")
        arr, dp, p = _stokes_call(dr, mesh, tuple(bcs) + (None,) * (3 - len(bcs)), keep, **kw)
        states[i] = dr.handle
        bcp[i] = C.cast(arr, _P(_BC))
        dps[i] = dp
        params[i] = p
    rc = _lib.fem_stokes_solve_batch(n, states, bcp, dps, params, stats, status, int(nthreads))
    out = []
    for i in range(n):
        d = _stats_dict(stats[i])
        d["status"] = status[i]
        if status[i] != FEM_OK:
            d["error"] = problems[i][0].ctx.error()
        out.append(d)
    if rc != FEM_OK and (raise_errors or not any(st != FEM_OK for st in status)):
        for i, (dr, *_rest) in enumerate(problems):
            if status[i] != FEM_OK:
                dr.ctx.check(status[i])
        # rejected before any solve ran (the message is on a context of the batch)
        for dr, *_rest in problems:
            if dr.ctx.error():
                dr.ctx.check(rc)
        raise FemError(rc, "batched solve rejected")
    return out


def update_stokes_current_stress(dr: StokesDR, mesh: MixedMesh, dt, plastic=None):
    """update_stokes_current_stress!(dr, mesh, τ, Δt; plastic) (DR.jl:820-897):
    writes dr.tau at the integration points."""
    d, dp = _dp_ptr(plastic)
    if _is_cap(plastic):
        dr.ctx.check(_lib.fem_stokes_update_current_stress_cap(dr.handle, C.cast(dp, _P(_Cap)), float(dt)))
    else:
        dr.ctx.check(_lib.fem_stokes_update_current_stress(dr.handle, dp, float(dt)))
    return dr.tau


def commit_stokes_plastic_history(dr: StokesDR, mesh: MixedMesh, dt, *, plastic):
    """commit_stokes_plastic_history!(dr, mesh, Δt; plastic) (DR.jl:899-958):
    γ += Δt γdot, θ += Δt θdot in dr.plastic_history from the tensile cap's
    local update at the current dr.v, dr.P (a no-op for other yield models)."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    if not _is_cap(plastic):
        return dr
    d, dp = _dp_ptr(plastic)
    dr.ctx.check(_lib.fem_stokes_commit_plastic_history(dr.handle, C.cast(dp, _P(_Cap)), float(dt)))
    return dr


def residual(dr: StokesDR, dt, plastic=None, with_pnum=False, finite_K=False, **inputs):
    """Element residual sweep: dr.RP (and dr.Pnum) and dr.Rv at the current
    state; body_force, viscosity and P_old as solve! takes them (the loads are
    subtracted from dr.Rv before the Dirichlet rows)."""
    d, dp = _dp_ptr(plastic)
    kw = _pop_inputs(inputs)
    if inputs:
        raise TypeError(f"unexpected keywords {sorted(inputs)}")
    with _Inputs(dr, **kw):
        if _is_cap(plastic):
            dr.ctx.check(_lib.fem_stokes_residual_cap(dr.handle, C.cast(dp, _P(_Cap)), float(dt), int(with_pnum),
                                                      int(finite_K)))
        else:
            dr.ctx.check(_lib.fem_stokes_residual(dr.handle, dp, float(dt), int(with_pnum), int(finite_K)))


def jacobian_diagnostics(dr: StokesDR, dt, plastic=None, finite_K=False, viscosity=None):
    """Augmented-Jacobian diagnostics dr.dRvdv (row sums) and dr.PC_v (diagonal);
    a viscosity at the points enters as its element maximum."""
    d, dp = _dp_ptr(plastic)
    with _Inputs(dr, viscosity=viscosity):
        if _is_cap(plastic):
            dr.ctx.check(_lib.fem_stokes_jacobian_diagnostics_cap(dr.handle, C.cast(dp, _P(_Cap)), float(dt),
                                                                  int(finite_K)))
        else:
            dr.ctx.check(_lib.fem_stokes_jacobian_diagnostics(dr.handle, dp, float(dt), int(finite_K)))


def copyto(dst, src):
    """copyto!(dst, src) for device arrays and the field containers (dr.v,
    dr.tau, ...): e.g. copyto(dr.tau_old, dr.tau) and copyto(dr.P0, dr.P) at a
    time-step boundary, on the device."""
    dst.copyto(src)
    return dst


def rotate_stress(dr: StokesDR, mesh: MixedMesh, dt):
    """rotate_stress!(dr, mesh_stokes, Δt) (src/stokes/stress_rotation.jl:16-61):
    dr.tau_old becomes dr.tau rotated by the local vorticity over Δt (2-D:
    θ = ωΔt, ω = ½(∂vx/∂y − ∂vy/∂x); 3-D: the rotation vector curl(v)Δt/2)."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    dr.ctx.check(_lib.fem_stokes_rotate_stress(dr.handle, float(dt)))
    return dr.tau_old


def time_element_kernel(dr: StokesDR, dt, n=20, plastic=None) -> float:
    """Average device time (ms) of the fused element kernel over n launches
    (HIP events on the context stream). Advances bubble-node velocities."""
    if _is_cap(plastic):
        raise ValueError("the element-kernel timer takes DruckerPrager or no yield model")
    d, dp = _dp_ptr(plastic)
    out = D()
    dr.ctx.check(_lib.fem_stokes_time_element_kernel(dr.handle, dp, float(dt), int(n), C.byref(out)))
    return out.value


def time_correction_kernel(dr: StokesDR, dt, n=20, plastic=None) -> float:
    """Average device time (ms) of one FP32 correction step of the 3-D
    mixed-precision solve (element kernel and shared-node gather) over n
    sweeps on a zero correction; with plastic every element takes the FP64
    returns. Overwrites the solve's FP32 work arrays only."""
    if _is_cap(plastic):
        raise ValueError("the correction-step timer takes DruckerPrager or no yield model")
    d, dp = _dp_ptr(plastic)
    out = D()
    dr.ctx.check(_lib.fem_stokes_time_correction_kernel(dr.handle, dp, float(dt), int(n), C.byref(out)))
    return out.value


# ---------------------------------------------------------------------------
# thermal diffusion (src/heat_diffusion/**)
# ---------------------------------------------------------------------------
def dr_velocity_update(ctx: Context, rate: DeviceArray, v: DeviceArray, R: DeviceArray, PC: DeviceArray, alpha,
                       beta):
    """update_stokes_velocity! for one component (helpers.jl:31-41) with the
    solver's negated step: rate = beta rate + R / PC; v -= alpha rate."""
    ctx.check(_lib.fem_dr_velocity_update(ctx.handle, len(v), rate.ptr, v.ptr, R.ptr, PC.ptr, float(alpha),
                                          float(beta)))


def dp_return_map_batch(ctx: Context, dt, *, exx, eyy, exy, tau_old, P, eta, G, C_, phi, psi, eta_reg, Kb):
    """The plane-strain VEP stress with the Drucker-Prager return at
    independent points (rheology.jl:226-245, 272-334); host arrays in,
    (txx, tyy, txy, yields) out."""
    n = len(exx)
    ins = dict(exx=exx, eyy=eyy, exy=exy, txx_old=tau_old[0], tyy_old=tau_old[1], txy_old=tau_old[2], P=P, eta=eta,
               G=G, C=C_, phi=phi, psi=psi, eta_reg=eta_reg, Kb=Kb)
    dev = {k: device_array(ctx, np.broadcast_to(np.asarray(v, float), (n,))) for k, v in ins.items()}
    outs = {k: device_array(ctx, shape=(n,)) for k in ("txx", "tyy", "txy")}
    y = device_array(ctx, shape=(n,), dtype=np.int32)
    pts = _DPPoints(**{k: a.ptr for k, a in {**dev, **outs}.items()}, yields=y.ptr)
    ctx.check(_lib.fem_dp_return_map_batch(ctx.handle, n, float(dt), C.byref(pts)))
    ctx.sync()
    return outs["txx"].numpy(), outs["tyy"].numpy(), outs["txy"].numpy(), y.numpy()


class _OwnedBuf:
    """A fem_buf_alloc allocation, freed with its owner."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        buf = VP()
        ctx.check(_lib.fem_buf_alloc(ctx.handle, max(int(nbytes), 8), C.byref(buf)))
        self.ptr = buf.value

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.fem_buf_free(self.ctx.handle, VP(self.ptr))
            self.ptr = None


def device_array(ctx: Context, host=None, shape=None, dtype=np.float64) -> DeviceArray:
    """A new device array owned by the caller (freed when it is collected),
    filled from host when given."""
    if host is not None:
        host = np.ascontiguousarray(np.asarray(host, dtype=dtype))
        shape = host.shape
    own = _OwnedBuf(ctx, int(np.prod(shape)) * np.dtype(dtype).itemsize)
    a = DeviceArray(ctx, own.ptr, shape, dtype, owner=own)
    if host is not None:
        a.copyto(host)
    return a


def norms(ctx: Context, Rv_x, Rv_y, RP, M_P):
    """The residual norms of a DR check (DR.jl:556-557, 700) on device arrays
    (DeviceArray, FP64): returns (‖Rv_x‖₂, ‖Rv_y‖₂, ‖RP./M_P‖₂, err_v, err_P)."""
    out = (D * 5)()
    ctx.check(_lib.fem_norms(ctx.handle, Rv_x.ptr, Rv_y.ptr, len(Rv_x), RP.ptr, M_P.ptr, len(RP), out))
    return tuple(out)


class ThermalMaterial:
    """ThermalMaterial(; k=(1.0,), Cp=(1.0,), ρ0=(1.0,), α=(1.0,), K=(1.0,))
    (src/heat_diffusion/types/heat_diffusion_types.jl:15-41). The first tuple
    given fixes the phase count; scalars apply to every phase; omitted
    properties are unit-valued."""

    def __init__(self, k=None, Cp=None, rho0=None, alpha=None, K=None):
        given = [v for v in (k, Cp, rho0, alpha, K) if isinstance(v, (tuple, list, np.ndarray))]
        n = len(given[0]) if given else 1
        if n < 1 or n > MAXPH:
            raise ValueError(f"the phase count must be in 1..{MAXPH}")

        def prop(v):
            t = _tup(1.0 if v is None else v, n)
            if len(t) != n:
                raise DimensionMismatch("material property tuples must have the same length")
            return t

        self.k, self.Cp, self.rho0, self.alpha, self.K = prop(k), prop(Cp), prop(rho0), prop(alpha), prop(K)

    @property
    def nphases(self):
        return len(self.k)

    def _c(self):
        m = _ThermalMaterial()
        m.nphases = self.nphases
        for i in range(self.nphases):
            m.k[i], m.Cp[i], m.rho0[i], m.alpha[i], m.K[i] = self.k[i], self.Cp[i], self.rho0[i], self.alpha[i], \
                self.K[i]
        return m


class ThermalDiffusionDR:
    """ThermalDiffusionDR(mesh, material; CFL=0.98, c_fact=0.9, ϵ=1e-6)
    (heat_diffusion_types.jl:90-153): nodal device fields R, R0, dRdT (∂R∂T),
    PC, T, T0, dTdtau (∂T∂τ), P, source and phases (Int32, 1-based, all 1)."""

    _names = ("R", "R0", "dRdT", "PC", "T", "T0", "dTdtau", "P", "source", "phases")

    def __init__(self, mesh: "MixedMesh", material: ThermalMaterial, CFL=0.98, c_fact=0.9, eps=1e-6):
        self.ctx = mesh.ctx
        self.mesh = mesh
        self.material = material
        self.CFL, self.c_fact, self.eps = float(CFL), float(c_fact), float(eps)
        h = VP()
        m = material._c()
        o = _ThermalOpts(CFL=self.CFL, c_fact=self.c_fact, eps=self.eps)
        self.ctx.check(_lib.fem_thermal_create(self.ctx.handle, mesh.handle, C.byref(m), C.byref(o), C.byref(h)))
        self.handle = h
        for k, name in enumerate(self._names):
            p, n, dt = VP(), I64(), I32()
            self.ctx.check(_lib.fem_thermal_field(h, k, C.byref(p), C.byref(n), C.byref(dt)))
            setattr(self, name, DeviceArray(self.ctx, p.value, (n.value,), np.int32 if dt.value == 1 else np.float64,
                                            owner=self))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.fem_thermal_destroy(h)
            self.handle = None


def solve_thermal(dr: ThermalDiffusionDR, mesh: "MixedMesh", bc: "DirichletBoundaryCondition" = None, *, dt,
                  tolerance=None, max_iterations=10_000, check_interval=100, verbose=False, throw_on_failure=True,
                  Tref=273.0):
    """solve!(dr::ThermalDiffusionDR, mesh, bc; dt, tolerance=dr.ϵ,
    max_iterations=10_000, check_interval=100, throw_on_failure=true, Tref=273)
    (src/heat_diffusion/solvers/DR.jl:25-54). Returns (; converged,
    iterations, residual) as a dict, plus the wall time."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    keep = []
    bcp = C.pointer(bc._c(keep)) if bc is not None else C.cast(None, _P(_BC))
    p = _ThermalParams(dt=float(dt), tolerance=-1.0 if tolerance is None else float(tolerance),
                       max_iterations=int(max_iterations), check_interval=int(check_interval), Tref=float(Tref),
                       throw_on_failure=int(bool(throw_on_failure)), verbose=int(bool(verbose)))
    st = _ThermalStats()
    rc = _lib.fem_thermal_solve(dr.handle, bcp, C.byref(p), C.byref(st))
    dr.ctx.check(rc)
    return {"converged": bool(st.converged), "iterations": st.iterations, "residual": st.residual,
            "seconds": st.seconds}


def thermal_residual(dr: ThermalDiffusionDR, dt, Tref=273.0, compute_jacobian=False):
    """assemble_diffusion_matrices_atomix! (heat residual.jl:90-121): dr.R and,
    with compute_jacobian, dr.dRdT and dr.PC."""
    dr.ctx.check(_lib.fem_thermal_residual(dr.handle, float(dt), float(Tref), int(bool(compute_jacobian))))


# ---------------------------------------------------------------------------
# coupled thermal--Stokes (src/stokes/solvers/DR.jl:44-165, 514-661)
# ---------------------------------------------------------------------------
def assemble_shear_heating(dr: StokesDR, mesh: MixedMesh, dt, plastic=None) -> np.ndarray:
    """assemble_shear_heating! (src/stokes/assemblers/shear_heating.jl:18-103):
    Φ = τ : (ε̇ − (τ − τ_old)/(2GΔt)) at the velocity quadrature points for the
    current dr.v, dr.P and dr.tau_old, as an (nels, nq) array (nq = 7 on T7,
    27 on Hex27)."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    if _is_cap(plastic):
        raise ValueError("shear heating is assembled for DruckerPrager or no yield model")
    d, dp = _dp_ptr(plastic)
    buf = VP()
    nq = mesh.nq
    dr.ctx.check(_lib.fem_buf_alloc(dr.ctx.handle, mesh.nels * nq * 8, C.byref(buf)))
    try:
        phi = DeviceArray(dr.ctx, buf.value, (mesh.nels, nq))
        dr.ctx.check(_lib.fem_stokes_shear_heating(dr.handle, dp, float(dt), phi.ptr))
        return phi.numpy()
    finally:
        _lib.fem_buf_free(dr.ctx.handle, buf)


def solve_coupled(thermal: ThermalDiffusionDR, stokes: StokesDR, thermal_mesh: MixedMesh, stokes_mesh: MixedMesh,
                  bc_T: DirichletBoundaryCondition, bc_vx: DirichletBoundaryCondition,
                  bc_vy: DirichletBoundaryCondition, bc_vz: DirichletBoundaryCondition | None = None, *, dt,
                  Tref=273.0, shear_heating=False, **kwargs):
    """solve_coupled!(thermal, stokes, thermal_mesh, stokes_mesh, bc_T, bc_v;
    dt, Tref=273, shear_heating=false, kwargs...) (DR.jl:81-165): one coupled
    step of size dt, with two velocity conditions on a T7 mesh and three on a
    Hex27 mesh. Every inner Stokes iteration first advances one thermal DR
    iteration and gathers the thermal field at the velocity nodes that carry
    the pressure DoFs (the triangle vertices; the hexahedron's centre and +x,
    +y, +z face centres) into stokes.T. The remaining keywords are solve_stokes's
    (check_interval sets the check cadence of both relaxations). Returns the
    Stokes stats dict plus err_T and thermal_iterations; converged requires
    the thermal test (err_T < thermal.eps) too. thermal.source is unchanged;
    shear_heating adds τ : (ε̇ − ε̇ᵉˡ) at the quadrature points to it inside
    the thermal residual, refreshed at every thermal check."""
    if thermal_mesh.nnodes != stokes_mesh.nnodes:
        raise DimensionMismatch("thermal mesh nodes must match the Stokes velocity-node layout")
    if len(thermal.T) != thermal_mesh.nnodes:
        raise DimensionMismatch("thermal state size must match thermal mesh nodes")
    if len(stokes.T) != stokes_mesh.nnodesP:
        raise DimensionMismatch("Stokes temperature size must match pressure DoFs")
    if thermal_mesh is not thermal.mesh:
        raise ValueError("thermal was built on a different mesh")
    if thermal.ctx is not stokes.ctx:
        raise ValueError("the thermal and Stokes states must share one Context")
    if _is_cap(kwargs.get("plastic")):
        raise ValueError("coupled solves take DruckerPrager or no yield model, not the tensile cap")
    keep = []
    bcs, dp, p = _stokes_call(stokes, stokes_mesh, (bc_vx, bc_vy, bc_vz), keep, dt=dt, **kwargs)
    bct = C.pointer(bc_T._c(keep)) if bc_T is not None else C.cast(None, _P(_BC))
    cp = _CoupledParams(Tref=float(Tref), shear_heating=int(bool(shear_heating)))
    st, cst = _SolveStats(), _CoupledStats()
    rc = _lib.fem_coupled_solve(thermal.handle, stokes.handle, bct, bcs, dp, C.byref(p), C.byref(cp), C.byref(st),
                                C.byref(cst))
    stats = _stats_dict(st)
    stats["err_T"] = cst.err_T
    stats["thermal_iterations"] = cst.thermal_iterations
    stats["thermal_converged"] = bool(cst.thermal_converged)
    stokes.ctx.check(rc)
    return stats


# ---------------------------------------------------------------------------
# CellPressureStokesDR: viscous Hex27 with cell pressure modes (1, ξ, η, ζ)
# ---------------------------------------------------------------------------
class CellPressureStokesDR:
    """CellPressureStokesDR(mesh, material; phases=None) (stokes_types.jl:545-613):
    the viscous, incompressible 3-D Hex27 state with four cell-local pressure
    modes (1, ξ, η, ζ). Device fields: v, Rv, diagonal (VectorField3D), P, RP,
    pressure_mass ((nels, 4)); phases one per cell (1-based)."""

    def __init__(self, mesh: MixedMesh, material: StokesMaterial, phases=None):
        if mesh.dim != 3 or mesh.nv_local != 27:
            raise ValueError("CellPressureStokesDR requires a Hex27 mesh")
        if len(material.g) != 3:
            raise DimensionMismatch("material gravity must have 3 components")
        if not (all(math.isinf(x) for x in material.G) and all(math.isinf(x) for x in material.K)
                and all(x == 0 for x in material.alpha)):
            raise ValueError("CellPressureStokesDR is purely viscous and incompressible; G and K must be Inf and "
                             "α zero")
        n = len(material.eta)
        if phases is not None:
            ph = np.asarray(phases).astype(np.int64).ravel()
            if len(ph) != mesh.nels:
                raise DimensionMismatch(f"phases has {len(ph)} entries but the mesh has {mesh.nels} cells")
            lo, hi = int(ph.min()), int(ph.max())
            if not (1 <= lo and hi <= n):
                raise ValueError(f"phases must lie in 1:{n}, got {lo}:{hi}")
        self.mesh, self.material, self.ctx = mesh, material, mesh.ctx
        self._st = StokesDR(mesh, material)
        if phases is not None:
            self._st.set_phases("v", ph.astype(np.int32))
        self.handle = self._st.handle
        st = self._st
        self.v, self.Rv, self.diagonal = st.v, st.Rv, st.PC_v
        self.P = DeviceArray(self.ctx, st.P.ptr, (mesh.nels, 4), owner=st)
        self.RP = DeviceArray(self.ctx, st.RP.ptr, (mesh.nels, 4), owner=st)
        self.pressure_mass = DeviceArray(self.ctx, st.M_P.ptr, (mesh.nels, 4), owner=st)


def stokes_preconditioner_3d(dr: CellPressureStokesDR):
    """stokes_preconditioner_3d! (momentum_residuals.jl:715-745): fills
    dr.diagonal and dr.pressure_mass."""
    dr.ctx.check(_lib.fem_cell_stokes_preconditioner(dr.handle))
    return dr.diagonal, dr.pressure_mass


def assemble_stokes_pressure_residual_3d(dr: CellPressureStokesDR):
    """assemble_stokes_pressure_residual_3d! (pressure_residual.jl:337-366): dr.RP from dr.v."""
    dr.ctx.check(_lib.fem_cell_stokes_pressure_residual(dr.handle))
    return dr.RP


def assemble_stokes_momentum_residual_3d(dr: CellPressureStokesDR):
    """assemble_stokes_momentum_residual_3d! (momentum_residuals.jl:645-691): dr.Rv from dr.v, dr.P."""
    dr.ctx.check(_lib.fem_cell_stokes_momentum_residual(dr.handle))
    return dr.Rv


def solve_cell_pressure(dr: CellPressureStokesDR, mesh: MixedMesh, bc_vx, bc_vy, bc_vz, *, tolerance=1e-5,
                        max_iterations=3000, check_interval=100, velocity_step=0.6, pressure_step=0.2, verbose=False,
                        collect_history=False, throw_on_failure=True):
    """solve!(dr::CellPressureStokesDR, mesh, bc_v; tolerance=1e-5,
    max_iterations=3000, check_interval=100, velocity_step=0.6,
    pressure_step=0.2, verbose, collect_history, throw_on_failure)
    (DR.jl:980-1057). Returns (converged, iterations, residual, history,
    err_v, err_P, iter, err, reached_total_iter) as a dict."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    keep = []
    arr = (_BC * 3)(*(b._c(keep) for b in (bc_vx, bc_vy, bc_vz)))
    p = _CellParams(tolerance=float(tolerance), max_iterations=int(max_iterations),
                    check_interval=int(check_interval), velocity_step=float(velocity_step),
                    pressure_step=float(pressure_step), verbose=int(bool(verbose)),
                    throw_on_failure=int(bool(throw_on_failure)))
    st = _CellStats()
    cap = (int(max_iterations) // max(int(check_interval), 1) + 2) if collect_history else 0
    hist = np.zeros((max(cap, 1), 3))
    nh = I64(0)
    rc = _lib.fem_cell_stokes_solve(dr.handle, arr, C.byref(p), C.byref(st),
                                    hist.ctypes.data_as(_P(D)) if cap else C.cast(None, _P(D)), cap, C.byref(nh))
    out = {k: getattr(st, k) for k, _ in _CellStats._fields_}
    out["converged"] = bool(out["converged"])
    out["reached_total_iter"] = bool(out["reached_total_iter"])
    out["iter"] = out["iterations"]
    out["err"] = out["residual"]
    out["history"] = [dict(iter=int(h[0]), err_v=h[1], err_P=h[2]) for h in hist[:nh.value]]
    dr.ctx.check(rc)
    return out


def _cell_result(st, hist, nh):
    out = {k: getattr(st, k) for k, _ in _CellStats._fields_}
    out["converged"] = bool(out["converged"])
    out["reached_total_iter"] = bool(out["reached_total_iter"])
    out["iter"] = out["iterations"]
    out["err"] = out["residual"]
    out["history"] = [dict(iter=int(h[0]), err_v=h[1], err_P=h[2]) for h in hist[:nh.value]]
    return out


def solve_adjoint(dr, mesh, *bcs, **kwargs):
    """solve_adjoint!(dr, mesh, bc...; kwargs...), dispatched on the state type
    like the reference's methods: StokesDR -> solve_stokes_adjoint
    (DR_adjoint.jl:396-427), CellPressureStokesDR -> solve_cell_adjoint
    (DR_adjoint.jl:446-480)."""
    if isinstance(dr, CellPressureStokesDR):
        return solve_cell_adjoint(dr, mesh, *bcs, **kwargs)
    return solve_stokes_adjoint(dr, mesh, *bcs, **kwargs)


def solve_stokes_adjoint(dr: StokesDR, mesh: MixedMesh, bc_vx: DirichletBoundaryCondition,
                         bc_vy: DirichletBoundaryCondition, *, dt, objective_vx, objective_vy, lam_vx, lam_vy, lam_P,
                         tolerance=1e-6, max_iterations=50_000, check_interval=50, throw_on_failure=True,
                         iterMax=50_000, max_ph_iterations=100, rel_drop=0.1, measure_lmax=True, plastic=None,
                         phases_v=None, phases_P=None, operator="blocks", verbose=False):
    """solve_adjoint!(dr::StokesDR, mesh, bc_vx, bc_vy; dt, objective_vx,
    objective_vy, λvx, λvy, λP, ...) (DR_adjoint.jl:396-427) with the stored
    element blocks (operator = :blocks) on a converged 2-D forward state without
    a yield model or with the Drucker-Prager model (plastic; not the tensile
    cap). lam_* are DeviceArrays updated in place (their values are the
    initial iterate) or host arrays (then the results are in the returned
    dict as lam_vx, lam_vy, lam_P). Returns (; itPH, iter, err, err_v, err_P,
    converged, iterations, residual, lmax, lmax_gershgorin, lmax_iterations)."""
    if operator != "blocks":
        raise ValueError(f"operator must be 'blocks' in this build, got {operator!r}")
    if _is_cap(plastic):
        raise ValueError("the frozen-operator adjoint covers models without a yield model or with the "
                         "Drucker-Prager model")
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    if mesh.dim != 2:
        raise ValueError("the frozen-operator adjoint runs on 2-D (T7/P1-disc) states")
    if phases_v is not None:
        dr.set_phases("v", phases_v)
    if phases_P is not None:
        dr.set_phases("P", phases_P)
    nv, npp = mesh.nnodes, mesh.nnodesP
    for o in (objective_vx, objective_vy, lam_vx, lam_vy):
        if _size(o) != nv:
            raise DimensionMismatch("objective and velocity adjoint arrays must have mesh.nnodes entries")
    if _size(lam_P) != npp:
        raise DimensionMismatch("the pressure adjoint must have mesh.nnodesP entries")
    keep = []
    arr = (_BC * 2)(*(b._c(keep) for b in (bc_vx, bc_vy)))
    p = _AdjointParams(dt=float(dt), tolerance=float(tolerance), max_iterations=int(max_iterations),
                       check_interval=int(check_interval), iter_max=int(iterMax),
                       max_ph_iterations=int(max_ph_iterations), rel_drop=float(rel_drop),
                       measure_lmax=int(bool(measure_lmax)), throw_on_failure=int(bool(throw_on_failure)),
                       verbose=int(bool(verbose)))
    st = _AdjointStats()
    d, dp = _dp_ptr(plastic)
    keep.append(d)
    with _Inputs(dr) as io:
        ox, oy = io._dev(objective_vx, (nv,)), io._dev(objective_vy, (nv,))
        lams = [io._dev(a, (n,)) for a, n in ((lam_vx, nv), (lam_vy, nv), (lam_P, npp))]
        rc = _lib.fem_stokes_solve_adjoint(dr.handle, arr, dp, ox, oy, *lams, C.byref(p), C.byref(st))
        out = {k: getattr(st, k) for k, _ in _AdjointStats._fields_}
        out["converged"] = bool(out["converged"])
        out["iterations"], out["residual"] = out["iter"], out["err"]
        if rc == FEM_OK or rc == FEM_ENOCONV:
            for name, a, ptr_, n in (("lam_vx", lam_vx, lams[0], nv), ("lam_vy", lam_vy, lams[1], nv),
                                     ("lam_P", lam_P, lams[2], npp)):
                out[name] = a if isinstance(a, DeviceArray) else DeviceArray(dr.ctx, ptr_, (n,)).numpy()
    dr.ctx.check(rc)
    return out


def adjoint_blocks(dr: StokesDR, dt, plastic=None):
    """assemble_adjoint_operator (adjoint_operator.jl:412-475) at the state in
    dr: dict of host arrays A (nels, 14, 14), B (nels, 14, 3), C (nels, 3, 14),
    D (nels, 3, 3), W (nels, 14, 3); velocity DoFs stacked by component."""
    if _is_cap(plastic):
        raise ValueError("the frozen-operator adjoint covers models without a yield model or with the "
                         "Drucker-Prager model")
    d, dp = _dp_ptr(plastic)
    ne = dr.mesh.nels
    out = dict(A=np.zeros((ne, 14, 14)), B=np.zeros((ne, 14, 3)), C=np.zeros((ne, 3, 14)), D=np.zeros((ne, 3, 3)),
               W=np.zeros((ne, 14, 3)))
    dr.ctx.check(_lib.fem_stokes_adjoint_blocks(dr.handle, dp, float(dt), *(out[k].ctypes.data_as(_P(D))
                                                                            for k in "ABCDW")))
    return out


def solve_cell_adjoint(dr: CellPressureStokesDR, mesh: MixedMesh, bc_vx, bc_vy, bc_vz, *, objective_v, lam_v, lam_P,
                  tolerance=1e-5, max_iterations=3000, check_interval=100, velocity_step=0.6, pressure_step=0.2,
                  verbose=False, collect_history=False, throw_on_failure=True):
    """solve_adjoint!(dr::CellPressureStokesDR, mesh, bc_v; objective_v, λv, λP, ...)
    (DR_adjoint.jl:446-480): the viscous operator is symmetric, so the forward
    relaxation runs on (lam_v, lam_P) with objective_v as momentum load, zero
    density and gravity and homogeneous conditions on the nodes of the bcs.
    lam_v (three nodal arrays) and lam_P ((nels, 4)) are DeviceArrays updated
    in place (their values are the initial iterate), or host arrays (then
    the results are returned). dr.v and dr.P are unchanged."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    nv, ne = mesh.nnodes, mesh.nels
    if len(objective_v) != 3 or any(_size(o) != nv for o in objective_v):
        raise DimensionMismatch("objective_v must have mesh.nnodes entries per component")
    if len(lam_v) != 3 or any(_size(o) != nv for o in lam_v) or _size(lam_P) != 4 * ne:
        raise DimensionMismatch("velocity must have mesh.nnodes entries per component and pressure must be 4 × nels")
    keep = []
    arr = (_BC * 3)(*(b._c(keep) for b in (bc_vx, bc_vy, bc_vz)))
    p = _CellParams(tolerance=float(tolerance), max_iterations=int(max_iterations),
                    check_interval=int(check_interval), velocity_step=float(velocity_step),
                    pressure_step=float(pressure_step), verbose=int(bool(verbose)),
                    throw_on_failure=int(bool(throw_on_failure)))
    st = _CellStats()
    cap = (int(max_iterations) // max(int(check_interval), 1) + 2) if collect_history else 0
    hist = np.zeros((max(cap, 1), 3))
    nh = I64(0)
    with _Inputs(dr) as io:
        obj = (VP * 3)(*(io._dev(o, (nv,)) for o in objective_v))
        lv = [a if isinstance(a, DeviceArray) else None for a in lam_v]
        lvp = (VP * 3)(*(io._dev(a, (nv,)) for a in lam_v))
        lp = io._dev(lam_P, (ne, 4))
        rc = _lib.fem_cell_stokes_solve_adjoint(dr.handle, arr, obj, lvp, lp, C.byref(p), C.byref(st),
                                                hist.ctypes.data_as(_P(D)) if cap else C.cast(None, _P(D)), cap,
                                                C.byref(nh))
        out = _cell_result(st, hist, nh)
        if rc == FEM_OK or rc == FEM_ENOCONV:
            out["lam_v"] = tuple(a if a is not None else DeviceArray(dr.ctx, lvp[k], (nv,)).numpy()
                                 for k, a in enumerate(lv))
            out["lam_P"] = lam_P if isinstance(lam_P, DeviceArray) else DeviceArray(dr.ctx, lp, (ne, 4)).numpy()
    dr.ctx.check(rc)
    return out


def stokes_material_gradient_3d(dr: CellPressureStokesDR, mesh: MixedMesh, adjoint_velocity, forward_velocity=None):
    """stokes_material_gradient_3d(dr, mesh, adjoint_velocity) (DR_adjoint.jl:483-531):
    dict(density_gradient, viscosity_gradient), one entry per phase; the forward
    velocity is dr.v unless forward_velocity (three nodal arrays) is given."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    nv = mesh.nnodes
    nph = dr.material.nphases if hasattr(dr.material, "nphases") else len(dr.material.eta)
    dens = np.zeros(nph)
    visc = np.zeros(nph)
    with _Inputs(dr) as io:
        lam = (VP * 3)(*(io._dev(a, (nv,)) for a in adjoint_velocity))
        fwd = (VP * 3)(*(io._dev(a, (nv,)) for a in forward_velocity)) if forward_velocity is not None else None
        dr.ctx.check(_lib.fem_cell_stokes_material_gradient(dr.handle, fwd, lam, dens.ctypes.data_as(_P(D)),
                                                            visc.ctypes.data_as(_P(D))))
    return dict(density_gradient=tuple(dens), viscosity_gradient=tuple(visc))


_POST_KEYS = ("exx", "eyy", "ezz", "exy", "eII", "txx", "tyy", "tzz", "txy", "tII")
_POST_KEYS3 = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "eII", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tII")


def compute_strain_rate_stress_postprocess(dr: StokesDR, mesh: MixedMesh):
    """compute_strain_rate_stress_postprocess (postprocess.jl:84-100) on the
    device from dr.v and the point stresses dr.tau (as left by
    update_stokes_current_stress): element averages keyed exx, eyy, ezz, exy,
    eII, txx, tyy, tzz, txy, tII (the reference's εxx ... tauII); 3-D states
    (postprocess.jl:169-212) add exz, eyz, txz, tyz."""
    if mesh is not dr.mesh:
        raise ValueError("dr was built on a different mesh")
    keys = _POST_KEYS if mesh.dim == 2 else _POST_KEYS3
    buf = VP()
    n = len(keys) * mesh.nels
    dr.ctx.check(_lib.fem_buf_alloc(dr.ctx.handle, n * 8, C.byref(buf)))
    try:
        dr.ctx.check(_lib.fem_stokes_postprocess(dr.handle, buf))
        a = DeviceArray(dr.ctx, buf.value, (len(keys), mesh.nels)).numpy()
    finally:
        _lib.fem_buf_free(dr.ctx.handle, buf)
    return dict(zip(keys, a))


from .vtk import Mesh as VtkMesh, write_vtk, write_stokes_vtk  # noqa: E402
