/*
 * femdr.h -- C ABI of libfemdr.so, the MI355X (gfx950) implementation of
 * FEMTools.jl's matrix-free dynamic-relaxation (DR) Stokes hot path.
 *
 * Plain pointers and sizes only. "host_*" arguments are host memory; every
 * other pointer is device memory owned by the library. Connectivity and
 * node/DoF lists cross the boundary 1-based, as Julia holds them; the
 * library converts them once. Element-indexed matrices (Julia NV x nels or
 * nq x nels, column-major) are element-contiguous: C [nels][NV].
 *
 * Every function returns FEM_OK (0) or a negative FEM_E* code; the message
 * of the last failure on a context is fem_last_error(). One context owns one
 * HIP stream on one device; a context is not thread-safe.
 *
 * Reference citations are file:line in FEMTools.jl.
 */
#ifndef FEMDR_H
#define FEMDR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the Julia wrapper maps them to the reference's exceptions */
#define FEM_OK 0
#define FEM_EARG -1       /* ArgumentError                                      */
#define FEM_EDIM -2       /* DimensionMismatch                                  */
#define FEM_ENOCONV -3    /* "Stokes solve did not converge ..." DR.jl:332-335  */
#define FEM_ENAN -4       /* "NaN detected in outer/inner loop ..." DR.jl:574,711 */
#define FEM_EKABOOM -5    /* "Kaboom! Error > 1e10 ..." DR.jl:575,712           */
#define FEM_EPRECOND -6   /* "... preconditioner produced invalid λmax" utils.jl:155 */
#define FEM_EHIP -7       /* HIP runtime failure                                */
#define FEM_ENOMEM -8     /* device allocation failed                           */

#define FEM_MAX_PHASES 64

typedef struct fem_ctx fem_ctx;
typedef struct fem_mesh fem_mesh;
typedef struct fem_stokes fem_stokes;

/* ---------------------------------------------------------------- context
 * Replaces the KernelAbstractions backend object and TA(backend)
 * (src/FEMTools.jl:30-46, ext/FEMToolsCUDAExt.jl:5). */
int fem_ctx_create(int device, fem_ctx** out);
int fem_ctx_destroy(fem_ctx* ctx);

/* Launch choices of a context's solves. No counterpart in the reference (its
 * KernelAbstractions launches have no such parameters); the Python package's
 * autotuner (femtools_amd.autotune) measures them on a problem and writes
 * them as JSON. Changing them between solves is allowed; a batch or solve in
 * progress keeps the values it started with. */
typedef struct {
    int32_t tiles;          /* 1 (default): FP32 inner iterations of structured 2-D meshes run the
                               tile kernel where the cell counts allow; 0: element-slot gather */
    int32_t tile_rows;      /* cell rows per tile of 16 cells: 4, 8 (default) or 16 (blocks of
                               128, 256 or 512 threads); needs ny % tile_rows == 0 */
    int32_t graphs;         /* 1 (default): check intervals replayed from HIP graphs; 0: every
                               inner iteration launched from the host */
    int32_t fold_interval;  /* mixed precision with a yield model: δv folded into the FP64 state
                               every fold_interval iterations between checks (0: at checks only);
                               default 25 */
} fem_tuning;

/* Sets / reads the context's launch choices; FEM_EARG for values outside the
 * ranges above. */
int fem_ctx_set_tuning(fem_ctx* ctx, const fem_tuning* t);
int fem_ctx_get_tuning(fem_ctx* ctx, fem_tuning* t);
int fem_last_error(const fem_ctx* ctx, char* buf, size_t len);
int fem_sync(fem_ctx* ctx);

/* ---------------------------------------------------------------- buffers
 * Back the wrapper's device arrays (KernelAbstractions.zeros, copyto!,
 * Array(.), fill!; stokes_types.jl:268-279). */
int fem_buf_alloc(fem_ctx* ctx, size_t bytes, void** dptr);
int fem_buf_free(fem_ctx* ctx, void* dptr);
int fem_copy_h2d(fem_ctx* ctx, void* dst, const void* host_src, size_t bytes);
int fem_copy_d2h(fem_ctx* ctx, void* host_dst, const void* src, size_t bytes);
int fem_copy_d2d(fem_ctx* ctx, void* dst, const void* src, size_t bytes);
int fem_fill_f64(fem_ctx* ctx, double* dptr, double value, int64_t n);
int fem_fill_f32(fem_ctx* ctx, float* dptr, float value, int64_t n);
int fem_fill_i32(fem_ctx* ctx, int32_t* dptr, int32_t value, int64_t n);

/* ------------------------------------------------------------------- mesh */
typedef struct {
    int32_t dim;        /* 2 or 3                                    */
    int32_t nv_local;   /* 7 (T7), 9 (Q9) or 27 (Hex27)              */
    int32_t np_local;   /* 3 (P1-disc), 4 (P1-disc modes on Hex27), 0 (Q9) */
    int32_t nq;         /* 7 (Dunavant degree 5), 9 (3x3 Gauss) or 27 (3x3x3 Gauss) */
    int64_t nnodes;     /* velocity nodes                            */
    int64_t nels;       /* elements                                  */
    int64_t nnodesP;    /* pressure DoFs = 3 nels                    */
    int64_t n_shared;   /* velocity nodes shared by >1 element slot  */
    int64_t n_boundary; /* nodes on the box boundary (structured)    */
} fem_mesh_info;

/* Mesh(backend, (x0..x1)x(y0..y1), ReferenceElement(QuadraticElement{2,7}), (nx,ny))
 * followed by MixedMesh(mesh, ReferenceElement(LinearElement{2,3}))
 * (src/mesh/mesh.jl:67-101, 412-440; connectivity.jl:111-140; mixed_mesh.jl:173-211,
 * 436-457). Builds coordinates, connectivity, node->element lists, geometry
 * and pressure DoFs on the device. */
int fem_mesh_t7_create(fem_ctx* ctx, int32_t nx, int32_t ny, double x0, double x1,
                       double y0, double y1, fem_mesh** out);
/* Mesh(backend, (x0..x1)x(y0..y1)x(z0..z1), ReferenceElement(QuadraticElement{3,27}), (nx,ny,nz))
 * followed by MixedMesh(mesh, ReferenceElement(LinearElement{3,4})) (mesh.jl:340-398,
 * connectivity.jl:273-322, mixed_mesh.jl:358-417): structured Hex27 velocity mesh,
 * discontinuous pressure with the four modes (1-ξ-η-ζ, ξ, η, ζ) per cell (DoF 4e+a). */
int fem_mesh_hex27_create(fem_ctx* ctx, int32_t nx, int32_t ny, int32_t nz, double x0, double x1,
                          double y0, double y1, double z0, double z1, fem_mesh** out);
/* Mesh(backend, coords, el2n, ReferenceElement(QuadraticElement{3,27})) then
 * MixedMesh(mesh, ReferenceElement(LinearElement{3,4})) (mesh.jl:147-188,
 * mixed_mesh.jl:358-417) from host arrays: coords [nnodes][3], el2n [nels][27]
 * 1-based in the local node order of generate_element2node (connectivity.jl:273-322).
 * Every element must be an axis-aligned box (local ξ, η, ζ along +x, +y, +z) with
 * its nodes at the corner, edge, face and centre points; sizes may differ per
 * element and nodes and elements may come in any order. Γnodes are the nodes of
 * faces owned by one element. FEM_EARG names the first element that breaks this. */
int fem_mesh_hex27_upload(fem_ctx* ctx, const double* coords, int64_t nnodes,
                          const int32_t* el2n_1based, int64_t nels, fem_mesh** out);
/* Mesh(backend, (x0..x1)x(y0..y1), ReferenceElement(QuadraticElement{2,9}), (nx,ny))
 * (mesh.jl:340-398; connectivity.jl:193-227): a structured Q9 mesh on the
 * (2nx+1) x (2ny+1) grid for ThermalDiffusionDR states (heat_diffusion_types.jl:90-153
 * takes any Mesh). It has no pressure space (nnodesP = 0, np_local = 0) and
 * fem_stokes_create rejects it. */
int fem_mesh_q9_create(fem_ctx* ctx, int32_t nx, int32_t ny, double x0, double x1, double y0, double y1,
                       fem_mesh** out);
/* Unstructured T7 mesh: Mesh(backend, coords, el2n, element) (mesh.jl:106-115)
 * + MixedMesh. host_el2n_1based is [nels][7]; local node 7 must be the
 * element's private centroid (bubble) node. */
int fem_mesh_t7_upload(fem_ctx* ctx, const double* host_coords, int64_t nnodes,
                       const int32_t* host_el2n_1based, int64_t nels, fem_mesh** out);
int fem_mesh_get_info(const fem_mesh* mesh, fem_mesh_info* out);
int fem_mesh_get_coords(const fem_mesh* mesh, double* host_out /* [nnodes][dim] */);
int fem_mesh_get_el2n(const fem_mesh* mesh, int32_t* host_out_1based /* [nels][nv_local] */);
/* Γnodes of a structured box mesh (mesh.jl:79-80), 1-based */
int fem_mesh_get_boundary_nodes(const fem_mesh* mesh, int32_t* host_out_1based);
int fem_mesh_destroy(fem_mesh* mesh);

/* --------------------------------------------------------------- material */
/* StokesMaterial (src/stokes/types/stokes_types.jl:19-60) */
typedef struct {
    int32_t nphases;
    int32_t ndim;                      /* length of g: 2              */
    double eta[FEM_MAX_PHASES];
    double etab[FEM_MAX_PHASES];
    double G[FEM_MAX_PHASES];
    double alpha[FEM_MAX_PHASES];
    double rho0[FEM_MAX_PHASES];
    double K[FEM_MAX_PHASES];
    double g[3];
    double Tref;
} fem_material;

/* DruckerPrager (stokes_types.jl:388-423): cos(phi), sin(phi), sin(psi) stored */
typedef struct {
    int32_t nphases;
    double cos_phi[FEM_MAX_PHASES];
    double sin_phi[FEM_MAX_PHASES];
    double sin_psi[FEM_MAX_PHASES];
    double C[FEM_MAX_PHASES];
    double eta_reg[FEM_MAX_PHASES];
    double Kb[FEM_MAX_PHASES];
} fem_drucker_prager;

/* DruckerPragerCap (stokes_types.jl:449-544): Drucker-Prager closed on the
 * tensile side by the circular cap of Popov, Berlie & Kaus (2025). Stored as
 * cos(ϕ), sin(ϕ), sin(Ψ), C, pT (<= 0, compression-positive pressure), η_reg,
 * Kb (> 0, finite), C_min and H_C (<= 0: linear softening with γ). */
typedef struct {
    int32_t nphases;
    double cos_phi[FEM_MAX_PHASES];
    double sin_phi[FEM_MAX_PHASES];
    double sin_psi[FEM_MAX_PHASES];
    double C[FEM_MAX_PHASES];
    double pT[FEM_MAX_PHASES];
    double eta_reg[FEM_MAX_PHASES];
    double Kb[FEM_MAX_PHASES];
    double C_min[FEM_MAX_PHASES];
    double H_C[FEM_MAX_PHASES];
} fem_drucker_prager_cap;

/* ------------------------------------------------------------------ state */
/* StokesDR(mesh, material; CFL_v, CFL_P, c_fact, ϵ) (stokes_types.jl:244-298,
 * 359-369): all fields zero, phases 1, stress at integration points (nq x nels). */
typedef struct {
    double CFL_v;   /* 0.98 */
    double CFL_P;   /* 0.98 */
    double c_fact;  /* 0.9  */
    double eps;     /* 1e-6 */
} fem_state_opts;

int fem_stokes_create(fem_ctx* ctx, fem_mesh* mesh, const fem_material* mat,
                      const fem_state_opts* opts, fem_stokes** out);
int fem_stokes_destroy(fem_stokes* st);

/* StokesDR struct-of-arrays fields (stokes_types.jl:197-241); dr.v.x is
 * FEM_FIELD_V_X, dr.τ.xy is FEM_FIELD_TAU_XY, and so on. */
enum {
    FEM_FIELD_V_X = 0, FEM_FIELD_V_Y,
    FEM_FIELD_DVDTAU_X, FEM_FIELD_DVDTAU_Y,
    FEM_FIELD_RV_X, FEM_FIELD_RV_Y,
    FEM_FIELD_RV0_X, FEM_FIELD_RV0_Y,
    FEM_FIELD_DRVDV_X, FEM_FIELD_DRVDV_Y,
    FEM_FIELD_PC_V_X, FEM_FIELD_PC_V_Y,
    FEM_FIELD_TAU_XX, FEM_FIELD_TAU_YY, FEM_FIELD_TAU_XY,
    FEM_FIELD_TAU_OLD_XX, FEM_FIELD_TAU_OLD_YY, FEM_FIELD_TAU_OLD_XY,
    FEM_FIELD_P, FEM_FIELD_P0, FEM_FIELD_GAMMA_P, FEM_FIELD_T, FEM_FIELD_T0, FEM_FIELD_Q,
    FEM_FIELD_PF, FEM_FIELD_RP, FEM_FIELD_RP0, FEM_FIELD_M_P, FEM_FIELD_PNUM,
    /* 3-D states only (zero length in 2-D): z components and the remaining
     * stress components (VectorField3D, SymmetricTensor3D) */
    FEM_FIELD_V_Z, FEM_FIELD_DVDTAU_Z, FEM_FIELD_RV_Z, FEM_FIELD_RV0_Z, FEM_FIELD_DRVDV_Z, FEM_FIELD_PC_V_Z,
    FEM_FIELD_TAU_ZZ, FEM_FIELD_TAU_XZ, FEM_FIELD_TAU_YZ,
    FEM_FIELD_TAU_OLD_ZZ, FEM_FIELD_TAU_OLD_XZ, FEM_FIELD_TAU_OLD_YZ,
    /* 2-D states (zero length in 3-D), at integration points: the tensile
     * cap's plastic history dr.plastic_history.γ / .θ (stokes_types.jl:83-96,
     * zero: no softening) and the pressure the momentum balance used at the
     * last stress update (the fourth output of update_stokes_current_stress!) */
    FEM_FIELD_GAMMA_PL, FEM_FIELD_THETA_PL, FEM_FIELD_P_IP,
    FEM_FIELD_COUNT
};
/* device pointer and length of one field; dtype: 0 = f64 */
int fem_stokes_field(fem_stokes* st, int32_t field, void** dptr, int64_t* len, int32_t* dtype);

/* Phase layouts accepted by the reference (src/utils.jl:70-74):
 * FEM_PHASE_CELL  1 x nels matrix, FEM_PHASE_ELEMENT  NV x nels matrix,
 * FEM_PHASE_NODAL  per-node (velocity) / per-DoF (pressure) vector. 1-based ids. */
#define FEM_PHASE_CELL 0
#define FEM_PHASE_ELEMENT 1
#define FEM_PHASE_NODAL 2
/* which: 0 = phases_v, 1 = phases_P (dr.phases_v / dr.phases_P, or the
 * phases_v= / phases_P= keywords of solve!, DR.jl:307-310) */
int fem_stokes_set_phases(fem_stokes* st, int32_t which, int32_t layout, const int32_t* host_phases,
                          int64_t n);

/* DirichletBoundaryCondition(DoFs, vals) (boundary_conditions.jl:30-42), 1-based */
typedef struct {
    int64_t n;
    const int32_t* host_dofs_1based;
    const double* host_vals;
} fem_bc;

/* solve!(dr, mesh, bc_vx, bc_vy; ...) keywords (DR.jl:302-319) and the
 * _solve_stokes_dyrel! keywords they forward (DR.jl:366-386). */
typedef struct {
    double dt;
    double pressure_factor;     /* 50                                   */
    double tolerance;           /* ϵ_tol, 1e-6                          */
    int64_t max_iterations;     /* total_iterMax, 50_000                */
    int64_t iter_max;           /* iterMax (inner per PH), 50_000       */
    int64_t check_interval;     /* ncheck, 50                           */
    int64_t max_ph_iterations;  /* 1000                                 */
    double rel_drop0;           /* 1e-2                                 */
    int32_t finite_K;           /* pressure storage modulus K (1) or ηb (0) */
    int32_t freeze_jacobian;    /* -1 = reference default (plastic === nothing) */
    int32_t throw_on_failure;   /* 1: non-convergence returns FEM_ENOCONV */
    int32_t use_tau_old;        /* 1: read dr.τ_old (default), 0: τ_old = nothing */
    int32_t skip_scaling;       /* 1: γP, M_P already assembled (solve_stokes_dyrel!) */
    int32_t verbose;
    int32_t precision;          /* 0 = FP64 throughout; 1 = FP64 state, outer
                                   residuals, norms and convergence test, inner
                                   DR iterations in FP32 on a correction δ of the
                                   linearised visco-elastic problem R(x) + L δ = 0,
                                   folded into x at every outer iteration (with a
                                   Drucker-Prager model: two FP64 returns per point
                                   near yield); 2 = as 1, with the 2-D return
                                   linearised about x at every fold (3-D: as 1) */
    int32_t time_kernels;       /* 1: HIP events around every element / gather launch */
    int32_t measure_lmax;       /* measure_λmax (DR.jl:210-211, 375-378): 1 = power iteration on the
                                   Jacobi-scaled velocity operator at every Jacobian assembly,
                                   λmax = min(Gershgorin, safety · measured); 2-D FP64 solves only */
    int32_t lmax_power_iterations;  /* λmax_power_iterations, 12 (0 reads as 12)  */
    double lmax_power_rtol;     /* λmax_power_rtol, 1e-2 (0 reads as 1e-2)           */
    double lmax_safety;         /* λmax_safety, 1.1 (0 reads as 1.1); ≥ 1.5 with a yield model */
} fem_solve_params;

/* stats NamedTuple (DR.jl:783-801) + device timing */
typedef struct {
    int64_t itPH, iter;
    double err, err_abs, err_rel, err_v, err_P;
    int32_t converged, reached_total_iter;
    double lambda_max, lambda_max_gershgorin;
    int64_t lambda_max_iterations, jacobian_assemblies;
    double seconds;             /* wall time of the solve (host clock, after a device sync) */
    double kernel_ms_element;   /* summed device time of the element kernel over the timed
                                   launches (time_kernels: every 16th inner iteration) */
    double kernel_ms_gather;    /* summed device time of the gather/update kernel          */
    int64_t element_launches;   /* inner-iteration element launches                        */
    int64_t timed_launches;     /* launches covered by kernel_ms_*                          */
    double kernel_ms_element_rm;  /* mixed precision: the part of kernel_ms_element timed on
                                     check intervals with the return-map correction operator */
    double kernel_ms_gather_rm;
    int64_t timed_launches_rm;
} fem_solve_stats;

/* solve!(::StokesDR, ::MixedMesh, bc_v; dt, ...) DR.jl:302-337: assembles
 * M_P and γP (pressure_scaling.jl:8-149) and runs the Powell-Hestenes/DYREL
 * loop of DR.jl:348-802 on the device. plastic may be NULL. */
int fem_stokes_solve(fem_stokes* st, const fem_bc bcs[2], const fem_drucker_prager* plastic,
                     const fem_solve_params* params, fem_solve_stats* stats);

/* The optional array keywords of solve! / solve_stokes_dyrel! (DR.jl:183-186,
 * 217-226, 302-337, 366-418, 556-561, 722-727). Device arrays (f64, NULL when
 * absent) stay attached to the state until the next fem_stokes_set_solve_inputs
 * call (NULL clears them) and are read by fem_stokes_solve, fem_stokes_solve_cap,
 * fem_stokes_residual(_cap), fem_stokes_jacobian_diagnostics(_cap) and
 * fem_stokes_pressure_scaling, so a binding sets them around one call:
 *   body_force[c]      nodal loads f_i = ∫N_i b dΩ (nnodes each), subtracted from
 *                      the momentum residual before the Dirichlet rows (2-D, 3-D)
 *   viscosity          nq x nels = [nels][nq] η at the velocity quadrature points
 *                      in place of the phase viscosities (2-D, purely viscous:
 *                      G = Inf, no yield model, FP64, not coupled)
 *   scaling_viscosity  [nels][nq] viscosity of the pressure scale only; NULL:
 *                      `viscosity` (the phase viscosities when that is NULL too)
 *   P_old              accepted previous-step pressure in the continuity rate in
 *                      place of dr.P0: nodal (nnodesP) or, with P_old_at_points,
 *                      [nels][nq] at the quadrature points (2-D)
 *   history            host buffer of history_cap rows (iter, err_v, err_v_x,
 *                      err_v_y, err_v_z, err_P) filled at every outer and inner
 *                      check (collect_history); *history_len receives the count
 *                      (FP64 solves; rows beyond history_cap are dropped) */
typedef struct {
    const double* body_force[3];
    const double* viscosity;
    const double* scaling_viscosity;
    const double* P_old;
    int32_t P_old_at_points;
    double* history;
    int64_t history_cap;
    int64_t* history_len;
} fem_solve_inputs;
int fem_stokes_set_solve_inputs(fem_stokes* st, const fem_solve_inputs* inputs);

/* _set_pressure_source!(dr.Q, Qq, Q2D, ...) (pressure_residual.jl:32-77), run by
 * solve_stokes_dyrel! when Qq is given (DR.jl:451-457): dr.Q = Qq (device,
 * nnodesP) scaled so that ∫ Q dΩ = Q2D; has_Q2D = 0 copies Qq unscaled. Qq must
 * be finite and nonnegative; zero support with a nonzero Q2D is an error. 2-D. */
int fem_stokes_set_pressure_source(fem_stokes* st, const double* Qq, double Q2D, int32_t has_Q2D);

/* n independent solve! calls (DR.jl:302-337) on one context (stream). No
 * counterpart in the reference, which solves one state per call; for callers
 * with many problems (parameter sweeps, ensembles). Mixed-precision 2-D
 * solves without a yield model on structured box meshes whose cell counts are
 * multiples of 16 x 8 (and check_interval >= 2) run together: the inner
 * iterations of all of them are grouped launches (one element and one
 * grid-node update launch per iteration for every problem), a check interval
 * captured as one HIP graph; each problem keeps the decisions and iteration
 * counts of its single solve, and leaves the group when it finishes. Other
 * problem kinds are solved one after the other. All states must share one
 * fem_ctx; nthreads is ignored (no host threads). bcs[i] holds the state's dim
 * conditions, plastic may be NULL or hold NULL entries, params/stats/status
 * are arrays of n; status[i] is the return code of solve i; the call returns
 * the first nonzero status in state order, FEM_OK if all succeed. */
int fem_stokes_solve_batch(int32_t n, fem_stokes* const* states, const fem_bc* const* bcs,
                           const fem_drucker_prager* const* plastic, const fem_solve_params* params,
                           fem_solve_stats* stats, int32_t* status, int32_t nthreads);

/* assemble_viscosity_weighted_pressure_scaling! (pressure_scaling.jl:118-216):
 * fills M_P and γP. */
int fem_stokes_pressure_scaling(fem_stokes* st, double pressure_factor, double dt, int32_t finite_K);

/* update_stokes_current_stress! (DR.jl:820-897): τ at integration points
 * from the current v, P and τ_old; overwrites Rv like the reference. */
int fem_stokes_update_current_stress(fem_stokes* st, const fem_drucker_prager* plastic, double dt);

/* Element-level entry points used by parity tests.
 * fem_stokes_residual: RP (pressure_residual.jl:256-294) and Rv
 * (momentum_residuals.jl:417-470) at the current state; with_pnum=1 forms
 * Pnum = γP RP / M_P first (DR.jl:674) and uses it in the momentum residual. */
int fem_stokes_residual(fem_stokes* st, const fem_drucker_prager* plastic, double dt,
                        int32_t with_pnum, int32_t finite_K);
/* assemble_augmented_momentum_jacobian_matrices_atomix! (momentum_residuals.jl
 * :998-1117): ∂Rv∂v row sums and PC_v diagonal. */
int fem_stokes_jacobian_diagnostics(fem_stokes* st, const fem_drucker_prager* plastic, double dt,
                                    int32_t finite_K);

/* ------------------------------------------------ frozen-operator adjoint
 * solve_adjoint!(dr::StokesDR, mesh, bc_vx, bc_vy; dt, objective_vx, objective_vy,
 * λvx, λvy, λP, tolerance=1e-6, max_iterations=50_000, check_interval=50,
 * throw_on_failure=true, iterMax, max_ph_iterations, rel_drop, measure_λmax)
 * (src/stokes/solvers/DR_adjoint.jl:396-427 -> _solve_stokes_adjoint_dyrel!
 * :1-268) with operator = :blocks (adjoint_operator.jl:1-660), on a 2-D state
 * without a yield model or with the Drucker-Prager model `plastic` (nullable;
 * the tensile cap is not covered): solves (∂R/∂u)ᵀλ = -∂J/∂u at the converged
 * forward state in the state (v, P, τ_old, dr.γP and dr.M_P of the forward
 * solve, the phases and material), with homogeneous conditions on the bc nodes
 * (values unused). CFL_v and c_fact are the state's. */
typedef struct {
    double dt;
    double tolerance;           /* 1e-6                                         */
    int64_t max_iterations;     /* total inner-iteration budget, 50_000         */
    int64_t check_interval;     /* 50                                           */
    int64_t iter_max;           /* inner iterations per Powell-Hestenes step, 50_000 */
    int64_t max_ph_iterations;  /* 100                                          */
    double rel_drop;            /* 0.1                                          */
    int32_t measure_lmax;       /* 1: power iteration (estimate_adjoint_λmax); 0: Gershgorin */
    int32_t throw_on_failure;   /* 1: non-convergence returns FEM_ENOCONV       */
    int32_t verbose;
} fem_adjoint_params;

/* (; itPH, iter, err, err_v, err_P, converged, λmax, λmax_gershgorin,
 * λmax_iterations) of the adjoint solve (DR_adjoint.jl:253-267) + wall time */
typedef struct {
    int32_t converged;
    int64_t itPH, iter, lmax_iterations;
    double err, err_v, err_P, lmax, lmax_gershgorin, seconds;
} fem_adjoint_stats;

/* objective_x/y: ∂J/∂v (device, nnodes each); lam_x/y (nnodes), lam_P (nnodesP):
 * device arrays holding the initial iterate, overwritten with the adjoint.
 * FEM_EARG for a 3-D state, an unassembled dr.γP (the forward solve assembles
 * it) or a tangent that breaks the symmetric layout; FEM_ENOCONV as solve!. */
int fem_stokes_solve_adjoint(fem_stokes* st, const fem_bc* bc_v /* [2] */, const fem_drucker_prager* plastic,
                             const double* objective_x, const double* objective_y, double* lam_x, double* lam_y,
                             double* lam_P, const fem_adjoint_params* params, fem_adjoint_stats* stats);
/* assemble_adjoint_operator (adjoint_operator.jl:412-475): the element blocks at
 * the current state, to host arrays A [nels][14][14] (∂Rv/∂v with the
 * augmentation), B [nels][14][3] (∂Rv/∂P), C [nels][3][14] (∂RP/∂v), D [nels][3][3]
 * (∂RP/∂P, the ηb storage), W [nels][14][3] (∂Rv/∂Pnum · γP/M_P); velocity DoFs
 * stacked by component (vx 0..6, vy 7..13). Also fills ∂Rv∂v and PC_v. */
int fem_stokes_adjoint_blocks(fem_stokes* st, const fem_drucker_prager* plastic, double dt, double* host_A,
                              double* host_B, double* host_C, double* host_D, double* host_W);

/* Sweeps the element kernel alone (no control flow) n times; used by the
 * bench to time the dominant kernel with device events. Returns the average
 * kernel time in milliseconds through avg_ms. */
int fem_stokes_time_element_kernel(fem_stokes* st, const fem_drucker_prager* plastic, double dt,
                                   int32_t n, double* avg_ms);

/* 3-D states: the FP32 correction step of the mixed-precision solve
 * (k_hex27_lin and the shared-node gather k_gather3f) swept n times on a zero
 * correction, every element taking the FP64 returns when plastic is given;
 * the average device time in milliseconds through avg_ms. Not a reference
 * entry point: the bench's device-event timing of that solve's hot kernel. */
int fem_stokes_time_correction_kernel(fem_stokes* st, const fem_drucker_prager* plastic, double dt, int32_t n,
                                      double* avg_ms);

/* ------------------------------------------------------- thermal diffusion
 * ThermalMaterial (src/heat_diffusion/types/heat_diffusion_types.jl:15-41) */
typedef struct {
    int32_t nphases;
    double k[FEM_MAX_PHASES];      /* conductivity          */
    double Cp[FEM_MAX_PHASES];     /* heat capacity         */
    double rho0[FEM_MAX_PHASES];   /* reference density     */
    double alpha[FEM_MAX_PHASES];  /* thermal expansivity   */
    double K[FEM_MAX_PHASES];      /* bulk modulus (Inf ok) */
} fem_thermal_material;

/* ThermalDiffusionDR(mesh, material; CFL, c_fact, ϵ) (heat_diffusion_types.jl:
 * 90-153, mesh constructor :141-153): nodal fields, zero; phases all 1. The
 * mesh may be T7 (generated or uploaded), a Q9 box or a Hex27 box. */
typedef struct {
    double CFL;     /* 0.98 */
    double c_fact;  /* 0.9  */
    double eps;     /* 1e-6 (default tolerance) */
} fem_thermal_opts;

typedef struct fem_thermal fem_thermal;
int fem_thermal_create(fem_ctx* ctx, fem_mesh* mesh, const fem_thermal_material* mat,
                       const fem_thermal_opts* opts, fem_thermal** out);
int fem_thermal_destroy(fem_thermal* th);

/* dr.R, dr.R0, dr.∂R∂T, dr.PC, dr.T, dr.T0, dr.∂T∂τ, dr.P, dr.source (f64),
 * dr.phases (i32, 1-based per node) */
enum {
    FEM_TFIELD_R = 0, FEM_TFIELD_R0, FEM_TFIELD_DRDT, FEM_TFIELD_PC, FEM_TFIELD_T, FEM_TFIELD_T0,
    FEM_TFIELD_DTDTAU, FEM_TFIELD_P, FEM_TFIELD_SOURCE, FEM_TFIELD_PHASES, FEM_TFIELD_COUNT
};
/* device pointer and length of one field; dtype: 0 = f64, 1 = i32 */
int fem_thermal_field(fem_thermal* th, int32_t field, void** dptr, int64_t* len, int32_t* dtype);

/* solve!(dr::ThermalDiffusionDR, mesh, bc; dt, tolerance=dr.ϵ,
 * max_iterations=10_000, check_interval=100, throw_on_failure=true, Tref=273)
 * (heat_diffusion/solvers/DR.jl:25-54) */
typedef struct {
    double dt;
    double tolerance;           /* < 0: dr.ϵ                                  */
    int64_t max_iterations;     /* 10_000                                     */
    int64_t check_interval;     /* 100                                        */
    double Tref;                /* 273                                        */
    int32_t throw_on_failure;   /* 1: non-convergence returns FEM_ENOCONV     */
    int32_t verbose;
} fem_thermal_params;

/* (; converged, iterations, residual) of solve_dynamic_relaxation!
 * (dynamic_relaxation/solver.jl:61-122) + wall time */
typedef struct {
    int32_t converged;
    int64_t iterations;
    double residual;            /* last relative residual ‖R‖/‖R⁰‖ */
    double seconds;
} fem_thermal_stats;

/* solve_dynamic_relaxation! over the thermal residual; bc: one Dirichlet set
 * (may be NULL or empty). */
int fem_thermal_solve(fem_thermal* th, const fem_bc* bc, const fem_thermal_params* params,
                      fem_thermal_stats* stats);

/* assemble_diffusion_matrices_atomix! (heat_diffusion/assembly/residual.jl:90-121):
 * dr.R, and dr.∂R∂T / dr.PC (|J| row sums / |J_ii|) when compute_jacobian. */
int fem_thermal_residual(fem_thermal* th, double dt, double Tref, int32_t compute_jacobian);

/* compute_strain_rate_stress_postprocess(vx, vy, el2n_v, geo_v, τ_ip, element_v)
 * (src/postprocess/postprocess.jl:17-100) on the state's v and τ (the point
 * stresses of the last update_stokes_current_stress!): per element the
 * dΩ-weighted averages of (ε̇xx, ε̇yy, ε̇zz, ε̇xy, ε̇II, τxx, τyy, τzz, τxy, τII)
 * into out (device, [10][nels]) for 2-D states; for 3-D states (postprocess.jl:
 * 169-212) the 14 averages (ε̇xx, ε̇yy, ε̇zz, ε̇xy, ε̇xz, ε̇yz, ε̇II, τxx, τyy, τzz,
 * τxy, τxz, τyz, τII) into [14][nels]. */
int fem_stokes_postprocess(fem_stokes* st, double* out);

/* rotate_stress!(dr, mesh_stokes, Δt) (src/stokes/stress_rotation.jl:16-61):
 * τ_old at every velocity quadrature point becomes the current τ rotated by
 * the local vorticity over Δt (GeoParams.rotate_elastic_stress: θ = ωΔt with
 * ω = ½(∂vx/∂y − ∂vy/∂x) in 2-D; the rotation vector curl(v)Δt/2 in 3-D).
 * 2-D and 3-D states. */
int fem_stokes_rotate_stress(fem_stokes* st, double dt);

/* ------------------------------------------------------ tensile cap (2-D) */
/* The entry points above with plastic::DruckerPragerCap (rheology.jl:510-787):
 * the visco-elastic trial stress returned by the local Newton solve of the
 * cap (Popov et al. 2025, Eq. 42), the momentum balance with the corrected
 * pressure, the Jacobian diagnostics with the return's implicit derivative.
 * The cohesion is softened by FEM_FIELD_GAMMA_PL. 2-D states, FP64
 * (precision 0). solve!'s default finite_K for the cap is 1 (DR.jl:312). */
int fem_stokes_solve_cap(fem_stokes* st, const fem_bc bcs[2], const fem_drucker_prager_cap* cap,
                         const fem_solve_params* params, fem_solve_stats* stats);
int fem_stokes_residual_cap(fem_stokes* st, const fem_drucker_prager_cap* cap, double dt, int32_t with_pnum,
                            int32_t finite_K);
int fem_stokes_jacobian_diagnostics_cap(fem_stokes* st, const fem_drucker_prager_cap* cap, double dt,
                                        int32_t finite_K);
/* τ at the points and the corrected pressure into FEM_FIELD_P_IP */
int fem_stokes_update_current_stress_cap(fem_stokes* st, const fem_drucker_prager_cap* cap, double dt);
/* commit_stokes_plastic_history! (DR.jl:899-958): γ += Δt γdot, θ += Δt θdot
 * at every point from the local update at the current v, P (softened by the
 * γ before the call); FEM_ENAN if a committed value is non-finite. */
int fem_stokes_commit_plastic_history(fem_stokes* st, const fem_drucker_prager_cap* cap, double dt);

/* ------------------------------------------- CellPressureStokesDR (Hex27) */
/* CellPressureStokesDR(mesh, material; phases) (stokes_types.jl:545-613) is a
 * 3-D fem_stokes on a Hex27 mesh with a purely viscous, incompressible
 * material (G, K infinite, α zero; cell phases): dr.v is the velocity, dr.P
 * and dr.RP hold the cell modes (1, ξ, η, ζ) [nels][4], dr.PC_v the Jacobi
 * diagonal and dr.M_P the mode mass. */
/* stokes_preconditioner_3d! (momentum_residuals.jl:715-745): PC_v, M_P */
int fem_cell_stokes_preconditioner(fem_stokes* st);
/* assemble_stokes_pressure_residual_3d! (pressure_residual.jl:337-366): RP from v */
int fem_cell_stokes_pressure_residual(fem_stokes* st);
/* assemble_stokes_momentum_residual_3d! (momentum_residuals.jl:645-691): Rv from v, P */
int fem_cell_stokes_momentum_residual(fem_stokes* st);

/* solve!(dr::CellPressureStokesDR, mesh, bc_v; ...) keywords (DR.jl:960-996) */
typedef struct {
    double tolerance;           /* 1e-5  */
    int64_t max_iterations;     /* 3000  */
    int64_t check_interval;     /* 100   */
    double velocity_step;       /* 0.6   */
    double pressure_step;       /* 0.2   */
    int32_t verbose;
    int32_t throw_on_failure;   /* 1: non-convergence returns FEM_ENOCONV */
} fem_cell_solve_params;

/* (; converged, iterations, residual, err_v, err_P, reached_total_iter) + wall time */
typedef struct {
    int32_t converged, reached_total_iter;
    int64_t iterations;
    double residual, err_v, err_P;
    double seconds;
} fem_cell_solve_stats;

/* the Arrow-Hurwicz / Jacobi relaxation of _relax_cell_pressure_stokes!
 * (DR.jl:997-1057) from the state's v and P; bcs[3] one Dirichlet set per
 * component. history (nullable, history_cap rows of (iter, err_v, err_P))
 * records every check (collect_history); *history_len receives the count. */
int fem_cell_stokes_solve(fem_stokes* st, const fem_bc bcs[3], const fem_cell_solve_params* params,
                          fem_cell_solve_stats* stats, double* history, int64_t history_cap, int64_t* history_len);

/* solve_adjoint!(dr::CellPressureStokesDR, mesh, bc_v; objective_v, λv, λP, ...)
 * (src/stokes/solvers/DR_adjoint.jl:446-480): the same relaxation on the
 * caller's device arrays lam_v (three nodal) and lam_P ([nels][4]), whose
 * values are the initial iterate, with objective (three nodal device arrays)
 * as momentum load, zero density and gravity and homogeneous conditions on
 * the nodes of bcs (their values are not used). dr.v and dr.P are untouched;
 * the state's scratch (Rv, RP, PC_v, M_P) is overwritten. */
int fem_cell_stokes_solve_adjoint(fem_stokes* st, const fem_bc bcs[3], const double* const objective[3],
                                  double* const lam_v[3], double* lam_P, const fem_cell_solve_params* params,
                                  fem_cell_solve_stats* stats, double* history, int64_t history_cap,
                                  int64_t* history_len);
/* stokes_material_gradient_3d (DR_adjoint.jl:483-531): per phase p,
 * density_gradient[p] = −λ·R(v = 0, P = 0, ρ = e_p) and viscosity_gradient[p] =
 * −λ·R(v_fwd, P = 0, η = e_p, ρ = 0) with R the assembled momentum residual;
 * v_fwd (three nodal device arrays) NULL: dr.v. Host outputs of nphases. */
int fem_cell_stokes_material_gradient(fem_stokes* st, const double* const v_fwd[3], const double* const lam_v[3],
                                      double* density_gradient, double* viscosity_gradient);

/* ------------------------------------------------- coupled thermal--Stokes */
/* assemble_shear_heating! (src/stokes/assemblers/shear_heating.jl:18-103):
 * phi (device, nq x nels f64: [nels][7] on T7, [nels][27] on Hex27) receives
 * τ : (ε̇ − (τ − τ_old)/(2GΔt)) at the velocity quadrature points for the
 * state's current v, P, τ_old (plane strain with τzz = −(τxx + τyy), or 3-D). */
int fem_stokes_shear_heating(fem_stokes* st, const fem_drucker_prager* plastic, double dt, double* phi);

/* solve_coupled! keywords beyond those of solve! (DR.jl:81-91) */
typedef struct {
    double Tref;                /* thermal reference temperature, 273            */
    int32_t shear_heating;      /* 1: τ : (ε̇ − ε̇ᵉˡ) added to the heat source,
                                   refreshed at every thermal check             */
} fem_coupled_params;

/* the fields solve_coupled! merges into the Stokes stats (DR.jl:801) */
typedef struct {
    double err_T;               /* last thermal ‖R‖/‖R⁰‖ (inf before the first check) */
    int64_t thermal_iterations;
    int32_t thermal_converged;
} fem_coupled_stats;

/* solve_coupled!(thermal, stokes, thermal_mesh, stokes_mesh, bc_T, bc_v; dt, ...)
 * (src/stokes/solvers/DR.jl:81-131, coupled loop DR.jl:514-661): the thermal
 * field is transferred to the Stokes pressure DoFs (T and T0), then every
 * inner Stokes iteration first advances one thermal DR iteration (tolerance
 * th's ϵ, check cadence params->check_interval) and transfers T again. Both
 * loop tests also wait for the thermal test; stats->converged requires it.
 * Both states must share one fem_ctx and a node layout (a T7 or a Hex27
 * mesh); bc_v holds one condition per velocity component (2 or 3);
 * precision 0. The thermal source field is left unchanged. */
int fem_coupled_solve(fem_thermal* th, fem_stokes* st, const fem_bc* bc_T, const fem_bc* bc_v,
                      const fem_drucker_prager* plastic, const fem_solve_params* params,
                      const fem_coupled_params* coupled, fem_solve_stats* stats, fem_coupled_stats* cstats);

/* Residual norms of a DR convergence check, kernel level (DR.jl:556-557,
 * 700: err_P = norm(RP ./ M_P)/sqrt(nnodesP), err_v = maximum(norm, Rv)/(2
 * sqrt(nnodes))). Device inputs Rv_x, Rv_y (nv each), RP, M_P (np each);
 * host_out[5] = (‖Rv_x‖₂, ‖Rv_y‖₂, ‖RP./M_P‖₂, err_v, err_P). The sums are
 * FP64 in a fixed block partition (run-to-run reproducible). Runs on the
 * context's stream and synchronises it. */
int fem_norms(fem_ctx* ctx, const double* Rv_x, const double* Rv_y, int64_t nv, const double* RP,
              const double* M_P, int64_t np, double* host_out);

/* fem_norms without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (work under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The block
 * partials live in a static device array of 2048 x 3 doubles per GPU (48 KB,
 * held until the process ends), so nv and np are limited to 2048 * 4096 =
 * 8,388,608 entries each (larger inputs return FEM_EARG; fem_norms takes
 * them). Same partition and summation order as fem_norms; dev_out[5] (device)
 * receives (‖Rv_x‖₂, ‖Rv_y‖₂, ‖RP./M_P‖₂, err_v, err_P). Enqueued on the
 * context's stream, not synchronised; every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (several streams or host threads)
 * race. */
int fem_norms_noalloc(fem_ctx* ctx, const double* Rv_x, const double* Rv_y, int64_t nv, const double* RP,
                      const double* M_P, int64_t np, double* dev_out);

/* ---------------------------------------------------------------------------
 * Learned warm start (no counterpart in the reference, which takes any initial
 * guess and lets the DR iteration decide the answer,
 * examples/miniapps/stokes/stokes_2D_pure_shear_triangle/
 * stokes_2D_pure_shear_triangle.jl:159-163). A 4-level encoder-decoder of 14
 * 3x3 convolutions (channels 32, 64, 128, 128; ReLU after all but the head)
 * on H x W images of 16 input features (NHWC, FP32 on the device), three
 * outputs per pixel. BF16 weights and activations, FP32 accumulation on the
 * matrix cores, FP32 master weights and Adam moments. Layer l's weight is
 * [Co][3][3][Ci] (tap (ky, kx) reads pixel (y + ky - 1, x + kx - 1)); host
 * weight arrays concatenate the 14 layers in forward order
 * (e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a head).
 * --------------------------------------------------------------------------- */
typedef struct fem_cnn fem_cnn;
/* H, W multiples of 8; buffers for batches up to max_batch. Weights start at zero. */
int fem_cnn_create(fem_ctx* ctx, int32_t max_batch, int32_t H, int32_t W, fem_cnn** out);
int fem_cnn_destroy(fem_cnn* net);
/* total weight and bias counts of the 14 layers */
int fem_cnn_parameter_count(int64_t* weights, int64_t* biases);
/* FP32 master weights/biases to/from the host (set_weights also resets the Adam moments) */
int fem_cnn_set_weights(fem_cnn* net, const float* host_W, const float* host_b);
int fem_cnn_get_weights(fem_cnn* net, float* host_W, float* host_b);
/* gradients of the last fem_cnn_train_step */
int fem_cnn_get_grads(fem_cnn* net, float* host_gW, float* host_gb);
/* y[B][H][W][3] = net(x[B][H][W][16]), device pointers, on the context stream (not synchronised) */
int fem_cnn_forward(fem_cnn* net, const float* x, int32_t B, float* y);
/* forward, loss = mean over B, pixels and 3 channels of (y - t)^2 with t
 * [B][H][W][3] (device), backward, and with lr > 0 one Adam step (step >= 1
 * is the step count of the bias correction). Synchronises; *loss on the host. */
int fem_cnn_train_step(fem_cnn* net, const float* x, const float* t, int32_t B, double lr, int32_t step,
                       double* loss);
/* C[M][N] = sum_k A[m][k] Bt[n][k] of bfloat16 (raw bits) device matrices, FP32
 * out: the matrix-core product every convolution of the network uses. K a
 * multiple of 16. Not synchronised. */
int fem_cnn_gemm_bf16(fem_ctx* ctx, int64_t M, int64_t N, int64_t K, const uint16_t* A, const uint16_t* Bt, float* C);

/* update_stokes_velocity_kernel! (src/stokes/helpers.jl:31-41) for one
 * velocity component with the negated step of the solver (DR.jl:690-695):
 * rate = beta rate + R / PC; v -= alpha rate. Device arrays of n; enqueued on
 * the context's stream. */
int fem_dr_velocity_update(fem_ctx* ctx, int64_t n, double* rate, double* v, const double* R, const double* PC,
                           double alpha, double beta);

/* Points of fem_dp_return_map_batch: device arrays of n, inputs then outputs. */
typedef struct {
    const double *exx, *eyy, *exy;              /* strain rate (ε̇xy the tensor shear)   */
    const double *txx_old, *tyy_old, *txy_old;  /* stress of the previous time step      */
    const double* P;                            /* pressure of the yield function        */
    const double *eta, *G, *C, *phi, *psi, *eta_reg, *Kb;   /* the point's material (angles
                                                   in radians; G = Inf: viscous)         */
    double *txx, *tyy, *txy;                    /* out: the returned stress              */
    int32_t* yields;                            /* out, nullable: 1 where the point yields */
} fem_dp_points;

/* The plane-strain visco-elasto-plastic stress with the Drucker-Prager
 * return at n independent points (deviatoric_stress + the return of
 * _deviatoric_stress_with_multiplier, src/stokes/assemblers/rheology.jl:35-42,
 * 226-245, 272-334), the computation every quadrature point of the element
 * kernel does. FP64; enqueued on the context's stream. */
int fem_dp_return_map_batch(fem_ctx* ctx, int64_t n, double dt, const fem_dp_points* points);

#ifdef __cplusplus
}
#endif
#endif
