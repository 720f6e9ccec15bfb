/*
 * femdr_oracle.h -- CPU FP64 restatement of FEMTools.jl's matrix-free
 * dynamic-relaxation (DR) Stokes hot path, used ONLY as a test checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library. The product path (libfemdr.so)
 * never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates
 * (paths relative to the FEMTools.jl repository root).
 *
 * Conventions: 0-based indices; element-node connectivity el2n is [nels][NV]
 * (element-contiguous, the C view of Julia's NV x nels column-major matrix);
 * integration-point arrays are [nels][nq]; pressure DoF a of element e is
 * 3*e + a (src/mesh/mixed_mesh.jl:447-454).
 */
#ifndef FEMDR_ORACLE_H
#define FEMDR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_PHASES 64

/* StokesMaterial (src/stokes/types/stokes_types.jl:19-60). */
typedef struct {
    int32_t nphases;
    double eta[ORC_MAX_PHASES];
    double etab[ORC_MAX_PHASES];
    double G[ORC_MAX_PHASES];
    double alpha[ORC_MAX_PHASES];
    double rho0[ORC_MAX_PHASES];
    double K[ORC_MAX_PHASES];
    double g[2];
    double Tref;
} orc_material;

/* DruckerPrager (src/stokes/types/stokes_types.jl:388-423): stored as
 * cos(phi), sin(phi), sin(psi), C, eta_reg, Kb per phase. */
typedef struct {
    int32_t nphases;
    double cos_phi[ORC_MAX_PHASES];
    double sin_phi[ORC_MAX_PHASES];
    double sin_psi[ORC_MAX_PHASES];
    double C[ORC_MAX_PHASES];
    double eta_reg[ORC_MAX_PHASES];
    double Kb[ORC_MAX_PHASES];
} orc_dp;

/* DruckerPragerCap (src/stokes/types/stokes_types.jl:449-544): the
 * Drucker-Prager shear line closed on the tensile side by the circular cap of
 * Popov, Berlie & Kaus (2025); pT <= 0 (compression-positive pressure). */
typedef struct {
    int32_t nphases;
    double cos_phi[ORC_MAX_PHASES];
    double sin_phi[ORC_MAX_PHASES];
    double sin_psi[ORC_MAX_PHASES];
    double C[ORC_MAX_PHASES];
    double pT[ORC_MAX_PHASES];
    double eta_reg[ORC_MAX_PHASES];
    double Kb[ORC_MAX_PHASES];
    double C_min[ORC_MAX_PHASES];
    double H_C[ORC_MAX_PHASES];
} orc_cap;

/* Phase layouts (src/utils.jl:70-74): */
#define ORC_PHASE_CELL 0      /* one phase per element (1 x nels matrix)        */
#define ORC_PHASE_ELEMENT 1   /* one phase per element-local node (NV x nels)  */
#define ORC_PHASE_NODAL 2     /* one phase per global node / pressure DoF       */

/* Whole Stokes problem on a 2-D T7/P1-disc mixed mesh. All arrays are host
 * memory owned by the caller. Nullable pointers are marked. */
typedef struct {
    int64_t nnodes, nels, nnodesP;
    const double* coords;   /* [nnodes][2] */
    const int32_t* el2n;    /* [nels][7], 0-based */
    const double* invJ;     /* [nels][7][4] row-major J^-1 per quadrature point */
    const double* dOmega;   /* [nels][7] |det J| w (velocity geometry)         */
    const double* dOmegaP;  /* [nels][7] |det J| w (pressure geometry)         */
    const int32_t* phases_v; int32_t phase_mode_v;  /* 1-based phase ids */
    const int32_t* phases_P; int32_t phase_mode_P;
    orc_material mat;
    const orc_dp* dp;       /* nullable: purely visco-elastic                  */
    /* state (StokesDR fields, src/stokes/types/stokes_types.jl:197-298) */
    double *vx, *vy, *rate_x, *rate_y, *Rvx, *Rvy, *Rv0x, *Rv0y;
    double *dRdvx, *dRdvy, *PCx, *PCy;
    double *P, *P0, *gammaP, *T, *T0, *Q, *RP, *MP, *Pnum;
    const double* Pf;       /* nullable (zero fluid pressure)                  */
    const double* tau_old[3]; /* nullable each: [nels][7] xx, yy, xy           */
    double* tau_store[3];     /* nullable: receives [nels][7] current stress   */
    /* Dirichlet conditions per component */
    int64_t nbc[2];
    const int32_t* bc_dofs[2];
    const double* bc_vals[2];
    /* tensile-cap yield model (nullable; replaces dp when set) and its
     * integration-point history (rheology.jl:689-760, DR.jl:899-958) */
    const orc_cap* cap;
    double* gamma_hist;       /* nullable [nels][7]: accepted plastic strain γ (softens C) */
    double* theta_hist;       /* nullable [nels][7]: accepted volumetric plastic strain θ */
    double* P_store;          /* nullable [nels][7]: pressure of the stress update (cap-corrected) */
    /* array keywords of solve! (DR.jl:183-186, 217-226, 302-337, 556-561, 722-727), nullable */
    const double* body_force[2];  /* nodal loads subtracted from Rv before the Dirichlet rows */
    const double* eta_ip;         /* [nels][7] viscosity at the points (purely viscous flow) */
    const double* eta_scale_ip;   /* [nels][7] viscosity of the pressure scale; null: eta_ip */
    const double* Pold_ip;        /* [nels][7] previous-step pressure at the points (continuity) */
    double* history;              /* rows (iter, err_v, err_v_x, err_v_y, 0, err_P) of every check */
    int64_t history_cap, history_len;
} orc_stokes;

/* solve!/_solve_stokes_dyrel! keyword parameters (src/stokes/solvers/DR.jl:302-386) */
typedef struct {
    double dt;
    double pressure_factor;    /* gamma fact, default 50            */
    double tolerance;          /* eps_tol                           */
    int64_t max_iterations;    /* total_iterMax                     */
    int64_t iter_max;          /* iterMax (inner per PH)            */
    int64_t check_interval;    /* ncheck                            */
    int64_t max_ph_iterations; /* default 1000                      */
    double rel_drop0;          /* default 1e-2                      */
    double CFL_v, c_fact;      /* StokesDR defaults 0.98, 0.9       */
    int32_t finite_K;          /* pressure bulk = K if 1 else etab  */
    int32_t freeze_jacobian;   /* -1: default (plastic == nothing)  */
    int32_t skip_scaling;      /* 1: gammaP/MP already assembled    */
    int32_t verbose;
    int32_t measure_lmax;      /* measure_λmax (DR.jl:375-378): power iteration on A */
    int32_t lmax_power_iterations;  /* 12 (0 reads as 12)         */
    double lmax_power_rtol, lmax_safety;  /* 1e-2, 1.1 (0 reads as the default) */
} orc_params;

typedef struct {
    int64_t itPH, iter;
    double err, err_abs, err_rel, err_v, err_P;
    int32_t converged, reached_total_iter;
    double lambda_max, lambda_max_gershgorin;
    int64_t jacobian_assemblies;
    int32_t status;            /* 0 ok, <0 error (see orc_last_error) */
    int64_t lambda_max_iterations;
} orc_stats;

const char* orc_last_error(void);
int orc_num_threads(void);
void orc_set_threads(int n);

/* --- reference elements (src/elements/shape_functions.jl:115-171, 231-239;
 *     src/elements/integration_points.jl:188-202) --- */
void orc_t7_basis(double xi, double eta, double* N /*7*/, double* dN /*7x2*/);
void orc_t7_rule(double* xi /*7*/, double* eta /*7*/, double* w /*7*/);
void orc_t7_tables(double* Nq /*7x7*/, double* NqP /*7x3*/, double* dNdxi /*7x7x2*/, double* w /*7*/);

/* --- structured T7 mesh (src/mesh/mesh.jl:412-440, src/mesh/connectivity.jl:111-140) --- */
void orc_t7_mesh_sizes(int32_t nx, int32_t ny, int64_t* nnodes, int64_t* nels);
void orc_t7_mesh(int32_t nx, int32_t ny, double x0, double x1, double y0, double y1,
                 double* coords, int32_t* el2n);
/* boundary nodes of the box (mesh.jl:79-80); returns count, fills nodes if non-null */
int64_t orc_box_boundary_nodes(int64_t nnodes, const double* coords, double x0, double x1,
                               double y0, double y1, int32_t* nodes);

/* --- geometry (src/mesh/mesh.jl:550-553, 674-682; mixed_mesh.jl:297-328) --- */
void orc_t7_geometry(int64_t nels, const double* coords, const int32_t* el2n,
                     double* invJ, double* dOmega, double* dOmegaP);

/* --- element-level kernels on caller tables (for closed-form checks) ---
 * dNdx: [nq][NV][2], dOm: [nq], Nq: [nq][NV], NqP: [nq][NP] */
void orc_pressure_element(int32_t NV, int32_t NP, int32_t nq,
                          const double* dNdx, const double* dOm, const double* dOmP,
                          const double* NqP, const double* vx, const double* vy,
                          const double* P, const double* P0, const double* T, const double* T0,
                          const double* Q /*nullable*/, const int32_t* phaseP /*NP*/,
                          const double* alpha, const double* bulk, double dt, double* RP_out);
void orc_momentum_element(int32_t NV, int32_t NP, int32_t nq,
                          const double* dNdx, const double* dOm, const double* Nq, const double* NqP,
                          const double* vx, const double* vy, const double* P,
                          const double* Pnum /*nullable*/, const double* T,
                          const int32_t* phase /*NV*/, const orc_material* mat,
                          const double* eta_override /*nullable: per-phase eta*/,
                          const orc_dp* dp /*nullable*/, double dt,
                          const double* tau_old /*nullable [3][nq]*/,
                          double* Rx, double* Ry, double* tau_out /*nullable [3][nq]*/);
/* one quadrature-point stress (rheology.jl:226-334): returns lambda */
double orc_stress_at_ip(const double* gradvx /*2*/, const double* gradvy /*2*/,
                        const double* Nv, int32_t NV, const int32_t* phase,
                        const orc_material* mat, const orc_dp* dp, double dt,
                        const double* tau_old /*3*/, double Pq, double* tau /*3*/);

/* --- assemblers over the whole mesh --- */
void orc_pressure_scaling(orc_stokes* s, double gamma_fact, double dt, int32_t finite_K);
void orc_pressure_residual(orc_stokes* s, double dt, int32_t finite_K, const double* P_old);
void orc_momentum_residual(orc_stokes* s, double dt, const double* Pnum /*nullable*/);
void orc_augmented_jacobian(orc_stokes* s, double dt, int32_t finite_K);
void orc_dr_update(orc_stokes* s, const double* alpha /*2*/, const double* beta /*2*/);

void orc_dr_iterations(orc_stokes* s, double dt, int32_t finite_K, const double* alpha, const double* beta,
                       int64_t k);

/* --- full solve (src/stokes/solvers/DR.jl:302-337, 348-802) --- */
int orc_stokes_solve(orc_stokes* s, const orc_params* p, orc_stats* out);
/* _set_pressure_source! (pressure_residual.jl:32-77): s->Q = Qq scaled to
 * ∫ Q dΩ = Q2D (has_Q2D = 0: copied); returns 0, or -1 with orc_last_error() */
int orc_set_pressure_source(orc_stokes* s, const double* Qq, double Q2D, int32_t has_Q2D);
/* update_stokes_current_stress! (DR.jl:820-897): tau_store must be set */
void orc_update_current_stress(orc_stokes* s, double dt);

/* --- 3-D: Hex27 velocity, P1-disc pressure with four modes per cell --- */
/* Whole 3-D Stokes problem. Cell phases (1-based, [nels]) serve velocity and
 * pressure; an optional temperature field (T = 0 without), no source or
 * fluid-pressure fields. Pressure
 * DoF a of element e is 4 e + a; stress components (xx, yy, zz, xy, xz, yz). */
typedef struct {
    int64_t nnodes, nels, nnodesP;
    const double* coords;   /* [nnodes][3] */
    const int32_t* el2n;    /* [nels][27], 0-based */
    const double* invJ;     /* [nels][27][9], (J^-1)[ξ-index][x-index] */
    const double* dOmega;   /* [nels][27] |det J| w */
    const int32_t* phases;  /* [nels] */
    orc_material mat;       /* g and Tref of mat are not used: g below */
    double g[3];
    const orc_dp* dp;       /* nullable */
    double *v[3], *rate[3], *Rv[3], *Rv0[3], *dRdv[3], *PC[3];
    double *P, *P0, *gammaP, *RP, *MP, *Pnum;
    const double* tau_old[6];   /* nullable (all or none): [nels][27] */
    double* tau_store[6];       /* for orc_update_current_stress3 */
    int64_t nbc[3];
    const int32_t* bc_dofs[3];
    const double* bc_vals[3];
    /* nullable: a momentum load subtracted in the cell-pressure relaxation
     * (_subtract_body_force!, DR.jl:1034; the adjoint's objective) */
    const double* body_force[3];
    /* nullable (both or none): temperature and its previous-step value at the
     * pressure DoFs [nels][4] (the thermal continuity term and the density) */
    const double* T;
    const double* T0;
} orc_stokes3;

/* reference Hex27 (shape_functions.jl:411-446, integration_points.jl:321-337) */
void orc_hex27_basis(double xi, double eta, double zeta, double* N /*27*/, double* dN /*27x3*/);
void orc_hex27_rule(double* xi, double* eta, double* zeta, double* w /*27 each*/);
void orc_hex27_tables(double* Nq /*27x27*/, double* NqP /*27x4*/, double* dNdxi /*27x27x3*/, double* w /*27*/);
/* structured Hex27 mesh (mesh.jl:340-398, connectivity.jl:273-322) */
void orc_hex27_mesh_sizes(int32_t nx, int32_t ny, int32_t nz, int64_t* nnodes, int64_t* nels);
void orc_hex27_mesh(int32_t nx, int32_t ny, int32_t nz, double x0, double x1, double y0, double y1, double z0,
                    double z1, double* coords, int32_t* el2n);
void orc_hex27_geometry(int64_t nels, const double* coords, const int32_t* el2n, double* invJ, double* dOmega);
/* element momentum residual on caller tables (momentum_residuals.jl:209-258):
 * dNdx [nq][NV][3], Nq [nq][NV], NqP [nq][NP], phase [NV] 1-based, g [3],
 * tau_old [6][nq] nullable, tau_out [6][nq] nullable */
void orc_momentum_element3(int32_t NV, int32_t NP, int32_t nq, const double* dNdx, const double* dOm, const double* Nq,
                           const double* NqP, const double* vx, const double* vy, const double* vz, const double* P,
                           const double* Pnum, const int32_t* phase, const orc_material* mat, const double* g,
                           const orc_dp* dp, double dt, const double* tau_old, double* Rx, double* Ry, double* Rz,
                           double* tau_out);
void orc_pressure_scaling3(orc_stokes3* s, double gamma_fact, double dt, int32_t finite_K);
void orc_pressure_residual3(orc_stokes3* s, double dt, int32_t finite_K);
void orc_momentum_residual3(orc_stokes3* s, double dt, const double* Pnum /*nullable*/);
void orc_augmented_jacobian3(orc_stokes3* s, double dt, int32_t finite_K);
int orc_stokes_solve3(orc_stokes3* s, const orc_params* p, orc_stats* out);
void orc_update_current_stress3(orc_stokes3* s, double dt);

/* --- CellPressureStokesDR (stokes_types.jl:545-613, DR.jl:960-1057): viscous
 *     Hex27 with the cell-local pressure modes (1, ξ, η, ζ); an orc_stokes3
 *     whose PC holds the Jacobi diagonal and MP the mode mass --- */
typedef struct {
    double tolerance;          /* 1e-5 */
    int64_t max_iterations;    /* 3000 */
    int64_t check_interval;    /* 100  */
    double velocity_step;      /* 0.6  */
    double pressure_step;      /* 0.2  */
    int32_t verbose;
} orc_cp3_params;
typedef struct {
    int32_t converged, reached_total_iter;
    int64_t iterations;
    double residual, err_v, err_P;
    int32_t status;
} orc_cp3_stats;
void orc_cp3_pressure_residual(const orc_stokes3* s, double* RP /* [nels][4] */);
void orc_cp3_momentum_residual(const orc_stokes3* s, double* const* R /* 3 x [nnodes] */);
void orc_cp3_preconditioner(const orc_stokes3* s, double* const* diag, double* M);
/* history (nullable): [history_cap][3] (iter, err_v, err_P) per check */
int orc_cp3_solve(orc_stokes3* s, const orc_cp3_params* p, orc_cp3_stats* out, double* history,
                  int64_t history_cap, int64_t* history_len);

/* --- thermal diffusion (src/heat_diffusion/, src/dynamic_relaxation/solver.jl) --- */

/* ThermalMaterial (src/heat_diffusion/types/heat_diffusion_types.jl:15-41) */
typedef struct {
    int32_t nphases;
    double k[ORC_MAX_PHASES];
    double Cp[ORC_MAX_PHASES];
    double rho0[ORC_MAX_PHASES];
    double alpha[ORC_MAX_PHASES];
    double K[ORC_MAX_PHASES];
} orc_thermal_material;

/* ThermalDiffusionDR (heat_diffusion_types.jl:90-153) on a T7 mesh. */
typedef struct {
    int64_t nnodes, nels;
    const int32_t* el2n;    /* [nels][7], 0-based */
    const double* invJ;     /* [nels][7][4] */
    const double* dOmega;   /* [nels][7]    */
    const int32_t* phases;  /* [nnodes], 1-based per-node phases */
    orc_thermal_material mat;
    double *R, *R0, *dRdT, *PC, *T, *T0, *rate;
    const double* P;        /* [nnodes] */
    const double* source;   /* [nnodes] */
    int64_t nbc;
    const int32_t* bc_dofs; /* 0-based */
    const double* bc_vals;
    double CFL, c_fact;     /* 0.98, 0.9 */
    const double* source_ip; /* nullable [nels][nq]: heat source at the quadrature points
                                (source_ip of residual.jl:7-47, the shear heating of a coupled solve) */
    int32_t dim;            /* 2 (T7; 0 reads as 2) or 3: Hex27 with el2n [nels][27],
                               invJ [nels][27][9] and dOmega [nels][27] (orc_hex27_geometry) */
    int32_t nloc;           /* nodes per element: 0 reads as 7 (dim 2) or 27 (dim 3); 9 with dim 2
                               is Q9 with el2n [nels][9], invJ [nels][9][4] and dOmega [nels][9]
                               (orc_q9_geometry) */
} orc_thermal;

typedef struct {
    double dt, tolerance, Tref;        /* tolerance defaults to dr.ϵ = 1e-6, Tref 273 */
    int64_t max_iterations, check_interval;
    int32_t verbose;
} orc_thermal_params;

typedef struct {
    int32_t converged;
    int64_t iterations;
    double residual;                  /* last relative residual ‖R‖/‖R⁰‖ */
    int32_t status;                   /* 0 ok, <0 error (orc_last_error)  */
} orc_thermal_stats;

/* integrate_residual (src/heat_diffusion/assembly/residual.jl:58-87) on caller
 * tables N [nq][nn], dNdx [nq][nn][2], dOm [nq]; phases 1-based per node.
 * J (nullable) receives dRe_i/dT_j [nn][nn], the matrix the reference takes
 * with ForwardDiff (residual.jl:30-47). */
void orc_thermal_integrate(int32_t nn, int32_t nq, const double* N, const double* dNdx, const double* dOm,
                           const double* T, const double* T0, const double* s, const int32_t* phases,
                           const orc_thermal_material* m, const double* P, double dt, double Tref,
                           double* Re, double* J);
/* assemble_diffusion_matrices_atomix! (residual.jl:90-121): R, and ∂R∂T (row
 * sums of |J|) and PC (|J_ii|) when compute_jacobian */
void orc_thermal_residual(orc_thermal* t, double dt, double Tref, int32_t compute_jacobian);
/* solve!(::ThermalDiffusionDR, ...) -> solve_dynamic_relaxation!
 * (heat_diffusion/solvers/DR.jl:50-54, dynamic_relaxation/solver.jl:61-122) */
int orc_thermal_solve(orc_thermal* t, const orc_thermal_params* p, orc_thermal_stats* out);

/* --- frozen-operator adjoint of StokesDR (src/stokes/solvers/DR_adjoint.jl:1-531,
 * src/stokes/assemblers/adjoint_operator.jl:1-700), operator = :blocks, 2-D T7/P1-disc,
 * models without a yield model or with the Drucker-Prager model --- */
typedef struct {
    double dt, tolerance, rel_drop, CFL_v, c_fact;  /* tolerance 1e-6, rel_drop 0.1 */
    int64_t max_iterations;     /* total_iterMax, 50_000 */
    int64_t iter_max;           /* iterMax (inner), 50_000 */
    int64_t check_interval;     /* ncheck, 50 */
    int64_t max_ph_iterations;  /* 100 */
    int32_t measure_lmax;       /* 1: power iteration (estimate_adjoint_λmax), 0: Gershgorin */
    int32_t verbose;
} orc_adjoint_params;
typedef struct {
    int64_t itPH, iter, lmax_iterations;
    double err, err_v, err_P, lmax, lmax_gershgorin;
    int32_t converged, status;
} orc_adjoint_stats;
/* Element blocks (element_adjoint_operator_blocks, adjoint_operator.jl:318-373) at
 * the state in s: A [nels][14][14] = ∂Rv/∂v with the augmentation, B [nels][14][3]
 * = ∂Rv/∂P (Pnum held), C [nels][3][14] = ∂RP/∂v, D [nels][3][3] = ∂RP/∂P (ηb
 * storage), W [nels][14][3] = ∂Rv/∂Pnum diag(γP/M_P); velocity DoFs stacked by
 * component (vx 0..6, vy 7..13). defects (nullable [2]) = max over elements of
 * ‖A−Aᵀ‖/‖A‖ and ‖C−Bᵀ‖/‖C‖. Also fills s->dRdvx/y and s->PCx/y from A.
 * Returns -1 (orc_last_error) for a state with the tensile cap. */
int orc_adjoint_blocks(orc_stokes* s, double dt, double* A, double* B, double* C, double* D, double* W,
                       double* defects);
/* apply_adjoint_operator! (adjoint_operator.jl:631-660): dv = Aᵀλv + CᵀλP (B λP and the
 * packed upper triangle of A when symmetric), dP = Bᵀλv + Dᵀ(Wᵀλv + λP) (Bᵀλv alone
 * unless stores) */
void orc_adjoint_apply(const orc_stokes* s, const double* A, const double* B, const double* C, const double* D,
                       const double* W, int32_t symmetric, int32_t stores, const double* lx, const double* ly,
                       const double* lP, double* dvx, double* dvy, double* dP);
/* _solve_stokes_adjoint_dyrel! (DR_adjoint.jl:1-268) with the stored blocks; objective
 * and λ are nodal [nnodes] (x, y) and [nnodesP]; λ holds the initial iterate. */
int orc_adjoint_solve(orc_stokes* s, const orc_adjoint_params* p, const double* obj_x, const double* obj_y,
                      double* lam_x, double* lam_y, double* lam_P, orc_adjoint_stats* st);

/* --- Q9: QuadraticElement{2, 9} (shape_functions.jl:259-283, integration_points.jl:218-241) --- */
/* N [9], dN [9][2] at (ξ, η); local nodes (-1,-1),(1,-1),(1,1),(-1,1),(0,-1),(1,0),(0,1),(-1,0),(0,0) */
void orc_q9_basis(double xi, double eta, double* N, double* dN);
/* the 3x3 Gauss rule, ξ fastest, weights 25/81, 40/81, 64/81 */
void orc_q9_rule(double* xi, double* eta, double* w);
void orc_q9_mesh_sizes(int32_t nx, int32_t ny, int64_t* nnodes, int64_t* nels);
/* generate_coordinates (mesh.jl:340-398) and generate_element2node (connectivity.jl:193-227):
 * coords [(2nx+1)(2ny+1)][2], x fastest; el2n [nx ny][9], 0-based */
void orc_q9_mesh(int32_t nx, int32_t ny, double x0, double x1, double y0, double y1, double* coords,
                 int32_t* el2n);
/* precompute_geometry (mesh.jl:592-628): the isoparametric J = Σ_a x_a ⊗ ∂N_a/∂ξ at each point;
 * invJ [nels][9][4] row-major [ξ-index][x-index], dOmega [nels][9] = |det J| w */
void orc_q9_geometry(int64_t nels, const double* coords, const int32_t* el2n, double* invJ, double* dOmega);

/* --- Drucker-Prager tensile cap (src/stokes/assemblers/rheology.jl:510-787) --- */
/* cap_geometry: out = (a, b, p_y, R_y, p_d, τ_d, p_q) (rheology.jl:546-557) */
void orc_cap_geometry(double k, double kq, double c, double pT, double* out);
double orc_cap_yield_function(double tII, double P, double k, double kq, double c, double pT);
/* cap_invariants: out = (F, Aτ, Ap) (rheology.jl:600-609) */
void orc_cap_invariants(double s, double p, double k, double kq, double c, double pT, double* out);
/* cap_return_map: out = (τII, P, λ, Aτ, Ap); returns converged (rheology.jl:652-690) */
int orc_cap_return_map(double s_trial, double p_trial, double eta_ve, double KDt, double k, double kq, double c,
                       double pT, double eta_reg, int32_t maxiter, double* out);
/* cap_local_update: tau_out (3), out = (P, γdot, θdot) (rheology.jl:708-717) */
void orc_cap_local_update(const double* tau_trial, double P_trial, double eta_ve, double KDt, double k, double kq,
                          double c, double pT, double eta_reg, double* tau_out, double* out);
/* commit_stokes_plastic_history! (DR.jl:899-958, momentum_residuals.jl:570-610):
 * γ += Δt γdot, θ += Δt θdot at every point; returns -1 if a value is non-finite */
int orc_commit_plastic_history(orc_stokes* s, double dt);

/* --- rotate_stress! (src/stokes/stress_rotation.jl:28-61) and the GeoParams
 *     0.7.21 rotate_elastic_stress it calls (restated, not vendored) --- */
void orc_rotate_elastic_stress2(double omega, const double* tau /*xx,yy,xy*/, double dt, double* out);
void orc_rotate_elastic_stress3(const double* omega /*curl v*/, const double* tau /*xx,yy,zz,xy,xz,yz*/, double dt,
                                double* out);
/* tau, tau_old: [3][nels*7] (2-D) and [6][nels*27] (3-D, assembler order) */
void orc_rotate_stress(const orc_stokes* s, double dt, const double* const* tau, double* const* tau_old);
void orc_rotate_stress3(const orc_stokes3* s, double dt, const double* const* tau, double* const* tau_old);

/* --- coupled thermal--Stokes step (src/stokes/solvers/DR.jl:44-131, 514-661) --- */

/* assemble_shear_heating! (src/stokes/assemblers/shear_heating.jl:18-92):
 * phi [nels][7] = τ : (ε̇ − (τ − τ_old)/(2GΔt)) at the velocity quadrature
 * points, τ the stress of the momentum residual for the current v and P */
void orc_shear_heating(const orc_stokes* s, double dt, double* phi);

typedef struct {
    double Tref;              /* thermal Tref keyword (default 273)            */
    double tolerance_T;       /* thermal dr.ϵ                                  */
    int32_t shear_heating;
} orc_coupled_params;

typedef struct {
    double err_T;             /* last thermal ‖R‖/‖R⁰‖                         */
    int64_t thermal_iterations;
    int32_t thermal_converged;
} orc_coupled_stats;

/* solve_coupled!: one thermal DR iteration per Stokes inner iteration, the
 * thermal field transferred to the Stokes pressure DoFs (T and T0), the
 * convergence tests of both loops extended by the thermal test. The
 * thermal state's R, T and ∂T∂τ (rate) carry the iteration; t->source_ip is
 * set by the call when shear heating is on. Stokes statistics in out;
 * out->converged requires the thermal test too. */
int orc_coupled_solve(orc_thermal* t, orc_stokes* s, const orc_params* p, const orc_coupled_params* cp,
                      orc_stats* out, orc_coupled_stats* cst);
/* the same on a Hex27 state (s->T and s->T0 set; thermal mesh: the same
 * nodes), with the 3-D shear dissipation [nels][27] */
void orc_shear_heating3(const orc_stokes3* s, double dt, double* phi);
int orc_coupled_solve3(orc_thermal* t, orc_stokes3* s, const orc_params* p, const orc_coupled_params* cp,
                       orc_stats* out, orc_coupled_stats* cst);

#ifdef __cplusplus
}
#endif
#endif
