// femdr_host.hpp -- host-side internals shared by the translation units of
// libfemdr.so: the context and mesh objects, kernel timing, error and
// allocation helpers, launch-shape helpers.
#pragma once
#include "../../include/femdr.h"
#include "femdr_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <vector>


using namespace fem;

struct fem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::set<void*> bufs;
    double* red_partial = nullptr;   // [RED_GRID * RED_MAXK]
    double* red_out = nullptr;       // [RED_MAXK]
    double* h_red = nullptr;         // pinned [RED_MAXK]
    double* h_cheb = nullptr;        // pinned [8]: Chebyshev scalars staged for graph replays
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> tev;     // kernel-timing events (3 per timed iteration)
    size_t tev_used = 0;
    fem_tuning tune{1, 8, 1, 25};    // launch choices (fem_ctx_set_tuning)
};

// Kernel timing: events e0 | element | e1 | gather | e2 around every
// STRIDE-th iteration (sampled, so that the event records do not slow the
// solve they measure), read back at the next host synchronisation.
struct KernelTimer {
    static constexpr int64_t STRIDE = 16;
    fem_ctx* c;
    bool on;
    bool take(int64_t iter) const { return on && iter % STRIDE == 1; }
    double ms_elem = 0, ms_gather = 0;
    int64_t n = 0;
    // the subset of the above timed on return-map correction intervals of the
    // mixed-precision solve (kind 1)
    double ms_elem_rm = 0, ms_gather_rm = 0;
    int64_t n_rm = 0;
    std::vector<int64_t> weight;   // inner iterations covered by each timed triple
    std::vector<int8_t> kind;      // 1: a return-map correction segment
    KernelTimer(fem_ctx* c_, bool on_) : c(c_), on(on_) { c->tev_used = 0; }
    hipEvent_t ev() {
        if (c->tev_used == c->tev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            c->tev.push_back(e);
        }
        return c->tev[c->tev_used++];
    }
    void mark(hipEvent_t e) {
        if (e) (void)hipEventRecord(e, c->stream);
    }
    // call after the stream has been synchronised
    void collect() {
        if (!on) return;
        for (size_t i = 0; i + 2 < c->tev_used; i += 3) {
            float a = 0, b = 0;
            if (hipEventElapsedTime(&a, c->tev[i], c->tev[i + 1]) == hipSuccess &&
                hipEventElapsedTime(&b, c->tev[i + 1], c->tev[i + 2]) == hipSuccess) {
                const int64_t w = i / 3 < weight.size() ? weight[i / 3] : 1;
                ms_elem += a;
                ms_gather += b;
                n += w;
                if (i / 3 < kind.size() && kind[i / 3]) {
                    ms_elem_rm += a;
                    ms_gather_rm += b;
                    n_rm += w;
                }
            }
        }
        c->tev_used = 0;
        weight.clear();
        kind.clear();
    }
};

struct fem_mesh {
    fem_ctx* ctx = nullptr;
    int32_t dim = 2, nloc = 7;         // T7 (2-D), Hex27 (3-D) or, for thermal states only, Q9 (2-D, nloc 9)
    int64_t nnodes = 0, nels = 0, nnodesP = 0, ns = 0;
    bool structured = false;
    double box[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> h_coords;      // [nnodes][dim]
    std::vector<int32_t> h_el2n;       // [nels][nloc] 0-based
    std::vector<int32_t> h_boundary;   // 0-based
    int32_t* d_el2n = nullptr;         // [7][nels]
    double* d_geo = nullptr;           // [5][nels]
    int32_t* d_off = nullptr;          // [ns+1]
    int32_t* d_ent = nullptr;
    int32_t* d_node_of = nullptr;      // nullable
    int32_t nx = 0, ny = 0, nz = 0;    // structured box mesh cell counts
};

struct fem_stokes {
    fem_ctx* ctx = nullptr;
    fem_mesh* mesh = nullptr;
    int32_t dim = 2;
    fem_material mat;
    fem_state_opts opts;
    double* f[FEM_FIELD_COUNT] = {};
    int64_t len[FEM_FIELD_COUNT] = {};
    int32_t* d_phase_v = nullptr;   // cell [nels] or element-node [7][nels], 0-based
    int32_t* d_phase_P = nullptr;   // cell [nels] or element-node [3][nels]
    int32_t layout_v = FEM_PHASE_CELL, layout_P = FEM_PHASE_CELL;
    bool phase_same = true;         // cell phases of v and P identical
    double* d_pscale = nullptr;
    double2* d_bpart64 = nullptr;   // FP64 EF_TILE perimeter partial sums (structured 2-D meshes)
    double2* d_Rn64 = nullptr;      // FP64 EF_TILE interior-node sums
    double2* d_Re = nullptr;
    double2* d_Jrow = nullptr;
    double2* d_Jdiag = nullptr;
    uint8_t* d_bcflag = nullptr;
    double* d_bcval_x = nullptr;
    double* d_bcval_y = nullptr;
    double* d_bcval_z = nullptr;      // 3-D
    double* d_slot3 = nullptr;        // 3-D element-slot residuals [nels][26][3]
    double* d_jslot3 = nullptr;       // 3-D Jacobian slots [nels][26][6]: |J| row sums, |J_ii|
    int32_t* d_bc_stage = nullptr;
    double* d_bc_stage_v = nullptr;
    int64_t bc_stage_cap = 0;
    double* d_ptab = nullptr;
    int32_t* d_flag = nullptr;      // [0] scratch yes/no flag, [1] yielding-element count
    double* d_cheb = nullptr;       // [8]: the Chebyshev scalars read by graph-replayed launches
    uint8_t* d_yflag = nullptr;     // near-yield flag per element (mixed-precision solve)
    bool pf_nonzero = false;        // dr.Pf holds a nonzero entry (checked at each entry point)
    bool q_nonzero = false;         // dr.Q holds a nonzero entry
    fem_solve_inputs in{};          // array keywords of the next calls (fem_stokes_set_solve_inputs)
    // FP32 work arrays of the mixed-precision solve (allocated by its first call):
    // the correction (δv, δP), its DR rates/residuals, the right-hand sides R(x)
    // and FP32 copies of the scaling, preconditioner and geometry
    struct W32 {
        float *vx, *vy, *rate_x, *rate_y, *Rv_x, *Rv_y, *Rv0_x, *Rv0_y, *PC_x, *PC_y;
        float *P, *RP, *Pnum, *gammaP, *MP, *pscale, *rhs_vx, *rhs_vy, *rhs_P, *zero_v, *geo;
        float2* Re;
        float2* bpart;   // per-tile perimeter partial sums (structured meshes in 16-cell-wide tiles)
        float2* Rn;      // summed residuals of the tiles' interior grid nodes
        // stress-history copy for the correction operator with the return map
        float *tau_old0, *tau_old1, *tau_old2;
        // [6][7][nels] linearisations of the return about x (precision 2)
        float* tangent;
    };
    W32* w32 = nullptr;
    // FP32 work arrays of the 3-D mixed-precision solve: the correction, its
    // rates and residuals, the right-hand sides R(x), FP32 copies of the
    // scaling and preconditioner, and the element-slot residuals [nels][26][3]
    struct W3 {
        float *vx, *vy, *vz, *rate_x, *rate_y, *rate_z, *Rv_x, *Rv_y, *Rv_z, *Rv0_x, *Rv0_y, *Rv0_z;
        float *PC_x, *PC_y, *PC_z, *rhs_vx, *rhs_vy, *rhs_vz;
        float *P, *RP, *Pnum, *gammaP, *MP, *pscale, *rhs_P;
        float* slot;
    };
    W3* w3 = nullptr;
};

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static int set_err(fem_ctx* c, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}
// The array keywords attached by fem_stokes_set_solve_inputs that a call with
// this yield model may use: a viscosity at the points is for purely viscous
// 2-D flow without a yield model (DR.jl:388-395)
static inline int check_solve_inputs(fem_stokes* s, bool plastic) {
    const fem_solve_inputs& in = s->in;
    if (in.viscosity) {
        bool ginf = true;
        for (int p = 0; p < s->mat.nphases; ++p) ginf = ginf && std::isinf(s->mat.G[p]);
        if (s->dim != 2 || !ginf)
            return set_err(s->ctx, FEM_EARG, "quadrature viscosity currently requires purely viscous 2-D flow");
        if (plastic) return set_err(s->ctx, FEM_EARG, "quadrature viscosity requires plastic=nothing");
    }
    if (s->dim != 2 && (in.scaling_viscosity || (in.P_old && in.P_old_at_points)))
        return set_err(s->ctx, FEM_EARG, "integration-point P_old and scaling_viscosity are built for 2-D states");
    return FEM_OK;
}
#define HIPCHK(ctx, x)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess)                                                                    \
            return set_err(ctx, FEM_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),   \
                           __FILE__, __LINE__);                                                  \
    } while (0)
#define CHK(x)                 \
    do {                       \
        int r_ = (x);          \
        if (r_ != FEM_OK) return r_; \
    } while (0)

static int dalloc(fem_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return set_err(c, FEM_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return FEM_OK;
}
template <class T>
static int dalloc_t(fem_ctx* c, T** p, size_t n) { return dalloc(c, (void**)p, n * sizeof(T)); }

static int check_launch(fem_ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, FEM_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return FEM_OK;
}

// Tensor-product quadratic elements (Hex27, Q9). The 3-point Gauss rule
// (integration_points.jl:218-241, 321-337): ±√(3/5), 0; weights 5/9, 8/9
__host__ __device__ inline double gauss_x(int p) {
    return p == 0 ? -0x1.8c97ef43f7248p-1 : (p == 1 ? 0.0 : 0x1.8c97ef43f7248p-1);
}
__host__ __device__ inline double gauss_w(int p) { return p == 1 ? 8.0 / 9.0 : 5.0 / 9.0; }
// 1-D quadratic Lagrange basis on the nodes r = 0, 1, 2 at -1, 0, +1 and its
// derivative (shape_functions.jl:173-196)
__host__ __device__ inline double lq(double x, int r) {
    return r == 0 ? x * (x - 1) * 0.5 : (r == 1 ? 1 - x * x : x * (x + 1) * 0.5);
}
__host__ __device__ inline double dlq(double x, int r) { return r == 0 ? x - 0.5 : (r == 1 ? -2 * x : x + 0.5); }

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }
static inline unsigned grid_xcd(int64_t n, int block) {
    unsigned g = grid_for(n, block);
    return (g + 7u) & ~7u;
}

// shared between the 2-D and 3-D translation units
int upload_bcs(fem_stokes* s, const fem_bc* bcs);   // one fem_bc per velocity component
// 3-D (Hex27) entry points behind the fem_stokes_* C ABI
int hex27_pressure_scaling(fem_stokes* s, double gamma_fact, double dt, int finite_K);
int hex27_residual(fem_stokes* s, const fem_drucker_prager* dp, double dt, bool with_pnum, int finite_K);
int hex27_jacobian(fem_stokes* s, const fem_drucker_prager* dp, double dt, int finite_K);
int hex27_update_current_stress(fem_stokes* s, const fem_drucker_prager* dp, double dt);
struct CoupledSide;
// with cp, the coupled loop of solve_coupled! (DR.jl:514-661) on a Hex27 state
int hex27_solve(fem_stokes* s, const fem_bc* bcs, const fem_drucker_prager* dp, const fem_solve_params* p,
                fem_solve_stats* st, CoupledSide* cp = nullptr);
int hex27_time_element(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t n, double* avg_ms);
int hex27_time_lin(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t n, double* avg_ms);
int hex27_rotate_stress(fem_stokes* s, double dt);
int hex27_shear_heating(fem_stokes* s, const fem_drucker_prager* dp, double dt, double* phi);   // [nels][27]
int hex27_postprocess(fem_stokes* s, double* out);

// eps(x) of Julia (spacing at |x|) and the Chebyshev step/momentum (DR.jl:39-42)
static inline double julia_eps(double x) {
    const double a = std::fabs(x);
    if (!std::isfinite(a)) return std::numeric_limits<double>::quiet_NaN();
    return std::nextafter(a, std::numeric_limits<double>::infinity()) - a;
}

static inline void cheb(double dtau, double lmin, double c_fact, double* a, double* b) {
    // _stokes_cheb, DR.jl:39-42
    const double c = 2 * std::sqrt(lmin) * c_fact;
    *a = 2 * dtau * dtau / (2 + c * dtau);
    *b = (2 - c * dtau) / (2 + c * dtau);
}

int upload_ptab(fem_stokes* s, const fem_drucker_prager* dp, double dt, int finite_K, bool* bulk_inf,
                bool* thermal, bool* eos_p, bool* elastic, const fem_drucker_prager_cap* cap = nullptr);

// per-phase table columns of the thermal state (heat_diffusion_types.jl:15-41);
// β = 1/K is interpolated rather than K (residual.jl:62-64)
enum ThermCol { TC_K = 0, TC_CP, TC_RHO0, TC_ALPHA, TC_BETA, TC_NCOL };

// arguments of the thermal element kernels on structured Hex27 and Q9 meshes
struct ThermBoxArgs {
    const double *T, *T0, *src, *P;
    const int32_t* phases;                 // per node, 1-based
    const double* tab;
    double dt, Tref;
    double* Rs; double2* Js;               // [nels][26] (Hex27) or [nels][8] (Q9) slots of the shared nodes
    double *R, *R0, *dRdT, *PC, *rate, *Tw;
    const uint8_t* bcflag; const double* bcval;
    double alpha, beta;
    int32_t update;
    const double* phi;                     // nullable: source at the quadrature points [nels][nq]
};
int hex27_thermal_element(fem_ctx* c, const fem_mesh* m, const ThermBoxArgs& a, bool jac);
// the same arguments on a structured Q9 mesh (slots [nels][8])
int q9_thermal_element(fem_ctx* c, const fem_mesh* m, const ThermBoxArgs& a, bool jac);

// The thermal side of a coupled thermal-Stokes solve (DR.jl:514-529,
// 604-661): one thermal DR iteration per Stokes inner iteration, on the
// thermal state's stream, with the Chebyshev scalars and the convergence
// test carried between iterations.
struct ThermalCoupling {
    fem_thermal* t = nullptr;
    double dt = 1, Tref = 273;
    int64_t nout = 1;                // check cadence: the Stokes check_interval
    const double* phi = nullptr;     // nullable: heat source at the velocity quadrature points [nels][nq]
    int64_t iter = 0;
    double alpha = 0, beta = 0, nr0 = 0, err_T = INFINITY;
    bool converged = false;
    bool checked = false;            // the last iteration was a check
};
// The thermal half of a coupled solve (solve_coupled!, DR.jl:81-131)
struct CoupledSide {
    ThermalCoupling tc;
    const fem_bc* bc_T = nullptr;
    bool shear = false;
    double* d_phi = nullptr;   // [nels][nq] shear dissipation, with shear heating
};
int thermal_coupling_begin(ThermalCoupling& tc, const fem_bc* bc);   // tables, Dirichlet set, ∂T∂τ = R0 = 0
int thermal_coupling_step(ThermalCoupling& tc);
fem_mesh* thermal_mesh(const fem_thermal* t);
fem_ctx* thermal_ctx(const fem_thermal* t);
double* thermal_field_ptr(fem_thermal* t, int32_t field);
