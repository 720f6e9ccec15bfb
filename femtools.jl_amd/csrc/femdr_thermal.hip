// femdr_thermal.hip -- ThermalDiffusionDR on the device (T7 meshes here, Hex27
// and Q9 boxes through hex27_thermal_element and q9_thermal_element): the backward-Euler
// heat-diffusion residual of src/heat_diffusion/assembly/residual.jl:58-87,
// its Jacobian diagnostics (residual.jl:30-47), and the Chebyshev-accelerated
// pseudo-transient loop of src/dynamic_relaxation/solver.jl:61-122.
//
// Layout and scatter follow the Stokes path: one thread per T7 element, the
// six shared nodes' contributions written to element slots and summed per
// node in ascending element order (deterministic, no atomics), the private
// bubble node updated by its element. Between checks an iteration is two
// launches (element residual + bubble update, gather + update); a check
// iteration assembles the Jacobian diagnostics, checks λmax on the host
// before the update, as the reference does, and reads three reduced sums.
#include "femdr_host.hpp"

namespace {


struct ThermArgs {
    int64_t nels;
    const int32_t* el2n; const double* geo;                  // generic mesh
    int32_t nx, sx; int64_t nT6; double idx, idy, detc;      // structured box mesh
    const double* T; const double* T0; const double* src; const double* P;
    const int32_t* phases;                                   // per node, 1-based
    const double* tab;
    double dt, Tref;
    double* Rs; double2* Js;                                 // [nels][NSH] slots
    double* R; double* R0; double* dRdT; double* PC; double* rate; double* Tw;
    const uint8_t* bcflag; const double* bcval;
    double alpha, beta;
    int32_t update;
    const double* phi;                                       // nullable: source at the quadrature points [nels][7]
};

template <bool STRUCT, bool JAC>
__global__ __launch_bounds__(128) void k_thermal_element(ThermArgs A) {
    const int64_t e = xcd_tile() * 128 + threadIdx.x;
    if (e >= A.nels) return;
    const int64_t ne = A.nels;
    int32_t n[NV];
    double i00, i01, i10, i11, det;
    if (STRUCT) {
        t7_struct_element<double>(e, A.nx, A.sx, A.nT6, A.idx, A.idy, n, i00, i01, i10, i11);
        det = A.detc;
    } else {
#pragma unroll
        for (int a = 0; a < NV; ++a) n[a] = A.el2n[a * ne + e];
        i00 = A.geo[0 * ne + e]; i01 = A.geo[1 * ne + e];
        i10 = A.geo[2 * ne + e]; i11 = A.geo[3 * ne + e];
        det = A.geo[4 * ne + e];
    }
    double T[NV], T0[NV], s[NV], P[NV], kk[NV], cp[NV], r0[NV], al[NV], be[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        const int32_t g = n[a];
        T[a] = A.T[g]; T0[a] = A.T0[g]; s[a] = A.src[g]; P[a] = A.P[g];
        const int p = A.phases[g] - 1;
        kk[a] = A.tab[TC_K * MAXPH + p]; cp[a] = A.tab[TC_CP * MAXPH + p];
        r0[a] = A.tab[TC_RHO0 * MAXPH + p]; al[a] = A.tab[TC_ALPHA * MAXPH + p];
        be[a] = A.tab[TC_BETA * MAXPH + p];
    }
    double Re[NV];
    double J[JAC ? NV * NV : 1];
#pragma unroll
    for (int a = 0; a < NV; ++a) Re[a] = 0;
    if (JAC) {
#pragma unroll
        for (int i = 0; i < NV * NV; ++i) J[i] = 0;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double G[NV][2];
        double Tq = 0, Pq = 0, T0q = 0, sq = 0, gx = 0, gy = 0;
        double kq = 0, aq = 0, bq = 0, rq = 0, cq = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const double d0 = kT.dN[q][a][0], d1 = kT.dN[q][a][1], Na = kT.N[q][a];
            G[a][0] = d0 * i00 + d1 * i10;
            G[a][1] = d0 * i01 + d1 * i11;
            Tq += Na * T[a]; Pq += Na * P[a]; T0q += Na * T0[a]; sq += Na * s[a];
            gx += G[a][0] * T[a]; gy += G[a][1] * T[a];
            kq += Na * kk[a]; aq += Na * al[a]; bq += Na * be[a]; rq += Na * r0[a]; cq += Na * cp[a];
        }
        const double rho = rq * (1 - aq * (Tq - A.Tref) + bq * Pq);
        const double D = A.dt / (rho * cq);
        double rate = T0q - Tq + D * sq;
        const double phq = A.phi ? A.phi[e * NQ + q] : 0.0;
        if (A.phi) rate += D * phq;   // Δt/(ρ Cp) Φ_q (residual.jl:75-79)
        const double dOm = kT.w[q] * det;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double KTi = kq * (G[i][0] * gx + G[i][1] * gy);
            Re[i] += rate * kT.N[q][i] * dOm - D * KTi * dOm;
        }
        if (JAC) {
            const double c = D * rq * aq / rho;
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const double KTi = kq * (G[i][0] * gx + G[i][1] * gy);
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    const double dD = c * kT.N[q][j];
                    const double Kij = kq * (G[i][0] * G[j][0] + G[i][1] * G[j][1]);
                    double dr = -kT.N[q][j] + dD * sq;
                    if (A.phi) dr += dD * phq;
                    J[i * NV + j] += (dr * kT.N[q][i] - dD * KTi - D * Kij) * dOm;
                }
            }
        }
    }
    double rs[JAC ? NV : 1], dg[JAC ? NV : 1];
    if (JAC) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double t = 0;
#pragma unroll
            for (int j = 0; j < NV; ++j) t += fabs(J[i * NV + j]);
            rs[i] = t;
            dg[i] = fabs(J[i * NV + i]);
        }
    }
#pragma unroll
    for (int a = 0; a < NSH; ++a) {
        A.Rs[e * NSH + a] = Re[a];
        if (JAC) A.Js[e * NSH + a] = make_double2(rs[a], dg[a]);
    }
    // private bubble node
    const int32_t b = n[NSH];
    const bool fixed = A.bcflag[b] & 1;
    const double r = fixed ? 0.0 : Re[NSH];
    if (JAC) {
        A.R0[b] = A.R[b];           // R0 <- R before the assembly (solver.jl:95)
        A.dRdT[b] = rs[NSH];
        A.PC[b] = dg[NSH];
    }
    A.R[b] = r;
    if (A.update) {
        const double nr = fixed ? 0.0 : r / A.PC[b] + A.beta * A.rate[b];
        A.rate[b] = nr;
        A.Tw[b] = fixed ? A.bcval[b] : T[NSH] + A.alpha * nr;
    }
}

struct ThermGatherArgs {
    int64_t ns;
    const int32_t* node_of; const int32_t* off; const int32_t* ent;
    const double* Rs; const double2* Js;
    double* R; double* R0; double* dRdT; double* PC; double* rate; double* T;
    const uint8_t* bcflag; const double* bcval;
    double alpha, beta;
    int32_t jac, update;
};

__global__ __launch_bounds__(256) void k_thermal_gather(ThermGatherArgs A) {
    const int64_t i = xcd_tile() * 256 + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    double r = 0, rs = 0, dg = 0;
    for (int32_t k = A.off[i]; k < A.off[i + 1]; ++k) {
        r += A.Rs[A.ent[k]];
        if (A.jac) {
            const double2 j = A.Js[A.ent[k]];
            rs += j.x;
            dg += j.y;
        }
    }
    const bool fixed = A.bcflag[node] & 1;
    if (fixed) r = 0;
    if (A.jac) {
        A.R0[node] = A.R[node];
        A.dRdT[node] = rs;
        A.PC[node] = dg;
    }
    A.R[node] = r;
    if (A.update) {
        const double nr = fixed ? 0.0 : r / A.PC[node] + A.beta * A.rate[node];
        A.rate[node] = nr;
        A.T[node] = fixed ? A.bcval[node] : A.T[node] + A.alpha * nr;
    }
}

// rate and temperature update of every node (update_rate_kernel!,
// update_variable_kernel!, Dirichlet pinning: solver.jl:98-105)
__global__ void k_thermal_update(int64_t n, const double* R, const double* PC, double* rate, double* T,
                                 const uint8_t* bcflag, const double* bcval, double alpha, double beta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (bcflag[i] & 1) {
        rate[i] = 0;
        T[i] = bcval[i];
    } else {
        const double nr = R[i] / PC[i] + beta * rate[i];
        rate[i] = nr;
        T[i] += alpha * nr;
    }
}

__global__ void k_thermal_set_bc(int64_t n, const int32_t* dofs, const double* vals, uint8_t* flag, double* bcval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        flag[dofs[i]] = 1;
        bcval[dofs[i]] = vals[i];
    }
}

}  // namespace

struct fem_thermal {
    fem_ctx* ctx = nullptr;
    fem_mesh* mesh = nullptr;
    fem_thermal_material mat;
    fem_thermal_opts opts;
    double* f[FEM_TFIELD_PHASES] = {};
    int32_t* phases = nullptr;
    double* d_tab = nullptr;
    double* d_Rs = nullptr;
    double2* d_Js = nullptr;
    uint8_t* d_bcflag = nullptr;
    double* d_bcval = nullptr;
    int32_t* d_stage = nullptr;
    double* d_stage_v = nullptr;
    int64_t stage_cap = 0;
};

namespace {

int thermal_upload_bc(fem_thermal* t, const fem_bc* bc) {
    fem_ctx* c = t->ctx;
    const int64_t nn = t->mesh->nnodes;
    HIPCHK(c, hipMemsetAsync(t->d_bcflag, 0, nn, c->stream));
    if (!bc || bc->n == 0) return FEM_OK;
    if (bc->n < 0 || !bc->host_dofs_1based || !bc->host_vals) return set_err(c, FEM_EARG, "invalid boundary condition");
    // a serial v[dofs[i]] = vals[i] lets a later duplicate win; keep that
    std::vector<int64_t> last(nn, -1);
    for (int64_t i = 0; i < bc->n; ++i) {
        const int32_t d = bc->host_dofs_1based[i] - 1;
        if (d < 0 || d >= nn) return set_err(c, FEM_EARG, "Dirichlet DoF %d out of range", d + 1);
        last[d] = i;
    }
    std::vector<int32_t> dofs;
    std::vector<double> vals;
    for (int64_t i = 0; i < bc->n; ++i) {
        const int32_t d = bc->host_dofs_1based[i] - 1;
        if (last[d] == i) { dofs.push_back(d); vals.push_back(bc->host_vals[i]); }
    }
    const int64_t nb = (int64_t)dofs.size();
    if (nb > t->stage_cap) {
        hipFree(t->d_stage);
        hipFree(t->d_stage_v);
        t->d_stage = nullptr;
        t->d_stage_v = nullptr;
        CHK(dalloc_t(c, &t->d_stage, nb));
        CHK(dalloc_t(c, &t->d_stage_v, nb));
        t->stage_cap = nb;
    }
    HIPCHK(c, hipMemcpyAsync(t->d_stage, dofs.data(), nb * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(t->d_stage_v, vals.data(), nb * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_thermal_set_bc, dim3(grid_for(nb, 256)), dim3(256), 0, c->stream, nb,
                       (const int32_t*)t->d_stage, (const double*)t->d_stage_v, t->d_bcflag, t->d_bcval);
    CHK(check_launch(c, "k_thermal_set_bc"));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}

// per-phase table and a range check of dr.phases
int thermal_prepare(fem_thermal* t) {
    fem_ctx* c = t->ctx;
    const fem_thermal_material& m = t->mat;
    std::vector<double> tab((size_t)TC_NCOL * MAXPH, 0.0);
    for (int p = 0; p < m.nphases; ++p) {
        tab[TC_K * MAXPH + p] = m.k[p];
        tab[TC_CP * MAXPH + p] = m.Cp[p];
        tab[TC_RHO0 * MAXPH + p] = m.rho0[p];
        tab[TC_ALPHA * MAXPH + p] = m.alpha[p];
        tab[TC_BETA * MAXPH + p] = 1.0 / m.K[p];
    }
    HIPCHK(c, hipMemcpyAsync(t->d_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c->stream));
    const int64_t nn = t->mesh->nnodes;
    std::vector<int32_t> ph(nn);
    HIPCHK(c, hipMemcpyAsync(ph.data(), t->phases, nn * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < nn; ++i)
        if (ph[i] < 1 || ph[i] > m.nphases)
            return set_err(c, FEM_EARG, "phase id %d of node %lld out of range 1..%d", ph[i], (long long)(i + 1),
                           m.nphases);
    return FEM_OK;
}

ThermArgs thermal_args(fem_thermal* t, double dt, double Tref) {
    fem_mesh* m = t->mesh;
    ThermArgs a;
    std::memset(&a, 0, sizeof a);
    a.nels = m->nels;
    a.el2n = m->d_el2n;
    a.geo = m->d_geo;
    if (m->nx > 0) {
        const double dx = (m->box[1] - m->box[0]) / m->nx, dy = (m->box[3] - m->box[2]) / m->ny;
        a.nx = m->nx;
        a.sx = 2 * m->nx + 1;
        a.nT6 = (int64_t)(2 * m->nx + 1) * (2 * m->ny + 1);
        a.idx = 1.0 / dx;
        a.idy = 1.0 / dy;
        a.detc = dx * dy;
    }
    a.T = t->f[FEM_TFIELD_T];
    a.T0 = t->f[FEM_TFIELD_T0];
    a.src = t->f[FEM_TFIELD_SOURCE];
    a.P = t->f[FEM_TFIELD_P];
    a.phases = t->phases;
    a.tab = t->d_tab;
    a.dt = dt;
    a.Tref = Tref;
    a.Rs = t->d_Rs;
    a.Js = t->d_Js;
    a.R = t->f[FEM_TFIELD_R];
    a.R0 = t->f[FEM_TFIELD_R0];
    a.dRdT = t->f[FEM_TFIELD_DRDT];
    a.PC = t->f[FEM_TFIELD_PC];
    a.rate = t->f[FEM_TFIELD_DTDTAU];
    a.Tw = t->f[FEM_TFIELD_T];
    a.bcflag = t->d_bcflag;
    a.bcval = t->d_bcval;
    return a;
}

// one assembly (+ the fused update when a.update), element then gather
int thermal_assemble(fem_thermal* t, ThermArgs& a, bool jac) {
    fem_ctx* c = t->ctx;
    fem_mesh* m = t->mesh;
    if (m->dim == 3 || m->nloc == 9) {
        ThermBoxArgs h;
        h.T = a.T; h.T0 = a.T0; h.src = a.src; h.P = a.P; h.phases = a.phases; h.tab = a.tab;
        h.dt = a.dt; h.Tref = a.Tref; h.Rs = a.Rs; h.Js = a.Js;
        h.R = a.R; h.R0 = a.R0; h.dRdT = a.dRdT; h.PC = a.PC; h.rate = a.rate; h.Tw = a.Tw;
        h.bcflag = a.bcflag; h.bcval = a.bcval; h.alpha = a.alpha; h.beta = a.beta; h.update = a.update;
        h.phi = a.phi;
        CHK(m->dim == 3 ? hex27_thermal_element(c, m, h, jac) : q9_thermal_element(c, m, h, jac));
    } else {
        const dim3 ge(grid_xcd(m->nels, 128));
        const bool strct = m->nx > 0;
        if (strct && jac) hipLaunchKernelGGL((k_thermal_element<true, true>), ge, dim3(128), 0, c->stream, a);
        else if (strct) hipLaunchKernelGGL((k_thermal_element<true, false>), ge, dim3(128), 0, c->stream, a);
        else if (jac) hipLaunchKernelGGL((k_thermal_element<false, true>), ge, dim3(128), 0, c->stream, a);
        else hipLaunchKernelGGL((k_thermal_element<false, false>), ge, dim3(128), 0, c->stream, a);
        CHK(check_launch(c, "k_thermal_element"));
    }
    ThermGatherArgs g;
    std::memset(&g, 0, sizeof g);
    g.ns = m->ns;
    g.node_of = m->d_node_of;
    g.off = m->d_off;
    g.ent = m->d_ent;
    g.Rs = a.Rs;
    g.Js = a.Js;
    g.R = a.R; g.R0 = a.R0; g.dRdT = a.dRdT; g.PC = a.PC; g.rate = a.rate; g.T = a.Tw;
    g.bcflag = a.bcflag;
    g.bcval = a.bcval;
    g.alpha = a.alpha;
    g.beta = a.beta;
    g.jac = jac;
    g.update = a.update;
    hipLaunchKernelGGL(k_thermal_gather, dim3(grid_xcd(m->ns, 256)), dim3(256), 0, c->stream, g);
    return check_launch(c, "k_thermal_gather");
}

// The rest of a check iteration after its Jacobian assembly: λmax before the
// update (solver.jl:97), the rate and temperature update with the current
// Chebyshev scalars (a.alpha, a.beta), then ‖R‖ and the two sums of the
// Rayleigh-quotient λmin (solver.jl:107-117), with Δτ = 2/√λmax CFL.
int thermal_check(fem_thermal* t, const ThermArgs& a, double* nr, double* den, double* num, double* dtau) {
    fem_ctx* c = t->ctx;
    const int64_t nn = t->mesh->nnodes;
    const unsigned rg = std::min<unsigned>(RED_GRID, grid_for(nn, RED_BLOCK));
    hipLaunchKernelGGL(k_reduce_lmax<double>, dim3(rg), dim3(RED_BLOCK), 0, c->stream, nn,
                       (const double*)a.dRdT, (const double*)a.dRdT, (const double*)a.PC, (const double*)a.PC,
                       c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)rg, 3, 1,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_reduce_lmax"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double lmax = c->h_red[2] > 0 ? std::numeric_limits<double>::quiet_NaN() : c->h_red[0];
    if (!(std::isfinite(lmax) && lmax > 0))
        return set_err(c, FEM_EPRECOND,
                       "thermal diffusion preconditioner produced invalid λmax = %g; check for zero or "
                       "non-finite preconditioner entries", lmax);
    hipLaunchKernelGGL(k_thermal_update, dim3(grid_for(nn, 256)), dim3(256), 0, c->stream, nn,
                       (const double*)a.R, (const double*)a.PC, a.rate, a.Tw, (const uint8_t*)a.bcflag,
                       (const double*)a.bcval, a.alpha, a.beta);
    CHK(check_launch(c, "k_thermal_update"));
    *dtau = 2 / std::sqrt(lmax) * t->opts.CFL;
    hipLaunchKernelGGL(k_reduce_check<double>, dim3(rg), dim3(RED_BLOCK), 0, c->stream, nn, (const double*)a.R,
                       (const double*)a.R, (const double*)a.R0, (const double*)a.R0, (const double*)a.rate,
                       (const double*)a.rate, (const double*)a.PC, (const double*)a.PC, *dtau, *dtau,
                       c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)rg, 6, 0,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_reduce_check"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 6 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *nr = std::sqrt(c->h_red[0]);
    *den = c->h_red[2];
    *num = c->h_red[3];
    return FEM_OK;
}

}  // namespace

// One thermal iteration of the coupled inner loop (DR.jl:604-661): R0 <- R
// and the Jacobian diagnostics at checks (thermal_iter == 1 or a multiple of
// the check cadence), the rate/temperature update, and at checks the
// relative residual, λmin and the next Chebyshev scalars.
int thermal_coupling_begin(ThermalCoupling& tc, const fem_bc* bc) {
    fem_thermal* t = tc.t;
    fem_ctx* c = t->ctx;
    const int64_t nn = t->mesh->nnodes;
    CHK(thermal_prepare(t));
    CHK(thermal_upload_bc(t, bc));
    HIPCHK(c, hipMemsetAsync(t->f[FEM_TFIELD_DTDTAU], 0, nn * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(t->f[FEM_TFIELD_R0], 0, nn * 8, c->stream));
    tc.iter = 0;
    tc.alpha = tc.beta = tc.nr0 = 0;
    tc.err_T = INFINITY;
    tc.converged = tc.checked = false;
    return FEM_OK;
}

int thermal_coupling_step(ThermalCoupling& tc) {
    fem_thermal* t = tc.t;
    fem_ctx* c = t->ctx;
    ++tc.iter;
    const bool chk = tc.iter == 1 || tc.iter % tc.nout == 0;
    tc.checked = chk;
    ThermArgs a = thermal_args(t, tc.dt, tc.Tref);
    a.phi = tc.phi;
    a.alpha = tc.alpha;
    a.beta = tc.beta;
    a.update = !chk;
    CHK(thermal_assemble(t, a, chk));
    if (!chk) return FEM_OK;
    double nr, den, num, dtau;
    CHK(thermal_check(t, a, &nr, &den, &num, &dtau));
    if (tc.iter == 1) tc.nr0 = std::max(nr, julia_eps(nr));
    tc.err_T = nr / tc.nr0;
    if (!std::isfinite(tc.err_T))
        return set_err(c, FEM_ENAN, "non-finite thermal residual at coupled iteration %lld", (long long)tc.iter);
    const double lmin = (tc.iter == 1 || den == 0) ? 0.0 : std::fabs(num) / den;
    cheb(dtau, lmin, t->opts.c_fact, &tc.alpha, &tc.beta);
    tc.converged = tc.err_T < t->opts.eps;
    return FEM_OK;
}

fem_mesh* thermal_mesh(const fem_thermal* t) { return t->mesh; }
fem_ctx* thermal_ctx(const fem_thermal* t) { return t->ctx; }
double* thermal_field_ptr(fem_thermal* t, int32_t field) { return t->f[field]; }

extern "C" {

int fem_thermal_create(fem_ctx* c, fem_mesh* m, const fem_thermal_material* mat, const fem_thermal_opts* opts,
                       fem_thermal** out) {
    if (!c || !m || !mat || !out) return FEM_EARG;
    *out = nullptr;
    if (mat->nphases < 1 || mat->nphases > FEM_MAX_PHASES)
        return set_err(c, FEM_EARG, "nphases must be in 1..%d", FEM_MAX_PHASES);
    HIPCHK(c, hipSetDevice(c->device));
    fem_thermal* t = new fem_thermal();
    t->ctx = c;
    t->mesh = m;
    t->mat = *mat;
    t->opts = opts ? *opts : fem_thermal_opts{0.98, 0.9, 1e-6};
    const int64_t nn = m->nnodes;
    int r = FEM_OK;
    for (int k = 0; k < FEM_TFIELD_PHASES && r == FEM_OK; ++k) r = dalloc_t(c, &t->f[k], nn);
    if (r == FEM_OK) r = dalloc_t(c, &t->phases, nn);
    if (r == FEM_OK) r = dalloc_t(c, &t->d_tab, TC_NCOL * MAXPH);
    // shared-node slots per element: six for T7, eight for Q9, 26 for Hex27
    const int64_t nsh = m->dim == 3 ? 26 : (m->nloc == 9 ? 8 : NSH);
    if (r == FEM_OK) r = dalloc_t(c, &t->d_Rs, nsh * m->nels);
    if (r == FEM_OK) r = dalloc_t(c, &t->d_Js, nsh * m->nels);
    if (r == FEM_OK) r = dalloc_t(c, &t->d_bcflag, nn);
    if (r == FEM_OK) r = dalloc_t(c, &t->d_bcval, nn);
    if (r != FEM_OK) {
        fem_thermal_destroy(t);
        return r;
    }
    for (int k = 0; k < FEM_TFIELD_PHASES; ++k) HIPCHK(c, hipMemsetAsync(t->f[k], 0, nn * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(t->d_bcflag, 0, nn, c->stream));
    // phases default to 1 (heat_diffusion_types.jl:126-127)
    hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(nn, 256)), dim3(256), 0, c->stream, nn, t->phases, (int32_t)1);
    CHK(check_launch(c, "k_fill"));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = t;
    return FEM_OK;
}

int fem_thermal_destroy(fem_thermal* t) {
    if (!t) return FEM_OK;
    hipStreamSynchronize(t->ctx->stream);
    for (int k = 0; k < FEM_TFIELD_PHASES; ++k) hipFree(t->f[k]);
    hipFree(t->phases); hipFree(t->d_tab); hipFree(t->d_Rs); hipFree(t->d_Js); hipFree(t->d_bcflag);
    hipFree(t->d_bcval); hipFree(t->d_stage); hipFree(t->d_stage_v);
    delete t;
    return FEM_OK;
}

int fem_thermal_field(fem_thermal* t, int32_t field, void** dptr, int64_t* len, int32_t* dtype) {
    if (!t || field < 0 || field >= FEM_TFIELD_COUNT) return FEM_EARG;
    if (dptr) *dptr = field == FEM_TFIELD_PHASES ? (void*)t->phases : (void*)t->f[field];
    if (len) *len = t->mesh->nnodes;
    if (dtype) *dtype = field == FEM_TFIELD_PHASES ? 1 : 0;
    return FEM_OK;
}

int fem_thermal_residual(fem_thermal* t, double dt, double Tref, int32_t compute_jacobian) {
    if (!t) return FEM_EARG;
    fem_ctx* c = t->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    CHK(thermal_prepare(t));
    HIPCHK(c, hipMemsetAsync(t->d_bcflag, 0, t->mesh->nnodes, c->stream));
    ThermArgs a = thermal_args(t, dt, Tref);
    CHK(thermal_assemble(t, a, compute_jacobian != 0));
    return fem_sync(c);
}

// solve!(::ThermalDiffusionDR, mesh, bc; dt, ...) -> solve_dynamic_relaxation!
// (heat_diffusion/solvers/DR.jl:50-54, dynamic_relaxation/solver.jl:61-122)
int fem_thermal_solve(fem_thermal* t, const fem_bc* bc, const fem_thermal_params* p, fem_thermal_stats* st) {
    if (!t || !p || !st) return FEM_EARG;
    fem_ctx* c = t->ctx;
    std::memset(st, 0, sizeof *st);
    HIPCHK(c, hipSetDevice(c->device));
    if (p->check_interval < 1) return set_err(c, FEM_EARG, "check_interval must be positive");
    if (!(p->dt > 0) || !std::isfinite(p->dt)) return set_err(c, FEM_EARG, "dt must be positive and finite");
    const auto t0 = std::chrono::steady_clock::now();
    CHK(thermal_prepare(t));
    CHK(thermal_upload_bc(t, bc));
    const double tol = p->tolerance >= 0 ? p->tolerance : t->opts.eps;
    ThermArgs a = thermal_args(t, p->dt, p->Tref);
    double a_dr = 0, b_dr = 0, nr0 = 0, last_rel = std::numeric_limits<double>::quiet_NaN();
    bool converged = false;
    int64_t it_done = 0;
    for (int64_t it = 1; it <= p->max_iterations; ++it) {
        it_done = it;
        const bool do_jac = (it % p->check_interval == 0) || it == 1;
        a.alpha = a_dr;
        a.beta = b_dr;
        a.update = !do_jac;
        CHK(thermal_assemble(t, a, do_jac));
        if (!do_jac) continue;
        double nr, den, num, dtau;
        CHK(thermal_check(t, a, &nr, &den, &num, &dtau));
        if (it == 1) {
            const double a_ = std::fabs(nr);
            nr0 = std::max(nr, std::nextafter(a_, std::numeric_limits<double>::infinity()) - a_);
        }
        if (std::isnan(nr / nr0)) return set_err(c, FEM_ENAN, "NaNs at PT iter %lld", (long long)it);
        const double lmin = (it == 1 || den == 0) ? 0.0 : std::fabs(num) / den;
        const double cc = 2 * std::sqrt(lmin) * t->opts.c_fact;
        a_dr = 2 * dtau * dtau / (2 + cc * dtau);
        b_dr = (2 - cc * dtau) / (2 + cc * dtau);
        last_rel = nr / nr0;
        if (p->verbose) printf("  PT %05lld  res = %6.2e\n", (long long)it, last_rel);
        if (last_rel < tol) { converged = true; break; }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    st->converged = converged;
    st->iterations = converged ? it_done : p->max_iterations;
    st->residual = last_rel;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!converged && p->throw_on_failure)
        return set_err(c, FEM_ENOCONV,
                       "Thermal diffusion DR solver did not converge after %lld pseudo-transient iterations "
                       "(relative residual = %g)", (long long)p->max_iterations, last_rel);
    return FEM_OK;
}

}  // extern "C"
