// femdr_kernels.hpp -- gfx950 device code of the T7/P1-disc DR Stokes path.
//
// One work-item per element computes the whole element update of one DR
// pseudo-iteration: continuity residual (pressure_residual.jl:132-158),
// Arrow-Hurwicz pressure Pnum = γP RP / M_P (DR.jl:674), momentum residual
// with the visco-elasto-plastic Drucker-Prager stress (momentum_residuals.jl
// :136-189, rheology.jl:226-334) and, for the element-private centroid
// (bubble) node, the Chebyshev velocity update (helpers.jl:31-41). The six
// shared local nodes are written to an element-slot buffer; a node-centric
// kernel gathers them over node->element lists in a fixed order (no atomics,
// run-to-run bitwise reproducible) and applies the same update.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fem {

constexpr int NV = 7;      // T7 velocity nodes
constexpr int NSH = 6;     // local nodes that may be shared (bubble is private)
constexpr int NP = 3;      // P1-disc pressure DoFs
constexpr int NQ = 7;      // Dunavant degree-5 points
constexpr int MAXPH = 64;

// Per-phase table columns (device global, copied to LDS per block).
enum PhaseCol {
    PC_ETA = 0, PC_INVG, PC_RHO0, PC_ALPHA, PC_BETA, PC_BULK,
    PC_COSPHI, PC_SINPHI, PC_SINPSI, PC_C, PC_EREG, PC_KB,
    PC_ETAVE, PC_INV2GDT, PC_YIELD0, PC_LAMDEN,   // dt-dependent, cell-phase fast path
    PC_NCOL
};

// Reference-element tables (shape_functions.jl:115-171, 231-239;
// integration_points.jl:188-202), filled from the host at load.
struct Tables {
    double N[NQ][NV];
    double NP1[NQ][NP];
    double dN[NQ][NV][2];
    double w[NQ];
    double TB[NP][NV][2];   // Σ_q w_q NP1_q[j] dN_q[b][k] (pressure coupling of the Jacobian)
};
extern __constant__ Tables c_tab;

template <class F> struct Vec2;
template <> struct Vec2<double> { typedef double2 T; };
template <> struct Vec2<float> { typedef float2 T; };

// ---------------------------------------------------------------------------
// forward-mode tangent over the three strain-rate components
// ---------------------------------------------------------------------------
template <class F>
struct D3 {
    F v, d0, d1, d2;
    __device__ D3() {}
    __device__ D3(F x) : v(x), d0(0), d1(0), d2(0) {}
    __device__ D3(F x, F a, F b, F c) : v(x), d0(a), d1(b), d2(c) {}
};
template <class F> __device__ inline D3<F> operator+(D3<F> a, D3<F> b) { return D3<F>(a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2); }
template <class F> __device__ inline D3<F> operator-(D3<F> a, D3<F> b) { return D3<F>(a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2); }
template <class F> __device__ inline D3<F> operator-(D3<F> a) { return D3<F>(-a.v, -a.d0, -a.d1, -a.d2); }
template <class F> __device__ inline D3<F> operator*(D3<F> a, D3<F> b) { return D3<F>(a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2); }
template <class F> __device__ inline D3<F> operator/(D3<F> a, D3<F> b) {
    F iv = F(1) / b.v; F q = a.v * iv;
    return D3<F>(q, (a.d0 - q * b.d0) * iv, (a.d1 - q * b.d1) * iv, (a.d2 - q * b.d2) * iv);
}
template <class F> __device__ inline D3<F> operator+(D3<F> a, F b) { a.v += b; return a; }
template <class F> __device__ inline D3<F> operator-(D3<F> a, F b) { a.v -= b; return a; }
template <class F> __device__ inline D3<F> operator*(D3<F> a, F b) { return D3<F>(a.v * b, a.d0 * b, a.d1 * b, a.d2 * b); }
template <class F> __device__ inline D3<F> operator*(F b, D3<F> a) { return a * b; }
template <class F> __device__ inline D3<F> operator/(D3<F> a, F b) { return D3<F>(a.v / b, a.d0 / b, a.d1 / b, a.d2 / b); }
template <class F> __device__ inline D3<F> dsqrt(D3<F> a) { F s = sqrt(a.v); F h = F(0.5) / s; return D3<F>(s, a.d0 * h, a.d1 * h, a.d2 * h); }
__device__ inline double dsqrt(double a) { return sqrt(a); }
__device__ inline float dsqrt(float a) { return sqrtf(a); }
template <class F> __device__ inline F val(D3<F> a) { return a.v; }
__device__ inline double val(double a) { return a; }
__device__ inline float val(float a) { return a; }

template <class F> __device__ inline F eps2();
template <> __device__ inline double eps2<double>() { return 4.930380657631324e-32; }   // eps(Float64)^2
template <> __device__ inline float eps2<float>() { return 1.4210855e-14f; }           // eps(Float32)^2

// Per-quadrature-point material coefficients (rheology.jl:35-42) and
// Drucker-Prager parameters interpolated at the point (rheology.jl:296-302).
template <class F>
struct IpMat {
    F etave, inv2Gdt;
    F cphi, C, sphi, spsi, ereg, Kb;
};

// Deviatoric stress with the plane-strain DP return (rheology.jl:226-245, 272-334).
template <class S, class F, bool PLASTIC>
__device__ inline void stress2d(S exx, S eyy, S exy, const IpMat<F>& m, F dt, F txo, F tyo, F txyo, F Pq,
                                S& txx, S& tyy, S& txy) {
    S tr = (exx + eyy) / F(3);
    F two_eve = F(2) * m.etave;
    txx = two_eve * ((exx - tr) + txo * m.inv2Gdt);
    tyy = two_eve * ((eyy - tr) + tyo * m.inv2Gdt);
    txy = two_eve * (exy + txyo * m.inv2Gdt);
    if (PLASTIC) {
        S tzz = -txx - tyy;
        S tII = dsqrt((txx * txx + tyy * tyy + tzz * tzz) / F(2) + txy * txy + eps2<F>());
        S F_ = (tII - m.cphi * m.C) - m.sphi * Pq;
        if (val(F_) > F(0)) {
            S tIIs = tII + eps2<F>();
            F den = (m.etave + m.ereg) + ((m.Kb * dt) * (-m.spsi)) * (-m.sphi);
            S lam = F_ / den;
            if (val(lam) > F(0)) {
                S fac = two_eve * lam;
                S dqxx = (F(2) * txx + tyy) / (F(2) * tIIs);
                S dqyy = (txx + F(2) * tyy) / (F(2) * tIIs);
                S dqxy = txy / tIIs;
                txx = txx - fac * dqxx;
                tyy = tyy - fac * dqyy;
                txy = txy - fac * dqxy;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Element kernel arguments
// ---------------------------------------------------------------------------
template <class F>
struct ElemArgs {
    int64_t nels;
    const int32_t* el2n;       // [7][nels] (SoA, 0-based)
    const F* geo;              // [5][nels]: J^-1 (00,01,10,11), |det J|
    F* vx; F* vy;              // velocity (bubble entries updated in place)
    const F* P; const F* Pold; // Pold nullable (finite bulk only)
    const F* pscale;           // γP / M_P per pressure DoF
    const F* T; const F* T0; const F* Q; const F* Pf; // nullable
    const F* tau_old0; const F* tau_old1; const F* tau_old2; // [nels][7] nullable
    F* tau_out0; F* tau_out1; F* tau_out2;                   // nullable
    const int32_t* phase_v;    // cell [nels] or element-node [7][nels] (0-based)
    const int32_t* phase_P;    // cell [nels] or element-node [3][nels] (0-based)
    const double* ptab;        // [PC_NCOL][MAXPH]
    int32_t nphases;
    int32_t bulk_all_inf;      // every phase has infinite pressure storage modulus
    int32_t thermal;           // α != 0 somewhere: read T, T0
    int32_t eos_p;             // finite K somewhere: density depends on P
    F gx, gy, Tref, dt;
    F* RP_out;                 // nullable
    F* Pnum_out;               // nullable
    typename Vec2<F>::T* Re;   // [nels][6] element-slot residuals of shared nodes
    // bubble-node DR update (mode 1) / residual store
    F* rate_x; F* rate_y; const F* PCx; const F* PCy;
    const uint8_t* bcflag; const F* bcval_x; const F* bcval_y;
    F* Rv_out_x; F* Rv_out_y;      // nullable: store residual (after Dirichlet)
    F* Rv0_out_x; F* Rv0_out_y;    // nullable
    F alpha_x, alpha_y, beta_x, beta_y;  // positive Chebyshev step and momentum
    int32_t use_pnum;
    int32_t do_update;
};

// Workgroups are dealt round-robin over the 8 XCDs (one L2 each). The grid
// is padded to a multiple of 8 and block b is given logical tile
// (b % 8) * (grid / 8) + b / 8, so each XCD works on one contiguous range of
// elements / nodes and neighbouring tiles share that XCD's L2.
__device__ inline int64_t xcd_tile() {
    const int64_t per = gridDim.x >> 3;
    return (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

template <class F, bool ELASTIC, bool PLASTIC, bool NODEPH, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_t7_element(ElemArgs<F> A) {
    __shared__ double s_tab[PC_NCOL * MAXPH];
    const int nph = A.nphases;
    for (int i = threadIdx.x; i < PC_NCOL * MAXPH; i += BLOCK)
        if ((i % MAXPH) < nph) s_tab[i] = A.ptab[i];
    __syncthreads();
    const int64_t e = xcd_tile() * BLOCK + threadIdx.x;
    if (e >= A.nels) return;
    const int64_t ne = A.nels;
#define PT(col, p) ((F)s_tab[(col) * MAXPH + (p)])

    int32_t n[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) n[a] = A.el2n[a * ne + e];
    const F i00 = A.geo[0 * ne + e], i01 = A.geo[1 * ne + e];
    const F i10 = A.geo[2 * ne + e], i11 = A.geo[3 * ne + e];
    const F det = A.geo[4 * ne + e];
    F vx[NV], vy[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) { vx[a] = A.vx[n[a]]; vy[a] = A.vy[n[a]]; }
    F P[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) P[i] = A.P[e * NP + i];

    int pv[NV];
    int pc = 0;
    if (NODEPH) {
#pragma unroll
        for (int a = 0; a < NV; ++a) pv[a] = A.phase_v[a * ne + e];
    } else {
        pc = A.phase_v[e];
    }

    // ---- velocity gradients at the quadrature points (affine element)
    F gxx[NQ], gxy[NQ], gyx[NQ], gyy[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        F ax = 0, ay = 0, bx = 0, by = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const F d0 = (F)c_tab.dN[q][a][0], d1 = (F)c_tab.dN[q][a][1];
            ax += d0 * vx[a]; ay += d1 * vx[a];
            bx += d0 * vy[a]; by += d1 * vy[a];
        }
        gxx[q] = ax * i00 + ay * i10;  // ∂vx/∂x
        gxy[q] = ax * i01 + ay * i11;  // ∂vx/∂y
        gyx[q] = bx * i00 + by * i10;  // ∂vy/∂x
        gyy[q] = bx * i01 + by * i11;  // ∂vy/∂y
    }

    // ---- continuity residual (pressure_residual.jl:132-158)
    F RP[NP] = {0, 0, 0};
    {
        int pp[NP];
        if (NODEPH) {
#pragma unroll
            for (int i = 0; i < NP; ++i) pp[i] = A.phase_P[i * ne + e];
        } else {
            const int p = A.phase_P[e];
#pragma unroll
            for (int i = 0; i < NP; ++i) pp[i] = p;
        }
        F dP[NP] = {0, 0, 0}, dT[NP] = {0, 0, 0}, Qv[NP] = {0, 0, 0};
        if (A.Pold) {
#pragma unroll
            for (int i = 0; i < NP; ++i) dP[i] = P[i] - A.Pold[e * NP + i];
        }
        if (A.thermal) {
#pragma unroll
            for (int i = 0; i < NP; ++i) dT[i] = A.T[e * NP + i] - A.T0[e * NP + i];
        }
        if (A.Q) {
#pragma unroll
            for (int i = 0; i < NP; ++i) Qv[i] = A.Q[e * NP + i];
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const F dOm = (F)c_tab.w[q] * det;
            F src = 0;
            if (A.Pold || A.thermal || A.Q) {
                F bulk = 0, alq = 0, dPq = 0, dTq = 0, Qq = 0;
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const F Ni = (F)c_tab.NP1[q][i];
                    bulk += Ni * PT(PC_BULK, pp[i]);
                    alq += Ni * PT(PC_ALPHA, pp[i]);
                    dPq += Ni * dP[i];
                    dTq += Ni * dT[i];
                    Qq += Ni * Qv[i];
                }
                if (A.bulk_all_inf) bulk = (F)INFINITY;
                src = -(dPq / (bulk * A.dt)) + alq * dTq / A.dt + Qq;
            }
            const F r = ((-(gxx[q] + gyy[q])) + src) * dOm;
#pragma unroll
            for (int i = 0; i < NP; ++i) RP[i] += (F)c_tab.NP1[q][i] * r;
        }
    }
    F Pn[NP] = {0, 0, 0};
    if (A.use_pnum) {
#pragma unroll
        for (int i = 0; i < NP; ++i) Pn[i] = A.pscale[e * NP + i] * RP[i];
    }
    if (A.RP_out) {
#pragma unroll
        for (int i = 0; i < NP; ++i) A.RP_out[e * NP + i] = RP[i];
    }
    if (A.Pnum_out) {
#pragma unroll
        for (int i = 0; i < NP; ++i) A.Pnum_out[e * NP + i] = Pn[i];
    }

    // ---- momentum residual (momentum_residuals.jl:136-189)
    IpMat<F> m;
    F rho_c = 0, alpha_c = 0, beta_c = 0;
    if (!NODEPH) {
        m.etave = PT(PC_ETAVE, pc);
        m.inv2Gdt = PT(PC_INV2GDT, pc);
        if (PLASTIC) {
            m.cphi = PT(PC_COSPHI, pc); m.C = PT(PC_C, pc); m.sphi = PT(PC_SINPHI, pc);
            m.spsi = PT(PC_SINPSI, pc); m.ereg = PT(PC_EREG, pc); m.Kb = PT(PC_KB, pc);
        }
        rho_c = PT(PC_RHO0, pc); alpha_c = PT(PC_ALPHA, pc); beta_c = PT(PC_BETA, pc);
    }
    F Rx[NV], Ry[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) { Rx[a] = 0; Ry[a] = 0; }
    const bool grav = (A.gx != F(0)) || (A.gy != F(0));
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (NODEPH) {
            F eta = 0, invG = 0;
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const F Na = (F)c_tab.N[q][a];
                eta += Na * PT(PC_ETA, pv[a]);
                invG += Na * PT(PC_INVG, pv[a]);
            }
            m.etave = F(1) / (F(1) / eta + invG / A.dt);
            m.inv2Gdt = invG / (F(2) * A.dt);
            if (PLASTIC) {
                F cphi = 0, C = 0, sphi = 0, spsi = 0, ereg = 0, Kb = 0;
#pragma unroll
                for (int a = 0; a < NV; ++a) {
                    const F Na = (F)c_tab.N[q][a];
                    cphi += Na * PT(PC_COSPHI, pv[a]); C += Na * PT(PC_C, pv[a]);
                    sphi += Na * PT(PC_SINPHI, pv[a]); spsi += Na * PT(PC_SINPSI, pv[a]);
                    ereg += Na * PT(PC_EREG, pv[a]); Kb += Na * PT(PC_KB, pv[a]);
                }
                m.cphi = cphi; m.C = C; m.sphi = sphi; m.spsi = spsi; m.ereg = ereg; m.Kb = Kb;
            }
        }
        F txo = 0, tyo = 0, txyo = 0;
        if (ELASTIC) {
            txo = A.tau_old0[e * NQ + q];
            tyo = A.tau_old1[e * NQ + q];
            txyo = A.tau_old2[e * NQ + q];
        }
        F Pq = 0, Pnq = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Pq += (F)c_tab.NP1[q][i] * P[i];
            Pnq += (F)c_tab.NP1[q][i] * Pn[i];
        }
        F Pfq = 0;
        if (A.Pf) {
#pragma unroll
            for (int i = 0; i < NP; ++i) Pfq += (F)c_tab.NP1[q][i] * A.Pf[e * NP + i];
        }
        F txx, tyy, txy;
        stress2d<F, F, PLASTIC>(gxx[q], gyy[q], (gxy[q] + gyx[q]) / F(2), m, A.dt, txo, tyo, txyo, Pq - Pfq,
                                txx, tyy, txy);
        if (A.tau_out0) {
            A.tau_out0[e * NQ + q] = txx;
            A.tau_out1[e * NQ + q] = tyy;
            A.tau_out2[e * NQ + q] = txy;
        }
        const F dOm = (F)c_tab.w[q] * det;
        const F Ptot = Pq + Pnq;
        const F sxx = txx - Ptot, syy = tyy - Ptot;
        // reference-space stress: R_i[a] = Σ_j ∂N_a/∂ξ_j s_ji
        const F s0x = (i00 * sxx + i01 * txy) * dOm;
        const F s1x = (i10 * sxx + i11 * txy) * dOm;
        const F s0y = (i00 * txy + i01 * syy) * dOm;
        const F s1y = (i10 * txy + i11 * syy) * dOm;
        F bx = 0, by = 0;
        if (grav) {
            F rho;
            if (NODEPH) {
                F r0 = 0, al = 0, be = 0;
#pragma unroll
                for (int a = 0; a < NV; ++a) {
                    const F Na = (F)c_tab.N[q][a];
                    r0 += Na * PT(PC_RHO0, pv[a]); al += Na * PT(PC_ALPHA, pv[a]); be += Na * PT(PC_BETA, pv[a]);
                }
                rho_c = r0; alpha_c = al; beta_c = be;
            }
            F Tq = 0;
            if (A.thermal) {
#pragma unroll
                for (int i = 0; i < NP; ++i) Tq += (F)c_tab.NP1[q][i] * A.T[e * NP + i];
            }
            rho = rho_c * ((F(1) - alpha_c * (Tq - A.Tref)) + beta_c * Pq);
            bx = rho * A.gx * dOm;
            by = rho * A.gy * dOm;
        }
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const F d0 = (F)c_tab.dN[q][a][0], d1 = (F)c_tab.dN[q][a][1];
            Rx[a] += d0 * s0x + d1 * s1x;
            Ry[a] += d0 * s0y + d1 * s1y;
            if (grav) {
                const F Na = (F)c_tab.N[q][a];
                Rx[a] -= Na * bx;
                Ry[a] -= Na * by;
            }
        }
    }

    // ---- shared nodes: element-slot residuals for the gather
    typedef typename Vec2<F>::T V2;
#pragma unroll
    for (int a = 0; a < NSH; ++a) {
        V2 r; r.x = Rx[a]; r.y = Ry[a];
        A.Re[e * NSH + a] = r;
    }
    // ---- private bubble node: Dirichlet, residual store, Chebyshev update
    {
        const int32_t b = n[NSH];
        const uint8_t fl = A.bcflag[b];
        F rx = (fl & 1) ? F(0) : Rx[NSH];
        F ry = (fl & 2) ? F(0) : Ry[NSH];
        if (A.Rv_out_x) { A.Rv_out_x[b] = rx; A.Rv_out_y[b] = ry; }
        if (A.Rv0_out_x) { A.Rv0_out_x[b] = rx; A.Rv0_out_y[b] = ry; }
        if (A.do_update) {
            F nrx = A.beta_x * A.rate_x[b] + rx / A.PCx[b];
            F nry = A.beta_y * A.rate_y[b] + ry / A.PCy[b];
            F nvx = vx[NSH] + (-A.alpha_x) * nrx;
            F nvy = vy[NSH] + (-A.alpha_y) * nry;
            if (fl & 1) { nrx = 0; nvx = A.bcval_x[b]; }
            if (fl & 2) { nry = 0; nvy = A.bcval_y[b]; }
            A.rate_x[b] = nrx; A.rate_y[b] = nry;
            A.vx[b] = nvx; A.vy[b] = nvy;
        }
    }
#undef PT
}

// ---------------------------------------------------------------------------
// Node-centric gather + Dirichlet + Chebyshev update (helpers.jl:31-41,
// dynamic_relaxation/solver.jl:139-166; DR.jl:686-697)
// ---------------------------------------------------------------------------
template <class F>
struct GatherArgs {
    int64_t ns;                 // shared nodes
    const int32_t* node_of;     // nullable: shared node i is node i
    const int32_t* off;         // [ns+1]
    const int32_t* ent;         // element slot e*6+a, ascending e
    const typename Vec2<F>::T* Re;
    F* vx; F* vy; F* rate_x; F* rate_y; const F* PCx; const F* PCy;
    const uint8_t* bcflag; const F* bcval_x; const F* bcval_y;
    F* Rv_out_x; F* Rv_out_y; F* Rv0_out_x; F* Rv0_out_y;
    F alpha_x, alpha_y, beta_x, beta_y;
    int32_t do_update;
};

template <class F, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_gather_update(GatherArgs<F> A) {
    const int64_t i = xcd_tile() * BLOCK + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    const int32_t k0 = A.off[i], k1 = A.off[i + 1];
    F rx = 0, ry = 0;
    for (int32_t k = k0; k < k1; ++k) {
        const typename Vec2<F>::T r = A.Re[A.ent[k]];
        rx += r.x;
        ry += r.y;
    }
    const uint8_t fl = A.bcflag[node];
    if (fl & 1) rx = 0;
    if (fl & 2) ry = 0;
    if (A.Rv_out_x) { A.Rv_out_x[node] = rx; A.Rv_out_y[node] = ry; }
    if (A.Rv0_out_x) { A.Rv0_out_x[node] = rx; A.Rv0_out_y[node] = ry; }
    if (A.do_update) {
        F nrx = A.beta_x * A.rate_x[node] + rx / A.PCx[node];
        F nry = A.beta_y * A.rate_y[node] + ry / A.PCy[node];
        F nvx = A.vx[node] + (-A.alpha_x) * nrx;
        F nvy = A.vy[node] + (-A.alpha_y) * nry;
        if (fl & 1) { nrx = 0; nvx = A.bcval_x[node]; }
        if (fl & 2) { nry = 0; nvy = A.bcval_y[node]; }
        A.rate_x[node] = nrx; A.rate_y[node] = nry;
        A.vx[node] = nvx; A.vy[node] = nvy;
    }
}

// ---------------------------------------------------------------------------
// Augmented Jacobian diagnostics (momentum_residuals.jl:929-1117): absolute
// row sums over both velocity blocks and the same-direction diagonal of
//   J = Σ_q dΩ_q (∂R/∂τ)(∂τ/∂ε̇)(∂ε̇/∂v) + B^T diag(γP/M_P) B,
// with ∂τ/∂ε̇ the exact tangent of the (DP) stress at the current state and
// B the element divergence operator of the P1-disc continuity residual. The
// element viscosity is the element maximum (momentum_residuals.jl:36-44).
// ---------------------------------------------------------------------------
template <class F>
struct JacArgs {
    int64_t nels;
    const int32_t* el2n; const F* geo;
    const F* vx; const F* vy; const F* P; const F* pscale;
    const F* tau_old0; const F* tau_old1; const F* tau_old2;
    const F* Pf;
    const int32_t* phase_v;
    const double* ptab; int32_t nphases;
    F dt;
    typename Vec2<F>::T* Jrow;   // [nels][6] (row sums x,y)
    typename Vec2<F>::T* Jdiag;  // [nels][6] (diag x,y)
    F* dRdv_x; F* dRdv_y; F* PC_x; F* PC_y;   // bubble entries written directly
};

template <class F, bool ELASTIC, bool PLASTIC, bool NODEPH, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_t7_jacobian(JacArgs<F> A) {
    __shared__ double s_tab[PC_NCOL * MAXPH];
    const int nph = A.nphases;
    for (int i = threadIdx.x; i < PC_NCOL * MAXPH; i += BLOCK)
        if ((i % MAXPH) < nph) s_tab[i] = A.ptab[i];
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= A.nels) return;
    const int64_t ne = A.nels;
#define PT(col, p) ((F)s_tab[(col) * MAXPH + (p)])
    int32_t n[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) n[a] = A.el2n[a * ne + e];
    const F iJ[2][2] = {{A.geo[0 * ne + e], A.geo[1 * ne + e]}, {A.geo[2 * ne + e], A.geo[3 * ne + e]}};
    const F det = A.geo[4 * ne + e];
    F vx[NV], vy[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) { vx[a] = A.vx[n[a]]; vy[a] = A.vy[n[a]]; }
    F P[NP], ps[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) { P[i] = A.P[e * NP + i]; ps[i] = A.pscale[e * NP + i]; }
    int pv[NV];
    F etamax;
    if (NODEPH) {
#pragma unroll
        for (int a = 0; a < NV; ++a) pv[a] = A.phase_v[a * ne + e];
        etamax = PT(PC_ETA, pv[0]);
#pragma unroll
        for (int a = 1; a < NV; ++a) etamax = fmax(etamax, PT(PC_ETA, pv[a]));
    } else {
        const int p = A.phase_v[e];
#pragma unroll
        for (int a = 0; a < NV; ++a) pv[a] = p;
        etamax = PT(PC_ETA, p);
    }
    // tangents at the quadrature points: C[q][t][m] = ∂τ_t/∂ε_m
    F Cq[NQ][3][3];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        F ax = 0, ay = 0, bx = 0, by = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const F d0 = (F)c_tab.dN[q][a][0], d1 = (F)c_tab.dN[q][a][1];
            ax += d0 * vx[a]; ay += d1 * vx[a]; bx += d0 * vy[a]; by += d1 * vy[a];
        }
        const F gxx = ax * iJ[0][0] + ay * iJ[1][0], gxy = ax * iJ[0][1] + ay * iJ[1][1];
        const F gyx = bx * iJ[0][0] + by * iJ[1][0], gyy = bx * iJ[0][1] + by * iJ[1][1];
        IpMat<F> m;
        F invG = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) invG += (F)c_tab.N[q][a] * PT(PC_INVG, pv[a]);
        m.etave = F(1) / (F(1) / etamax + invG / A.dt);
        m.inv2Gdt = invG / (F(2) * A.dt);
        if (PLASTIC) {
            F cphi = 0, C = 0, sphi = 0, spsi = 0, ereg = 0, Kb = 0;
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const F Na = (F)c_tab.N[q][a];
                cphi += Na * PT(PC_COSPHI, pv[a]); C += Na * PT(PC_C, pv[a]);
                sphi += Na * PT(PC_SINPHI, pv[a]); spsi += Na * PT(PC_SINPSI, pv[a]);
                ereg += Na * PT(PC_EREG, pv[a]); Kb += Na * PT(PC_KB, pv[a]);
            }
            m.cphi = cphi; m.C = C; m.sphi = sphi; m.spsi = spsi; m.ereg = ereg; m.Kb = Kb;
        }
        F txo = 0, tyo = 0, txyo = 0;
        if (ELASTIC) { txo = A.tau_old0[e * NQ + q]; tyo = A.tau_old1[e * NQ + q]; txyo = A.tau_old2[e * NQ + q]; }
        F Pq = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i) Pq += (F)c_tab.NP1[q][i] * P[i];
        if (A.Pf) {
#pragma unroll
            for (int i = 0; i < NP; ++i) Pq -= (F)c_tab.NP1[q][i] * A.Pf[e * NP + i];
        }
        D3<F> exx(gxx, F(1), F(0), F(0)), eyy(gyy, F(0), F(1), F(0)), exy((gxy + gyx) / F(2), F(0), F(0), F(1));
        D3<F> txx, tyy, txy;
        stress2d<D3<F>, F, PLASTIC>(exx, eyy, exy, m, A.dt, txo, tyo, txyo, Pq, txx, tyy, txy);
        Cq[q][0][0] = txx.d0; Cq[q][0][1] = txx.d1; Cq[q][0][2] = txx.d2;
        Cq[q][1][0] = tyy.d0; Cq[q][1][1] = tyy.d1; Cq[q][1][2] = tyy.d2;
        Cq[q][2][0] = txy.d0; Cq[q][2][1] = txy.d1; Cq[q][2][2] = txy.d2;
    }
    typedef typename Vec2<F>::T V2;
    for (int a = 0; a < NV; ++a) {
        F Jx[2][NV], Jy[2][NV];   // row (x,a) and row (y,a); columns (c, b)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < NV; ++b) { Jx[c][b] = 0; Jy[c][b] = 0; }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const F da0 = (F)c_tab.dN[q][a][0], da1 = (F)c_tab.dN[q][a][1];
            const F nax = da0 * iJ[0][0] + da1 * iJ[1][0];
            const F nay = da0 * iJ[0][1] + da1 * iJ[1][1];
            const F dOm = (F)c_tab.w[q] * det;
            // R_x[a] = nax τxx + nay τxy ; R_y[a] = nax τxy + nay τyy
            F wx[3], wy[3];
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) {
                wx[mm] = (nax * Cq[q][0][mm] + nay * Cq[q][2][mm]) * dOm;
                wy[mm] = (nay * Cq[q][1][mm] + nax * Cq[q][2][mm]) * dOm;
            }
            // ∂ε/∂vx_b = (∂N_b/∂x, 0, ½∂N_b/∂y); ∂ε/∂vy_b = (0, ∂N_b/∂y, ½∂N_b/∂x)
            // Σ_m w_m ∂ε_m/∂v = Σ_k dN_b,k z_k with z from J^-1
            const F zxx0 = wx[0] * iJ[0][0] + F(0.5) * wx[2] * iJ[0][1];
            const F zxx1 = wx[0] * iJ[1][0] + F(0.5) * wx[2] * iJ[1][1];
            const F zxy0 = wx[1] * iJ[0][1] + F(0.5) * wx[2] * iJ[0][0];
            const F zxy1 = wx[1] * iJ[1][1] + F(0.5) * wx[2] * iJ[1][0];
            const F zyx0 = wy[0] * iJ[0][0] + F(0.5) * wy[2] * iJ[0][1];
            const F zyx1 = wy[0] * iJ[1][0] + F(0.5) * wy[2] * iJ[1][1];
            const F zyy0 = wy[1] * iJ[0][1] + F(0.5) * wy[2] * iJ[0][0];
            const F zyy1 = wy[1] * iJ[1][1] + F(0.5) * wy[2] * iJ[1][0];
#pragma unroll
            for (int b = 0; b < NV; ++b) {
                const F d0 = (F)c_tab.dN[q][b][0], d1 = (F)c_tab.dN[q][b][1];
                Jx[0][b] += d0 * zxx0 + d1 * zxx1;
                Jx[1][b] += d0 * zxy0 + d1 * zxy1;
                Jy[0][b] += d0 * zyx0 + d1 * zyx1;
                Jy[1][b] += d0 * zyy0 + d1 * zyy1;
            }
        }
        // Powell-Hestenes coupling: + Σ_j Bm[j][i][a] ps_j Bm[j][c][b], with
        // Bm[j][c][b] = Σ_q NP1_q[j] ∂N_b/∂x_c(q) dΩ_q = det Σ_k TB[j][b][k] J^-1[k][c]
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const F ta0 = (F)c_tab.TB[j][a][0], ta1 = (F)c_tab.TB[j][a][1];
            const F fx = (ta0 * iJ[0][0] + ta1 * iJ[1][0]) * det * ps[j];
            const F fy = (ta0 * iJ[0][1] + ta1 * iJ[1][1]) * det * ps[j];
#pragma unroll
            for (int b = 0; b < NV; ++b) {
                const F t0 = (F)c_tab.TB[j][b][0], t1 = (F)c_tab.TB[j][b][1];
                const F bx = (t0 * iJ[0][0] + t1 * iJ[1][0]) * det;
                const F by = (t0 * iJ[0][1] + t1 * iJ[1][1]) * det;
                Jx[0][b] += fx * bx; Jx[1][b] += fx * by;
                Jy[0][b] += fy * bx; Jy[1][b] += fy * by;
            }
        }
        F sx = 0, sy = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < NV; ++b) { sx += fabs(Jx[c][b]); sy += fabs(Jy[c][b]); }
        F dxv = 0, dyv = 0;
#pragma unroll
        for (int b = 0; b < NV; ++b) if (b == a) { dxv = fabs(Jx[0][b]); dyv = fabs(Jy[1][b]); }
        if (a < NSH) {
            V2 r; r.x = sx; r.y = sy;
            V2 d; d.x = dxv; d.y = dyv;
            A.Jrow[e * NSH + a] = r;
            A.Jdiag[e * NSH + a] = d;
        } else {
            const int32_t bn = n[a];
            A.dRdv_x[bn] = sx; A.dRdv_y[bn] = sy;
            A.PC_x[bn] = dxv; A.PC_y[bn] = dyv;
        }
    }
#undef PT
}

template <class F>
struct GatherJacArgs {
    int64_t ns;
    const int32_t* node_of; const int32_t* off; const int32_t* ent;
    const typename Vec2<F>::T* Jrow; const typename Vec2<F>::T* Jdiag;
    F* dRdv_x; F* dRdv_y; F* PC_x; F* PC_y;
};

template <class F, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_gather_jac(GatherJacArgs<F> A) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    F jx = 0, jy = 0, dx = 0, dy = 0;
    for (int32_t k = A.off[i]; k < A.off[i + 1]; ++k) {
        const typename Vec2<F>::T r = A.Jrow[A.ent[k]];
        const typename Vec2<F>::T d = A.Jdiag[A.ent[k]];
        jx += r.x; jy += r.y; dx += d.x; dy += d.y;
    }
    A.dRdv_x[node] = jx; A.dRdv_y[node] = jy; A.PC_x[node] = dx; A.PC_y[node] = dy;
}

// ---------------------------------------------------------------------------
// Pressure mass and scale (pressure_scaling.jl:166-216); DoFs are element
// private so no atomics are needed.
// ---------------------------------------------------------------------------
template <class F, bool NODEPH, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_pressure_scaling(int64_t nels, const F* geo, const int32_t* phase_v,
                                                           const double* ptab, int32_t nphases, double gamma_fact,
                                                           double dt, F* MP, F* gammaP, F* pscale) {
    __shared__ double s_tab[PC_NCOL * MAXPH];
    for (int i = threadIdx.x; i < PC_NCOL * MAXPH; i += BLOCK)
        if ((i % MAXPH) < nphases) s_tab[i] = ptab[i];
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= nels) return;
    const double det = geo[4 * nels + e];
    int pv[NV];
    if (NODEPH) {
        for (int a = 0; a < NV; ++a) pv[a] = phase_v[a * nels + e];
    } else {
        for (int a = 0; a < NV; ++a) pv[a] = phase_v[e];
    }
    double M[NP] = {0, 0, 0}, g[NP] = {0, 0, 0};
    for (int q = 0; q < NQ; ++q) {
        double eta = 0, bq = 0;
        for (int a = 0; a < NV; ++a) {
            eta += c_tab.N[q][a] * s_tab[PC_ETA * MAXPH + pv[a]];
            bq += c_tab.N[q][a] * (1.0 / s_tab[PC_BULK * MAXPH + pv[a]]);
        }
        const double gnum = gamma_fact * eta;
        const double gphy = bq == 0.0 ? gnum : dt / bq;
        const double gq = gphy * gnum / (gphy + gnum);
        const double dO = c_tab.w[q] * det;
        for (int i = 0; i < NP; ++i) {
            const double w = c_tab.NP1[q][i] * dO;
            M[i] += w;
            g[i] += w * gq;
        }
    }
    for (int i = 0; i < NP; ++i) {
        const double gp = M[i] == 0.0 ? 0.0 : g[i] / M[i];
        MP[e * NP + i] = (F)M[i];
        gammaP[e * NP + i] = (F)gp;
        pscale[e * NP + i] = (F)(M[i] == 0.0 ? 0.0 : gp / M[i]);
    }
}

// P += γP RP / M_P (DR.jl:777)
template <class F>
__global__ void k_pressure_update(int64_t np, F* P, const F* gammaP, const F* RP, const F* MP) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) P[i] += gammaP[i] * RP[i] / MP[i];
}

// ---------------------------------------------------------------------------
// Deterministic reductions: fixed grid, fixed per-thread ranges, wave
// shuffles, then one block sums the partials in a fixed order (FP64).
// ---------------------------------------------------------------------------
constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 512;
constexpr int RED_MAXK = 8;

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}

template <int K>
__device__ inline void block_sum_store(double (&acc)[K], double* partial) {
    __shared__ double s[RED_BLOCK / 64][RED_MAXK];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double v = wave_sum(acc[k]);
        if (lane == 0) s[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double t = 0;
        for (int j = 0; j < RED_BLOCK / 64; ++j) t += s[j][threadIdx.x];
        partial[blockIdx.x * K + threadIdx.x] = t;
    }
}

// sums at a check (DR.jl:700, 32-37): |Rx|², |Ry|², λmin num/den per component
template <class F>
__global__ __launch_bounds__(RED_BLOCK) void k_reduce_check(int64_t n, const F* Rx, const F* Ry, const F* R0x,
                                                            const F* R0y, const F* rx, const F* ry, const F* PCx,
                                                            const F* PCy, double ax, double ay, double* partial) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += stride) {
        const double Rxi = Rx[i], Ryi = Ry[i];
        acc[0] += Rxi * Rxi;
        acc[1] += Ryi * Ryi;
        const double tx = ax * (double)rx[i], ty = ay * (double)ry[i];
        acc[2] += tx * tx;
        acc[3] += tx * ((Rxi - (double)R0x[i]) / (double)PCx[i]);
        acc[4] += ty * ty;
        acc[5] += ty * ((Ryi - (double)R0y[i]) / (double)PCy[i]);
    }
    block_sum_store<6>(acc, partial);
}

// sums at an outer iteration (DR.jl:556-557): |RP/M_P|², |Rx|², |Ry|²
template <class F>
__global__ __launch_bounds__(RED_BLOCK) void k_reduce_outer(int64_t nv, int64_t np, const F* Rx, const F* Ry,
                                                            const F* RP, const F* MP, double* partial) {
    double acc[3] = {0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < np; i += stride) {
        const double r = (double)RP[i] / (double)MP[i];
        acc[0] += r * r;
    }
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < nv; i += stride) {
        const double a = Rx[i], b = Ry[i];
        acc[1] += a * a;
        acc[2] += b * b;
    }
    block_sum_store<3>(acc, partial);
}

// λmax per component = max(∂Rv∂v ./ PC) with NaN propagation (utils.jl:152-158)
template <class F>
__global__ __launch_bounds__(RED_BLOCK) void k_reduce_lmax(int64_t n, const F* Jx, const F* Jy, const F* PCx,
                                                           const F* PCy, double* partial) {
    double mx = -INFINITY, my = -INFINITY, nan = 0;
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += stride) {
        const double a = (double)Jx[i] / (double)PCx[i], b = (double)Jy[i] / (double)PCy[i];
        if (a != a || b != b) nan = 1;
        mx = fmax(mx, a);
        my = fmax(my, b);
    }
    __shared__ double s[RED_BLOCK / 64][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    mx = wave_max(mx); my = wave_max(my); nan = wave_max(nan);
    if (lane == 0) { s[w][0] = mx; s[w][1] = my; s[w][2] = nan; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = s[0][0], b = s[0][1], c = s[0][2];
        for (int j = 1; j < RED_BLOCK / 64; ++j) { a = fmax(a, s[j][0]); b = fmax(b, s[j][1]); c = fmax(c, s[j][2]); }
        partial[blockIdx.x * 3 + 0] = a; partial[blockIdx.x * 3 + 1] = b; partial[blockIdx.x * 3 + 2] = c;
    }
}

// second stage: one block, K values; op 0 = sum, 1 = max
__global__ __launch_bounds__(RED_BLOCK) void k_reduce_final(int nparts, int K, int op, const double* partial,
                                                           double* out) {
    __shared__ double s[RED_BLOCK];
    for (int k = 0; k < K; ++k) {
        double v = op == 0 ? 0.0 : -INFINITY;
        for (int j = threadIdx.x; j < nparts; j += RED_BLOCK) {
            const double x = partial[j * K + k];
            v = op == 0 ? v + x : fmax(v, x);
        }
        s[threadIdx.x] = v;
        __syncthreads();
        for (int h = RED_BLOCK / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) s[threadIdx.x] = op == 0 ? s[threadIdx.x] + s[threadIdx.x + h]
                                                          : fmax(s[threadIdx.x], s[threadIdx.x + h]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = s[0];
        __syncthreads();
    }
}

template <class T>
__global__ void k_fill(int64_t n, T* p, T v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace fem
