// femdr_hex27.hip -- the 3-D path of the Stokes DR solve on structured Hex27
// velocity / P1-disc pressure meshes (QuadraticElement{3,27} with the four
// modes (1-ξ-η-ζ, ξ, η, ζ) per cell; src/mesh/mixed_mesh.jl:358-417).
//
// Work layout: each 64-lane wave serves two elements, 32 lanes each, 27 of
// them active. A kernel runs in phases separated by block barriers, with the
// element's data in LDS: lanes load the 27 nodal velocities (lane = node),
// evaluate strain rate, continuity, stress and return map at the 27 Gauss
// points (lane = point; the continuity residual is a lane reduction), and
// finally assemble the nodal residuals (lane = node) from the point fluxes.
// Shape functions are tensor products of the 1-D quadratic basis, so a lane
// holds the 1-D values at its point or node and forms each 3-D value with two
// multiplications. Structured boxes have one affine map: ∂ξ/∂x =
// diag(2/hx, 2/hy, 2/hz), |det J| = hx hy hz / 8. As in 2-D the 26 shared
// nodes of an element are summed by a gather over a node -> element-slot list
// in ascending element order; the centre node is private and is updated by
// its element.
#include "femdr_host.hpp"
#include <algorithm>
#include <array>
#include <map>

namespace {

constexpr int HNV = 27, HNQ = 27, HNP = 4, HNSH = 26;
constexpr int HBLOCK = 256, HEPB = HBLOCK / 32;   // elements per block

// QuadraticElement{3, 27} node order (src/elements/shape_functions.jl:411-446):
// 1-D node index per axis (0: -1, 1: 0, 2: +1)
constexpr int8_t kHexR[HNV][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2},
    {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2},
    {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {1, 1, 0}, {1, 0, 1}, {2, 1, 1}, {1, 2, 1},
    {0, 1, 1}, {1, 1, 2}, {1, 1, 1}};

// local node number of the node at tensor position t = i + 3 j + 9 k (the
// inverse of kHexR); the centre node 27 sits at t = 13
constexpr int8_t kHexA[HNV] = {0, 8, 1, 11, 20, 9, 3, 10, 2, 16, 21, 17, 24, 26, 22,
                               19, 23, 18, 4, 12, 5, 15, 25, 13, 7, 14, 6};
constexpr int HCENTRE_T = 13;

struct HexGeo {
    int64_t nels;
    int32_t nx, ny, nz;
    int64_t sx, sxy;                 // refined-grid strides
    double jac[3], det;              // ∂ξ/∂x per axis, |det J|
    // uploaded meshes (null on generated boxes): the connectivity [nels][27]
    // in local node order and the per-element metric [nels][4] = (∂ξ/∂x,
    // ∂η/∂y, ∂ζ/∂z, |det J|) of axis-aligned box elements of any size
    const int32_t* el2n;
    const double* metric;
};

HexGeo hex_geo(const fem_mesh* m) {
    HexGeo g;
    g.nels = m->nels;
    g.nx = m->nx; g.ny = m->ny; g.nz = m->nz;
    g.sx = 2 * (int64_t)m->nx + 1;
    g.sxy = g.sx * (2 * (int64_t)m->ny + 1);
    const double hx = (m->box[1] - m->box[0]) / m->nx, hy = (m->box[3] - m->box[2]) / m->ny,
                 hz = (m->box[5] - m->box[4]) / m->nz;
    g.jac[0] = 2 / hx; g.jac[1] = 2 / hy; g.jac[2] = 2 / hz;
    g.det = (hx / 2) * (hy / 2) * (hz / 2);
    g.el2n = nullptr;
    g.metric = nullptr;
    if (!m->structured) {
        g.nx = g.ny = g.nz = 1;
        g.sx = g.sxy = 0;
        g.jac[0] = g.jac[1] = g.jac[2] = g.det = 0;
        g.el2n = m->d_el2n;
        g.metric = m->d_geo;
    }
    return g;
}

// arguments of the residual / DR-step kernel (FP64 throughout)
struct Hex27Args {
    HexGeo g;
    const double *vx, *vy, *vz;            // written for the centre node when do_update
    const double *P, *Pold, *pscale;       // [nels][4]
    const double* tau_old[6];              // [6][nels * 27]
    double* tau_out[6];                    // nullable
    const int32_t* phase;                  // cell phases, 0-based
    const double* ptab;
    double gx, gy, gz, Tref, dt;
    double* slot;                          // [nels][26][3]
    double *rate_x, *rate_y, *rate_z;
    const double *PCx, *PCy, *PCz;
    const uint8_t* bcflag;
    const double *bcval_x, *bcval_y, *bcval_z;
    double ax, ay, az, bx, by, bz;
    int32_t do_update;
    double *RP_out, *Pnum_out;             // nullable
    double *Rv_x, *Rv_y, *Rv_z;            // nullable
    double *Rv0_x, *Rv0_y, *Rv0_z;         // nullable
    const double *bf_x, *bf_y, *bf_z;      // nullable: nodal loads subtracted from the residual
    // nullable (residual evaluations with plasticity): per element, a point at
    // or near the yield surface; the number of elements with a yielding point
    uint8_t* yield_flag;
    int32_t* yield_count;
    // nullable (a finite thermal expansivity): the temperature and its
    // previous-step value at the pressure DoFs, [nels][4]
    const double *T, *T0;
    // nullable, residual evaluations only: the shear dissipation
    // τ : (ε̇ − (τ − τ_old)/(2GΔt)) per point, [nels][27]
    double* phi_out;
};

// global node of element e at tensor position t = i + 3 j + 9 k
__device__ inline int64_t hex_node(const HexGeo& g, int64_t e, int t) {
    if (g.el2n) return g.el2n[e * HNV + kHexA[t]];
    const int64_t ex = e % g.nx, ey = (e / g.nx) % g.ny, ez = e / ((int64_t)g.nx * g.ny);
    return (2 * ez + t / 9) * g.sxy + (2 * ey + (t / 3) % 3) * g.sx + (2 * ex + t % 3);
}

// element e's geometry: the box's own on generated meshes, the element's
// metric on uploaded ones. UP = false (kernels instantiated for generated
// meshes only) keeps the box's metric uniform, out of per-lane registers
template <bool UP = true>
__device__ inline HexGeo hex_at(const HexGeo& g, int64_t e) {
    if constexpr (!UP) return g;
    HexGeo r = g;
    if (g.metric && e >= 0 && e < g.nels) {
        const double* m = g.metric + e * 4;
        r.jac[0] = m[0]; r.jac[1] = m[1]; r.jac[2] = m[2]; r.det = m[3];
    }
    return r;
}

// forward-mode dual with six partials (the strain-rate inputs of stress3d)
struct D6 {
    double v, d[6];
    __device__ D6() : v(0) { for (int i = 0; i < 6; ++i) d[i] = 0; }
    __device__ D6(double x) : v(x) { for (int i = 0; i < 6; ++i) d[i] = 0; }
};
__device__ inline D6 operator+(D6 a, const D6& b) { a.v += b.v; for (int i = 0; i < 6; ++i) a.d[i] += b.d[i]; return a; }
__device__ inline D6 operator-(D6 a, const D6& b) { a.v -= b.v; for (int i = 0; i < 6; ++i) a.d[i] -= b.d[i]; return a; }
__device__ inline D6 operator*(const D6& a, const D6& b) {
    D6 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
__device__ inline D6 operator+(D6 a, double b) { a.v += b; return a; }
__device__ inline D6 operator-(D6 a, double b) { a.v -= b; return a; }
__device__ inline D6 operator*(D6 a, double b) { a.v *= b; for (int i = 0; i < 6; ++i) a.d[i] *= b; return a; }
__device__ inline D6 operator*(double b, D6 a) { return a * b; }
__device__ inline D6 dsqrt(const D6& a) {
    D6 r;
    r.v = sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] * h;
    return r;
}
__device__ inline D6 drcp(const D6& a) {
    D6 r;
    r.v = 1.0 / a.v;
    const double h = -r.v * r.v;
    for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] * h;
    return r;
}
__device__ inline double val(const D6& a) { return a.v; }
using fem::dsqrt;
using fem::drcp;
using fem::val;

// 3-D deviatoric stress with the Drucker-Prager return (rheology.jl:413-472);
// in = (∂vx/∂x, ∂vy/∂y, ∂vz/∂z, ε̇xy, ε̇xz, ε̇yz); tau = (xx, yy, zz, xy, xz, yz)
// Returns whether the return was applied; fy (nullable) receives the yield
// function of the trial stress.
template <class S, bool PLASTIC>
__device__ inline bool stress3d(const S in[6], const IpMat<double>& m, const double to[6], double Pq, S tau[6],
                                double* fy = nullptr) {
    const S tr = (in[0] + in[1] + in[2]) * (1.0 / 3.0);
    const S e[6] = {in[0] - tr, in[1] - tr, in[2] - tr, in[3], in[4], in[5]};
    const double two_eve = 2 * m.etave;
#pragma unroll
    for (int k = 0; k < 6; ++k) tau[k] = two_eve * (e[k] + to[k] * m.inv2Gdt);
    if (PLASTIC) {
        const S tII = dsqrt((tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]) * 0.5 + tau[3] * tau[3] +
                            tau[4] * tau[4] + tau[5] * tau[5] + eps2<double>());
        const S F_ = (tII - m.yield0) - m.sphi * Pq;
        if (fy) *fy = val(F_);
        if (val(F_) > 0) {
            const S lam = F_ * m.invden;
            if (val(lam) > 0) {
                // ∂Q/∂τ = (τ_nn / (2 τII_s), τ_shear / τII_s)
                const S tIIs = tII + eps2<double>();
                const S h = two_eve * lam;
                const S r2 = drcp(tIIs * 2.0), r1 = drcp(tIIs);
#pragma unroll
                for (int k = 0; k < 3; ++k) tau[k] = tau[k] - h * (tau[k] * r2);
#pragma unroll
                for (int k = 3; k < 6; ++k) tau[k] = tau[k] - h * (tau[k] * r1);
                return true;
            }
        }
    }
    return false;
}

// the plastic correction c of stress3d's return for a visco-elastic trial
// stress t (returned stress = t - c), FP64; false where the point does not yield
__device__ inline bool dp_correction3(const double t[6], const IpMat<double>& m, double Pq, double c[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = 0;
    const double tII = sqrt((t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) * 0.5 + t[3] * t[3] + t[4] * t[4] +
                            t[5] * t[5] + eps2<double>());
    const double F_ = (tII - m.yield0) - m.sphi * Pq;
    if (!(F_ > 0)) return false;
    const double lam = F_ * m.invden;
    if (!(lam > 0)) return false;
    const double tIIs = tII + eps2<double>();
    const double h = (2 * m.etave) * lam;
    const double r2 = 1.0 / (tIIs * 2.0), r1 = 1.0 / tIIs;
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = h * (t[k] * r2);
#pragma unroll
    for (int k = 3; k < 6; ++k) c[k] = h * (t[k] * r1);
    return true;
}

__device__ inline IpMat<double> hex_ipmat(const double* ptab, int p) {
    IpMat<double> m;
    m.etave = ptab[PC_ETAVE * MAXPH + p];
    m.inv2Gdt = ptab[PC_INV2GDT * MAXPH + p];
    m.yield0 = ptab[PC_YIELD0 * MAXPH + p];
    m.sphi = ptab[PC_SINPHI * MAXPH + p];
    m.invden = ptab[PC_INVDEN * MAXPH + p];
    return m;
}

// sum over the 32 lanes of a lane group
template <class F>
__device__ inline F group_sum(F v) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 32);
    return v;
}

// 1-D basis values at a point (lane = point) or of a node (lane = node)
template <class F>
struct Lane1DT {
    F L0[3], L1[3], L2[3], d0[3], d1[3], d2[3];
};
typedef Lane1DT<double> Lane1D;
// lane = point q = qi + 3 qj + 9 qk: L0[r] = L_r(x_qi) etc.
template <class F = double>
__device__ inline Lane1DT<F> point_1d(int q) {
    Lane1DT<F> b;
    const double x0 = gauss_x(q % 3), x1 = gauss_x((q / 3) % 3), x2 = gauss_x(q / 9);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        b.L0[r] = lq(x0, r); b.L1[r] = lq(x1, r); b.L2[r] = lq(x2, r);
        b.d0[r] = dlq(x0, r); b.d1[r] = dlq(x1, r); b.d2[r] = dlq(x2, r);
    }
    return b;
}
// lane = node at tensor position t = i + 3 j + 9 k: L0[p] = L_i(x_p) etc.
template <class F = double>
__device__ inline Lane1DT<F> node_1d(int t) {
    Lane1DT<F> b;
    const int i = t % 3, j = (t / 3) % 3, k = t / 9;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const double x = gauss_x(p);
        b.L0[p] = lq(x, i); b.L1[p] = lq(x, j); b.L2[p] = lq(x, k);
        b.d0[p] = dlq(x, i); b.d1[p] = dlq(x, j); b.d2[p] = dlq(x, k);
    }
    return b;
}

// velocity gradient at the lane's point, gv[c][d] = ∂v_c/∂x_d, from the
// tensor-ordered nodal values sv[t][3], factorised over the three axes:
// ∂/∂ξ = Σ_jk L_j L_k Σ_i L'_i v_ijk, ∂/∂η and ∂/∂ζ from Σ_i L_i v_ijk
// (the ζ loop stays rolled, with its basis values recomputed, so that the
// compiler does not hold all 81 nodal values in registers at once)
template <class F>
__device__ inline void point_gradient(const HexGeo& g, const Lane1DT<F>& b, int q, const F* sv, F gv[3][3]) {
    F a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    const F x2 = (F)gauss_x(q / 9);
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        const F L2k = (F)lq(x2, k), d2k = (F)dlq(x2, k);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const F wLL = b.L1[j] * L2k, wdL = b.d1[j] * L2k, wLd = b.L1[j] * d2k;
            const F* v = sv + (3 * j + 9 * k) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const F s0 = b.d0[0] * v[c] + b.d0[1] * v[3 + c] + b.d0[2] * v[6 + c];
                const F s1 = b.L0[0] * v[c] + b.L0[1] * v[3 + c] + b.L0[2] * v[6 + c];
                a0[c] += wLL * s0;
                a1[c] += wdL * s1;
                a2[c] += wLd * s1;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gv[c][0] = a0[c] * (F)g.jac[0];
        gv[c][1] = a1[c] * (F)g.jac[1];
        gv[c][2] = a2[c] * (F)g.jac[2];
    }
}

// P1 modes at a point: (1 - ξ - η - ζ, ξ, η, ζ), and the point weight |det J| w
template <class F>
__device__ inline void point_modes(int q, F NP[HNP], F& dOm, double det) {
    const int qi = q % 3, qj = (q / 3) % 3, qk = q / 9;
    const double xi = gauss_x(qi), eta = gauss_x(qj), zeta = gauss_x(qk);
    NP[0] = (F)(1 - xi - eta - zeta); NP[1] = (F)xi; NP[2] = (F)eta; NP[3] = (F)zeta;
    dOm = (F)(det * ((gauss_w(qi) * gauss_w(qj)) * gauss_w(qk)));
}

// nodal residual of the lane's node from the point fluxes in LDS
// (accumulate_momentum, momentum_residuals.jl:28-33), factorised over the
// axes like point_gradient. sf[q][0..6] = dΩ·(τxx - P, τyy - P, τzz - P,
// τxy, τxz, τyz, ρ)
template <class F>
__device__ inline void node_residual(const HexGeo& g, const Lane1DT<F>& b, int t, const F* sf, F gx, F gy, F gz,
                                     F R[3]) {
    F rx[3] = {0, 0, 0}, ry[3] = {0, 0, 0}, rz[3] = {0, 0, 0}, rr = 0;   // ξ, η, ζ parts per row
    const int kn = t / 9;
#pragma unroll 1
    for (int r = 0; r < 3; ++r) {
        const double xr = gauss_x(r);
        const F L2r = (F)lq(xr, kn), d2r = (F)dlq(xr, kn);
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) {
            const F* f0 = sf + (3 * qq + 9 * r) * 8;
            const F* f1 = f0 + 8;
            const F* f2 = f0 + 16;
            auto dsum = [&](int k) { return b.d0[0] * f0[k] + b.d0[1] * f1[k] + b.d0[2] * f2[k]; };
            auto lsum = [&](int k) { return b.L0[0] * f0[k] + b.L0[1] * f1[k] + b.L0[2] * f2[k]; };
            const F wLL = b.L1[qq] * L2r, wdL = b.d1[qq] * L2r, wLd = b.L1[qq] * d2r;
            rx[0] += wLL * dsum(0);
            ry[0] += wLL * dsum(3);
            rz[0] += wLL * dsum(4);
            const F lxy = lsum(3), lyz = lsum(5);
            rx[1] += wdL * lxy;
            ry[1] += wdL * lsum(1);
            rz[1] += wdL * lyz;
            rx[2] += wLd * lsum(4);
            ry[2] += wLd * lyz;
            rz[2] += wLd * lsum(2);
            rr += wLL * lsum(6);
        }
    }
    const F j0 = (F)g.jac[0], j1 = (F)g.jac[1], j2 = (F)g.jac[2];
    R[0] = (rx[0] * j0 + rx[1] * j1 + rx[2] * j2) - gx * rr;
    R[1] = (ry[0] * j0 + ry[1] * j1 + ry[2] * j2) - gy * rr;
    R[2] = (rz[0] * j0 + rz[1] * j1 + rz[2] * j2) - gz * rr;
}

// element-group bookkeeping of a 256-thread block: 8 elements, 32 lanes each
struct HexLane {
    int es, l;
    int64_t e;
    bool valid;    // element exists
    bool active;   // lane serves a node / point (l < 27)
};
__device__ inline HexLane hex_lane(int64_t nels) {
    HexLane h;
    h.es = threadIdx.x >> 5;
    h.l = threadIdx.x & 31;
    h.e = xcd_tile() * HEPB + h.es;
    h.valid = h.e < nels;
    h.active = h.valid && h.l < HNV;
    return h;
}

// ---- residual: continuity (pressure_residual.jl:132-158), Pnum, momentum
// (momentum_residuals.jl:209-258), Dirichlet and Chebyshev update of the
// centre node
template <int FL>
__global__ __launch_bounds__(HBLOCK) void k_hex27_element(Hex27Args A) {
    constexpr bool ELASTIC = FL & 1, PLASTIC = FL & 2, BULK = FL & 4, ITER = FL & 8, UP = FL & 64;
    __shared__ double s_v[HEPB][HNV * 3];
    __shared__ double s_f[HEPB][HNQ * 8];
    const HexLane h = hex_lane(A.g.nels);
    const int64_t e = h.valid ? h.e : A.g.nels - 1;
    const HexGeo g = hex_at<UP>(A.g, e);
    // phase 0: nodal velocities (lane = node at tensor position h.l)
    int64_t node = 0;
    if (h.l < HNV) {
        node = hex_node(g, e, h.l);
        s_v[h.es][h.l * 3 + 0] = A.vx[node];
        s_v[h.es][h.l * 3 + 1] = A.vy[node];
        s_v[h.es][h.l * 3 + 2] = A.vz[node];
    }
    __syncthreads();
    // phase A: point quantities (lane = point)
    const int pc = A.phase[e];
    double P[HNP], Pn[HNP] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < HNP; ++i) P[i] = A.P[e * HNP + i];
    {
        const int q = h.l < HNQ ? h.l : 0;
        const Lane1D b = point_1d(q);
        double gv[3][3];
        point_gradient(g, b, q, s_v[h.es], gv);
        double NP[HNP], dOm;
        point_modes(q, NP, dOm, g.det);
        // continuity contributions of this point, then the lane-group sum
        const double div = gv[0][0] + gv[1][1] + gv[2][2];
        double dPdt = 0;
        if (BULK) {
            double dP = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i) dP += NP[i] * (P[i] - A.Pold[e * HNP + i]);
            const double bulk = A.ptab[PC_BULK * MAXPH + pc];
            double bq = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i) bq += NP[i] * bulk;
            dPdt = dP / (bq * A.dt);
        }
        // the thermal term α (T − T0) / Δt and the temperature of the density
        // (pressure_residual.jl:132-158, momentum_residuals.jl:11-16), both
        // interpolated with the pressure modes
        double dTdt = 0, Tq = 0;
        if (A.T) {
            double dT = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i) {
                dT += NP[i] * (A.T[e * HNP + i] - A.T0[e * HNP + i]);
                Tq += NP[i] * A.T[e * HNP + i];
            }
            dTdt = A.ptab[PC_ALPHA * MAXPH + pc] * dT / A.dt;
        }
        double RP[HNP];
#pragma unroll
        for (int i = 0; i < HNP; ++i) RP[i] = group_sum(h.l < HNQ ? NP[i] * (((-div) - dPdt) + dTdt) * dOm : 0.0);
        if (ITER) {
#pragma unroll
            for (int i = 0; i < HNP; ++i) Pn[i] = A.pscale[e * HNP + i] * RP[i];
        }
        if (h.valid && h.l < HNP) {
            if (A.RP_out) A.RP_out[e * HNP + h.l] = RP[h.l];
            if (A.Pnum_out) A.Pnum_out[e * HNP + h.l] = Pn[h.l];
        }
        // stress, return map, total pressure, density
        const IpMat<double> m = hex_ipmat(A.ptab, pc);
        double to[6] = {0, 0, 0, 0, 0, 0};
        if (ELASTIC && h.l < HNQ) {
#pragma unroll
            for (int k = 0; k < 6; ++k) to[k] = A.tau_old[k][e * HNQ + q];
        }
        double Pq = 0, Pnq = 0;
#pragma unroll
        for (int i = 0; i < HNP; ++i) { Pq += NP[i] * P[i]; Pnq += NP[i] * Pn[i]; }
        const double in[6] = {gv[0][0], gv[1][1], gv[2][2], (gv[0][1] + gv[1][0]) * 0.5, (gv[0][2] + gv[2][0]) * 0.5,
                              (gv[1][2] + gv[2][1]) * 0.5};
        double tau[6], fy = -INFINITY;
        const bool yel = stress3d<double, PLASTIC>(in, m, to, Pq, tau, &fy) && h.l < HNQ;
        if (PLASTIC && !ITER && A.yield_flag) {
            // lanes 0-31 and 32-63 of a wave serve two elements
            const bool near = h.l < HNQ && fy > -0.2 * fabs(m.yield0 + m.sphi * Pq);
            const int sh = 32 * (h.es & 1);
            const uint64_t bn = __ballot(near), by = __ballot(yel);
            const bool en = ((bn >> sh) & 0xffffffffull) != 0, ey = ((by >> sh) & 0xffffffffull) != 0;
            if (h.l == 0 && h.valid) A.yield_flag[e] = en;
            const uint64_t ew = __ballot(h.l == 0 && h.valid && ey);
            if ((threadIdx.x & 63) == 0 && ew) atomicAdd(A.yield_count, (int32_t)__popcll(ew));
        }
        if (A.tau_out[0] && h.active) {
#pragma unroll
            for (int k = 0; k < 6; ++k) A.tau_out[k][e * HNQ + q] = tau[k];
        }
        const double Ptot = Pq + Pnq;
        const double rho0 = A.ptab[PC_RHO0 * MAXPH + pc], al = A.ptab[PC_ALPHA * MAXPH + pc],
                     be = A.ptab[PC_BETA * MAXPH + pc];
        // eos_density (momentum_residuals.jl:11-16); T = 0 without a temperature field
        const double rho = rho0 * ((1 - al * (Tq - A.Tref)) + be * Pq);
        if (A.phi_out && h.active) {
            // shear_dissipation, 3-D (shear_heating.jl:91-103): the stress power
            // less the elastic power, shear terms counted twice
            const double power = tau[0] * gv[0][0] + tau[1] * gv[1][1] + tau[2] * gv[2][2] +
                                 tau[3] * (gv[0][1] + gv[1][0]) + tau[4] * (gv[0][2] + gv[2][0]) +
                                 tau[5] * (gv[1][2] + gv[2][1]);
            const double elastic = (tau[0] * (tau[0] - to[0]) + tau[1] * (tau[1] - to[1]) + tau[2] * (tau[2] - to[2]) +
                                    2 * (tau[3] * (tau[3] - to[3]) + tau[4] * (tau[4] - to[4]) +
                                         tau[5] * (tau[5] - to[5]))) * m.inv2Gdt;
            A.phi_out[e * HNQ + q] = power - elastic;
        }
        if (h.l < HNQ) {
            double* f = s_f[h.es] + q * 8;
            f[0] = (tau[0] - Ptot) * dOm; f[1] = (tau[1] - Ptot) * dOm; f[2] = (tau[2] - Ptot) * dOm;
            f[3] = tau[3] * dOm; f[4] = tau[4] * dOm; f[5] = tau[5] * dOm;
            f[6] = rho * dOm;
        }
    }
    __syncthreads();
    // phase B: nodal residuals (lane = node at tensor position t)
    if (!h.active) return;
    const int t = h.l;
    const Lane1D b = node_1d(t);
    double R[3];
    node_residual(g, b, t, s_f[h.es], A.gx, A.gy, A.gz, R);
    if (t != HCENTRE_T) {
        double* s = A.slot + ((e * HNSH + kHexA[t]) * 3);
        s[0] = R[0]; s[1] = R[1]; s[2] = R[2];
        return;
    }
    // centre node: private to the element
    const uint8_t fl = A.bcflag[node];
    if (A.bf_x) { R[0] -= A.bf_x[node]; R[1] -= A.bf_y[node]; R[2] -= A.bf_z[node]; }
    const double rx = (fl & 1) ? 0.0 : R[0], ry = (fl & 2) ? 0.0 : R[1], rz = (fl & 4) ? 0.0 : R[2];
    if (A.Rv_x) { A.Rv_x[node] = rx; A.Rv_y[node] = ry; A.Rv_z[node] = rz; }
    if (A.Rv0_x) { A.Rv0_x[node] = rx; A.Rv0_y[node] = ry; A.Rv0_z[node] = rz; }
    if (ITER && A.do_update) {
        double nrx = A.bx * A.rate_x[node] + rx / A.PCx[node];
        double nry = A.by * A.rate_y[node] + ry / A.PCy[node];
        double nrz = A.bz * A.rate_z[node] + rz / A.PCz[node];
        double nvx = s_v[h.es][t * 3 + 0] + (-A.ax) * nrx;
        double nvy = s_v[h.es][t * 3 + 1] + (-A.ay) * nry;
        double nvz = s_v[h.es][t * 3 + 2] + (-A.az) * nrz;
        if (fl & 1) { nrx = 0; nvx = A.bcval_x[node]; }
        if (fl & 2) { nry = 0; nvy = A.bcval_y[node]; }
        if (fl & 4) { nrz = 0; nvz = A.bcval_z[node]; }
        A.rate_x[node] = nrx; A.rate_y[node] = nry; A.rate_z[node] = nrz;
        ((double*)A.vx)[node] = nvx; ((double*)A.vy)[node] = nvy; ((double*)A.vz)[node] = nvz;
    }
}

struct Gather3Args {
    int64_t ns;
    const int32_t* node_of; const int32_t* off; const int32_t* ent;
    const double* slot;
    double *vx, *vy, *vz, *rate_x, *rate_y, *rate_z;
    const double *PCx, *PCy, *PCz;
    const uint8_t* bcflag;
    const double *bcval_x, *bcval_y, *bcval_z;
    double *Rv_x, *Rv_y, *Rv_z, *Rv0_x, *Rv0_y, *Rv0_z;
    double ax, ay, az, bx, by, bz;
    int32_t do_update;
    const double *bf_x, *bf_y, *bf_z;   // nullable: nodal loads subtracted from the gathered sum
};

__global__ __launch_bounds__(256) void k_gather3(Gather3Args A) {
    const int64_t i = xcd_tile() * 256 + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    double rx = 0, ry = 0, rz = 0;
    for (int32_t k = A.off[i]; k < A.off[i + 1]; ++k) {
        const double* s = A.slot + (int64_t)A.ent[k] * 3;
        rx += s[0]; ry += s[1]; rz += s[2];
    }
    if (A.bf_x) { rx -= A.bf_x[node]; ry -= A.bf_y[node]; rz -= A.bf_z[node]; }
    const uint8_t fl = A.bcflag[node];
    if (fl & 1) rx = 0;
    if (fl & 2) ry = 0;
    if (fl & 4) rz = 0;
    if (A.Rv_x) { A.Rv_x[node] = rx; A.Rv_y[node] = ry; A.Rv_z[node] = rz; }
    if (A.Rv0_x) { A.Rv0_x[node] = rx; A.Rv0_y[node] = ry; A.Rv0_z[node] = rz; }
    if (A.do_update) {
        double nrx = A.bx * A.rate_x[node] + rx / A.PCx[node];
        double nry = A.by * A.rate_y[node] + ry / A.PCy[node];
        double nrz = A.bz * A.rate_z[node] + rz / A.PCz[node];
        double nvx = A.vx[node] + (-A.ax) * nrx;
        double nvy = A.vy[node] + (-A.ay) * nry;
        double nvz = A.vz[node] + (-A.az) * nrz;
        if (fl & 1) { nrx = 0; nvx = A.bcval_x[node]; }
        if (fl & 2) { nry = 0; nvy = A.bcval_y[node]; }
        if (fl & 4) { nrz = 0; nvz = A.bcval_z[node]; }
        A.rate_x[node] = nrx; A.rate_y[node] = nry; A.rate_z[node] = nrz;
        A.vx[node] = nvx; A.vy[node] = nvy; A.vz[node] = nvz;
    }
}

// ---- FP32 correction operator of the mixed-precision solve. With x the FP64
// state and δ = (δv, δP) the correction, the kernel evaluates the residual
// R(x) + L δ, L the operator linearised at x: continuity with storage on δP,
// Pnum, 2 ηve ε̇'(δ) − (δP + Pnum) in the momentum balance, the density term
// ρ0 β δP g, Dirichlet nodes held at δ = 0, and the Chebyshev update of the
// centre node. With DELTA, elements flagged hot (a point at or near yield in
// the last FP64 evaluation) take the stress term τ(x + δ) − τ(x): x's strain
// rate, trial stress and both Drucker-Prager returns in FP64 from the FP64
// state; elsewhere neither state yields and the visco-elastic term is exact.
struct Lin3Args {
    HexGeo g;
    float *vx, *vy, *vz;                     // δv (centre nodes updated here)
    const float *P, *pscale, *rhs_P;         // δP, Pnum scaling, R_P(x): [nels][4]
    const float *rhs_vx, *rhs_vy, *rhs_vz;   // R_v(x): centre nodes here, shared nodes in the gather
    const int32_t* phase;
    const double* ptab;
    float gx, gy, gz, dt;
    float* slot;                             // [nels][26][3]
    float *rate_x, *rate_y, *rate_z;
    const float *PCx, *PCy, *PCz;
    const uint8_t* bcflag;
    float ax, ay, az, bx, by, bz;
    float *RP_out, *Pnum_out;                // nullable
    float *Rv_x, *Rv_y, *Rv_z;               // nullable
    float *Rv0_x, *Rv0_y, *Rv0_z;            // nullable
    const double *xvx, *xvy, *xvz, *xP;      // DELTA: the FP64 state
    const double* tau_old[6];                // DELTA and ELASTIC: [6][nels * 27]
    const uint8_t* hot;                      // DELTA: per element
};

template <int FL>
__global__ __launch_bounds__(HBLOCK) void k_hex27_lin(Lin3Args A) {
    constexpr bool BULK = FL & 1, DELTA = FL & 2, ELASTIC = FL & 4, UP = FL & 64;
    __shared__ float s_v[HEPB][HNV * 3];
    __shared__ float s_f[HEPB][HNQ * 8];
    __shared__ double s_x[DELTA ? HEPB : 1][HNV * 3];
    const HexLane h = hex_lane(A.g.nels);
    const int64_t e = h.valid ? h.e : A.g.nels - 1;
    const HexGeo g = hex_at<UP>(A.g, e);
    const bool hot_e = DELTA && A.hot[e] != 0;   // uniform over the element's 32 lanes
    int64_t node = 0;
    if (h.l < HNV) {
        node = hex_node(g, e, h.l);
        s_v[h.es][h.l * 3 + 0] = A.vx[node];
        s_v[h.es][h.l * 3 + 1] = A.vy[node];
        s_v[h.es][h.l * 3 + 2] = A.vz[node];
        if (DELTA && hot_e) {
            s_x[h.es][h.l * 3 + 0] = A.xvx[node];
            s_x[h.es][h.l * 3 + 1] = A.xvy[node];
            s_x[h.es][h.l * 3 + 2] = A.xvz[node];
        }
    }
    __syncthreads();
    const int pc = A.phase[e];
    float P[HNP];
#pragma unroll
    for (int i = 0; i < HNP; ++i) P[i] = A.P[e * HNP + i];
    {
        const int q = h.l < HNQ ? h.l : 0;
        const Lane1DT<float> b = point_1d<float>(q);
        float gv[3][3];
        point_gradient<float>(g, b, q, s_v[h.es], gv);
        float NP[HNP], dOm;
        point_modes<float>(q, NP, dOm, g.det);
        const float div = gv[0][0] + gv[1][1] + gv[2][2];
        float dPdt = 0;
        if (BULK) {
            float dP = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i) dP += NP[i] * P[i];
            const float bulk = (float)A.ptab[PC_BULK * MAXPH + pc];
            float bq = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i) bq += NP[i] * bulk;
            dPdt = dP / (bq * A.dt);
        }
        float RP[HNP], Pn[HNP];
#pragma unroll
        for (int i = 0; i < HNP; ++i) {
            RP[i] = group_sum(h.l < HNQ ? NP[i] * ((-div) - dPdt) * dOm : 0.0f) + A.rhs_P[e * HNP + i];
            Pn[i] = A.pscale[e * HNP + i] * RP[i];
        }
        if (h.valid && h.l < HNP) {
            if (A.RP_out) A.RP_out[e * HNP + h.l] = RP[h.l];
            if (A.Pnum_out) A.Pnum_out[e * HNP + h.l] = Pn[h.l];
        }
        const float eve2 = 2.0f * (float)A.ptab[PC_ETAVE * MAXPH + pc];
        const float tr = div * (1.0f / 3.0f);
        float tau[6] = {eve2 * (gv[0][0] - tr), eve2 * (gv[1][1] - tr), eve2 * (gv[2][2] - tr),
                        eve2 * ((gv[0][1] + gv[1][0]) * 0.5f), eve2 * ((gv[0][2] + gv[2][0]) * 0.5f),
                        eve2 * ((gv[1][2] + gv[2][1]) * 0.5f)};
        float Pq = 0, Pnq = 0;
#pragma unroll
        for (int i = 0; i < HNP; ++i) { Pq += NP[i] * P[i]; Pnq += NP[i] * Pn[i]; }
        if (DELTA && hot_e && h.l < HNQ) {
            // x's strain rate and trial stress as the FP64 residual forms them,
            // then the returns of x + δ and of x
            const Lane1D bd = point_1d<double>(q);
            double gx[3][3];
            point_gradient<double>(g, bd, q, s_x[h.es], gx);
            double NPd[HNP], dOmd;
            point_modes<double>(q, NPd, dOmd, g.det);
            double Pqx = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i) Pqx += NPd[i] * A.xP[e * HNP + i];
            const IpMat<double> m = hex_ipmat(A.ptab, pc);
            double to[6] = {0, 0, 0, 0, 0, 0};
            if (ELASTIC) {
#pragma unroll
                for (int k = 0; k < 6; ++k) to[k] = A.tau_old[k][e * HNQ + q];
            }
            const double in[6] = {gx[0][0], gx[1][1], gx[2][2], (gx[0][1] + gx[1][0]) * 0.5,
                                  (gx[0][2] + gx[2][0]) * 0.5, (gx[1][2] + gx[2][1]) * 0.5};
            const double trx = (in[0] + in[1] + in[2]) * (1.0 / 3.0);
            const double ex[6] = {in[0] - trx, in[1] - trx, in[2] - trx, in[3], in[4], in[5]};
            const double two_eve = 2 * m.etave;
            double tB[6], tA[6], cA[6], cB[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                tB[k] = two_eve * (ex[k] + to[k] * m.inv2Gdt);
                tA[k] = tB[k] + (double)tau[k];
            }
            const bool yA = dp_correction3(tA, m, Pqx + (double)Pq, cA);
            const bool yB = dp_correction3(tB, m, Pqx, cB);
            if (yA || yB) {
#pragma unroll
                for (int k = 0; k < 6; ++k) tau[k] = (float)((double)tau[k] + (cB[k] - cA[k]));
            }
        }
        const float Ptot = Pq + Pnq;
        const float rb = (float)(A.ptab[PC_RHO0 * MAXPH + pc] * A.ptab[PC_BETA * MAXPH + pc]);
        if (h.l < HNQ) {
            float* f = s_f[h.es] + q * 8;
            f[0] = (tau[0] - Ptot) * dOm; f[1] = (tau[1] - Ptot) * dOm; f[2] = (tau[2] - Ptot) * dOm;
            f[3] = tau[3] * dOm; f[4] = tau[4] * dOm; f[5] = tau[5] * dOm;
            f[6] = (rb * Pq) * dOm;
        }
    }
    __syncthreads();
    if (!h.active) return;
    const int t = h.l;
    const Lane1DT<float> b = node_1d<float>(t);
    float R[3];
    node_residual<float>(g, b, t, s_f[h.es], A.gx, A.gy, A.gz, R);
    if (t != HCENTRE_T) {
        float* s = A.slot + ((e * HNSH + kHexA[t]) * 3);
        s[0] = R[0]; s[1] = R[1]; s[2] = R[2];
        return;
    }
    const uint8_t fl = A.bcflag[node];
    R[0] += A.rhs_vx[node]; R[1] += A.rhs_vy[node]; R[2] += A.rhs_vz[node];
    const float rx = (fl & 1) ? 0.0f : R[0], ry = (fl & 2) ? 0.0f : R[1], rz = (fl & 4) ? 0.0f : R[2];
    if (A.Rv_x) { A.Rv_x[node] = rx; A.Rv_y[node] = ry; A.Rv_z[node] = rz; }
    if (A.Rv0_x) { A.Rv0_x[node] = rx; A.Rv0_y[node] = ry; A.Rv0_z[node] = rz; }
    float nrx = A.bx * A.rate_x[node] + rx / A.PCx[node];
    float nry = A.by * A.rate_y[node] + ry / A.PCy[node];
    float nrz = A.bz * A.rate_z[node] + rz / A.PCz[node];
    float nvx = s_v[h.es][t * 3 + 0] + (-A.ax) * nrx;
    float nvy = s_v[h.es][t * 3 + 1] + (-A.ay) * nry;
    float nvz = s_v[h.es][t * 3 + 2] + (-A.az) * nrz;
    if (fl & 1) { nrx = 0; nvx = 0; }
    if (fl & 2) { nry = 0; nvy = 0; }
    if (fl & 4) { nrz = 0; nvz = 0; }
    A.rate_x[node] = nrx; A.rate_y[node] = nry; A.rate_z[node] = nrz;
    A.vx[node] = nvx; A.vy[node] = nvy; A.vz[node] = nvz;
}

// the shared nodes of the correction: slot sums plus R_v(x), Dirichlet, update
struct Gather3fArgs {
    int64_t ns;
    const int32_t* node_of; const int32_t* off; const int32_t* ent;
    const float* slot;
    float *vx, *vy, *vz, *rate_x, *rate_y, *rate_z;
    const float *PCx, *PCy, *PCz;
    const uint8_t* bcflag;
    const float *rhs_vx, *rhs_vy, *rhs_vz;
    float *Rv_x, *Rv_y, *Rv_z, *Rv0_x, *Rv0_y, *Rv0_z;
    float ax, ay, az, bx, by, bz;
};

__global__ __launch_bounds__(256) void k_gather3f(Gather3fArgs A) {
    const int64_t i = xcd_tile() * 256 + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of[i];
    float rx = 0, ry = 0, rz = 0;
    for (int32_t k = A.off[i]; k < A.off[i + 1]; ++k) {
        const float* s = A.slot + (int64_t)A.ent[k] * 3;
        rx += s[0]; ry += s[1]; rz += s[2];
    }
    rx += A.rhs_vx[node]; ry += A.rhs_vy[node]; rz += A.rhs_vz[node];
    const uint8_t fl = A.bcflag[node];
    if (fl & 1) rx = 0;
    if (fl & 2) ry = 0;
    if (fl & 4) rz = 0;
    if (A.Rv_x) { A.Rv_x[node] = rx; A.Rv_y[node] = ry; A.Rv_z[node] = rz; }
    if (A.Rv0_x) { A.Rv0_x[node] = rx; A.Rv0_y[node] = ry; A.Rv0_z[node] = rz; }
    float nrx = A.bx * A.rate_x[node] + rx / A.PCx[node];
    float nry = A.by * A.rate_y[node] + ry / A.PCy[node];
    float nrz = A.bz * A.rate_z[node] + rz / A.PCz[node];
    float nvx = A.vx[node] + (-A.ax) * nrx;
    float nvy = A.vy[node] + (-A.ay) * nry;
    float nvz = A.vz[node] + (-A.az) * nrz;
    if (fl & 1) { nrx = 0; nvx = 0; }
    if (fl & 2) { nry = 0; nvy = 0; }
    if (fl & 4) { nrz = 0; nvz = 0; }
    A.rate_x[node] = nrx; A.rate_y[node] = nry; A.rate_z[node] = nrz;
    A.vx[node] = nvx; A.vy[node] = nvy; A.vz[node] = nvz;
}

// ---- augmented Jacobian diagnostics (momentum_residuals.jl:919-1117): per
// column (comp j, node b) of the element operator, the directional
// derivative of the momentum residual with Pnum(v) inline, from the exact
// tangent of the (DP) stress at each point (6-partial duals)
struct Jac3Args {
    HexGeo g;
    const double *vx, *vy, *vz, *P, *pscale;
    const double* tau_old[6];
    const int32_t* phase;
    const double* ptab;
    double* jslot;                   // [nels][26][6]: |J| row sums (x, y, z), |J_ii| (x, y, z)
    double *dRdv_x, *dRdv_y, *dRdv_z, *PC_x, *PC_y, *PC_z;
};

template <bool ELASTIC, bool PLASTIC>
__global__ __launch_bounds__(HBLOCK) void k_hex27_jacobian(Jac3Args A) {
    __shared__ double s_v[HEPB][HNV * 3];
    __shared__ double s_f[HEPB][HNQ * 8];
    const HexLane h = hex_lane(A.g.nels);
    const int64_t e = h.valid ? h.e : A.g.nels - 1;
    const HexGeo g = hex_at(A.g, e);
    int64_t node = 0;
    if (h.l < HNV) {
        node = hex_node(g, e, h.l);
        s_v[h.es][h.l * 3 + 0] = A.vx[node];
        s_v[h.es][h.l * 3 + 1] = A.vy[node];
        s_v[h.es][h.l * 3 + 2] = A.vz[node];
    }
    __syncthreads();
    const int pc = A.phase[e];
    const int q = h.l < HNQ ? h.l : 0;
    const Lane1D bq = point_1d(q);
    double NP[HNP], dOm;
    point_modes(q, NP, dOm, g.det);
    // tangent of the stress at this point: T[k][m] = ∂τ_k/∂in_m
    double T[6][6];
    {
        double gv[3][3];
        point_gradient(g, bq, q, s_v[h.es], gv);
        const IpMat<double> m = hex_ipmat(A.ptab, pc);
        double to[6] = {0, 0, 0, 0, 0, 0};
        if (ELASTIC && h.l < HNQ) {
#pragma unroll
            for (int k = 0; k < 6; ++k) to[k] = A.tau_old[k][e * HNQ + q];
        }
        double Pq = 0;
#pragma unroll
        for (int i = 0; i < HNP; ++i) Pq += NP[i] * A.P[e * HNP + i];
        D6 in[6] = {D6(gv[0][0]), D6(gv[1][1]), D6(gv[2][2]), D6((gv[0][1] + gv[1][0]) * 0.5),
                    D6((gv[0][2] + gv[2][0]) * 0.5), D6((gv[1][2] + gv[2][1]) * 0.5)};
#pragma unroll
        for (int k = 0; k < 6; ++k) in[k].d[k] = 1.0;
        D6 tau[6];
        stress3d<D6, PLASTIC>(in, m, to, Pq, tau);
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int mm = 0; mm < 6; ++mm) T[k][mm] = tau[k].d[mm];
    }
    double ps[HNP];
#pragma unroll
    for (int i = 0; i < HNP; ++i) ps[i] = A.pscale[e * HNP + i];
    const int t = h.l < HNV ? h.l : 0;   // lane's node (tensor position) in phase B
    const Lane1D bn = node_1d(t);
    double rs[3] = {0, 0, 0}, dg[3] = {0, 0, 0};
    // columns (velocity component j, node at tensor position tb)
    for (int col = 0; col < 3 * HNV; ++col) {
        const int j = col / HNV, tb = col % HNV;
        // phase A (lane = point): derivative of the point flux
        {
            const int r0 = tb % 3, r1 = (tb / 3) % 3, r2 = tb / 9;
            const double Gb[3] = {(bq.d0[r0] * bq.L1[r1]) * bq.L2[r2] * g.jac[0],
                                  (bq.L0[r0] * bq.d1[r1]) * bq.L2[r2] * g.jac[1],
                                  (bq.L0[r0] * bq.L1[r1]) * bq.d2[r2] * g.jac[2]};
            double dgv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int k = 0; k < 3; ++k) dgv[j][k] = Gb[k];
            const double din[6] = {dgv[0][0], dgv[1][1], dgv[2][2], (dgv[0][1] + dgv[1][0]) * 0.5,
                                   (dgv[0][2] + dgv[2][0]) * 0.5, (dgv[1][2] + dgv[2][1]) * 0.5};
            const double ddiv = dgv[0][0] + dgv[1][1] + dgv[2][2];
            double dPn = 0;
#pragma unroll
            for (int i = 0; i < HNP; ++i)
                dPn += NP[i] * (ps[i] * group_sum(h.l < HNQ ? NP[i] * (-ddiv) * dOm : 0.0));
            double dt_[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                double t = 0;
#pragma unroll
                for (int mm = 0; mm < 6; ++mm) t += T[k][mm] * din[mm];
                dt_[k] = t;
            }
            if (h.l < HNQ) {
                double* f = s_f[h.es] + q * 8;
                f[0] = (dt_[0] - dPn) * dOm; f[1] = (dt_[1] - dPn) * dOm; f[2] = (dt_[2] - dPn) * dOm;
                f[3] = dt_[3] * dOm; f[4] = dt_[4] * dOm; f[5] = dt_[5] * dOm;
                f[6] = 0.0;
            }
        }
        __syncthreads();
        // phase B (lane = node): entries J[(i, a), (j, b)]
        double Jc[3];
        node_residual(g, bn, t, s_f[h.es], 0.0, 0.0, 0.0, Jc);
#pragma unroll
        for (int i = 0; i < 3; ++i) rs[i] += fabs(Jc[i]);
        if (t == tb) dg[j] = fabs(Jc[j]);
        __syncthreads();
    }
    if (!h.active) return;
    if (t != HCENTRE_T) {
        double* s = A.jslot + (e * HNSH + kHexA[t]) * 6;
        s[0] = rs[0]; s[1] = rs[1]; s[2] = rs[2];
        s[3] = dg[0]; s[4] = dg[1]; s[5] = dg[2];
        return;
    }
    A.dRdv_x[node] = rs[0]; A.dRdv_y[node] = rs[1]; A.dRdv_z[node] = rs[2];
    A.PC_x[node] = dg[0]; A.PC_y[node] = dg[1]; A.PC_z[node] = dg[2];
}

struct GatherJ3Args {
    int64_t ns;
    const int32_t* node_of; const int32_t* off; const int32_t* ent;
    const double* jslot;
    double *dRdv_x, *dRdv_y, *dRdv_z, *PC_x, *PC_y, *PC_z;
};

__global__ __launch_bounds__(256) void k_gather_jac3(GatherJ3Args A) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t k = A.off[i]; k < A.off[i + 1]; ++k) {
        const double* s = A.jslot + (int64_t)A.ent[k] * 6;
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c] += s[c];
    }
    A.dRdv_x[node] = t[0]; A.dRdv_y[node] = t[1]; A.dRdv_z[node] = t[2];
    A.PC_x[node] = t[3]; A.PC_y[node] = t[4]; A.PC_z[node] = t[5];
}

// ---- assemble_viscosity_weighted_pressure_scaling! with the Hex27 Jacobi
// weight N_a² dΩ (pressure_scaling.jl:166-216); one thread per element
__global__ void k_hex27_scaling(HexGeo g0, const int32_t* phase, const double* ptab, double gamma_fact, double dt,
                                double* MP, double* gammaP, double* pscale) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g0.nels) return;
    const HexGeo g = hex_at(g0, e);
    const int pc = phase[e];
    const double eta = ptab[PC_ETA * MAXPH + pc], bulk = ptab[PC_BULK * MAXPH + pc];
    const double invK = 1.0 / bulk;
    double M[HNP] = {0, 0, 0, 0}, G[HNP] = {0, 0, 0, 0};
    for (int q = 0; q < HNQ; ++q) {
        const double x0 = gauss_x(q % 3), x1 = gauss_x((q / 3) % 3), x2 = gauss_x(q / 9);
        // interp2ip_phase with the velocity basis (partition of unity), in
        // the local node order
        double etaq = 0, bq = 0;
#pragma unroll
        for (int a = 0; a < HNV; ++a) {
            const double N = (lq(x0, kHexR[a][0]) * lq(x1, kHexR[a][1])) * lq(x2, kHexR[a][2]);
            etaq += N * eta;
            bq += N * invK;
        }
        const double gnum = gamma_fact * etaq;
        const double gphy = bq == 0 ? gnum : dt / bq;
        const double gq = gphy * gnum / (gphy + gnum);
        double NP[HNP], dOm;
        point_modes(q, NP, dOm, g.det);
#pragma unroll
        for (int i = 0; i < HNP; ++i) {
            const double w = NP[i] * NP[i] * dOm;
            M[i] += w;
            G[i] += w * gq;
        }
    }
#pragma unroll
    for (int i = 0; i < HNP; ++i) {
        MP[e * HNP + i] = M[i];
        const double gp = M[i] == 0 ? 0.0 : G[i] / M[i];
        gammaP[e * HNP + i] = gp;
        pscale[e * HNP + i] = M[i] == 0 ? 0.0 : gp / M[i];
    }
}

__global__ void k_pscale3(int64_t n, const double* g, const double* M, double* ps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ps[i] = M[i] == 0.0 ? 0.0 : g[i] / M[i];
}

// ---- reductions for three components
template <class T>
struct Red3T {
    const T *R[3], *R0[3], *rate[3], *PC[3], *J[3];
    double al[3];
};
typedef Red3T<double> Red3;

// inner check sums (FP64 accumulation) over the FP64 state or the FP32 correction
template <class T>
__global__ __launch_bounds__(RED_BLOCK) void k_red3_check(int64_t n, Red3T<T> r, double* partial) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double Ri = r.R[c][i];
            acc[c] += Ri * Ri;
            const double t = r.al[c] * (double)r.rate[c][i];
            acc[3 + 2 * c] += t * t;
            acc[4 + 2 * c] += t * ((Ri - (double)r.R0[c][i]) / (double)r.PC[c][i]);
        }
    }
    block_sum_store<9>(acc, partial);
}

__global__ __launch_bounds__(RED_BLOCK) void k_red3_outer(int64_t nv, int64_t np, Red3 r, const double* RP,
                                                          const double* MP, double* partial) {
    double acc[4] = {0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < np; i += stride) {
        const double x = RP[i] / MP[i];
        acc[0] += x * x;
    }
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < nv; i += stride)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[1 + c] += r.R[c][i] * r.R[c][i];
    block_sum_store<4>(acc, partial);
}

__global__ __launch_bounds__(RED_BLOCK) void k_red3_lmax(int64_t n, Red3 r, double* partial) {
    double mx[3] = {-INFINITY, -INFINITY, -INFINITY}, nan = 0;
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += stride)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double a = r.J[c][i] / r.PC[c][i];
            if (a != a) nan = 1;
            mx[c] = fmax(mx[c], a);
        }
    __shared__ double s[RED_BLOCK / 64][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) mx[c] = wave_max(mx[c]);
    nan = wave_max(nan);
    if (lane == 0) { s[w][0] = mx[0]; s[w][1] = mx[1]; s[w][2] = mx[2]; s[w][3] = nan; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {s[0][0], s[0][1], s[0][2], s[0][3]};
        for (int j = 1; j < RED_BLOCK / 64; ++j)
            for (int c = 0; c < 4; ++c) t[c] = fmax(t[c], s[j][c]);
        for (int c = 0; c < 4; ++c) partial[blockIdx.x * 4 + c] = t[c];
    }
}

// ---- host side
struct Opts3 {
    bool elastic = false, plastic = false, bulk = false;
    bool thermal = false;            // a finite expansivity: the T, T0 fields enter the residual
    double* phi_out = nullptr;       // residual evaluations: the shear dissipation per point
    double dt = 1;
    bool use_pnum = false, do_update = false, store_RP = false, store_Pnum = false, store_Rv = false, store_Rv0 = false,
         store_tau = false;
    double al[3] = {0, 0, 0}, be[3] = {0, 0, 0};
    uint8_t* yield_flag = nullptr;   // residual evaluations with plasticity: near-yield elements
    int32_t* yield_count = nullptr;  // and the count of yielding ones (added to)
};

int check3(fem_ctx* c, const char* what) { return check_launch(c, what); }

template <int FL>
void launch_h27(fem_stokes* s, const Hex27Args& a) {
    if (s->mesh->structured)
        hipLaunchKernelGGL((k_hex27_element<FL>), dim3(grid_xcd(s->mesh->nels, HEPB)), dim3(HBLOCK), 0,
                           s->ctx->stream, a);
    else
        hipLaunchKernelGGL((k_hex27_element<FL | 64>), dim3(grid_xcd(s->mesh->nels, HEPB)), dim3(HBLOCK), 0,
                           s->ctx->stream, a);
}
template <int FL>
void dispatch_h27(int key, fem_stokes* s, const Hex27Args& a) {
    if (key == FL) launch_h27<FL>(s, a);
    else if constexpr (FL > 0) dispatch_h27<FL - 1>(key, s, a);
}

int hex_element(fem_stokes* s, const Opts3& o, bool gather = true) {
    fem_mesh* m = s->mesh;
    double** f = s->f;
    Hex27Args a;
    std::memset(&a, 0, sizeof a);
    a.g = hex_geo(m);
    a.vx = f[FEM_FIELD_V_X]; a.vy = f[FEM_FIELD_V_Y]; a.vz = f[FEM_FIELD_V_Z];
    a.P = f[FEM_FIELD_P];
    a.Pold = (s->in.P_old && !s->in.P_old_at_points) ? s->in.P_old : f[FEM_FIELD_P0];
    a.pscale = s->d_pscale;
    a.bf_x = s->in.body_force[0]; a.bf_y = s->in.body_force[1]; a.bf_z = s->in.body_force[2];
    const int tf[6] = {FEM_FIELD_TAU_OLD_XX, FEM_FIELD_TAU_OLD_YY, FEM_FIELD_TAU_OLD_ZZ,
                       FEM_FIELD_TAU_OLD_XY, FEM_FIELD_TAU_OLD_XZ, FEM_FIELD_TAU_OLD_YZ};
    const int tc[6] = {FEM_FIELD_TAU_XX, FEM_FIELD_TAU_YY, FEM_FIELD_TAU_ZZ, FEM_FIELD_TAU_XY, FEM_FIELD_TAU_XZ,
                       FEM_FIELD_TAU_YZ};
    for (int k = 0; k < 6; ++k) {
        a.tau_old[k] = f[tf[k]];
        a.tau_out[k] = o.store_tau ? f[tc[k]] : nullptr;
    }
    a.phase = s->d_phase_v;
    a.ptab = s->d_ptab;
    a.gx = s->mat.g[0]; a.gy = s->mat.g[1]; a.gz = s->mat.g[2];
    a.Tref = s->mat.Tref;
    a.dt = o.dt;
    a.slot = s->d_slot3;
    a.rate_x = f[FEM_FIELD_DVDTAU_X]; a.rate_y = f[FEM_FIELD_DVDTAU_Y]; a.rate_z = f[FEM_FIELD_DVDTAU_Z];
    a.PCx = f[FEM_FIELD_PC_V_X]; a.PCy = f[FEM_FIELD_PC_V_Y]; a.PCz = f[FEM_FIELD_PC_V_Z];
    a.bcflag = s->d_bcflag;
    a.bcval_x = s->d_bcval_x; a.bcval_y = s->d_bcval_y; a.bcval_z = s->d_bcval_z;
    a.ax = o.al[0]; a.ay = o.al[1]; a.az = o.al[2];
    a.bx = o.be[0]; a.by = o.be[1]; a.bz = o.be[2];
    a.do_update = o.do_update;
    a.RP_out = o.store_RP ? f[FEM_FIELD_RP] : nullptr;
    a.Pnum_out = o.store_Pnum ? f[FEM_FIELD_PNUM] : nullptr;
    if (o.store_Rv) { a.Rv_x = f[FEM_FIELD_RV_X]; a.Rv_y = f[FEM_FIELD_RV_Y]; a.Rv_z = f[FEM_FIELD_RV_Z]; }
    if (o.store_Rv0) { a.Rv0_x = f[FEM_FIELD_RV0_X]; a.Rv0_y = f[FEM_FIELD_RV0_Y]; a.Rv0_z = f[FEM_FIELD_RV0_Z]; }
    if (o.plastic && !o.use_pnum && o.yield_flag && o.yield_count) {
        a.yield_flag = o.yield_flag;
        a.yield_count = o.yield_count;
    }
    if (o.thermal) {
        a.T = f[FEM_FIELD_T];
        a.T0 = f[FEM_FIELD_T0];
    }
    a.phi_out = o.use_pnum ? nullptr : o.phi_out;
    const int key = (o.elastic ? 1 : 0) | (o.plastic ? 2 : 0) | (o.bulk ? 4 : 0) | (o.use_pnum ? 8 : 0);
    dispatch_h27<15>(key, s, a);
    CHK(check3(s->ctx, "k_hex27_element"));
    if (!gather) return FEM_OK;
    Gather3Args g;
    std::memset(&g, 0, sizeof g);
    g.ns = m->ns;
    g.node_of = m->d_node_of; g.off = m->d_off; g.ent = m->d_ent;
    g.slot = s->d_slot3;
    g.vx = f[FEM_FIELD_V_X]; g.vy = f[FEM_FIELD_V_Y]; g.vz = f[FEM_FIELD_V_Z];
    g.rate_x = a.rate_x; g.rate_y = a.rate_y; g.rate_z = a.rate_z;
    g.PCx = a.PCx; g.PCy = a.PCy; g.PCz = a.PCz;
    g.bcflag = s->d_bcflag;
    g.bcval_x = s->d_bcval_x; g.bcval_y = s->d_bcval_y; g.bcval_z = s->d_bcval_z;
    g.Rv_x = a.Rv_x; g.Rv_y = a.Rv_y; g.Rv_z = a.Rv_z;
    g.Rv0_x = a.Rv0_x; g.Rv0_y = a.Rv0_y; g.Rv0_z = a.Rv0_z;
    g.ax = a.ax; g.ay = a.ay; g.az = a.az; g.bx = a.bx; g.by = a.by; g.bz = a.bz;
    g.do_update = o.do_update;
    g.bf_x = a.bf_x; g.bf_y = a.bf_y; g.bf_z = a.bf_z;
    hipLaunchKernelGGL(k_gather3, dim3(grid_xcd(m->ns, 256)), dim3(256), 0, s->ctx->stream, g);
    return check3(s->ctx, "k_gather3");
}

int hex_jacobian_impl(fem_stokes* s, const Opts3& o) {
    fem_mesh* m = s->mesh;
    double** f = s->f;
    Jac3Args a;
    std::memset(&a, 0, sizeof a);
    a.g = hex_geo(m);
    a.vx = f[FEM_FIELD_V_X]; a.vy = f[FEM_FIELD_V_Y]; a.vz = f[FEM_FIELD_V_Z];
    a.P = f[FEM_FIELD_P]; a.pscale = s->d_pscale;
    const int tf[6] = {FEM_FIELD_TAU_OLD_XX, FEM_FIELD_TAU_OLD_YY, FEM_FIELD_TAU_OLD_ZZ,
                       FEM_FIELD_TAU_OLD_XY, FEM_FIELD_TAU_OLD_XZ, FEM_FIELD_TAU_OLD_YZ};
    for (int k = 0; k < 6; ++k) a.tau_old[k] = f[tf[k]];
    a.phase = s->d_phase_v;
    a.ptab = s->d_ptab;
    a.jslot = s->d_jslot3;
    a.dRdv_x = f[FEM_FIELD_DRVDV_X]; a.dRdv_y = f[FEM_FIELD_DRVDV_Y]; a.dRdv_z = f[FEM_FIELD_DRVDV_Z];
    a.PC_x = f[FEM_FIELD_PC_V_X]; a.PC_y = f[FEM_FIELD_PC_V_Y]; a.PC_z = f[FEM_FIELD_PC_V_Z];
    const dim3 grid(grid_xcd(m->nels, HEPB)), block(HBLOCK);
    if (o.elastic && o.plastic) hipLaunchKernelGGL((k_hex27_jacobian<true, true>), grid, block, 0, s->ctx->stream, a);
    else if (o.elastic) hipLaunchKernelGGL((k_hex27_jacobian<true, false>), grid, block, 0, s->ctx->stream, a);
    else if (o.plastic) hipLaunchKernelGGL((k_hex27_jacobian<false, true>), grid, block, 0, s->ctx->stream, a);
    else hipLaunchKernelGGL((k_hex27_jacobian<false, false>), grid, block, 0, s->ctx->stream, a);
    CHK(check3(s->ctx, "k_hex27_jacobian"));
    GatherJ3Args g;
    g.ns = m->ns;
    g.node_of = m->d_node_of; g.off = m->d_off; g.ent = m->d_ent;
    g.jslot = s->d_jslot3;
    g.dRdv_x = a.dRdv_x; g.dRdv_y = a.dRdv_y; g.dRdv_z = a.dRdv_z;
    g.PC_x = a.PC_x; g.PC_y = a.PC_y; g.PC_z = a.PC_z;
    hipLaunchKernelGGL(k_gather_jac3, dim3(grid_for(m->ns, 256)), dim3(256), 0, s->ctx->stream, g);
    return check3(s->ctx, "k_gather_jac3");
}

Red3 red3(fem_stokes* s) {
    Red3 r;
    double** f = s->f;
    r.R[0] = f[FEM_FIELD_RV_X]; r.R[1] = f[FEM_FIELD_RV_Y]; r.R[2] = f[FEM_FIELD_RV_Z];
    r.R0[0] = f[FEM_FIELD_RV0_X]; r.R0[1] = f[FEM_FIELD_RV0_Y]; r.R0[2] = f[FEM_FIELD_RV0_Z];
    r.rate[0] = f[FEM_FIELD_DVDTAU_X]; r.rate[1] = f[FEM_FIELD_DVDTAU_Y]; r.rate[2] = f[FEM_FIELD_DVDTAU_Z];
    r.PC[0] = f[FEM_FIELD_PC_V_X]; r.PC[1] = f[FEM_FIELD_PC_V_Y]; r.PC[2] = f[FEM_FIELD_PC_V_Z];
    r.J[0] = f[FEM_FIELD_DRVDV_X]; r.J[1] = f[FEM_FIELD_DRVDV_Y]; r.J[2] = f[FEM_FIELD_DRVDV_Z];
    r.al[0] = r.al[1] = r.al[2] = 0;
    return r;
}

// launches a two-stage reduction and copies K results to host
template <class L>
int reduce3(fem_stokes* s, int K, int op, int64_t n, L launch, double* out) {
    fem_ctx* c = s->ctx;
    const unsigned g = std::min<unsigned>(RED_GRID, grid_for(n, RED_BLOCK));
    launch(g);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, K, op,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_red3"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, K * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_red, K * 8);
    return FEM_OK;
}

int checked_lmax3(fem_stokes* s, double lmax[3]) {
    fem_ctx* c = s->ctx;
    const Red3 r = red3(s);
    const int64_t nv = s->mesh->nnodes;
    double out[4];
    CHK(reduce3(s, 4, 1, nv, [&](unsigned g) {
        hipLaunchKernelGGL(k_red3_lmax, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, r, c->red_partial);
    }, out));
    const char* labels[3] = {"stokes vx", "stokes vy", "stokes vz"};
    for (int k = 0; k < 3; ++k) {
        lmax[k] = out[3] > 0 ? std::numeric_limits<double>::quiet_NaN() : out[k];
        if (!(std::isfinite(lmax[k]) && lmax[k] > 0))
            return set_err(c, FEM_EPRECOND,
                           "%s preconditioner produced invalid λmax = %g; check for zero or non-finite preconditioner entries",
                           labels[k], lmax[k]);
    }
    return FEM_OK;
}

int prepare3(fem_stokes* s, const fem_drucker_prager* dp, double dt, int finite_K, Opts3& o, bool use_tau_old) {
    bool bulk_inf, thermal, eos_p, elastic;
    if (!(dt > 0) || !std::isfinite(dt)) return set_err(s->ctx, FEM_EARG, "dt must be positive and finite");
    if (s->layout_v != FEM_PHASE_CELL || s->layout_P != FEM_PHASE_CELL || !s->phase_same)
        return set_err(s->ctx, FEM_EARG, "the 3-D path takes one phase per cell (a 1 x nels phase matrix)");
    CHK(check_solve_inputs(s, dp != nullptr));
    CHK(upload_ptab(s, dp, dt, finite_K, &bulk_inf, &thermal, &eos_p, &elastic));
    o.thermal = thermal;
    o.plastic = dp != nullptr;
    o.elastic = elastic && use_tau_old;
    o.bulk = !bulk_inf;
    o.dt = dt;
    return FEM_OK;
}

// ---------------------------------------------------------------------------
// ThermalDiffusionDR on a Hex27 box: the backward-Euler heat residual of
// heat_diffusion/assembly/residual.jl:58-87 with per-node phases interpolated
// at the 27 Gauss points, and on check iterations the Jacobian row sums
// Σ_j |J_ij| and |J_ii| of the exact derivative (residual.jl:30-47 takes it with
// ForwardDiff): with D = Δt/(ρ Cp), ρ = ρ0 (1 - α (T - Tref) + β P),
// ∂D/∂T_j = D ρ0 α N_j / ρ. 32 lanes per element as in the Stokes kernel:
// lane = node (gather), lane = point (point quantities), lane = node (nodal
// rows); the 26 shared nodes write element slots, the private centre node is
// finished (and with `update` relaxed) by its element.
// ---------------------------------------------------------------------------
template <bool JAC>
__global__ __launch_bounds__(HBLOCK) void k_thermal_hex27(HexGeo g0, ThermBoxArgs A) {
    const int es = threadIdx.x >> 5, l = threadIdx.x & 31;
    const int64_t e = xcd_tile() * HEPB + es;
    const HexGeo g = hex_at(g0, e);
    const bool ev = e < g.nels;
    __shared__ double sT[HEPB][HNV], sT0[HEPB][HNV], sS[HEPB][HNV], sP[HEPB][HNV];
    __shared__ double sK[HEPB][HNV], sCp[HEPB][HNV], sR0[HEPB][HNV], sAl[HEPB][HNV], sBe[HEPB][HNV];
    __shared__ double qRate[HEPB][HNQ], qD[HEPB][HNQ], qC[HEPB][HNQ], qK[HEPB][HNQ], qS[HEPB][HNQ];
    __shared__ double qW[HEPB][HNQ], qG[HEPB][HNQ][3];
    if (ev && l < HNV) {
        const int64_t n = hex_node(g, e, l);
        sT[es][l] = A.T[n]; sT0[es][l] = A.T0[n]; sS[es][l] = A.src[n]; sP[es][l] = A.P[n];
        const int p = A.phases[n] - 1;
        sK[es][l] = A.tab[TC_K * MAXPH + p]; sCp[es][l] = A.tab[TC_CP * MAXPH + p];
        sR0[es][l] = A.tab[TC_RHO0 * MAXPH + p]; sAl[es][l] = A.tab[TC_ALPHA * MAXPH + p];
        sBe[es][l] = A.tab[TC_BETA * MAXPH + p];
    }
    __syncthreads();
    if (ev && l < HNQ) {
        const int q = l;
        const double x0 = gauss_x(q % 3), x1 = gauss_x((q / 3) % 3), x2 = gauss_x(q / 9);
        double Tq = 0, T0q = 0, sq = 0, Pq = 0, kq = 0, cq = 0, rq = 0, aq = 0, bq = 0;
        double gT[3] = {0, 0, 0};
        for (int t = 0; t < HNV; ++t) {
            const int i = t % 3, j = (t / 3) % 3, k = t / 9;
            const double L0 = lq(x0, i), L1 = lq(x1, j), L2 = lq(x2, k);
            const double N = L0 * L1 * L2;
            const double v = sT[es][t];
            Tq += N * v; T0q += N * sT0[es][t]; sq += N * sS[es][t]; Pq += N * sP[es][t];
            kq += N * sK[es][t]; cq += N * sCp[es][t]; rq += N * sR0[es][t]; aq += N * sAl[es][t];
            bq += N * sBe[es][t];
            gT[0] += dlq(x0, i) * L1 * L2 * v;
            gT[1] += L0 * dlq(x1, j) * L2 * v;
            gT[2] += L0 * L1 * dlq(x2, k) * v;
        }
        for (int d = 0; d < 3; ++d) gT[d] *= g.jac[d];
        const double rho = rq * (1 - aq * (Tq - A.Tref) + bq * Pq);
        const double D = A.dt / (rho * cq);
        const double phq = A.phi ? A.phi[e * HNQ + q] : 0.0;
        double rate = T0q - Tq + D * sq;
        if (A.phi) rate += D * phq;   // Δt/(ρ Cp) Φ_q (residual.jl:75-79)
        const double dOm = g.det * gauss_w(q % 3) * gauss_w((q / 3) % 3) * gauss_w(q / 9);
        qRate[es][q] = rate;
        qD[es][q] = D;
        qC[es][q] = D * rq * aq / rho;
        qK[es][q] = kq;
        qS[es][q] = sq + phq;
        qW[es][q] = dOm;
        qG[es][q][0] = gT[0]; qG[es][q][1] = gT[1]; qG[es][q][2] = gT[2];
    }
    __syncthreads();
    if (!ev || l >= HNV) return;
    const int t = l, ti = t % 3, tj = (t / 3) % 3, tk = t / 9;
    double r = 0;
    for (int q = 0; q < HNQ; ++q) {
        const double x0 = gauss_x(q % 3), x1 = gauss_x((q / 3) % 3), x2 = gauss_x(q / 9);
        const double L0 = lq(x0, ti), L1 = lq(x1, tj), L2 = lq(x2, tk);
        const double N = L0 * L1 * L2;
        const double g0 = dlq(x0, ti) * L1 * L2 * g.jac[0], g1 = L0 * dlq(x1, tj) * L2 * g.jac[1],
                     g2 = L0 * L1 * dlq(x2, tk) * g.jac[2];
        const double KTi = qK[es][q] * (g0 * qG[es][q][0] + g1 * qG[es][q][1] + g2 * qG[es][q][2]);
        r += (qRate[es][q] * N - qD[es][q] * KTi) * qW[es][q];
    }
    double rs = 0, dg = 0;
    if (JAC) {
        for (int u = 0; u < HNV; ++u) {
            const int ui = u % 3, uj = (u / 3) % 3, uk = u / 9;
            double J = 0;
            for (int q = 0; q < HNQ; ++q) {
                const double x0 = gauss_x(q % 3), x1 = gauss_x((q / 3) % 3), x2 = gauss_x(q / 9);
                const double L0 = lq(x0, ti), L1 = lq(x1, tj), L2 = lq(x2, tk);
                const double M0 = lq(x0, ui), M1 = lq(x1, uj), M2 = lq(x2, uk);
                const double Nt = L0 * L1 * L2, Nu = M0 * M1 * M2;
                const double gt0 = dlq(x0, ti) * L1 * L2 * g.jac[0], gt1 = L0 * dlq(x1, tj) * L2 * g.jac[1],
                             gt2 = L0 * L1 * dlq(x2, tk) * g.jac[2];
                const double gu0 = dlq(x0, ui) * M1 * M2 * g.jac[0], gu1 = M0 * dlq(x1, uj) * M2 * g.jac[1],
                             gu2 = M0 * M1 * dlq(x2, uk) * g.jac[2];
                const double k = qK[es][q];
                const double KTi = k * (gt0 * qG[es][q][0] + gt1 * qG[es][q][1] + gt2 * qG[es][q][2]);
                const double Kij = k * (gt0 * gu0 + gt1 * gu1 + gt2 * gu2);
                const double dD = qC[es][q] * Nu;
                J += ((-Nu + dD * qS[es][q]) * Nt - dD * KTi - qD[es][q] * Kij) * qW[es][q];
            }
            rs += fabs(J);
            if (u == t) dg = fabs(J);
        }
    }
    if (t != HCENTRE_T) {
        const int64_t sl = e * HNSH + kHexA[t];
        A.Rs[sl] = r;
        if (JAC) A.Js[sl] = make_double2(rs, dg);
        return;
    }
    // the private centre node
    const int64_t b = hex_node(g, e, t);
    const bool fixed = A.bcflag[b] & 1;
    const double rr = fixed ? 0.0 : r;
    if (JAC) {
        A.R0[b] = A.R[b];           // R0 <- R before the assembly (solver.jl:95)
        A.dRdT[b] = rs;
        A.PC[b] = dg;
    }
    A.R[b] = rr;
    if (A.update) {
        const double nr = fixed ? 0.0 : rr / A.PC[b] + A.beta * A.rate[b];
        A.rate[b] = nr;
        A.Tw[b] = fixed ? A.bcval[b] : sT[es][t] + A.alpha * nr;
    }
}

}  // namespace

int hex27_thermal_element(fem_ctx* c, const fem_mesh* m, const ThermBoxArgs& a, bool jac) {
    const HexGeo g = hex_geo(m);
    const dim3 grid(grid_xcd(m->nels, HEPB));
    if (jac) hipLaunchKernelGGL(k_thermal_hex27<true>, grid, dim3(HBLOCK), 0, c->stream, g, a);
    else hipLaunchKernelGGL(k_thermal_hex27<false>, grid, dim3(HBLOCK), 0, c->stream, g, a);
    return check_launch(c, "k_thermal_hex27");
}

namespace {

}  // namespace

// ---------------------------------------------------------------------------
// entry points (behind fem_stokes_* for 3-D states)
// ---------------------------------------------------------------------------
// the scaling kernel alone, with the phase table already on the device
static int hex_scaling_launch(fem_stokes* s, double gamma_fact, double dt) {
    fem_mesh* m = s->mesh;
    hipLaunchKernelGGL(k_hex27_scaling, dim3(grid_for(m->nels, 128)), dim3(128), 0, s->ctx->stream, hex_geo(m),
                       (const int32_t*)s->d_phase_v, (const double*)s->d_ptab, gamma_fact, dt, s->f[FEM_FIELD_M_P],
                       s->f[FEM_FIELD_GAMMA_P], s->d_pscale);
    return check_launch(s->ctx, "k_hex27_scaling");
}

int hex27_pressure_scaling(fem_stokes* s, double gamma_fact, double dt, int finite_K) {
    Opts3 o;
    CHK(prepare3(s, nullptr, dt, finite_K, o, true));
    return hex_scaling_launch(s, gamma_fact, dt);
}

static int hex_pscale(fem_stokes* s) {
    const int64_t np = s->mesh->nnodesP;
    hipLaunchKernelGGL(k_pscale3, dim3(grid_for(np, 256)), dim3(256), 0, s->ctx->stream, np,
                       (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    return check_launch(s->ctx, "k_pscale3");
}

int hex27_residual(fem_stokes* s, const fem_drucker_prager* dp, double dt, bool with_pnum, int finite_K) {
    Opts3 o;
    CHK(prepare3(s, dp, dt, finite_K, o, true));
    CHK(hex_pscale(s));
    HIPCHK(s->ctx, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, s->ctx->stream));
    o.use_pnum = with_pnum;
    o.store_RP = true;
    o.store_Pnum = with_pnum;
    o.store_Rv = true;
    return hex_element(s, o);
}

int hex27_jacobian(fem_stokes* s, const fem_drucker_prager* dp, double dt, int finite_K) {
    Opts3 o;
    CHK(prepare3(s, dp, dt, finite_K, o, true));
    CHK(hex_pscale(s));
    return hex_jacobian_impl(s, o);
}

// rotate_stress! (src/stokes/stress_rotation.jl:28-61) in 3-D: at every
// Gauss point ω = curl v and τ_old = R τ Rᵀ with R the rotation by the
// vector ωΔt/2 (Rodrigues' formula; no rotation where ω = 0). One thread per
// (element, point): the 27 nodal velocities are read per point through the
// caches; the step runs once per time step, outside the DR loop.
struct RotArgs3 {
    HexGeo g;
    const double *vx, *vy, *vz;
    const double* tau[6];   // assembler order xx, yy, zz, xy, xz, yz
    double* out[6];
    double dt;
};

__global__ __launch_bounds__(256) void k_hex27_rotate(RotArgs3 A) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.g.nels * HNQ) return;
    const int64_t e = i / HNQ;
    const int q = (int)(i - e * HNQ);
    const HexGeo g = hex_at(A.g, e);
    const double x0 = gauss_x(q % 3), x1 = gauss_x((q / 3) % 3), x2 = gauss_x(q / 9);
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // G[c][k] = ∂v_c/∂x_k
    for (int t = 0; t < HNV; ++t) {
        const int r0 = t % 3, r1 = (t / 3) % 3, r2 = t / 9;
        const double L0 = lq(x0, r0), L1 = lq(x1, r1), L2 = lq(x2, r2);
        const double d[3] = {dlq(x0, r0) * L1 * L2 * g.jac[0], L0 * dlq(x1, r1) * L2 * g.jac[1],
                             L0 * L1 * dlq(x2, r2) * g.jac[2]};
        const int64_t n = hex_node(g, e, t);
        const double v[3] = {A.vx[n], A.vy[n], A.vz[n]};
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k) G[c][k] += d[k] * v[c];
    }
    const double r[3] = {0.5 * A.dt * (G[2][1] - G[1][2]), 0.5 * A.dt * (G[0][2] - G[2][0]),
                         0.5 * A.dt * (G[1][0] - G[0][1])};
    const double th = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double t[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) t[c] = A.tau[c][i];
    if (th == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c) A.out[c][i] = t[c];
        return;
    }
    const double n0 = r[0] / th, n1 = r[1] / th, n2 = r[2] / th;
    const double sn = sin(th), cm = 1 - cos(th);
    // R = I + sin θ K + (1 − cos θ) K², K the cross-product matrix of n (K² = n nᵀ − I)
    const double R[3][3] = {{1 + cm * (n0 * n0 - 1), -sn * n2 + cm * n0 * n1, sn * n1 + cm * n0 * n2},
                            {sn * n2 + cm * n1 * n0, 1 + cm * (n1 * n1 - 1), -sn * n0 + cm * n1 * n2},
                            {-sn * n1 + cm * n2 * n0, sn * n0 + cm * n2 * n1, 1 + cm * (n2 * n2 - 1)}};
    const double T[3][3] = {{t[0], t[3], t[4]}, {t[3], t[1], t[5]}, {t[4], t[5], t[2]}};
    double RT[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) RT[a][b] = R[a][0] * T[0][b] + R[a][1] * T[1][b] + R[a][2] * T[2][b];
    auto M = [&](int a, int b) { return RT[a][0] * R[b][0] + RT[a][1] * R[b][1] + RT[a][2] * R[b][2]; };
    A.out[0][i] = M(0, 0);
    A.out[1][i] = M(1, 1);
    A.out[2][i] = M(2, 2);
    A.out[3][i] = M(0, 1);
    A.out[4][i] = M(0, 2);
    A.out[5][i] = M(1, 2);
}

int hex27_rotate_stress(fem_stokes* s, double dt) {
    RotArgs3 a;
    a.g = hex_geo(s->mesh);
    a.vx = s->f[FEM_FIELD_V_X];
    a.vy = s->f[FEM_FIELD_V_Y];
    a.vz = s->f[FEM_FIELD_V_Z];
    const int tc[6] = {FEM_FIELD_TAU_XX, FEM_FIELD_TAU_YY, FEM_FIELD_TAU_ZZ, FEM_FIELD_TAU_XY, FEM_FIELD_TAU_XZ,
                       FEM_FIELD_TAU_YZ};
    const int tf[6] = {FEM_FIELD_TAU_OLD_XX, FEM_FIELD_TAU_OLD_YY, FEM_FIELD_TAU_OLD_ZZ,
                       FEM_FIELD_TAU_OLD_XY, FEM_FIELD_TAU_OLD_XZ, FEM_FIELD_TAU_OLD_YZ};
    for (int k = 0; k < 6; ++k) {
        a.tau[k] = s->f[tc[k]];
        a.out[k] = s->f[tf[k]];
    }
    a.dt = dt;
    const int64_t n = a.g.nels * HNQ;
    hipLaunchKernelGGL(k_hex27_rotate, dim3(grid_for(n, 256)), dim3(256), 0, s->ctx->stream, a);
    return check_launch(s->ctx, "k_hex27_rotate");
}

// compute_strain_rate_stress_postprocess(v, el2n_v, geo_v, τ_ip, element_v)
// (postprocess.jl:169-212), one thread per cell: the dΩ-weighted cell averages
// of ε̇ (xx, yy, zz, xy, xz, yz), ε̇II of its deviatoric part, the six point
// stresses and τII, the invariants with the eps² floor of second_invariant
// (rheology.jl:395-399); out is [14][nels]
struct Post3Args {
    HexGeo g;
    const double *vx, *vy, *vz;
    const double* tau[6];   // [nels][27] each, (xx, yy, zz, xy, xz, yz)
    double* out;
};

__global__ __launch_bounds__(128) void k_postprocess3(Post3Args A) {
    const int64_t e = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (e >= A.g.nels) return;
    const HexGeo g = hex_at(A.g, e);
    double sv[HNV * 3];
    for (int t = 0; t < HNV; ++t) {
        const int64_t n = hex_node(g, e, t);
        sv[t * 3 + 0] = A.vx[n]; sv[t * 3 + 1] = A.vy[n]; sv[t * 3 + 2] = A.vz[n];
    }
    constexpr double eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
    double tot[14];
    for (int k = 0; k < 14; ++k) tot[k] = 0;
    double vol = 0;
    for (int q = 0; q < HNQ; ++q) {
        const Lane1D b = point_1d(q);
        double gv[3][3];
        point_gradient(g, b, q, sv, gv);
        const double ex = gv[0][0], ey = gv[1][1], ez = gv[2][2];
        const double exy = (gv[0][1] + gv[1][0]) / 2, exz = (gv[0][2] + gv[2][0]) / 2, eyz = (gv[1][2] + gv[2][1]) / 2;
        const double tr = (ex + ey + ez) / 3;
        const double dx = ex - tr, dy = ey - tr, dz = ez - tr;
        double t6[6];
        for (int c = 0; c < 6; ++c) t6[c] = A.tau[c][e * HNQ + q];
        const double val[14] = {
            ex, ey, ez, exy, exz, eyz,
            sqrt((dx * dx + dy * dy + dz * dz) / 2 + exy * exy + exz * exz + eyz * eyz + eps2),
            t6[0], t6[1], t6[2], t6[3], t6[4], t6[5],
            sqrt((t6[0] * t6[0] + t6[1] * t6[1] + t6[2] * t6[2]) / 2 + t6[3] * t6[3] + t6[4] * t6[4] +
                 t6[5] * t6[5] + eps2)};
        const int qi = q % 3, qj = (q / 3) % 3, qk = q / 9;
        const double dOm = g.det * ((gauss_w(qi) * gauss_w(qj)) * gauss_w(qk));
        for (int k = 0; k < 14; ++k) tot[k] += val[k] * dOm;
        vol += dOm;
    }
    for (int k = 0; k < 14; ++k) A.out[k * g.nels + e] = tot[k] / vol;
}

int hex27_postprocess(fem_stokes* s, double* out) {
    Post3Args a;
    std::memset(&a, 0, sizeof a);
    a.g = hex_geo(s->mesh);
    double** f = s->f;
    a.vx = f[FEM_FIELD_V_X]; a.vy = f[FEM_FIELD_V_Y]; a.vz = f[FEM_FIELD_V_Z];
    const int tc[6] = {FEM_FIELD_TAU_XX, FEM_FIELD_TAU_YY, FEM_FIELD_TAU_ZZ, FEM_FIELD_TAU_XY, FEM_FIELD_TAU_XZ,
                       FEM_FIELD_TAU_YZ};
    for (int k = 0; k < 6; ++k) a.tau[k] = f[tc[k]];
    a.out = out;
    hipLaunchKernelGGL(k_postprocess3, dim3(grid_for(a.g.nels, 128)), dim3(128), 0, s->ctx->stream, a);
    return check_launch(s->ctx, "k_postprocess3");
}

// ===========================================================================
// CellPressureStokesDR (stokes_types.jl:545-613, DR.jl:960-1057): the viscous
// Hex27 system with the four cell-local pressure modes (1, ξ, η, ζ), relaxed
// by an Arrow-Hurwicz pressure step (the constant mode's mass-weighted mean
// removed) and a Jacobi velocity step. The state is a 3-D fem_stokes: P and
// RP hold the modes per cell, PC_v the Jacobi diagonal, M_P the mode mass.
// One iteration: k_cp3_pressure (RP, P += ps RP/M, partial sums of P_0 M_0),
// the deterministic reduction of those sums, k_cp3_momentum (the mean
// removed from P_0 as it is read, the residual, the centre node's step) and
// k_gather3 (the shared nodes' step, v -= vs R/D).
// ===========================================================================
struct CP3Args {
    HexGeo g;
    double *vx, *vy, *vz;
    double* P;                     // [nels][4]
    double* RP;                    // [nels][4]
    double* M;                     // [nels][4] mode mass (written by the preconditioner)
    const int32_t* phase;          // cell phases, 0-based
    const double* ptab;
    double gx, gy, gz;
    double ps, vs;                 // pressure_step, velocity_step
    double* part;                  // [tiles]: Σ P_0 M_0 of each tile's cells (pressure pass)
    const double* psum;            // nullable: Σ_cells P_0 M_0 (momentum pass); the mean is removed
    double inv_sumM;               // 1 / Σ_cells M_0
    double* slot;                  // [nels][26][3]
    double *Dx, *Dy, *Dz;          // Jacobi diagonal
    const uint8_t* bcflag;
    const double *bcval_x, *bcval_y, *bcval_z;
    double *Rv_x, *Rv_y, *Rv_z;    // nullable: the residual of the centre nodes
    int32_t update;                // pressure pass: P += ps RP/M; momentum pass: the centre node's step
    const double *bf_x, *bf_y, *bf_z;   // nullable: a momentum load subtracted before the Dirichlet mask
};

// the modes (1, ξ, η, ζ) at a point (pressure_mode_table, momentum_residuals.jl:618-621)
__device__ inline void point_modes_cp(int q, double NP[HNP], double& dOm, double det) {
    const int qi = q % 3, qj = (q / 3) % 3, qk = q / 9;
    NP[0] = 1; NP[1] = gauss_x(qi); NP[2] = gauss_x(qj); NP[3] = gauss_x(qk);
    dOm = det * ((gauss_w(qi) * gauss_w(qj)) * gauss_w(qk));
}

// stokes_preconditioner_3d_kernel! (momentum_residuals.jl:724-745): lane =
// node t, the Jacobi diagonal η (|∇N_t|² + (∂N_t/∂x_c)²/3) dΩ over the 27
// points into the element slot (the centre node directly); lanes 0-3 the
// mode mass Σ_q mode² dΩ
__global__ __launch_bounds__(HBLOCK) void k_cp3_precond(CP3Args A) {
    const HexLane h = hex_lane(A.g.nels);
    if (!h.active) return;
    const int64_t e = h.e;
    const HexGeo g = hex_at(A.g, e);
    const double eta = A.ptab[PC_ETA * MAXPH + A.phase[e]];
    const int t = h.l;
    const Lane1D b = node_1d(t);
    double D[3] = {0, 0, 0};
    for (int p2 = 0; p2 < 3; ++p2)
        for (int p1 = 0; p1 < 3; ++p1)
#pragma unroll
            for (int p0 = 0; p0 < 3; ++p0) {
                const double g0 = b.d0[p0] * b.L1[p1] * b.L2[p2] * g.jac[0];
                const double g1 = b.L0[p0] * b.d1[p1] * b.L2[p2] * g.jac[1];
                const double g2 = b.L0[p0] * b.L1[p1] * b.d2[p2] * g.jac[2];
                const double dOm = g.det * ((gauss_w(p0) * gauss_w(p1)) * gauss_w(p2));
                const double n2 = g0 * g0 + g1 * g1 + g2 * g2;
                D[0] += eta * (n2 + g0 * g0 / 3) * dOm;
                D[1] += eta * (n2 + g1 * g1 / 3) * dOm;
                D[2] += eta * (n2 + g2 * g2 / 3) * dOm;
            }
    if (t != HCENTRE_T) {
        double* sl = A.slot + ((e * HNSH + kHexA[t]) * 3);
        sl[0] = D[0]; sl[1] = D[1]; sl[2] = D[2];
    } else {
        const int64_t node = hex_node(g, e, t);
        A.Dx[node] = D[0]; A.Dy[node] = D[1]; A.Dz[node] = D[2];
    }
    if (t < HNP) {
        double m = 0;
        for (int q = 0; q < HNQ; ++q) {
            double NP[HNP], dOm;
            point_modes_cp(q, NP, dOm, g.det);
            m += NP[t] * NP[t] * dOm;
        }
        A.M[e * HNP + t] = m;
    }
}

// stokes_pressure_residual_3d_kernel! (pressure_residual.jl:337-366) and the
// Arrow-Hurwicz step of DR.jl:1025-1029: RP = −Σ_q modes (∇·v) dΩ per cell
// (a lane-group sum over the points), P += ps RP / M, and the tile's Σ P_0 M_0
__global__ __launch_bounds__(HBLOCK) void k_cp3_pressure(CP3Args A) {
    __shared__ double s_v[HEPB][HNV * 3];
    __shared__ double s_part[HEPB];
    const HexLane h = hex_lane(A.g.nels);
    const int64_t e = h.valid ? h.e : A.g.nels - 1;
    const HexGeo g = hex_at(A.g, e);
    if (h.l < HNV) {
        const int64_t node = hex_node(g, e, h.l);
        s_v[h.es][h.l * 3 + 0] = A.vx[node];
        s_v[h.es][h.l * 3 + 1] = A.vy[node];
        s_v[h.es][h.l * 3 + 2] = A.vz[node];
    }
    __syncthreads();
    const int q = h.l < HNQ ? h.l : 0;
    const Lane1D b = point_1d(q);
    double gv[3][3];
    point_gradient(g, b, q, s_v[h.es], gv);
    double NP[HNP], dOm;
    point_modes_cp(q, NP, dOm, g.det);
    const double div = gv[0][0] + gv[1][1] + gv[2][2];
    double RP[HNP];
#pragma unroll
    for (int i = 0; i < HNP; ++i) RP[i] = group_sum(h.l < HNQ ? -(NP[i] * (div * dOm)) : 0.0);
    if (h.l == 0) s_part[h.es] = 0;
    if (h.valid && h.l < HNP) {
        const int64_t k = e * HNP + h.l;
        A.RP[k] = RP[h.l];
        if (A.update) {
            const double Pn = A.P[k] + A.ps * RP[h.l] / A.M[k];
            A.P[k] = Pn;
            if (h.l == 0) s_part[h.es] = Pn * A.M[k];
        }
    }
    __syncthreads();
    if (A.update && threadIdx.x == 0) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < HEPB; ++i) acc += s_part[i];
        A.part[xcd_tile()] = acc;
    }
}

// stokes_momentum_residual_3d_kernel! (momentum_residuals.jl:645-691): the
// viscous residual ∫ ∇N·(2η ε̇′ − P I) − N ρ0 g with P from the modes (the
// constant mode's mean removed, and written back, when psum is set); the
// shared nodes' residuals go to the element slots, the centre node takes its
// Jacobi step v -= vs R / D here (DR.jl:1031-1038)
__global__ __launch_bounds__(HBLOCK) void k_cp3_momentum(CP3Args A) {
    __shared__ double s_v[HEPB][HNV * 3];
    __shared__ double s_f[HEPB][HNQ * 8];
    const HexLane h = hex_lane(A.g.nels);
    const int64_t e = h.valid ? h.e : A.g.nels - 1;
    const HexGeo g = hex_at(A.g, e);
    int64_t node = 0;
    if (h.l < HNV) {
        node = hex_node(g, e, h.l);
        s_v[h.es][h.l * 3 + 0] = A.vx[node];
        s_v[h.es][h.l * 3 + 1] = A.vy[node];
        s_v[h.es][h.l * 3 + 2] = A.vz[node];
    }
    double P[HNP];
#pragma unroll
    for (int i = 0; i < HNP; ++i) P[i] = A.P[e * HNP + i];
    if (A.psum) P[0] -= A.psum[0] * A.inv_sumM;
    __syncthreads();
    if (A.psum && h.valid && h.l == 0) A.P[e * HNP] = P[0];
    const int pc = A.phase[e];
    const double eta = A.ptab[PC_ETA * MAXPH + pc], rho0 = A.ptab[PC_RHO0 * MAXPH + pc];
    {
        const int q = h.l < HNQ ? h.l : 0;
        const Lane1D b = point_1d(q);
        double gv[3][3];
        point_gradient(g, b, q, s_v[h.es], gv);
        double NP[HNP], dOm;
        point_modes_cp(q, NP, dOm, g.det);
        const double tr = (gv[0][0] + gv[1][1] + gv[2][2]) / 3.0;
        double Pq = 0;
#pragma unroll
        for (int i = 0; i < HNP; ++i) Pq += NP[i] * P[i];
        const double two = 2 * eta;
        if (h.l < HNQ) {
            double* f = s_f[h.es] + q * 8;
            f[0] = (two * (gv[0][0] - tr) - Pq) * dOm;
            f[1] = (two * (gv[1][1] - tr) - Pq) * dOm;
            f[2] = (two * (gv[2][2] - tr) - Pq) * dOm;
            f[3] = eta * (gv[0][1] + gv[1][0]) * dOm;
            f[4] = eta * (gv[0][2] + gv[2][0]) * dOm;
            f[5] = eta * (gv[1][2] + gv[2][1]) * dOm;
            f[6] = rho0 * dOm;
        }
    }
    __syncthreads();
    if (!h.active) return;
    const int t = h.l;
    const Lane1D b = node_1d(t);
    double R[3];
    node_residual(g, b, t, s_f[h.es], A.gx, A.gy, A.gz, R);
    if (t != HCENTRE_T) {
        double* sl = A.slot + ((e * HNSH + kHexA[t]) * 3);
        sl[0] = R[0]; sl[1] = R[1]; sl[2] = R[2];
        return;
    }
    const uint8_t fl = A.bcflag[node];
    if (A.bf_x) { R[0] -= A.bf_x[node]; R[1] -= A.bf_y[node]; R[2] -= A.bf_z[node]; }
    const double rx = (fl & 1) ? 0.0 : R[0], ry = (fl & 2) ? 0.0 : R[1], rz = (fl & 4) ? 0.0 : R[2];
    if (A.Rv_x) { A.Rv_x[node] = rx; A.Rv_y[node] = ry; A.Rv_z[node] = rz; }
    if (A.update) {
        double nvx = s_v[h.es][t * 3 + 0] - A.vs * rx / A.Dx[node];
        double nvy = s_v[h.es][t * 3 + 1] - A.vs * ry / A.Dy[node];
        double nvz = s_v[h.es][t * 3 + 2] - A.vs * rz / A.Dz[node];
        if (fl & 1) nvx = A.bcval_x[node];
        if (fl & 2) nvy = A.bcval_y[node];
        if (fl & 4) nvz = A.bcval_z[node];
        A.vx[node] = nvx; A.vy[node] = nvy; A.vz[node] = nvz;
    }
}

// Σ RP², Σ Rx², Σ Ry², Σ Rz² (the norms of the check, DR.jl:1040-1042)
__global__ __launch_bounds__(RED_BLOCK) void k_cp3_norms(int64_t nv, int64_t np, const double* Rx, const double* Ry,
                                                         const double* Rz, const double* RP, double* partial) {
    double acc[4] = {0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < np; i += stride) acc[0] += RP[i] * RP[i];
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < nv; i += stride) {
        acc[1] += Rx[i] * Rx[i];
        acc[2] += Ry[i] * Ry[i];
        acc[3] += Rz[i] * Rz[i];
    }
    block_sum_store<4>(acc, partial);
}

__global__ void k_pin3(int64_t n, const uint8_t* flag, const double* bx, const double* by, const double* bz,
                       double* vx, double* vy, double* vz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        if (flag[i] & 1) vx[i] = bx[i];
        if (flag[i] & 2) vy[i] = by[i];
        if (flag[i] & 4) vz[i] = bz[i];
    }
}

namespace {

CP3Args cp3_args(fem_stokes* s) {
    CP3Args a;
    std::memset(&a, 0, sizeof a);
    double** f = s->f;
    a.g = hex_geo(s->mesh);
    a.vx = f[FEM_FIELD_V_X]; a.vy = f[FEM_FIELD_V_Y]; a.vz = f[FEM_FIELD_V_Z];
    a.P = f[FEM_FIELD_P];
    a.RP = f[FEM_FIELD_RP];
    a.M = f[FEM_FIELD_M_P];
    a.phase = s->d_phase_v;
    a.ptab = s->d_ptab;
    a.gx = s->mat.g[0]; a.gy = s->mat.g[1]; a.gz = s->mat.g[2];
    a.slot = s->d_slot3;
    a.Dx = f[FEM_FIELD_PC_V_X]; a.Dy = f[FEM_FIELD_PC_V_Y]; a.Dz = f[FEM_FIELD_PC_V_Z];
    a.bcflag = s->d_bcflag;
    a.bcval_x = s->d_bcval_x; a.bcval_y = s->d_bcval_y; a.bcval_z = s->d_bcval_z;
    return a;
}

// the state must be a purely viscous, incompressible 3-D one
// (CellPressureStokesDR, stokes_types.jl:581-613); the phase table carries η, ρ0
int cp3_prepare(fem_stokes* s) {
    fem_ctx* c = s->ctx;
    if (s->dim != 3 || s->mesh->nloc != 27)
        return set_err(c, FEM_EARG, "CellPressureStokesDR requires a Hex27 mesh");
    for (int p = 0; p < s->mat.nphases; ++p)
        if (!(std::isinf(s->mat.G[p]) && std::isinf(s->mat.K[p]) && s->mat.alpha[p] == 0))
            return set_err(c, FEM_EARG,
                           "CellPressureStokesDR is purely viscous and incompressible; G and K must be Inf and α zero");
    bool bulk_inf, thermal, eos_p, elastic;
    return upload_ptab(s, nullptr, 1.0, 0, &bulk_inf, &thermal, &eos_p, &elastic);
}

// the shared nodes' sums of the element slots: into the diagonal
// (preconditioner) or, with the Jacobi step, the velocity
int cp3_gather(fem_stokes* s, double* const out[3], bool step, double vs, bool store_R,
               double* const* v = nullptr, const double* const* load = nullptr) {
    fem_mesh* m = s->mesh;
    double** f = s->f;
    Gather3Args g;
    std::memset(&g, 0, sizeof g);
    g.ns = m->ns;
    g.node_of = m->d_node_of; g.off = m->d_off; g.ent = m->d_ent;
    g.slot = s->d_slot3;
    g.bcflag = s->d_bcflag;
    if (step) {
        g.vx = v ? v[0] : f[FEM_FIELD_V_X]; g.vy = v ? v[1] : f[FEM_FIELD_V_Y]; g.vz = v ? v[2] : f[FEM_FIELD_V_Z];
        if (load) { g.bf_x = load[0]; g.bf_y = load[1]; g.bf_z = load[2]; }
        g.rate_x = f[FEM_FIELD_DVDTAU_X]; g.rate_y = f[FEM_FIELD_DVDTAU_Y]; g.rate_z = f[FEM_FIELD_DVDTAU_Z];
        g.PCx = f[FEM_FIELD_PC_V_X]; g.PCy = f[FEM_FIELD_PC_V_Y]; g.PCz = f[FEM_FIELD_PC_V_Z];
        g.bcval_x = s->d_bcval_x; g.bcval_y = s->d_bcval_y; g.bcval_z = s->d_bcval_z;
        g.ax = g.ay = g.az = vs;
        g.bx = g.by = g.bz = 0;
        g.do_update = 1;
    }
    if (store_R) { g.Rv_x = out[0]; g.Rv_y = out[1]; g.Rv_z = out[2]; }
    hipLaunchKernelGGL(k_gather3, dim3(grid_xcd(m->ns, 256)), dim3(256), 0, s->ctx->stream, g);
    return check_launch(s->ctx, "k_gather3");
}

int cp3_preconditioner(fem_stokes* s) {
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, c->stream));
    CP3Args a = cp3_args(s);
    hipLaunchKernelGGL(k_cp3_precond, dim3(grid_xcd(s->mesh->nels, HEPB)), dim3(HBLOCK), 0, c->stream, a);
    CHK(check_launch(c, "k_cp3_precond"));
    double* D[3] = {s->f[FEM_FIELD_PC_V_X], s->f[FEM_FIELD_PC_V_Y], s->f[FEM_FIELD_PC_V_Z]};
    return cp3_gather(s, D, false, 0, true);
}

}  // namespace

extern "C" {

int fem_cell_stokes_preconditioner(fem_stokes* s) {
    if (!s) return FEM_EARG;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    CHK(cp3_prepare(s));
    CHK(cp3_preconditioner(s));
    return fem_sync(s->ctx);
}

int fem_cell_stokes_pressure_residual(fem_stokes* s) {
    if (!s) return FEM_EARG;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    CHK(cp3_prepare(s));
    CP3Args a = cp3_args(s);
    hipLaunchKernelGGL(k_cp3_pressure, dim3(grid_xcd(s->mesh->nels, HEPB)), dim3(HBLOCK), 0, s->ctx->stream, a);
    CHK(check_launch(s->ctx, "k_cp3_pressure"));
    return fem_sync(s->ctx);
}

int fem_cell_stokes_momentum_residual(fem_stokes* s) {
    if (!s) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    CHK(cp3_prepare(s));
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, c->stream));
    CP3Args a = cp3_args(s);
    double* R[3] = {s->f[FEM_FIELD_RV_X], s->f[FEM_FIELD_RV_Y], s->f[FEM_FIELD_RV_Z]};
    a.Rv_x = R[0]; a.Rv_y = R[1]; a.Rv_z = R[2];
    hipLaunchKernelGGL(k_cp3_momentum, dim3(grid_xcd(s->mesh->nels, HEPB)), dim3(HBLOCK), 0, c->stream, a);
    CHK(check_launch(c, "k_cp3_momentum"));
    CHK(cp3_gather(s, R, false, 0, true));
    return fem_sync(c);
}

}  // extern "C"

// Σ_i a_x b_x + a_y b_y + a_z b_z over the nodes (deterministic block partials)
__global__ __launch_bounds__(RED_BLOCK) void k_dot3(int64_t n, const double* ax, const double* ay, const double* az,
                                                    const double* bx, const double* by, const double* bz,
                                                    double* partial) {
    double acc[1] = {0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += stride)
        acc[0] += ax[i] * bx[i] + ay[i] * by[i] + az[i] * bz[i];
    block_sum_store<1>(acc, partial);
}

// _relax_cell_pressure_stokes! (DR.jl:997-1057) on the arrays v (3 nodal) and
// P ([nels][4]) with the state's viscosity, phases and scratch: gravity on or
// off (the adjoint runs with zero density), an optional momentum load
// subtracted before the Dirichlet rows, the Dirichlet values of bcs
static int cp3_relax(fem_stokes* s, double* const v[3], double* P, bool gravity, const double* const load[3],
                     const fem_bc bcs[3], const fem_cell_solve_params* p, fem_cell_solve_stats* st, double* history,
                     int64_t history_cap, int64_t* history_len, const char* what) {
    fem_ctx* c = s->ctx;
    std::memset(st, 0, sizeof *st);
    if (history_len) *history_len = 0;
    HIPCHK(c, hipSetDevice(c->device));
    if (p->check_interval < 1) return set_err(c, FEM_EARG, "check_interval must be positive");
    const auto t0 = std::chrono::steady_clock::now();
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, ne = m->nels, np = HNP * ne;
    CHK(cp3_prepare(s));
    CHK(cp3_preconditioner(s));
    CHK(upload_bcs(s, bcs));
    double** f = s->f;
    hipLaunchKernelGGL(k_pin3, dim3(grid_for(nv, 256)), dim3(256), 0, c->stream, nv, (const uint8_t*)s->d_bcflag,
                       (const double*)s->d_bcval_x, (const double*)s->d_bcval_y, (const double*)s->d_bcval_z,
                       v[0], v[1], v[2]);
    CHK(check_launch(c, "k_pin3"));
    // Σ_cells M_0 in cell order on the host (a constant of the solve)
    std::vector<double> hM(np);
    HIPCHK(c, hipMemcpyAsync(hM.data(), f[FEM_FIELD_M_P], np * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double sumM = 0;
    for (int64_t e = 0; e < ne; ++e) sumM += hM[e * HNP];
    const unsigned nb = grid_xcd(ne, HEPB);
    double* part = nullptr;
    CHK(dalloc_t(c, &part, nb + 1));
    CP3Args a = cp3_args(s);
    a.vx = v[0]; a.vy = v[1]; a.vz = v[2];
    a.P = P;
    if (!gravity) a.gx = a.gy = a.gz = 0;
    if (load) { a.bf_x = load[0]; a.bf_y = load[1]; a.bf_z = load[2]; }
    a.ps = p->pressure_step;
    a.vs = p->velocity_step;
    a.part = part;
    a.inv_sumM = 1.0 / sumM;
    double* R[3] = {f[FEM_FIELD_RV_X], f[FEM_FIELD_RV_Y], f[FEM_FIELD_RV_Z]};
    double err_v = INFINITY, err_P = INFINITY, err = INFINITY;
    int64_t iter = 0, nh = 0;
    int rc = FEM_OK;
    for (int64_t it = 1; it <= p->max_iterations && rc == FEM_OK; ++it) {
        iter = it;
        const bool chk = it % p->check_interval == 0 || it == p->max_iterations;
        a.update = 1;
        a.psum = nullptr;
        a.Rv_x = a.Rv_y = a.Rv_z = nullptr;
        hipLaunchKernelGGL(k_cp3_pressure, dim3(nb), dim3(HBLOCK), 0, c->stream, a);
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)nb, 1, 0, (const double*)part,
                           part + nb);
        a.psum = part + nb;
        if (chk) { a.Rv_x = R[0]; a.Rv_y = R[1]; a.Rv_z = R[2]; }
        hipLaunchKernelGGL(k_cp3_momentum, dim3(nb), dim3(HBLOCK), 0, c->stream, a);
        rc = check_launch(c, "k_cp3_momentum");
        if (rc == FEM_OK) rc = cp3_gather(s, R, true, p->velocity_step, chk, v, load);
        if (rc != FEM_OK || !chk) continue;
        const unsigned g = std::min<unsigned>(RED_GRID, grid_for(std::max(nv, np), RED_BLOCK));
        hipLaunchKernelGGL(k_cp3_norms, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, np, (const double*)R[0],
                           (const double*)R[1], (const double*)R[2], (const double*)f[FEM_FIELD_RP], c->red_partial);
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 4, 0,
                           (const double*)c->red_partial, c->red_out);
        rc = check_launch(c, "k_cp3_norms");
        if (rc != FEM_OK) break;
        HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 4 * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        err_P = std::sqrt(c->h_red[0]);
        err_v = std::max(std::sqrt(c->h_red[1]), std::max(std::sqrt(c->h_red[2]), std::sqrt(c->h_red[3])));
        err = std::max(err_v, err_P);
        if (!std::isfinite(err)) { rc = set_err(c, FEM_ENAN, "non-finite residual in %s", what); break; }
        if (history && nh < history_cap) {
            history[nh * 3] = (double)it; history[nh * 3 + 1] = err_v; history[nh * 3 + 2] = err_P;
            ++nh;
        }
        if (p->verbose) printf("iter = %06lld err = %.3e - norm[Rv=%.3e, Rp=%.3e]\n", (long long)it, err, err_v, err_P);
        if (err < p->tolerance) break;
    }
    hipStreamSynchronize(c->stream);
    hipFree(part);
    if (history_len) *history_len = nh;
    st->converged = err < p->tolerance;
    st->iterations = iter;
    st->residual = err;
    st->err_v = err_v;
    st->err_P = err_P;
    st->reached_total_iter = iter >= p->max_iterations && !st->converged;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != FEM_OK) return rc;
    if (!st->converged && p->throw_on_failure)
        return set_err(c, FEM_ENOCONV, "%s did not converge after %lld iterations (residual = %g, "
                       "tolerance = %g)", what, (long long)iter, err, p->tolerance);
    return FEM_OK;
}

extern "C" {

// solve!(dr::CellPressureStokesDR, mesh, bc_v; ...) (DR.jl:980-996)
int fem_cell_stokes_solve(fem_stokes* s, const fem_bc bcs[3], const fem_cell_solve_params* p,
                          fem_cell_solve_stats* st, double* history, int64_t history_cap, int64_t* history_len) {
    if (!s || !bcs || !p || !st) return FEM_EARG;
    double* v[3] = {s->f[FEM_FIELD_V_X], s->f[FEM_FIELD_V_Y], s->f[FEM_FIELD_V_Z]};
    return cp3_relax(s, v, s->f[FEM_FIELD_P], true, nullptr, bcs, p, st, history, history_cap, history_len,
                     "3-D Stokes solve");
}

// stokes_material_gradient_3d (DR_adjoint.jl:483-531): for every phase p,
// −λ·R(v = 0, P = 0, ρ = e_p, g) and −λ·R(v_fwd, P = 0, η = e_p, ρ = 0), R the
// assembled momentum residual without Dirichlet rows; v_fwd NULL: the state's v
int fem_cell_stokes_material_gradient(fem_stokes* s, const double* const v_fwd[3], const double* const lam_v[3],
                                      double* density_gradient, double* viscosity_gradient) {
    if (!s || !lam_v || !density_gradient || !viscosity_gradient) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    CHK(cp3_prepare(s));
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = HNP * m->nels;
    double** f = s->f;
    const double* vf[3] = {v_fwd ? v_fwd[0] : f[FEM_FIELD_V_X], v_fwd ? v_fwd[1] : f[FEM_FIELD_V_Y],
                           v_fwd ? v_fwd[2] : f[FEM_FIELD_V_Z]};
    double *zv = nullptr, *zP = nullptr;
    CHK(dalloc_t(c, &zv, nv));
    int rc = dalloc_t(c, &zP, np);
    if (rc == FEM_OK && (hipMemsetAsync(zv, 0, nv * 8, c->stream) != hipSuccess ||
                         hipMemsetAsync(zP, 0, np * 8, c->stream) != hipSuccess ||
                         hipMemsetAsync(s->d_bcflag, 0, nv, c->stream) != hipSuccess))
        rc = set_err(c, FEM_EHIP, "clearing the gradient work arrays failed");
    const fem_material saved = s->mat;
    double* R[3] = {f[FEM_FIELD_RV_X], f[FEM_FIELD_RV_Y], f[FEM_FIELD_RV_Z]};
    const int nph = saved.nphases;
    for (int which = 0; which < 2 && rc == FEM_OK; ++which)
        for (int p = 0; p < nph && rc == FEM_OK; ++p) {
            s->mat = saved;
            for (int q = 0; q < nph; ++q) {
                if (which == 0) s->mat.rho0[q] = q == p ? 1.0 : 0.0;
                else { s->mat.eta[q] = q == p ? 1.0 : 0.0; s->mat.rho0[q] = 0.0; }
            }
            rc = cp3_prepare(s);
            if (rc != FEM_OK) break;
            CP3Args a = cp3_args(s);
            a.P = zP;
            if (which == 0) { a.vx = zv; a.vy = zv; a.vz = zv; }
            else {
                a.vx = const_cast<double*>(vf[0]); a.vy = const_cast<double*>(vf[1]); a.vz = const_cast<double*>(vf[2]);
                a.gx = a.gy = a.gz = 0;
            }
            a.Rv_x = R[0]; a.Rv_y = R[1]; a.Rv_z = R[2];
            hipLaunchKernelGGL(k_cp3_momentum, dim3(grid_xcd(m->nels, HEPB)), dim3(HBLOCK), 0, c->stream, a);
            rc = check_launch(c, "k_cp3_momentum");
            if (rc == FEM_OK) rc = cp3_gather(s, R, false, 0, true);
            if (rc != FEM_OK) break;
            const unsigned g = std::min<unsigned>(RED_GRID, grid_for(nv, RED_BLOCK));
            hipLaunchKernelGGL(k_dot3, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, lam_v[0], lam_v[1], lam_v[2],
                               (const double*)R[0], (const double*)R[1], (const double*)R[2], c->red_partial);
            hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 1, 0,
                               (const double*)c->red_partial, c->red_out);
            rc = check_launch(c, "k_dot3");
            if (rc != FEM_OK) break;
            if (hipMemcpyAsync(c->h_red, c->red_out, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = set_err(c, FEM_EHIP, "reading the gradient reduction failed");
                break;
            }
            (which == 0 ? density_gradient : viscosity_gradient)[p] = -c->h_red[0];
        }
    s->mat = saved;
    const int rc2 = cp3_prepare(s);
    hipStreamSynchronize(c->stream);
    hipFree(zv);
    hipFree(zP);
    return rc != FEM_OK ? rc : rc2;
}

// solve_adjoint!(dr::CellPressureStokesDR, mesh, bc_v; objective_v, λv, λP) (DR_adjoint.jl:446-480)
int fem_cell_stokes_solve_adjoint(fem_stokes* s, const fem_bc bcs[3], const double* const objective[3],
                                  double* const lam_v[3], double* lam_P, const fem_cell_solve_params* p,
                                  fem_cell_solve_stats* st, double* history, int64_t history_cap,
                                  int64_t* history_len) {
    if (!s || !bcs || !objective || !lam_v || !lam_P || !p || !st) return FEM_EARG;
    for (int k = 0; k < 3; ++k)
        if (!objective[k] || !lam_v[k]) return set_err(s->ctx, FEM_EARG, "objective_v and λv take three arrays");
    // homogeneous conditions on the same nodes
    std::vector<double> zeros[3];
    fem_bc zb[3];
    for (int k = 0; k < 3; ++k) {
        zeros[k].assign((size_t)std::max<int64_t>(bcs[k].n, 0), 0.0);
        zb[k] = bcs[k];
        zb[k].host_vals = zeros[k].data();
    }
    return cp3_relax(s, lam_v, lam_P, false, objective, zb, p, st, history, history_cap, history_len,
                     "3-D Stokes adjoint solve");
}


}  // extern "C"

int hex27_update_current_stress(fem_stokes* s, const fem_drucker_prager* dp, double dt) {
    Opts3 o;
    CHK(prepare3(s, dp, dt, 0, o, true));
    HIPCHK(s->ctx, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, s->ctx->stream));
    o.store_tau = true;
    o.store_Rv = true;
    return hex_element(s, o);
}

// ---- mixed precision in 3-D (precision = 1): the 2-D solve's scheme on the
// Hex27 operator. The outer iterations evaluate R(x) in FP64 (no Pnum) and
// apply the reference's errors, criterion and rel_drop rule to it; the inner
// iterations run the reference's pseudo-transient loop in FP32 on a
// correction δ of R(x) + L δ = 0 (k_hex27_lin), with the reference's inner
// criterion, Chebyshev parameters and λmin estimates. δv is folded into the
// FP64 state at every check (and, with plasticity, every kHexFold iterations
// between checks), the correction then restarting from the FP64 residual
// there; δP is folded after the pressure update that ends an outer iteration.
// With a Drucker-Prager rheology every FP64 evaluation refreshes the per-
// element near-yield flags and the count of yielding elements, which selects
// the next iterations' operator (the return-map correction where any element
// yields, the exact visco-elastic one otherwise), and the Jacobian diagnostics
// are re-assembled at every check unless frozen (DR.jl:726-745).
namespace {
constexpr int64_t kHexFold = 25;

int cast32(fem_ctx* c, const double* x, float* y, int64_t n) {
    hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, x, y, 1.0);
    return check_launch(c, "k_cast");
}
int add32(fem_ctx* c, double* x, const float* d, int64_t n) {
    hipLaunchKernelGGL(k_add_correction, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, x, d, 1.0);
    return check_launch(c, "k_add_correction");
}

int ensure_w3(fem_stokes* s, bool plastic) {
    fem_mesh* m = s->mesh;
    fem_ctx* c = s->ctx;
    const int64_t nv = m->nnodes, np = m->nnodesP, ne = m->nels;
    if (!s->w3) {
        s->w3 = new fem_stokes::W3();
        std::memset(s->w3, 0, sizeof(fem_stokes::W3));
        fem_stokes::W3& w = *s->w3;
        float** vl[] = {&w.vx,    &w.vy,    &w.vz,    &w.rate_x, &w.rate_y, &w.rate_z, &w.Rv_x,   &w.Rv_y,   &w.Rv_z,
                        &w.Rv0_x, &w.Rv0_y, &w.Rv0_z, &w.PC_x,   &w.PC_y,   &w.PC_z,   &w.rhs_vx, &w.rhs_vy, &w.rhs_vz};
        for (float** q : vl) CHK(dalloc_t(c, q, nv));
        float** pl[] = {&w.P, &w.RP, &w.Pnum, &w.gammaP, &w.MP, &w.pscale, &w.rhs_P};
        for (float** q : pl) CHK(dalloc_t(c, q, np));
        CHK(dalloc_t(c, &w.slot, ne * HNSH * 3));
    }
    if (plastic && !s->d_yflag) {
        CHK(dalloc_t(c, &s->d_yflag, ne));
        HIPCHK(c, hipMemsetAsync(s->d_yflag, 1, ne, c->stream));
    }
    return FEM_OK;
}

// one FP32 correction iteration: element kernel, then the shared nodes
int lin3_step(fem_stokes* s, const Opts3& o, bool delta, const double al[3], const double be[3], bool store_end,
              bool store_Rv0) {
    fem_mesh* m = s->mesh;
    fem_ctx* c = s->ctx;
    fem_stokes::W3& w = *s->w3;
    double** f = s->f;
    Lin3Args a;
    std::memset(&a, 0, sizeof a);
    a.g = hex_geo(m);
    a.vx = w.vx; a.vy = w.vy; a.vz = w.vz;
    a.P = w.P; a.pscale = w.pscale; a.rhs_P = w.rhs_P;
    a.rhs_vx = w.rhs_vx; a.rhs_vy = w.rhs_vy; a.rhs_vz = w.rhs_vz;
    a.phase = s->d_phase_v;
    a.ptab = s->d_ptab;
    a.gx = (float)s->mat.g[0]; a.gy = (float)s->mat.g[1]; a.gz = (float)s->mat.g[2];
    a.dt = (float)o.dt;
    a.slot = w.slot;
    a.rate_x = w.rate_x; a.rate_y = w.rate_y; a.rate_z = w.rate_z;
    a.PCx = w.PC_x; a.PCy = w.PC_y; a.PCz = w.PC_z;
    a.bcflag = s->d_bcflag;
    a.ax = (float)al[0]; a.ay = (float)al[1]; a.az = (float)al[2];
    a.bx = (float)be[0]; a.by = (float)be[1]; a.bz = (float)be[2];
    if (store_end) {
        a.RP_out = w.RP; a.Pnum_out = w.Pnum;
        a.Rv_x = w.Rv_x; a.Rv_y = w.Rv_y; a.Rv_z = w.Rv_z;
    }
    if (store_Rv0) { a.Rv0_x = w.Rv0_x; a.Rv0_y = w.Rv0_y; a.Rv0_z = w.Rv0_z; }
    a.xvx = f[FEM_FIELD_V_X]; a.xvy = f[FEM_FIELD_V_Y]; a.xvz = f[FEM_FIELD_V_Z];
    a.xP = f[FEM_FIELD_P];
    const int tf[6] = {FEM_FIELD_TAU_OLD_XX, FEM_FIELD_TAU_OLD_YY, FEM_FIELD_TAU_OLD_ZZ,
                       FEM_FIELD_TAU_OLD_XY, FEM_FIELD_TAU_OLD_XZ, FEM_FIELD_TAU_OLD_YZ};
    for (int k = 0; k < 6; ++k) a.tau_old[k] = f[tf[k]];
    a.hot = s->d_yflag;
    const dim3 grid(grid_xcd(m->nels, HEPB)), block(HBLOCK);
    const int key = (o.bulk ? 1 : 0) | (delta ? 2 : 0) | (delta && o.elastic ? 4 : 0) | (m->structured ? 0 : 64);
    switch (key) {
        case 0: hipLaunchKernelGGL((k_hex27_lin<0>), grid, block, 0, c->stream, a); break;
        case 1: hipLaunchKernelGGL((k_hex27_lin<1>), grid, block, 0, c->stream, a); break;
        case 2: hipLaunchKernelGGL((k_hex27_lin<2>), grid, block, 0, c->stream, a); break;
        case 3: hipLaunchKernelGGL((k_hex27_lin<3>), grid, block, 0, c->stream, a); break;
        case 6: hipLaunchKernelGGL((k_hex27_lin<6>), grid, block, 0, c->stream, a); break;
        case 7: hipLaunchKernelGGL((k_hex27_lin<7>), grid, block, 0, c->stream, a); break;
        case 64: hipLaunchKernelGGL((k_hex27_lin<64>), grid, block, 0, c->stream, a); break;
        case 65: hipLaunchKernelGGL((k_hex27_lin<65>), grid, block, 0, c->stream, a); break;
        case 66: hipLaunchKernelGGL((k_hex27_lin<66>), grid, block, 0, c->stream, a); break;
        case 67: hipLaunchKernelGGL((k_hex27_lin<67>), grid, block, 0, c->stream, a); break;
        case 70: hipLaunchKernelGGL((k_hex27_lin<70>), grid, block, 0, c->stream, a); break;
        default: hipLaunchKernelGGL((k_hex27_lin<71>), grid, block, 0, c->stream, a); break;
    }
    CHK(check3(c, "k_hex27_lin"));
    Gather3fArgs g;
    std::memset(&g, 0, sizeof g);
    g.ns = m->ns;
    g.node_of = m->d_node_of; g.off = m->d_off; g.ent = m->d_ent;
    g.slot = w.slot;
    g.vx = w.vx; g.vy = w.vy; g.vz = w.vz;
    g.rate_x = w.rate_x; g.rate_y = w.rate_y; g.rate_z = w.rate_z;
    g.PCx = w.PC_x; g.PCy = w.PC_y; g.PCz = w.PC_z;
    g.bcflag = s->d_bcflag;
    g.rhs_vx = w.rhs_vx; g.rhs_vy = w.rhs_vy; g.rhs_vz = w.rhs_vz;
    g.Rv_x = a.Rv_x; g.Rv_y = a.Rv_y; g.Rv_z = a.Rv_z;
    g.Rv0_x = a.Rv0_x; g.Rv0_y = a.Rv0_y; g.Rv0_z = a.Rv0_z;
    g.ax = a.ax; g.ay = a.ay; g.az = a.az; g.bx = a.bx; g.by = a.by; g.bz = a.bz;
    hipLaunchKernelGGL(k_gather3f, dim3(grid_xcd(m->ns, 256)), dim3(256), 0, c->stream, g);
    return check3(c, "k_gather3f");
}

Red3T<float> red3f(fem_stokes* s, const double al[3]) {
    fem_stokes::W3& w = *s->w3;
    Red3T<float> r;
    std::memset(&r, 0, sizeof r);
    r.R[0] = w.Rv_x; r.R[1] = w.Rv_y; r.R[2] = w.Rv_z;
    r.R0[0] = w.Rv0_x; r.R0[1] = w.Rv0_y; r.R0[2] = w.Rv0_z;
    r.rate[0] = w.rate_x; r.rate[1] = w.rate_y; r.rate[2] = w.rate_z;
    r.PC[0] = w.PC_x; r.PC[1] = w.PC_y; r.PC[2] = w.PC_z;
    for (int k = 0; k < 3; ++k) r.al[k] = al[k];
    return r;
}

int hex27_mixed(fem_stokes* s, const fem_solve_params* p, const Opts3& o, bool freeze, fem_solve_stats* st) {
    fem_ctx* c = s->ctx;
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = m->nnodesP, ncheck = p->check_interval;
    const bool delta = o.plastic;
    CHK(ensure_w3(s, delta));
    fem_stokes::W3& w = *s->w3;
    double** f = s->f;
    // yielding elements of the last FP64 evaluation (device count, pinned copy)
    int32_t* d_yield = s->d_flag + 1;
    int32_t* h_yield = reinterpret_cast<int32_t*>(c->h_red + (RED_MAXK - 1));
    *h_yield = 1;
    CHK(cast32(c, f[FEM_FIELD_GAMMA_P], w.gammaP, np));
    CHK(cast32(c, f[FEM_FIELD_M_P], w.MP, np));
    CHK(cast32(c, s->d_pscale, w.pscale, np));
    hipLaunchKernelGGL(k_pin3, dim3(grid_for(nv, 256)), dim3(256), 0, c->stream, nv, (const uint8_t*)s->d_bcflag,
                       (const double*)s->d_bcval_x, (const double*)s->d_bcval_y, (const double*)s->d_bcval_z,
                       f[FEM_FIELD_V_X], f[FEM_FIELD_V_Y], f[FEM_FIELD_V_Z]);
    CHK(check_launch(c, "k_pin3"));
    auto cast_pc = [&]() -> int {
        CHK(cast32(c, f[FEM_FIELD_PC_V_X], w.PC_x, nv));
        CHK(cast32(c, f[FEM_FIELD_PC_V_Y], w.PC_y, nv));
        return cast32(c, f[FEM_FIELD_PC_V_Z], w.PC_z, nv);
    };
    CHK(hex_jacobian_impl(s, o));
    double lmax[3];
    CHK(checked_lmax3(s, lmax));
    double lmax_g = std::max(lmax[0], std::max(lmax[1], lmax[2]));
    int64_t jac = 1;
    bool jac_elastic = false;   // the last assembly was at a state without yielding
    CHK(cast_pc());
    for (float* z : {w.vx, w.vy, w.vz, w.rate_x, w.rate_y, w.rate_z, w.Rv0_x, w.Rv0_y, w.Rv0_z})
        HIPCHK(c, hipMemsetAsync(z, 0, nv * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(w.P, 0, np * 4, c->stream));
    if (p->verbose) printf("Initial momentum preconditioner λmax_v = (%g, %g, %g)\n", lmax[0], lmax[1], lmax[2]);
    double al[3], be[3];
    for (int k = 0; k < 3; ++k) cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, 0.0, s->opts.c_fact, &al[k], &be[k]);

    Opts3 oo = o;
    oo.store_RP = true;
    oo.store_Rv = true;
    oo.yield_flag = delta ? s->d_yflag : nullptr;
    oo.yield_count = delta ? d_yield : nullptr;
    // FP64 residual at x and the correction's right-hand sides; with
    // plasticity the fresh yield count picks the operator
    bool dlt = delta;
    auto residual64 = [&]() -> int {
        if (delta) HIPCHK(c, hipMemsetAsync(d_yield, 0, 8, c->stream));
        CHK(hex_element(s, oo));
        CHK(cast32(c, f[FEM_FIELD_RV_X], w.rhs_vx, nv));
        CHK(cast32(c, f[FEM_FIELD_RV_Y], w.rhs_vy, nv));
        CHK(cast32(c, f[FEM_FIELD_RV_Z], w.rhs_vz, nv));
        CHK(cast32(c, f[FEM_FIELD_RP], w.rhs_P, np));
        if (delta) {
            HIPCHK(c, hipMemcpyAsync(h_yield, d_yield, 4, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            dlt = *h_yield > 0;
        }
        return FEM_OK;
    };
    auto fold_v = [&]() -> int {
        CHK(add32(c, f[FEM_FIELD_V_X], w.vx, nv));
        CHK(add32(c, f[FEM_FIELD_V_Y], w.vy, nv));
        CHK(add32(c, f[FEM_FIELD_V_Z], w.vz, nv));
        for (float* z : {w.vx, w.vy, w.vz}) HIPCHK(c, hipMemsetAsync(z, 0, nv * 4, c->stream));
        return FEM_OK;
    };

    const double eps_tol = p->tolerance;
    double err_min = INFINITY, err = 2 * eps_tol, err_abs = INFINITY, err_rel = INFINITY;
    double err_v0 = 1.0, err_P0 = 1.0, err_v00 = 1.0, err_v = INFINITY, err_P = INFINITY;
    int64_t iter = 0, itPH_done = 0, launches = 0;
    bool converged = false;
    double rel_drop = p->rel_drop0;
    int rc = FEM_OK;
    const Red3 rbase = red3(s);
    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        CHK(residual64());
        double r4[4];
        CHK(reduce3(s, 4, 0, std::max(nv, np), [&](unsigned g) {
            hipLaunchKernelGGL(k_red3_outer, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, np, rbase,
                               (const double*)f[FEM_FIELD_RP], (const double*)f[FEM_FIELD_M_P], c->red_partial);
        }, r4));
        err_P = std::sqrt(r4[0]) / std::sqrt((double)np);
        err_v = std::max(std::sqrt(r4[1]), std::max(std::sqrt(r4[2]), std::sqrt(r4[3]))) / (2 * std::sqrt((double)nv));
        if (itPH == 1) { err_P0 = err_P + julia_eps(err_P); err_v0 = err_v + julia_eps(err_v); }
        else if (itPH == 2) { err_P0 = err_P + julia_eps(err_P); }
        err_abs = std::max(err_v, err_P);
        err_rel = std::max(err_v / err_v0, err_P / err_P0);
        err = std::min(err_abs, err_rel);
        if (std::isnan(err_v) || std::isnan(err_P) || std::isnan(err)) {
            rc = set_err(c, FEM_ENAN, "NaN detected in outer loop at PH=%lld", (long long)itPH);
            break;
        }
        if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in outer loop at PH=%lld", (long long)itPH); break; }
        if (p->verbose)
            printf("itPH = %02lld iter = %06lld err = %.3e abs = %.3e rel = %.3e\n", (long long)itPH, (long long)iter, err,
                   err_abs, err_rel);
        if (err < eps_tol) { converged = true; break; }
        if (err > err_min * 1.05) rel_drop = std::max(rel_drop * 0.1, 1.0e-3);
        err_min = std::min(err_min, err);
        const double eps_vel = err * rel_drop;
        int64_t itPT = 0;
        bool stop_total = false;
        while (err > eps_vel && itPT <= p->iter_max) {
            ++itPT;
            ++iter;
            const bool do_check = iter % ncheck == 0;
            const bool last_possible = do_check || itPT > p->iter_max || iter > p->max_iterations;
            if (do_check && ncheck == 1) {
                HIPCHK(c, hipMemcpyAsync(w.Rv0_x, w.Rv_x, nv * 4, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(w.Rv0_y, w.Rv_y, nv * 4, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(w.Rv0_z, w.Rv_z, nv * 4, hipMemcpyDeviceToDevice, c->stream));
            }
            CHK(lin3_step(s, o, dlt, al, be, last_possible, ncheck > 1 && (iter + 1) % ncheck == 0));
            ++launches;
            if (do_check) {
                const Red3T<float> r = red3f(s, al);
                double r9[9];
                CHK(reduce3(s, 9, 0, nv, [&](unsigned g) {
                    hipLaunchKernelGGL(k_red3_check<float>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, r,
                                       c->red_partial);
                }, r9));
                const double evi = std::max(std::sqrt(r9[0]), std::max(std::sqrt(r9[1]), std::sqrt(r9[2]))) /
                                   (2 * std::sqrt((double)nv));
                if (iter == ncheck) err_v00 = evi + julia_eps(evi);
                err = evi <= eps_tol ? evi : std::max(evi / err_v00, evi);
                if (std::isnan(err)) { rc = set_err(c, FEM_ENAN, "NaN detected in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (p->verbose > 1) printf("  it = %lld, iter = %lld, err = %.3e\n", (long long)itPT, (long long)iter, err);
                double lmin[3];
                for (int k = 0; k < 3; ++k) {
                    const double den = r9[3 + 2 * k], num = r9[4 + 2 * k];
                    lmin[k] = den == 0 ? 0.0 : std::fabs(num) / den;
                }
                CHK(fold_v());
                CHK(residual64());
                // without yielding the diagnostics are the visco-elastic ones:
                // re-assemble only when yielding was seen since the last assembly
                const bool skip_jac = !dlt && jac_elastic;
                if (delta && !freeze && !skip_jac) {
                    jac_elastic = !dlt;
                    ++jac;
                    CHK(hex_jacobian_impl(s, o));
                    rc = checked_lmax3(s, lmax);
                    if (rc != FEM_OK) break;
                    lmax_g = std::max(lmax[0], std::max(lmax[1], lmax[2]));
                    CHK(cast_pc());
                }
                for (int k = 0; k < 3; ++k)
                    cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, lmin[k], s->opts.c_fact, &al[k], &be[k]);
            } else if (delta && iter % kHexFold == 0) {
                CHK(fold_v());
                CHK(residual64());
            }
            if (iter > p->max_iterations) { stop_total = true; break; }
        }
        if (rc != FEM_OK) break;
        // P += γP RP / M_P on the correction, then fold δ into the FP64 state
        hipLaunchKernelGGL(k_pressure_update<float>, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np, w.P,
                           (const float*)w.gammaP, (const float*)w.RP, (const float*)w.MP);
        CHK(check_launch(c, "k_pressure_update"));
        CHK(fold_v());
        CHK(add32(c, f[FEM_FIELD_P], w.P, np));
        HIPCHK(c, hipMemsetAsync(w.P, 0, np * 4, c->stream));
        if (stop_total || iter > p->max_iterations) break;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    st->itPH = itPH_done;
    st->iter = iter;
    st->err = err;
    st->err_abs = err_abs;
    st->err_rel = err_rel;
    st->err_v = err_v;
    st->err_P = err_P;
    st->converged = converged;
    st->reached_total_iter = iter > p->max_iterations;
    st->lambda_max = std::max(lmax[0], std::max(lmax[1], lmax[2]));
    st->lambda_max_gershgorin = lmax_g;
    st->lambda_max_iterations = 0;
    st->jacobian_assemblies = jac;
    st->element_launches = launches;
    return rc;
}
}  // namespace

// solve! + _solve_stokes_dyrel! (DR.jl:302-337, 348-802) with D = 3
// _transfer_temperature_kernel! (DR.jl:44-47) on a Hex27 box: the four
// pressure DoFs of element e sit at its centre and its +x, +y and +z face
// centres (mixed_mesh.jl:399-417, local nodes 27, 23, 24, 26), tensor
// positions 13, 14, 16 and 22
__global__ void k_transfer_T3(HexGeo g, const double* T, double* TP) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * g.nels) return;
    const int64_t e = i >> 2;
    const int a = (int)(i & 3);
    const int t = a == 0 ? 13 : a == 1 ? 14 : a == 2 ? 16 : 22;
    TP[i] = T[hex_node(g, e, t)];
}

int hex27_transfer_T(fem_stokes* s, const double* T, double* TP) {
    const HexGeo g = hex_geo(s->mesh);
    hipLaunchKernelGGL(k_transfer_T3, dim3(grid_for(4 * g.nels, 256)), dim3(256), 0, s->ctx->stream, g, T, TP);
    return check3(s->ctx, "k_transfer_T3");
}

// the shear dissipation at the points: one residual evaluation (v, P without
// Pnum, τ_old) that stores Φ only
int hex27_shear_heating_launch(fem_stokes* s, const Opts3& o, double* phi) {
    Opts3 oo = o;
    oo.use_pnum = false;
    oo.do_update = false;
    oo.store_RP = oo.store_Pnum = oo.store_Rv = oo.store_Rv0 = oo.store_tau = false;
    oo.yield_flag = nullptr;
    oo.yield_count = nullptr;
    oo.phi_out = phi;
    return hex_element(s, oo, false);
}

int hex27_shear_heating(fem_stokes* s, const fem_drucker_prager* dp, double dt, double* phi) {
    Opts3 o;
    CHK(prepare3(s, dp, dt, 0, o, true));
    return hex27_shear_heating_launch(s, o, phi);
}

int hex27_solve(fem_stokes* s, const fem_bc* bcs, const fem_drucker_prager* dp, const fem_solve_params* p,
                fem_solve_stats* st, CoupledSide* cp) {
    fem_ctx* c = s->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = m->nnodesP, ncheck = p->check_interval;
    const double dt = p->dt;
    if (cp) {
        // the thermal field on the pressure DoFs before the Stokes set-up (DR.jl:123-124)
        CHK(hex27_transfer_T(s, thermal_field_ptr(cp->tc.t, FEM_TFIELD_T), s->f[FEM_FIELD_T]));
        CHK(hex27_transfer_T(s, thermal_field_ptr(cp->tc.t, FEM_TFIELD_T0), s->f[FEM_FIELD_T0]));
    }
    Opts3 o;
    CHK(prepare3(s, dp, dt, p->finite_K, o, p->use_tau_old != 0));
    const bool freeze = p->freeze_jacobian < 0 ? (dp == nullptr) : (p->freeze_jacobian != 0);
    CHK(upload_bcs(s, bcs));
    // the phase table uploaded by prepare3 (with the plastic parameters) is
    // the one the scaling kernel reads
    if (!p->skip_scaling) CHK(hex_scaling_launch(s, p->pressure_factor, dt));
    else CHK(hex_pscale(s));
    if (p->precision >= 1) {
        const int rcm = hex27_mixed(s, p, o, freeze, st);
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rcm != FEM_OK) return rcm;
        if (!st->converged && p->throw_on_failure)
            return set_err(c, FEM_ENOCONV, "Stokes solve did not converge after %lld iterations (residual = %g, tolerance = %g)",
                           (long long)st->iter, st->err_abs, p->tolerance);
        return FEM_OK;
    }
    double** f = s->f;
    for (int k : {FEM_FIELD_DVDTAU_X, FEM_FIELD_DVDTAU_Y, FEM_FIELD_DVDTAU_Z, FEM_FIELD_RV0_X, FEM_FIELD_RV0_Y,
                  FEM_FIELD_RV0_Z})
        HIPCHK(c, hipMemsetAsync(f[k], 0, nv * 8, c->stream));
    CHK(hex_jacobian_impl(s, o));
    double lmax[3];
    CHK(checked_lmax3(s, lmax));
    double lmax_g = std::max(lmax[0], std::max(lmax[1], lmax[2]));
    int64_t jac = 1;
    double al[3], be[3];
    for (int k = 0; k < 3; ++k) cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, 0.0, s->opts.c_fact, &al[k], &be[k]);
    if (p->verbose) printf("Initial momentum preconditioner λmax_v = (%g, %g, %g)\n", lmax[0], lmax[1], lmax[2]);

    if (cp) {
        // ∂T∂τ = R0 = 0 and the dissipation of the initial velocity (DR.jl:524-530)
        CHK(thermal_coupling_begin(cp->tc, cp->bc_T));
        if (cp->shear) CHK(hex27_shear_heating_launch(s, o, cp->d_phi));
    }
    auto thermal_ok = [&]() { return !cp || cp->tc.converged; };
    const double eps_tol = p->tolerance;
    double err_min = INFINITY, err = 2 * eps_tol, err_abs = INFINITY, err_rel = INFINITY;
    double err_v0 = 1.0, err_P0 = 1.0, err_v00 = 1.0, err_v = INFINITY, err_P = INFINITY;
    int64_t iter = 0, itPH_done = 0, launches = 0;
    bool converged = false;
    double rel_drop = p->rel_drop0;
    int rc = FEM_OK;
    const Red3 rbase = red3(s);
    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        Opts3 oo = o;
        oo.store_RP = true; oo.store_Rv = true;
        CHK(hex_element(s, oo));
        double r4[4];
        CHK(reduce3(s, 4, 0, std::max(nv, np), [&](unsigned g) {
            hipLaunchKernelGGL(k_red3_outer, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, np, rbase,
                               (const double*)f[FEM_FIELD_RP], (const double*)f[FEM_FIELD_M_P], c->red_partial);
        }, r4));
        err_P = std::sqrt(r4[0]) / std::sqrt((double)np);
        err_v = std::max(std::sqrt(r4[1]), std::max(std::sqrt(r4[2]), std::sqrt(r4[3]))) / (2 * std::sqrt((double)nv));
        if (itPH == 1) { err_P0 = err_P + julia_eps(err_P); err_v0 = err_v + julia_eps(err_v); }
        else if (itPH == 2) { err_P0 = err_P + julia_eps(err_P); }
        err_abs = std::max(err_v, err_P);
        err_rel = std::max(err_v / err_v0, err_P / err_P0);
        err = std::min(err_abs, err_rel);
        if (std::isnan(err_v) || std::isnan(err_P) || std::isnan(err)) {
            rc = set_err(c, FEM_ENAN, "NaN detected in outer loop at PH=%lld", (long long)itPH);
            break;
        }
        if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in outer loop at PH=%lld", (long long)itPH); break; }
        if (p->verbose)
            printf("itPH = %02lld iter = %06lld err = %.3e abs = %.3e rel = %.3e\n", (long long)itPH, (long long)iter, err,
                   err_abs, err_rel);
        if (err < eps_tol && thermal_ok()) { converged = true; break; }
        if (err > err_min * 1.05) rel_drop = std::max(rel_drop * 0.1, 1.0e-3);
        err_min = std::min(err_min, err);
        const double eps_vel = err * rel_drop;
        int64_t itPT = 0;
        bool stop_total = false;
        while ((err > eps_vel || !thermal_ok()) && itPT <= p->iter_max) {
            ++itPT;
            ++iter;
            const bool do_check = iter % ncheck == 0;
            const bool last_possible = do_check || itPT > p->iter_max || iter > p->max_iterations;
            if (do_check && ncheck == 1)
                for (int k = 0; k < 3; ++k)
                    HIPCHK(c, hipMemcpyAsync(f[FEM_FIELD_RV0_X + (k == 2 ? FEM_FIELD_RV0_Z - FEM_FIELD_RV0_X : k)],
                                             f[FEM_FIELD_RV_X + (k == 2 ? FEM_FIELD_RV_Z - FEM_FIELD_RV_X : k)], nv * 8,
                                             hipMemcpyDeviceToDevice, c->stream));
            Opts3 oi = o;
            oi.use_pnum = true;
            oi.do_update = true;
            oi.store_RP = last_possible;
            oi.store_Pnum = last_possible;
            oi.store_Rv = last_possible;
            oi.store_Rv0 = ncheck > 1 && ((iter + 1) % ncheck == 0);
            for (int k = 0; k < 3; ++k) { oi.al[k] = al[k]; oi.be[k] = be[k]; }
            if (cp) {
                // one thermal iteration ahead of the Stokes one (DR.jl:604-661)
                CHK(thermal_coupling_step(cp->tc));
                CHK(hex27_transfer_T(s, thermal_field_ptr(cp->tc.t, FEM_TFIELD_T), s->f[FEM_FIELD_T]));
                // refreshed after the thermal spectral estimate (DR.jl:659-661)
                if (cp->shear && cp->tc.checked) CHK(hex27_shear_heating_launch(s, o, cp->d_phi));
            }
            CHK(hex_element(s, oi));
            ++launches;
            if (do_check) {
                Red3 r = rbase;
                for (int k = 0; k < 3; ++k) r.al[k] = al[k];
                double r9[9];
                CHK(reduce3(s, 9, 0, nv, [&](unsigned g) {
                    hipLaunchKernelGGL(k_red3_check<double>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, r, c->red_partial);
                }, r9));
                const double evi = std::max(std::sqrt(r9[0]), std::max(std::sqrt(r9[1]), std::sqrt(r9[2]))) /
                                   (2 * std::sqrt((double)nv));
                if (iter == ncheck) err_v00 = evi + julia_eps(evi);
                err = evi <= eps_tol ? evi : std::max(evi / err_v00, evi);
                if (std::isnan(err)) { rc = set_err(c, FEM_ENAN, "NaN detected in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                double lmin[3];
                for (int k = 0; k < 3; ++k) {
                    const double den = r9[3 + 2 * k], num = r9[4 + 2 * k];
                    lmin[k] = den == 0 ? 0.0 : std::fabs(num) / den;
                }
                if (!freeze) {
                    ++jac;
                    CHK(hex_jacobian_impl(s, o));
                    rc = checked_lmax3(s, lmax);
                    if (rc != FEM_OK) break;
                    lmax_g = std::max(lmax[0], std::max(lmax[1], lmax[2]));
                }
                for (int k = 0; k < 3; ++k)
                    cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, lmin[k], s->opts.c_fact, &al[k], &be[k]);
            }
            if (iter > p->max_iterations) { stop_total = true; break; }
        }
        if (rc != FEM_OK) break;
        hipLaunchKernelGGL(k_pressure_update<double>, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np,
                           f[FEM_FIELD_P], (const double*)f[FEM_FIELD_GAMMA_P], (const double*)f[FEM_FIELD_RP],
                           (const double*)f[FEM_FIELD_M_P]);
        CHK(check_launch(c, "k_pressure_update"));
        if (stop_total || iter > p->max_iterations) break;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    st->itPH = itPH_done;
    st->iter = iter;
    st->err = err;
    st->err_abs = err_abs;
    st->err_rel = err_rel;
    st->err_v = err_v;
    st->err_P = err_P;
    st->converged = converged;
    st->reached_total_iter = iter > p->max_iterations;
    st->lambda_max = std::max(lmax[0], std::max(lmax[1], lmax[2]));
    st->lambda_max_gershgorin = lmax_g;
    st->lambda_max_iterations = 0;
    st->jacobian_assemblies = jac;
    st->element_launches = launches;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != FEM_OK) return rc;
    if (!converged && p->throw_on_failure)
        return set_err(c, FEM_ENOCONV, "Stokes solve did not converge after %lld iterations (residual = %g, tolerance = %g)",
                       (long long)iter, err_abs, p->tolerance);
    return FEM_OK;
}

// element kernel alone, n sweeps, for the bench's device-event timing
int hex27_time_element(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t n, double* avg_ms) {
    fem_ctx* c = s->ctx;
    Opts3 o;
    CHK(prepare3(s, dp, dt, 0, o, true));
    o.use_pnum = true;
    o.do_update = true;
    for (int k = 0; k < 3; ++k) { o.al[k] = 0.0; o.be[k] = 1.0; }
    CHK(hex_element(s, o, false));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < n; ++i) CHK(hex_element(s, o, false));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / n;
    return FEM_OK;
}

// the FP32 correction step alone, n sweeps, for the bench's device-event timing
int hex27_time_lin(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t n, double* avg_ms) {
    fem_ctx* c = s->ctx;
    Opts3 o;
    CHK(prepare3(s, dp, dt, 0, o, true));
    CHK(ensure_w3(s, o.plastic));
    fem_stokes::W3& w = *s->w3;
    const int64_t nv = s->mesh->nnodes, np = s->mesh->nnodesP, ne = s->mesh->nels;
    for (float* z : {w.vx, w.vy, w.vz, w.rate_x, w.rate_y, w.rate_z, w.rhs_vx, w.rhs_vy, w.rhs_vz})
        HIPCHK(c, hipMemsetAsync(z, 0, nv * 4, c->stream));
    for (float* z : {w.P, w.rhs_P, w.pscale}) HIPCHK(c, hipMemsetAsync(z, 0, np * 4, c->stream));
    for (float* z : {w.PC_x, w.PC_y, w.PC_z}) {
        hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(nv, 256)), dim3(256), 0, c->stream, nv, z, 1.0f);
        CHK(check_launch(c, "k_fill"));
    }
    if (o.plastic) HIPCHK(c, hipMemsetAsync(s->d_yflag, 1, ne, c->stream));
    const double al[3] = {0, 0, 0}, be[3] = {1, 1, 1};
    CHK(lin3_step(s, o, o.plastic, al, be, false, false));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < n; ++i) CHK(lin3_step(s, o, o.plastic, al, be, false, false));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / n;
    return FEM_OK;
}

// the 26 shared local nodes of every element: node -> element-slot lists in
// ascending element order (generate_node2element, connectivity.jl:341-349),
// over the nodes that some element uses and that are no cell centre (cell
// centres are private to their element and updated by the element kernel)
static int hex27_lists(fem_ctx* c, fem_mesh* m, const std::vector<uint8_t>& centre) {
    const int64_t nn = m->nnodes, ne = m->nels;
    std::vector<uint8_t> used(nn, 0);
    for (int64_t i = 0; i < ne * HNV; ++i) used[m->h_el2n[i]] = 1;
    std::vector<int64_t> sid(nn, -1);
    std::vector<int32_t> node_of;
    for (int64_t k = 0; k < nn; ++k) {
        if (centre[k] || !used[k]) continue;
        sid[k] = (int64_t)node_of.size();
        node_of.push_back((int32_t)k);
    }
    m->ns = (int64_t)node_of.size();
    std::vector<int32_t> off(m->ns + 1, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < HNSH; ++a) off[sid[m->h_el2n[e * HNV + a]] + 1]++;
    for (int64_t i = 0; i < m->ns; ++i) off[i + 1] += off[i];
    std::vector<int32_t> ent(off[m->ns]);
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < HNSH; ++a) ent[fill[sid[m->h_el2n[e * HNV + a]]]++] = (int32_t)(e * HNSH + a);
    CHK(dalloc_t(c, &m->d_off, off.size()));
    CHK(dalloc_t(c, &m->d_ent, ent.size()));
    CHK(dalloc_t(c, &m->d_node_of, node_of.size()));
    HIPCHK(c, hipMemcpy(m->d_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_ent, ent.data(), ent.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_node_of, node_of.data(), node_of.size() * 4, hipMemcpyHostToDevice));
    return FEM_OK;
}

extern "C" int fem_mesh_hex27_create(fem_ctx* c, int32_t nx, int32_t ny, int32_t nz, double x0, double x1, double y0,
                                     double y1, double z0, double z1, fem_mesh** out) {
    if (!c || !out) return FEM_EARG;
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1)
        return set_err(c, FEM_EARG, "element counts must be positive, got (%d, %d, %d)", nx, ny, nz);
    if (!(x1 > x0) || !(y1 > y0) || !(z1 > z0)) return set_err(c, FEM_EARG, "empty domain");
    const int64_t sx = 2 * (int64_t)nx + 1, sy = 2 * (int64_t)ny + 1, sz = 2 * (int64_t)nz + 1;
    const int64_t nn = sx * sy * sz, ne = (int64_t)nx * ny * nz;
    if (nn > std::numeric_limits<int32_t>::max() || ne * HNSH > std::numeric_limits<int32_t>::max())
        return set_err(c, FEM_EDIM, "mesh too large for 32-bit node and element-slot indices");
    HIPCHK(c, hipSetDevice(c->device));
    fem_mesh* m = new fem_mesh();
    m->ctx = c;
    m->dim = 3;
    m->nloc = HNV;
    m->structured = true;
    m->nx = nx; m->ny = ny; m->nz = nz;
    const double bx[6] = {x0, x1, y0, y1, z0, z1};
    std::memcpy(m->box, bx, sizeof bx);
    m->nnodes = nn;
    m->nels = ne;
    m->nnodesP = 4 * ne;
    // generate_coordinates (mesh.jl:340-398): lexicographic, x fastest
    m->h_coords.resize((size_t)nn * 3);
    for (int64_t iz = 0; iz < sz; ++iz)
        for (int64_t iy = 0; iy < sy; ++iy)
            for (int64_t ix = 0; ix < sx; ++ix) {
                const int64_t k = (iz * sy + iy) * sx + ix;
                m->h_coords[3 * k + 0] = x0 + (x1 - x0) * (double)ix / (double)(sx - 1);
                m->h_coords[3 * k + 1] = y0 + (y1 - y0) * (double)iy / (double)(sy - 1);
                m->h_coords[3 * k + 2] = z0 + (z1 - z0) * (double)iz / (double)(sz - 1);
                if (ix == 0 || iy == 0 || iz == 0 || ix == sx - 1 || iy == sy - 1 || iz == sz - 1)
                    m->h_boundary.push_back((int32_t)k);
            }
    // generate_element2node (connectivity.jl:273-322), elements x fastest
    m->h_el2n.resize((size_t)ne * HNV);
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / ((int64_t)nx * ny);
        for (int a = 0; a < HNV; ++a)
            m->h_el2n[e * HNV + a] =
                (int32_t)((2 * ez + kHexR[a][2]) * sx * sy + (2 * ey + kHexR[a][1]) * sx + (2 * ex + kHexR[a][0]));
    }
    std::vector<uint8_t> centre(nn, 0);
    for (int64_t e = 0; e < ne; ++e) centre[m->h_el2n[e * HNV + HNSH]] = 1;
    const int r = hex27_lists(c, m, centre);
    if (r != FEM_OK) {
        fem_mesh_destroy(m);
        return r;
    }
    *out = m;
    return FEM_OK;
}

// Mesh(backend, coords, el2n, Hex27) (mesh.jl:147-188) from host arrays:
// coords [nnodes][3], el2n [nels][27] 1-based in the local order of
// generate_element2node (connectivity.jl:273-322). The element kernels take
// one metric per element, so every element must be an axis-aligned box whose
// local ξ, η, ζ run along +x, +y, +z with its 27 nodes at the box's corners,
// edge, face and centre midpoints (to 1e-10 of its size); nodes and elements
// may come in any order and elements may differ in size (graded meshes).
// Γnodes are the nodes of the faces that belong to one element only.
extern "C" int fem_mesh_hex27_upload(fem_ctx* c, const double* host_coords, int64_t nnodes,
                                     const int32_t* host_el2n_1based, int64_t nels, fem_mesh** out) {
    if (!c || !out || !host_coords || !host_el2n_1based) return FEM_EARG;
    *out = nullptr;
    if (nnodes <= 0 || nels <= 0) return set_err(c, FEM_EARG, "empty mesh");
    if (nnodes > std::numeric_limits<int32_t>::max() || nels * HNSH > std::numeric_limits<int32_t>::max())
        return set_err(c, FEM_EDIM, "mesh too large for 32-bit node and element-slot indices");
    std::vector<int32_t> el((size_t)nels * HNV);
    std::vector<int32_t> val(nnodes, 0);
    for (int64_t i = 0; i < nels * HNV; ++i) {
        const int32_t n = host_el2n_1based[i] - 1;
        if (n < 0 || n >= nnodes)
            return set_err(c, FEM_EARG, "el2n entry %d out of range at element %lld", n + 1, (long long)(i / HNV) + 1);
        el[i] = n;
        val[n]++;
    }
    std::vector<double> metric((size_t)nels * 4);
    std::vector<uint8_t> centre(nnodes, 0);
    for (int64_t e = 0; e < nels; ++e) {
        const int32_t* en = &el[e * HNV];
        const double* lo = host_coords + 3 * (int64_t)en[kHexA[0]];
        const double* hi = host_coords + 3 * (int64_t)en[kHexA[HNV - 1]];
        double h[3];
        for (int d = 0; d < 3; ++d) h[d] = hi[d] - lo[d];
        if (!(h[0] > 0 && h[1] > 0 && h[2] > 0) || !std::isfinite(h[0] * h[1] * h[2]))
            return set_err(c, FEM_EARG, "Hex27 element %lld is inverted or degenerate: its local axes must run "
                           "along +x, +y, +z", (long long)e + 1);
        const double tol = 1e-10 * std::max(h[0], std::max(h[1], h[2]));
        for (int a = 0; a < HNV; ++a)
            for (int d = 0; d < 3; ++d) {
                const double want = lo[d] + 0.5 * kHexR[a][d] * h[d];
                if (!(std::fabs(host_coords[3 * (int64_t)en[a] + d] - want) <= tol))
                    return set_err(c, FEM_EARG, "Hex27 element %lld is not an axis-aligned box with its nodes at "
                                   "the corner, edge, face and centre points (local node %d); uploaded Hex27 "
                                   "meshes take box elements of any size", (long long)e + 1, a + 1);
            }
        const int32_t cn = en[HNSH];
        if (val[cn] != 1)
            return set_err(c, FEM_EARG, "Hex27 centre node %d of element %lld is shared", cn + 1, (long long)e + 1);
        centre[cn] = 1;
        metric[e * 4 + 0] = 2 / h[0];
        metric[e * 4 + 1] = 2 / h[1];
        metric[e * 4 + 2] = 2 / h[2];
        metric[e * 4 + 3] = (h[0] / 2) * (h[1] / 2) * (h[2] / 2);
    }
    // faces that one element owns: keyed by their sorted corner nodes
    std::map<std::array<int32_t, 4>, int> faces;
    auto face_key = [&](int64_t e, int d, int side) {
        std::array<int32_t, 4> k;
        int n = 0;
        for (int t = 0; t < HNV; ++t) {
            const int r[3] = {t % 3, (t / 3) % 3, t / 9};
            if (r[d] == side && r[(d + 1) % 3] != 1 && r[(d + 2) % 3] != 1) k[n++] = el[e * HNV + kHexA[t]];
        }
        std::sort(k.begin(), k.end());
        return k;
    };
    for (int64_t e = 0; e < nels; ++e)
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side <= 2; side += 2) faces[face_key(e, d, side)]++;
    std::vector<uint8_t> bnd(nnodes, 0);
    for (int64_t e = 0; e < nels; ++e)
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side <= 2; side += 2)
                if (faces[face_key(e, d, side)] == 1)
                    for (int t = 0; t < HNV; ++t) {
                        const int r[3] = {t % 3, (t / 3) % 3, t / 9};
                        if (r[d] == side) bnd[el[e * HNV + kHexA[t]]] = 1;
                    }
    HIPCHK(c, hipSetDevice(c->device));
    fem_mesh* m = new fem_mesh();
    m->ctx = c;
    m->dim = 3;
    m->nloc = HNV;
    m->structured = false;
    m->nnodes = nnodes;
    m->nels = nels;
    m->nnodesP = 4 * nels;
    m->h_coords.assign(host_coords, host_coords + 3 * nnodes);
    m->h_el2n = std::move(el);
    for (int64_t k = 0; k < nnodes; ++k)
        if (bnd[k]) m->h_boundary.push_back((int32_t)k);
    int r = hex27_lists(c, m, centre);
    if (r == FEM_OK) r = dalloc_t(c, &m->d_el2n, m->h_el2n.size());
    if (r == FEM_OK) r = dalloc_t(c, &m->d_geo, metric.size());
    if (r == FEM_OK) {
        hipError_t he = hipMemcpy(m->d_el2n, m->h_el2n.data(), m->h_el2n.size() * 4, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(m->d_geo, metric.data(), metric.size() * 8, hipMemcpyHostToDevice);
        if (he != hipSuccess) r = set_err(c, FEM_EHIP, "hipMemcpy of the Hex27 mesh failed: %s", hipGetErrorString(he));
    }
    if (r != FEM_OK) {
        fem_mesh_destroy(m);
        return r;
    }
    *out = m;
    return FEM_OK;
}
