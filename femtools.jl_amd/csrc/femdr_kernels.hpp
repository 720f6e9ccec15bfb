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

// Reference-element tables (shape_functions.jl:115-171, 231-239;
// integration_points.jl:188-202), evaluated at compile time so that, with the
// quadrature loops unrolled, every entry becomes an immediate operand.
struct Tables {
    double N[NQ][NV];
    double NP1[NQ][NP];
    double dN[NQ][NV][2];
    double w[NQ];
    double TB[NP][NV][2];   // Σ_q w_q NP1_q[j] dN_q[b][k]: the element divergence operator
    double MP[NP][NP];      // Σ_q w_q NP1_q[i] NP1_q[j]: the P1 mass matrix on the reference triangle
};

constexpr void t7_basis_ct(double xi, double eta, double* N, double (*dN)[2]) {
    const double L1 = 1 - xi - eta;
    const double b = 27.0 * (1 - xi - eta) * xi * eta;
    const double bx = 27.0 * eta * (1 - 2 * xi - eta);
    const double by = 27.0 * xi * (1 - xi - 2 * eta);
    N[0] = L1 * (2 * L1 - 1) + b / 9;
    N[1] = xi * (2 * xi - 1) + b / 9;
    N[2] = eta * (2 * eta - 1) + b / 9;
    N[3] = 4 * (1 - xi - eta) * xi - 4 * b / 9;
    N[4] = 4 * xi * eta - 4 * b / 9;
    N[5] = 4 * eta * (1 - xi - eta) - 4 * b / 9;
    N[6] = b;
    dN[0][0] = 1 - 4 * L1 + bx / 9;  dN[0][1] = 1 - 4 * L1 + by / 9;
    dN[1][0] = 4 * xi - 1 + bx / 9;  dN[1][1] = by / 9;
    dN[2][0] = bx / 9;               dN[2][1] = 4 * eta - 1 + by / 9;
    dN[3][0] = 4 * (1 - 2 * xi - eta) - 4 * bx / 9; dN[3][1] = -4 * xi - 4 * by / 9;
    dN[4][0] = 4 * eta - 4 * bx / 9; dN[4][1] = 4 * xi - 4 * by / 9;
    dN[5][0] = -4 * eta - 4 * bx / 9; dN[5][1] = 4 * (1 - xi - 2 * eta) - 4 * by / 9;
    dN[6][0] = bx;                   dN[6][1] = by;
}

constexpr Tables make_tables() {
    Tables t{};
    const double a1 = 0.101286507323456, a2 = 0.470142064105115;
    const double b1 = 1 - 2 * a1, b2 = 1 - 2 * a2;
    const double w1 = 0.062969590272414, w2 = 0.066197076394253, wc = 0.1125;
    const double X[NQ] = {a1, b1, a1, a2, b2, a2, 1.0 / 3};
    const double Y[NQ] = {a1, a1, b1, a2, a2, b2, 1.0 / 3};
    const double W[NQ] = {w1, w1, w1, w2, w2, w2, wc};
    for (int q = 0; q < NQ; ++q) {
        t7_basis_ct(X[q], Y[q], t.N[q], t.dN[q]);
        t.NP1[q][0] = 1 - X[q] - Y[q];
        t.NP1[q][1] = X[q];
        t.NP1[q][2] = Y[q];
        t.w[q] = W[q];
    }
    for (int j = 0; j < NP; ++j)
        for (int b = 0; b < NV; ++b)
            for (int k = 0; k < 2; ++k) {
                double s = 0;
                for (int q = 0; q < NQ; ++q) s += t.w[q] * t.NP1[q][j] * t.dN[q][b][k];
                t.TB[j][b][k] = s;
            }
    for (int i = 0; i < NP; ++i)
        for (int j = 0; j < NP; ++j) {
            double s = 0;
            for (int q = 0; q < NQ; ++q) s += t.w[q] * t.NP1[q][i] * t.NP1[q][j];
            t.MP[i][j] = s;
        }
    return t;
}
static constexpr Tables kT = make_tables();

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
template <class F> __device__ inline D3<F> operator+(D3<F> a, F b) { a.v += b; return a; }
template <class F> __device__ inline D3<F> operator-(D3<F> a, F b) { a.v -= b; return a; }
template <class F> __device__ inline D3<F> operator*(D3<F> a, F b) { return D3<F>(a.v * b, a.d0 * b, a.d1 * b, a.d2 * b); }
template <class F> __device__ inline D3<F> operator*(F b, D3<F> a) { return a * b; }
template <class F> __device__ inline D3<F> dsqrt(D3<F> a) { F s = sqrt(a.v); F h = F(0.5) / s; return D3<F>(s, a.d0 * h, a.d1 * h, a.d2 * h); }
template <class F> __device__ inline D3<F> drcp(D3<F> a) { F r = F(1) / a.v; F h = -r * r; return D3<F>(r, a.d0 * h, a.d1 * h, a.d2 * h); }
__device__ inline double dsqrt(double a) { return sqrt(a); }
__device__ inline float dsqrt(float a) { return sqrtf(a); }
__device__ inline double drcp(double a) { return 1.0 / a; }
__device__ inline float drcp(float a) { return 1.0f / a; }
template <class F> __device__ inline F val(D3<F> a) { return a.v; }
__device__ inline double val(double a) { return a; }
__device__ inline float val(float a) { return a; }

template <class F> __device__ inline F eps2();
template <> __device__ inline double eps2<double>() { return 4.930380657631324e-32; }   // eps(Float64)^2
template <> __device__ inline float eps2<float>() { return 1.4210855e-14f; }           // eps(Float32)^2

// Per-phase table columns (device global, copied to LDS per block).
enum PhaseCol {
    PC_ETA = 0, PC_INVG, PC_RHO0, PC_ALPHA, PC_BETA, PC_BULK,
    PC_COSPHI, PC_SINPHI, PC_SINPSI, PC_C, PC_EREG, PC_KB,
    PC_ETAVE, PC_INV2GDT, PC_YIELD0, PC_INVDEN,   // dt-dependent, element-constant phases
    PC_PT, PC_CMIN, PC_HC,                        // tensile cap: pT, C_min, H_C
    PC_NCOL
};

// Per-quadrature-point material coefficients (rheology.jl:35-42) and the
// Drucker-Prager parameters at the point (rheology.jl:296-302): yield0 =
// cos(ϕ)·C, invden = 1 / (ηve + η_reg + Kb Δt sin(Ψ) sin(ϕ)).
template <class F>
struct IpMat {
    F etave, inv2Gdt;
    F yield0, sphi, invden;
};

// Deviatoric stress with the plane-strain Drucker-Prager return
// (rheology.jl:226-245, 272-334). S is F, or D3<F> for the tangent. Returns
// whether the point yields.
// PQ is F, or S to carry the derivative with respect to the point pressure.
template <class S, class F, bool PLASTIC, class PQ = F>
__device__ inline bool stress2d(S exx, S eyy, S exy, const IpMat<F>& m, F txo, F tyo, F txyo, PQ Pq,
                                S& txx, S& tyy, S& txy, F* fyield = nullptr, float* tan = nullptr) {
    bool yields = false;
    S tr = (exx + eyy) * F(1.0 / 3.0);
    const F two_eve = F(2) * m.etave;
    txx = two_eve * ((exx - tr) + txo * m.inv2Gdt);
    tyy = two_eve * ((eyy - tr) + tyo * m.inv2Gdt);
    txy = two_eve * (exy + txyo * m.inv2Gdt);
    if (PLASTIC) {
        S tzz = -txx - tyy;
        S tII = dsqrt((txx * txx + tyy * tyy + tzz * tzz) * F(0.5) + txy * txy + eps2<F>());
        S F_ = (tII - m.yield0) - m.sphi * Pq;
        if (fyield) *fyield = val(F_);
        // whole waves without a yielding point skip the return (no if-conversion)
        if (__any(val(F_) > F(0)) && val(F_) > F(0)) {
            S lam = F_ * m.invden;
            if (val(lam) > F(0)) {
                // ∂Q/∂τ = ((2τxx+τyy), (τxx+2τyy), 2τxy) / (2 τII_s)
                S h = two_eve * lam * drcp((tII + eps2<F>()) * F(2));
                S nxx = (F(2) * txx + tyy) * h;
                S nyy = (txx + F(2) * tyy) * h;
                S nxy = (F(2) * txy) * h;
                if constexpr (std::is_same<S, F>::value) {
                    if (tan) {
                        const F tIIs = tII + eps2<F>();
                        tan[0] = (float)h;
                        tan[1] = (float)((m.etave / tIIs) * (m.invden - lam / tIIs) / (F(2) * tII));
                        tan[2] = (float)(-(m.etave * m.invden * m.sphi) / tIIs);
                        tan[3] = (float)(F(2) * txx + tyy);
                        tan[4] = (float)(txx + F(2) * tyy);
                        tan[5] = (float)(F(2) * txy);
                    }
                }
                txx = txx - nxx;
                tyy = tyy - nyy;
                txy = txy - nxy;
                yields = true;
            }
        }
    }
    return yields;
}

// Plastic correction c = τ_trial - τ_returned of the plane-strain
// Drucker-Prager return (rheology.jl:272-334) for a trial stress, zero where
// the point does not yield; returns whether it yields.
template <class F>
__device__ inline bool dp_correction(F txx, F tyy, F txy, const IpMat<F>& m, F Pq, F& cxx, F& cyy, F& cxy) {
    cxx = cyy = cxy = F(0);
    const F tzz = -txx - tyy;
    const F tII = dsqrt((txx * txx + tyy * tyy + tzz * tzz) * F(0.5) + txy * txy + eps2<F>());
    const F F_ = (tII - m.yield0) - m.sphi * Pq;
    if (!(F_ > F(0))) return false;
    const F lam = F_ * m.invden;
    if (!(lam > F(0))) return false;
    const F h = F(2) * m.etave * lam * drcp((tII + eps2<F>()) * F(2));
    cxx = (F(2) * txx + tyy) * h;
    cyy = (txx + F(2) * tyy) * h;
    cxy = (F(2) * txy) * h;
    return true;
}

// ---------------------------------------------------------------------------
// Drucker-Prager tensile cap (rheology.jl:510-760; Popov, Berlie & Kaus 2025,
// Eqs. 13-18, 42), FP64. Pressure is compression-positive, pT <= 0.
// ---------------------------------------------------------------------------
struct CapGeo { double a, b, p_y, R_y, p_d, tau_d, p_q; };

// cap_geometry (rheology.jl:546-557)
__device__ inline CapGeo cap_geo(double k, double kq, double c, double pT) {
    CapGeo g;
    g.a = sqrt(1 + k * k);
    g.b = sqrt(1 + kq * kq);
    g.p_y = (g.a * pT + c) / (g.a - k);
    g.R_y = g.p_y - pT;
    g.p_d = g.p_y - g.R_y * k / g.a;
    g.tau_d = g.R_y / g.a;
    g.p_q = g.p_d + kq * g.tau_d;
    return g;
}

// cap_yield_function and cap_invariants (rheology.jl:582-609) with their
// partials in (s, p): the yield and flow-potential branches are chosen
// independently by the half-plane tests of the reference
struct CapInv { double F, At, Ap, dF[2], dAt[2], dAp[2]; };
__device__ inline CapInv cap_inv(double s, double p, double k, double kq, double c, const CapGeo& g) {
    CapInv r;
    if (p + k * s >= g.p_y) {
        r.F = s - k * p - c;
        r.dF[0] = 1; r.dF[1] = -k;
    } else {
        const double dy = p - g.p_y;
        const double R = sqrt(s * s + dy * dy + eps2<double>());
        r.F = g.a * (R - g.R_y);
        r.dF[0] = g.a * s / R; r.dF[1] = g.a * dy / R;
    }
    r.dAt[0] = r.dAt[1] = r.dAp[0] = r.dAp[1] = 0;
    if (p + kq * s >= g.p_q) {
        r.At = 0.5; r.Ap = kq;
        return r;
    }
    const double d = p - g.p_q;
    const double Rq = sqrt(s * s + d * d);
    if (Rq == 0) { r.At = 0; r.Ap = 0; return r; }
    const double i3 = 1 / (Rq * Rq * Rq);
    r.At = g.b * s / (2 * Rq);
    r.Ap = -g.b * d / Rq;
    r.dAt[0] = 0.5 * g.b * d * d * i3; r.dAt[1] = -0.5 * g.b * s * d * i3;
    r.dAp[0] = g.b * d * s * i3;       r.dAp[1] = -g.b * s * s * i3;
    return r;
}

// 3x3 solve, Gaussian elimination with partial pivoting
__device__ inline bool cap_solve3(double A[3][3], double b[3], double x[3]) {
    int p0 = 0, p1 = 1, p2 = 2;
    int piv[3] = {p0, p1, p2};
#pragma unroll
    for (int col = 0; col < 3; ++col) {
        int best = col;
        for (int r = col + 1; r < 3; ++r)
            if (fabs(A[piv[r]][col]) > fabs(A[piv[best]][col])) best = r;
        const int t = piv[col]; piv[col] = piv[best]; piv[best] = t;
        const double dg = A[piv[col]][col];
        if (dg == 0 || !isfinite(dg)) return false;
        for (int r = col + 1; r < 3; ++r) {
            const double f = A[piv[r]][col] / dg;
            for (int cc = col; cc < 3; ++cc) A[piv[r]][cc] -= f * A[piv[col]][cc];
            b[piv[r]] -= f * b[piv[col]];
        }
    }
    for (int i = 2; i >= 0; --i) {
        double acc = b[piv[i]];
        for (int cc = i + 1; cc < 3; ++cc) acc -= A[piv[i]][cc] * x[cc];
        x[i] = acc / A[piv[i]][i];
    }
    return true;
}

struct CapPoint {
    double s_tr, p_tr, eve, KDt, k, kq, c, pT, ereg;
    CapGeo g;
    __device__ void residual(const double x[3], double r[3]) const {
        const CapInv v = cap_inv(x[0], x[1], k, kq, c, g);
        r[0] = x[0] - s_tr + 2 * eve * x[2] * v.At;
        r[1] = x[1] - p_tr - KDt * x[2] * v.Ap;
        r[2] = v.F - ereg * x[2];
    }
    __device__ void jacobian(const double x[3], double J[3][3]) const {
        const CapInv v = cap_inv(x[0], x[1], k, kq, c, g);
        const double lam = x[2];
        J[0][0] = 1 + 2 * eve * lam * v.dAt[0]; J[0][1] = 2 * eve * lam * v.dAt[1]; J[0][2] = 2 * eve * v.At;
        J[1][0] = -KDt * lam * v.dAp[0]; J[1][1] = 1 - KDt * lam * v.dAp[1]; J[1][2] = -KDt * v.Ap;
        J[2][0] = v.dF[0]; J[2][1] = v.dF[1]; J[2][2] = -ereg;
    }
};

// cap_return_map (rheology.jl:652-690): x = (s, p, λ); Newton on residuals
// scaled by max(|s_trial|, |p_trial|, |F_trial|, eps), halving backtracking
// (24 tries) under s >= 0, λ >= 0, a finite residual and the Armijo test,
// ‖r‖∞ <= 100 eps, one final full step. dx = ∂x/∂s_trial = J⁻¹ e₁ at the
// solution (the derivative the final step carries through dual numbers).
struct CapRes { double s, p, lam, At, Ap, dx[3]; bool ok; };
__device__ inline CapRes cap_return(const CapPoint& cp) {
    CapRes o;
    const CapInv v0 = cap_inv(cp.s_tr, cp.p_tr, cp.k, cp.kq, cp.c, cp.g);
    o.s = cp.s_tr; o.p = cp.p_tr; o.lam = 0; o.At = v0.At; o.Ap = v0.Ap;
    o.dx[0] = 1; o.dx[1] = 0; o.dx[2] = 0;
    o.ok = true;
    if (v0.F <= 0) return o;
    o.ok = false;
    if (!(isfinite(cp.KDt) && cp.KDt > 0 && isfinite(cp.eve) && cp.eve > 0)) return o;
    const double eps = 2.220446049250313e-16;
    const double scale = fmax(fmax(fabs(cp.s_tr), fabs(cp.p_tr)), fmax(fabs(v0.F), eps));
    const double isc = 1 / scale, tol = 100 * eps;
    double x[3] = {cp.s_tr, cp.p_tr, 0.0}, r[3];
    cp.residual(x, r);
    for (int i = 0; i < 3; ++i) r[i] *= isc;
    bool conv = fmax(fabs(r[0]), fmax(fabs(r[1]), fabs(r[2]))) <= tol;
    for (int it = 0; it < 40 && !conv; ++it) {
        double J[3][3], rhs[3] = {r[0], r[1], r[2]}, st[3];
        cp.jacobian(x, J);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] *= isc;
        if (!cap_solve3(J, rhs, st) || !(isfinite(st[0]) && isfinite(st[1]) && isfinite(st[2]))) return o;
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        double alpha = 1;
        bool acc = false;
        for (int ls = 0; ls < 24; ++ls) {
            const double cand[3] = {x[0] - alpha * st[0], x[1] - alpha * st[1], x[2] - alpha * st[2]};
            if (cand[0] >= 0 && cand[2] >= 0) {
                double rn[3];
                cp.residual(cand, rn);
                for (int i = 0; i < 3; ++i) rn[i] *= isc;
                if (isfinite(rn[0]) && isfinite(rn[1]) && isfinite(rn[2]) &&
                    rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2] <= (1 - alpha / 10000) * r2) {
                    for (int i = 0; i < 3; ++i) { x[i] = cand[i]; r[i] = rn[i]; }
                    acc = true;
                    break;
                }
            }
            alpha *= 0.5;
        }
        if (!acc) return o;
        conv = fmax(fabs(r[0]), fmax(fabs(r[1]), fabs(r[2]))) <= tol;
    }
    if (!conv) return o;
    double J[3][3], Jc[3][3], rhs[3] = {r[0], r[1], r[2]}, st[3];
    cp.jacobian(x, J);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { Jc[i][j] = J[i][j]; J[i][j] *= isc; }
    if (!cap_solve3(J, rhs, st)) return o;
    for (int i = 0; i < 3; ++i) x[i] -= st[i];
    double e1[3] = {1, 0, 0};
    if (!cap_solve3(Jc, e1, o.dx)) return o;
    const CapInv v = cap_inv(x[0], x[1], cp.k, cp.kq, cp.c, cp.g);
    o.s = x[0]; o.p = x[1]; o.lam = x[2]; o.At = v.At; o.Ap = v.Ap;
    o.ok = true;
    return o;
}

// cap_local_update (rheology.jl:708-717): radial deviatoric return of the
// trial stress and the corrected pressure P_trial + KΔt λ Ap. S = double, or
// D3<double> carrying the implicit derivative with respect to the strain
// rate (P_trial does not depend on the velocity). A failed local solve
// gives NaN, which the solver's residual checks report.
template <class S>
__device__ inline void cap_update(S txx, S tyy, S txy, const CapPoint& cp0, S& oxx, S& oyy, S& oxy, S& P,
                                  double* gdot = nullptr, double* thdot = nullptr) {
    const S tzz = -txx - tyy;
    const S s_tr = dsqrt((txx * txx + tyy * tyy + tzz * tzz) * 0.5 + txy * txy + eps2<double>());
    CapPoint cp = cp0;
    cp.s_tr = val(s_tr);
    const CapRes r = cap_return(cp);
    if (!r.ok) {
        const double nan = __builtin_nan("");
        oxx = oyy = oxy = P = S(nan);
        if (gdot) { *gdot = nan; *thdot = nan; }
        return;
    }
    const S ds = s_tr - val(s_tr);
    const S s = ds * r.dx[0] + r.s, p = ds * r.dx[1] + r.p, lam = ds * r.dx[2] + r.lam;
    S At = S(r.At), Ap = S(r.Ap);
    if (r.lam != 0) {
        const CapInv v = cap_inv(r.s, r.p, cp.k, cp.kq, cp.c, cp.g);
        At = (s - r.s) * v.dAt[0] + (p - r.p) * v.dAt[1] + r.At;
        Ap = (s - r.s) * v.dAp[0] + (p - r.p) * v.dAp[1] + r.Ap;
    }
    const S g = lam * At, th = lam * Ap;
    const S f = g * (2 * cp.eve) * drcp(s_tr);
    oxx = txx - f * txx;
    oyy = tyy - f * tyy;
    oxy = txy - f * txy;
    P = th * cp.KDt + cp.p_tr;
    if (gdot) { *gdot = val(g); *thdot = val(th); }
}

// ---------------------------------------------------------------------------
// Element kernel
// ---------------------------------------------------------------------------
// compile-time feature flags of the element kernel
enum ElemFlags {
    EF_ELASTIC = 1,   // finite shear modulus somewhere: read τ_old
    EF_PLASTIC = 2,   // Drucker-Prager return
    EF_BULK = 4,      // finite pressure storage modulus: read P_old
    EF_ITER = 8,      // inner pseudo-iteration: Pnum in the momentum, bubble update
    EF_GEN = 16,      // element-node phases, temperature, sources, fluid pressure
    EF_STRUCT = 32,   // structured box mesh: connectivity and geometry from the element index
    EF_LIN = 64,      // operator of a correction δ: R(x) + L δ with the right-hand side R(x)
                      // given; storage on δP alone, density linear in δP. Alone: L is the
                      // visco-elastic operator (no τ_old, no return map)
    EF_DELTA = 128,   // with EF_LIN|EF_PLASTIC: the stress term is τ(x + δ) - τ(x), both with
                      // the Drucker-Prager return about the FP64 state x (xvx, xvy, xP)
    EF_TILE = 256,    // with EF_STRUCT|EF_ITER: a block is a tile of 16 x (BLOCK/32) cells; the
                      // shared-node residuals are summed in LDS and stored per node for the
                      // tile's interior nodes, as per-tile partial sums for its perimeter
                      // nodes; k_tile_update finishes both
    EF_TANGENT = 512, // with EF_DELTA: the return's part of τ(x + δ) - τ(x) is its linearisation
                      // about x, read per point (tangent) instead of two FP64 returns
};

// Linearisation of the plane-strain Drucker-Prager return about a yielding
// trial stress t at pressure P (stress2d's formulas). With q = M t, M =
// [[2,1,0],[1,2,0],[0,0,2]], the correction c = t - τ is c = h q, h = ηve λ /
// τII_s, λ = F invden, F = τII - y0 - sinϕ P, and
//   ∂c/∂t u = h M u + kq q (q·u),   kq = (ηve/τII_s)(invden - λ/τII_s)/(2 τII),
//   ∂c/∂P   = pq q,                 pq = -ηve invden sinϕ / τII_s,
// from ∂τII/∂t = M t / (2 τII). Stored per point as (h, kq, pq, q0, q1, q2).
constexpr int NTAN = 6;

// Slot of the T6 grid node at offset pos = ox + 3 oy (ox, oy in 0..2) from a
// cell's lower-left node in triangle tri of a cell of parity par, or -1
// (the local node orders of t7_struct_element)
__device__ constexpr int8_t kT7Slot[2][2][9] = {
    // even cells: A = {BL, BR, TR, bot, rgt, ctr}, B = {BL, TR, TL, ctr, top, lft}
    {{0, 3, 1, -1, 5, 4, -1, -1, 2}, {0, -1, -1, 5, 3, -1, 2, 4, 1}},
    // odd cells:  A = {BL, BR, TL, bot, ctr, lft}, B = {BR, TR, TL, rgt, top, ctr}
    {{0, 3, 1, 5, 4, -1, 2, -1, -1}, {-1, -1, 0, -1, 5, 3, 2, 4, 1}},
};
// perimeter index of tile node (jx, jy) of a tile with TW x TH grid intervals:
// bottom row, top row, left column, right column
__device__ inline int tile_perim(int jx, int jy, int TW, int TH) {
    if (jy == 0) return jx;
    if (jy == TH) return (TW + 1) + jx;
    if (jx == 0) return 2 * (TW + 1) + (jy - 1);
    return 2 * (TW + 1) + (TH - 1) + (jy - 1);
}

// Structured T7 box mesh (connectivity.jl:111-140): element e = 2c + k of cell
// c = ey nx + ex; the cell's two triangles alternate diagonals with the cell
// parity. Returns the 7 local nodes (T6 grid nodes, then the centroid
// nT6 + e) and the affine map J^-1 of the triangle (|det J| = dx dy).
template <class F>
__device__ inline void t7_struct_element(int64_t e, int32_t nx, int32_t sx, int64_t nT6, F idx, F idy,
                                         int32_t n[7], F& i00, F& i01, F& i10, F& i11) {
    const int32_t c = (int32_t)(e >> 1), k = (int32_t)(e & 1);
    const int32_t ey = c / nx, ex = c - ey * nx;
    const int32_t base = 2 * ey * sx + 2 * ex;
    const int32_t BL = base, BR = base + 2, TL = base + 2 * sx, TR = base + 2 * sx + 2;
    const int32_t bot = base + 1, lft = base + sx, ctr = base + sx + 1, rgt = base + sx + 2, top = base + 2 * sx + 1;
    const bool even = ((ex + ey) & 1) == 0;
    if (even) {
        if (k == 0) { n[0] = BL; n[1] = BR; n[2] = TR; n[3] = bot; n[4] = rgt; n[5] = ctr; i00 = idx; i01 = -idy; i10 = 0; i11 = idy; }
        else        { n[0] = BL; n[1] = TR; n[2] = TL; n[3] = ctr; n[4] = top; n[5] = lft; i00 = idx; i01 = 0; i10 = -idx; i11 = idy; }
    } else {
        if (k == 0) { n[0] = BL; n[1] = BR; n[2] = TL; n[3] = bot; n[4] = ctr; n[5] = lft; i00 = idx; i01 = 0; i10 = 0; i11 = idy; }
        else        { n[0] = BR; n[1] = TR; n[2] = TL; n[3] = rgt; n[4] = top; n[5] = ctr; i00 = idx; i01 = idy; i10 = -idx; i11 = 0; }
    }
    n[6] = (int32_t)(nT6 + e);
}

template <class F>
struct ElemArgs {
    int64_t nels;
    const int32_t* el2n;       // [7][nels] (SoA, 0-based)
    const F* geo;              // [5][nels]: J^-1 (00,01,10,11), |det J|
    F* vx; F* vy;              // velocity (bubble entries updated in place)
    const F* P; const F* Pold; // Pold read with EF_BULK
    const F* pscale;           // γP / M_P per pressure DoF
    const F* tau_old0; const F* tau_old1; const F* tau_old2; // [nels][7], EF_ELASTIC
    const int32_t* phase_v;    // cell [nels] or, with EF_GEN, element-node [7][nels] (0-based)
    const int32_t* phase_P;    // cell [nels] or, with EF_GEN, element-node [3][nels]
    const double* ptab;        // [PC_NCOL][MAXPH]
    int32_t nphases;
    F gx, gy, Tref, dt;
    typename Vec2<F>::T* Re;   // [nels][6] element-slot residuals of shared nodes
    F* rate_x; F* rate_y; const F* PCx; const F* PCy;
    const uint8_t* bcflag; const F* bcval_x; const F* bcval_y;
    F alpha_x, alpha_y, beta_x, beta_y;  // positive Chebyshev step and momentum
    const F* cheb;             // nullable: (alpha_x, alpha_y, beta_x, beta_y) in device memory
                               // instead (launches replayed from a graph)
    int32_t do_update;         // EF_ITER launches inside the DR loop update the bubble node
    // optional outputs (null unless stored this launch)
    F* RP_out; F* Pnum_out;
    F* Rv_out_x; F* Rv_out_y; F* Rv0_out_x; F* Rv0_out_y;
    F* tau_out0; F* tau_out1; F* tau_out2;
    // EF_STRUCT: nx cells per row, sx = 2 nx + 1 grid nodes per row, nT6 grid nodes
    int32_t nx, sx; int64_t nT6; F idx, idy, detc;
    // EF_TILE: tiles per row and in all, per-tile partial sums of the perimeter
    // nodes, summed residuals of the interior nodes (by grid node)
    int32_t ntx; int64_t ntiles; typename Vec2<F>::T* bpart; typename Vec2<F>::T* Rn;
    int32_t phase_same;        // phase_P holds the same cell phases as phase_v
    // EF_LIN right-hand sides R(x) of the correction problem (element DoFs / nodes)
    const F* rhs_P; const F* rhs_vx; const F* rhs_vy;
    // EF_DELTA: the FP64 state x the correction is taken about (pressure,
    // nullable fluid pressure)
    const double* xvx; const double* xvy; const double* xP; const double* xPf;
    // nullable: receives the number of elements with a yielding point (one
    // atomic add per wave)
    int32_t* yield_count;
    // nullable: per element, 1 if a point is within 20 % of its yield stress
    // of the yield surface (or beyond); written with yield_count
    uint8_t* yield_flag;
    // EF_DELTA: those flags; blocks without a flagged element skip the return
    // map (their operator is the exact visco-elastic one)
    const uint8_t* hot;
    // [NTAN][NQ][nels] return linearisations: written by plastic residual
    // launches (tangent_out, nullable; zeros at points that do not yield), read
    // by EF_TANGENT launches at the points of hot elements (tangent)
    float* tangent_out;
    const float* tangent;
    // nullable, with tangent_out: receives the number of yielding points whose
    // linearisation may not be positive (the lower bound 1 - 3h - max(kq, 0)|q|^2
    // of its eigenvalues below 0.05: yield in tension, F > τII), for which the
    // caller keeps the two FP64 returns
    int32_t* tangent_bad;
    // solve! array keywords (FP64 launches; nullable): nodal loads subtracted
    // from the momentum residual, with EF_GEN the viscosity at the points
    // [nels][NQ] and the previous-step pressure at the points [nels][NQ]
    const F* bf_x; const F* bf_y;
    const F* eta_ip; const F* Pold_ip;
    // EF_GEN extras
    int32_t nodeph;            // phases stored per element node
    int32_t bulk_all_inf;
    int32_t thermal;
    const F* T; const F* T0; const F* Q; const F* Pf;   // nullable
    // nullable, residual launches only: the shear dissipation τ : (ε̇ − ε̇ᵉˡ)
    // per quadrature point [nels][7]
    F* phi_out;
    // tensile cap (FP64, EF_GEN|EF_PLASTIC launches): the yield model, the
    // accepted plastic strain γ that softens the cohesion (nullable), the
    // history written by a commit (γ += Δt γdot, θ += Δt θdot; nullable) and
    // the pressure of the momentum balance per point (nullable)
    int32_t cap;
    const F* gamma;
    F* hist_gamma; F* hist_theta;
    F* P_out;
};

// Workgroups are dealt round-robin over the 8 XCDs (one L2 each). The grid
// is padded to a multiple of 8 and block b is given logical tile
// (b % 8) * (grid / 8) + b / 8, so each XCD works on one contiguous range of
// elements / nodes and neighbouring tiles share that XCD's L2.
__device__ inline int64_t xcd_tile() {
    const int64_t per = gridDim.x >> 3;
    return (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// Occupancy target of the element kernel: 3 waves per SIMD (the FP64 kernel's
// register budget at which it runs fastest on gfx950)
#define FEM_ELEM_ATTR __attribute__((amdgpu_waves_per_eu(3)))

// gathered residual (rx, ry) of a shared node: right-hand side, Dirichlet
// pinning, optional stores and the Chebyshev update (helpers.jl:31-41,
// dynamic_relaxation/solver.jl:139-166; DR.jl:686-697)
template <class F>
__device__ inline void node_update(int64_t node, F rx, F ry, const F* rhs_vx, const F* rhs_vy, const uint8_t* bcflag,
                                   const F* bcval_x, const F* bcval_y, F* Rv_out_x, F* Rv_out_y, F* Rv0_out_x,
                                   F* Rv0_out_y, int32_t do_update, F alpha_x, F alpha_y, F beta_x, F beta_y,
                                   F* rate_x, F* rate_y, const F* PCx, const F* PCy, F* vx, F* vy) {
    if (rhs_vx) { rx += rhs_vx[node]; ry += rhs_vy[node]; }
    const uint8_t fl = bcflag[node];
    if (fl & 1) rx = 0;
    if (fl & 2) ry = 0;
    if (Rv_out_x) { Rv_out_x[node] = rx; Rv_out_y[node] = ry; }
    if (Rv0_out_x) { Rv0_out_x[node] = rx; Rv0_out_y[node] = ry; }
    if (do_update) {
        F nrx = beta_x * rate_x[node] + rx / PCx[node];
        F nry = beta_y * rate_y[node] + ry / PCy[node];
        F nvx = vx[node] + (-alpha_x) * nrx;
        F nvy = vy[node] + (-alpha_y) * nry;
        if (fl & 1) { nrx = 0; nvx = bcval_x[node]; }
        if (fl & 2) { nry = 0; nvy = bcval_y[node]; }
        rate_x[node] = nrx; rate_y[node] = nry;
        vx[node] = nvx; vy[node] = nvy;
    }
}

// The element kernel's body: block blk of the launch of one problem.
template <class F, int FL, int BLOCK>
__device__ __forceinline__ void t7_element_body(const ElemArgs<F>& A, const int64_t blk) {
    constexpr bool ELASTIC = FL & EF_ELASTIC, PLASTIC = FL & EF_PLASTIC, BULK = FL & EF_BULK;
    constexpr bool ITER = FL & EF_ITER, GEN = FL & EF_GEN, STRUCT = FL & EF_STRUCT, LIN = FL & EF_LIN;
    constexpr bool DELTA = FL & EF_DELTA, TILE = FL & EF_TILE, TANGENT = FL & EF_TANGENT;
    static_assert(!TANGENT || DELTA, "EF_TANGENT linearises the EF_DELTA return map");
    static_assert(!(LIN && !DELTA && (ELASTIC || PLASTIC)), "the linearised operator has no history or return map");
    static_assert(!DELTA || (LIN && PLASTIC), "EF_DELTA is the correction operator with the return map");
    static_assert(!TILE || (STRUCT && ITER && BLOCK % 64 == 0), "tiles are blocks of structured cells");
    // τ_old of the block's elements staged through LDS by direct global->LDS loads
    constexpr bool TAU_LDS = ELASTIC;
    constexpr int TCY = BLOCK / 32;                 // EF_TILE: rows of 16 cells per tile
    const int64_t ne = A.nels;
    int64_t e0, e, tile = 0, ttx = 0, tty = 0;
    bool valid;
    if (TILE) {
        // thread t: triangle t & 1 of cell (t >> 1) & 15 in row t >> 5 of the tile
        tile = blk;
        if (tile >= A.ntiles) return;                // whole block
        tty = tile / A.ntx;
        ttx = tile - tty * A.ntx;
        e0 = 2 * (tty * TCY * (int64_t)A.nx + ttx * 16);   // first element of the tile's first row
        e = e0 + 2 * (int64_t)(threadIdx.x >> 5) * A.nx + (threadIdx.x & 31);
        valid = true;
    } else {
        e0 = blk * BLOCK;                            // first element of this block
        valid = e0 + threadIdx.x < ne;
        e = valid ? e0 + threadIdx.x : ne - 1;       // idle lanes recompute the last element, store nothing
    }
    // per-phase coefficients: a few KB read through the caches
#define PT(col, p) ((F)A.ptab[(col) * MAXPH + (p)])
    const bool nodeph = GEN && A.nodeph;
    // τ_old of the block's elements is one contiguous chunk per component:
    // stream it straight into LDS (no registers held) while the velocity
    // gathers and the continuity residual proceed.
    // with EF_TILE the slot sums of the tile stage reuse this stage once every
    // τ_old read of the block is done (FP64: 168 B per thread instead of 264,
    // so three blocks of 128 fit a CU's LDS next to the register budget)
    constexpr bool SHARE_TAU = TAU_LDS && TILE;
    constexpr int TAU_N = TAU_LDS ? 3 * BLOCK * NQ : 1;
    constexpr int RE_N = (int)((BLOCK * NSH * sizeof(typename Vec2<F>::T) + sizeof(F) - 1) / sizeof(F));
    __shared__ __align__(16) F s_tau[SHARE_TAU && RE_N > TAU_N ? RE_N : TAU_N];
    // hot: the block stages τ_old (some element is near yield); hot_e: this
    // element takes the FP64 returns (an element away from yield has none in
    // either state, so its correction operator is the visco-elastic one)
    bool hot = true, hot_e = true;
    if (DELTA) {
        hot_e = valid && A.hot[e];
        hot = __syncthreads_or(hot_e);
    }
    if (TAU_LDS && TILE && hot) {
        // one contiguous chunk of 32 elements per tile row and component
        constexpr int bytes = 32 * NQ * (int)sizeof(F);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const F* src[3] = {A.tau_old0, A.tau_old1, A.tau_old2};
#pragma unroll
        for (int c = 0; c < 3; ++c)
            for (int r = wave; r < TCY; r += BLOCK / 64) {
                const char* g = (const char*)(src[c] + (e0 + 2 * (int64_t)r * A.nx) * NQ);
                char* l = (char*)(s_tau + c * BLOCK * NQ + r * 32 * NQ);
                for (int off0 = 0; off0 < bytes; off0 += 64 * 16)
                    if (off0 + lane * 16 < bytes)
                        __builtin_amdgcn_global_load_lds((const void*)(g + off0 + lane * 16),
                                                         (__attribute__((address_space(3))) void*)(l + off0), 16, 0, 0);
            }
    } else if (TAU_LDS && e0 < ne && hot) {
        const int64_t nvalid = ne - e0 < BLOCK ? ne - e0 : BLOCK;
        const int bytes = (int)(nvalid * NQ * sizeof(F));
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const F* src[3] = {A.tau_old0, A.tau_old1, A.tau_old2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const char* g = (const char*)(src[c] + e0 * NQ);
            char* l = (char*)(s_tau + c * BLOCK * NQ);
            for (int off0 = wave * 64 * 16; off0 < bytes; off0 += BLOCK * 16) {
                if (off0 + lane * 16 < bytes)
                    __builtin_amdgcn_global_load_lds((const void*)(g + off0 + lane * 16),
                                                     (__attribute__((address_space(3))) void*)(l + off0), 16, 0, 0);
            }
        }
    }

    int32_t n[NV];
    F i00, i01, i10, i11, det;
    if (STRUCT) {
        t7_struct_element<F>(e, A.nx, A.sx, A.nT6, A.idx, A.idy, n, i00, i01, i10, i11);
        det = A.detc;
    } else {
#pragma unroll
        for (int a = 0; a < NV; ++a) n[a] = A.el2n[a * ne + e];
        i00 = A.geo[0 * ne + e]; i01 = A.geo[1 * ne + e];
        i10 = A.geo[2 * ne + e]; i11 = A.geo[3 * ne + e];
        det = A.geo[4 * ne + e];
    }
    F vx[NV], vy[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) { vx[a] = A.vx[n[a]]; vy[a] = A.vy[n[a]]; }
    F P[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) P[i] = A.P[e * NP + i];
    const int pc = A.phase_v[e];

    // ---- continuity residual (pressure_residual.jl:132-158). The element is
    // affine, so -Σ_q N^P_i (∇·v)_q dΩ_q = -det Σ_a Σ_k TB[i][a][k] u_k(a)
    // with u(a) the nodal velocity rotated by J^-T; storage, thermal and
    // source terms are P1 functions that the rule integrates exactly (det M ·).
    F RP[NP] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        const F u0 = i00 * vx[a] + i01 * vy[a];
        const F u1 = i10 * vx[a] + i11 * vy[a];
#pragma unroll
        for (int i = 0; i < NP; ++i) RP[i] -= (F)kT.TB[i][a][0] * u0 + (F)kT.TB[i][a][1] * u1;
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) RP[i] *= det;
    if (BULK || GEN) {
        F dP[NP] = {0, 0, 0}, dT[NP] = {0, 0, 0}, Qv[NP] = {0, 0, 0};
        if (BULK) {
#pragma unroll
            for (int i = 0; i < NP; ++i) dP[i] = LIN ? P[i] : P[i] - A.Pold[e * NP + i];
        }
        if (GEN && !LIN && A.thermal) {
#pragma unroll
            for (int i = 0; i < NP; ++i) dT[i] = A.T[e * NP + i] - A.T0[e * NP + i];
        }
        if (GEN && !LIN && A.Q) {
#pragma unroll
            for (int i = 0; i < NP; ++i) Qv[i] = A.Q[e * NP + i];
        }
        const bool ipold = GEN && BULK && !LIN && A.Pold_ip != nullptr;
        if (!nodeph && !ipold) {
            const int p = A.phase_same ? pc : A.phase_P[e];
            const F bulk = (GEN && A.bulk_all_inf) || !BULK ? (F)INFINITY : PT(PC_BULK, p);
            const F al = PT(PC_ALPHA, p);
            F src[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) src[j] = -(dP[j] / (bulk * A.dt)) + al * dT[j] / A.dt + Qv[j];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                F acc = 0;
#pragma unroll
                for (int j = 0; j < NP; ++j) acc += (F)kT.MP[i][j] * src[j];
                RP[i] += det * acc;
            }
        } else {
            int pp[NP];
            const int pcP = A.phase_same ? pc : A.phase_P[e];
#pragma unroll
            for (int i = 0; i < NP; ++i) pp[i] = nodeph ? A.phase_P[i * ne + e] : pcP;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                F bulk = 0, alq = 0, dPq = 0, dTq = 0, Qq = 0, Pq = 0;
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const F Ni = (F)kT.NP1[q][i];
                    bulk += Ni * PT(PC_BULK, pp[i]);
                    alq += Ni * PT(PC_ALPHA, pp[i]);
                    dPq += Ni * dP[i];
                    dTq += Ni * dT[i];
                    Qq += Ni * Qv[i];
                    Pq += Ni * P[i];
                }
                // pressure_increment_at_ip (pressure_residual.jl:90-92): N·P − P_old[q]
                if (ipold) dPq = Pq - A.Pold_ip[e * NQ + q];
                if (A.bulk_all_inf || !BULK) bulk = (F)INFINITY;
                const F src = (-(dPq / (bulk * A.dt)) + alq * dTq / A.dt + Qq) * ((F)kT.w[q] * det);
#pragma unroll
                for (int i = 0; i < NP; ++i) RP[i] += (F)kT.NP1[q][i] * src;
            }
        }
    }
    if (LIN) {
#pragma unroll
        for (int i = 0; i < NP; ++i) RP[i] += A.rhs_P[e * NP + i];
    }
    F Pn[NP] = {0, 0, 0};
    if (ITER) {
#pragma unroll
        for (int i = 0; i < NP; ++i) Pn[i] = A.pscale[e * NP + i] * RP[i];
    }
    if (TAU_LDS) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    if (!valid) return;
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
    F rho_c = PT(PC_RHO0, pc), alpha_c = PT(PC_ALPHA, pc), beta_c = PT(PC_BETA, pc);
    m.etave = PT(PC_ETAVE, pc);
    m.inv2Gdt = PT(PC_INV2GDT, pc);
    if (PLASTIC) {
        m.yield0 = PT(PC_YIELD0, pc);
        m.sphi = PT(PC_SINPHI, pc);
        m.invden = PT(PC_INVDEN, pc);
    }
    int pv[NV];
    if (nodeph) {
#pragma unroll
        for (int a = 0; a < NV; ++a) pv[a] = A.phase_v[a * ne + e];
    }
    F Rx[NV], Ry[NV];
    bool yel = false, near = false;
    int32_t nbad = 0;   // points of this element with a non-positive return linearisation
#pragma unroll
    for (int a = 0; a < NV; ++a) { Rx[a] = 0; Ry[a] = 0; }
    const bool grav = (A.gx != F(0)) || (A.gy != F(0));
    // FP32 (correction operator): chained FMAs and the body force summed
    // once per element after the point loop (ρ dΩ per point in bw); the FP64
    // path keeps the reference's per-point summation order
    constexpr bool F32 = std::is_same<F, float>::value;
    F bw[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (nodeph) {
            F eta = 0, invG = 0;
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const F Na = (F)kT.N[q][a];
                eta += Na * PT(PC_ETA, pv[a]);
                invG += Na * PT(PC_INVG, pv[a]);
            }
            m.etave = F(1) / (F(1) / eta + invG / A.dt);
            m.inv2Gdt = invG / (F(2) * A.dt);
            if (PLASTIC) {
                F cphi = 0, C = 0, sphi = 0, spsi = 0, ereg = 0, Kb = 0;
#pragma unroll
                for (int a = 0; a < NV; ++a) {
                    const F Na = (F)kT.N[q][a];
                    cphi += Na * PT(PC_COSPHI, pv[a]); C += Na * PT(PC_C, pv[a]);
                    sphi += Na * PT(PC_SINPHI, pv[a]); spsi += Na * PT(PC_SINPSI, pv[a]);
                    ereg += Na * PT(PC_EREG, pv[a]); Kb += Na * PT(PC_KB, pv[a]);
                }
                m.yield0 = cphi * C;
                m.sphi = sphi;
                m.invden = F(1) / ((m.etave + ereg) + ((Kb * A.dt) * (-spsi)) * (-sphi));
            }
        }
        // a viscosity given at the points (purely viscous: G = Inf, so ηve = η)
        if (GEN && !LIN && A.eta_ip) m.etave = A.eta_ip[e * NQ + q];
        F txo = 0, tyo = 0, txyo = 0;
        if (TAU_LDS) {
            txo = s_tau[0 * BLOCK * NQ + threadIdx.x * NQ + q];
            tyo = s_tau[1 * BLOCK * NQ + threadIdx.x * NQ + q];
            txyo = s_tau[2 * BLOCK * NQ + threadIdx.x * NQ + q];
        } else if (ELASTIC) {
            txo = A.tau_old0[e * NQ + q];
            tyo = A.tau_old1[e * NQ + q];
            txyo = A.tau_old2[e * NQ + q];
        }
        F Pq = 0, Pnq = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Pq += (F)kT.NP1[q][i] * P[i];
            if (ITER) Pnq += (F)kT.NP1[q][i] * Pn[i];
        }
        F Pfq = 0;
        if (GEN && !LIN && A.Pf) {
#pragma unroll
            for (int i = 0; i < NP; ++i) Pfq += (F)kT.NP1[q][i] * A.Pf[e * NP + i];
        }
        F Pmom = Pq;   // the pressure of the momentum balance (the cap corrects it)
        F ax = 0, ay = 0, bx_ = 0, by_ = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const F d0 = (F)kT.dN[q][a][0], d1 = (F)kT.dN[q][a][1];
            ax += d0 * vx[a]; ay += d1 * vx[a];
            bx_ += d0 * vy[a]; by_ += d1 * vy[a];
        }
        const F gxx = ax * i00 + ay * i10;   // ∂vx/∂x
        const F gxy = ax * i01 + ay * i11;   // ∂vx/∂y
        const F gyx = bx_ * i00 + by_ * i10; // ∂vy/∂x
        const F gyy = bx_ * i01 + by_ * i11; // ∂vy/∂y
        F txx, tyy, txy;
        if (DELTA) {
            // τ(x + δ) - τ(x): the visco-elastic part is linear, 2 ηve ε̇'(δ), and
            // exact; the returns of both states add their plastic corrections
            const F two_eve = F(2) * m.etave;
            const F trd = (gxx + gyy) * F(1.0 / 3.0);
            txx = two_eve * (gxx - trd);
            tyy = two_eve * (gyy - trd);
            txy = two_eve * ((gxy + gyx) * F(0.5));
          if (TANGENT && hot_e) {
            // the return's linearisation about x: τ_δ = u - ∂c/∂t u - ∂c/∂P δP, u = 2 ηve ε̇'(δ)
            const float* tg = A.tangent + (int64_t)q * ne + e;
            const int64_t ks = (int64_t)NQ * ne;
            const F h = tg[0], kq = tg[ks], pq = tg[2 * ks], q0 = tg[3 * ks], q1 = tg[4 * ks], q2 = tg[5 * ks];
            const F sdot = q0 * txx + q1 * tyy + q2 * txy;
            const F wq = kq * sdot + pq * Pq;
            const F mxx = F(2) * txx + tyy, myy = txx + F(2) * tyy, mxy = F(2) * txy;
            txx = txx - h * mxx - q0 * wq;
            tyy = tyy - h * myy - q1 * wq;
            txy = txy - h * mxy - q2 * wq;
          } else if (!TANGENT && hot_e) {
            // x's strain rate and the two returns are evaluated in FP64 from the
            // FP64 state, read through the caches in blocks with points near
            // yield only: the stress of x then matches the FP64 residual's
            // (same yield decisions), and the difference of two stresses of
            // full size keeps an error relative to the correction, not to τ
            typedef double H;
            H hxx = 0, hxy = 0, hyx = 0, hyy = 0, Pqx = 0;
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const H d0 = kT.dN[q][a][0], d1 = kT.dN[q][a][1];
                const H ux = A.xvx[n[a]], uy = A.xvy[n[a]];
                const H cx = d0 * ux, cy = d1 * ux, dx = d0 * uy, dy = d1 * uy;
                hxx += cx * (H)i00 + cy * (H)i10; hxy += cx * (H)i01 + cy * (H)i11;
                hyx += dx * (H)i00 + dy * (H)i10; hyy += dx * (H)i01 + dy * (H)i11;
            }
#pragma unroll
            for (int i = 0; i < NP; ++i)
                Pqx += kT.NP1[q][i] * (A.xPf ? A.xP[e * NP + i] - A.xPf[e * NP + i] : A.xP[e * NP + i]);
            IpMat<H> mh;
            mh.etave = m.etave; mh.inv2Gdt = m.inv2Gdt; mh.yield0 = m.yield0; mh.sphi = m.sphi; mh.invden = m.invden;
            const H two_h = H(2) * mh.etave;
            const H trx = (hxx + hyy) * H(1.0 / 3.0);
            const H bxx = two_h * ((hxx - trx) + (H)txo * mh.inv2Gdt);
            const H byy = two_h * ((hyy - trx) + (H)tyo * mh.inv2Gdt);
            const H bxy = two_h * ((hxy + hyx) * H(0.5) + (H)txyo * mh.inv2Gdt);
            H cxx, cyy, cxy, dxx, dyy, dxy;
            const bool yA = dp_correction<H>(bxx + (H)txx, byy + (H)tyy, bxy + (H)txy, mh, Pqx + (H)Pq, cxx, cyy, cxy);
            const bool yB = dp_correction<H>(bxx, byy, bxy, mh, Pqx, dxx, dyy, dxy);
            if (yA || yB) {
                txx = (F)((H)txx + (dxx - cxx));
                tyy = (F)((H)tyy + (dyy - cyy));
                txy = (F)((H)txy + (dxy - cxy));
            }
          }
        } else {
            bool capped = false;
            if constexpr (GEN && PLASTIC && !LIN && std::is_same<F, double>::value) {
                if (A.cap) {
                    // the tensile cap: the visco-elastic trial stress, the cap
                    // parameters and softened cohesion at the point, the local
                    // update (rheology.jl:722-741); the momentum balance uses the
                    // corrected pressure
                    capped = true;
                    F ttxx, ttyy, ttxy;
                    stress2d<F, F, false>(gxx, gyy, (gxy + gyx) * F(0.5), m, txo, tyo, txyo, F(0), ttxx, ttyy, ttxy);
                    auto at = [&](int col) -> double {
                        if (!nodeph) return PT(col, pc);
                        double acc = 0;
#pragma unroll
                        for (int a = 0; a < NV; ++a) acc += kT.N[q][a] * PT(col, pv[a]);
                        return acc;
                    };
                    double C = at(PC_C);
                    const double gam = A.gamma ? A.gamma[e * NQ + q] : 0.0;
                    if (A.gamma) C = fmin(fmax(C + at(PC_HC) * gam, at(PC_CMIN)), C);
                    CapPoint cp;
                    cp.p_tr = Pq - Pfq;
                    cp.eve = m.etave;
                    cp.KDt = at(PC_KB) * A.dt;
                    cp.k = at(PC_SINPHI);
                    cp.kq = at(PC_SINPSI);
                    cp.c = C * at(PC_COSPHI);
                    cp.pT = at(PC_PT);
                    cp.ereg = at(PC_EREG);
                    cp.g = cap_geo(cp.k, cp.kq, cp.c, cp.pT);
                    F Pe;
                    double gd, thd;
                    cap_update<F>(ttxx, ttyy, ttxy, cp, txx, tyy, txy, Pe, &gd, &thd);
                    Pmom = Pe + Pfq;
                    if (PLASTIC && !ITER) yel |= gd != 0;
                    if (A.hist_gamma) {
                        A.hist_gamma[e * NQ + q] = gam + A.dt * gd;
                        A.hist_theta[e * NQ + q] += A.dt * thd;
                    }
                }
            }
            if (!capped) {
                // the yield statistics are taken by residual evaluations outside
                // the inner iteration only (no registers spent on them in EF_ITER)
                constexpr bool STATS = PLASTIC && !ITER;
                F fy = 0;
                float tg[NTAN] = {0, 0, 0, 0, 0, 0};
                const bool y = stress2d<F, F, PLASTIC>(gxx, gyy, (gxy + gyx) * F(0.5), m, txo, tyo, txyo, Pq - Pfq,
                                                       txx, tyy, txy, STATS ? &fy : nullptr,
                                                       STATS && A.tangent_out ? tg : nullptr);
                if (STATS) {
                    yel |= y;
                    near |= fy > F(-0.2) * fabs(m.yield0 + m.sphi * (Pq - Pfq));
                    if (A.tangent_out) {
#pragma unroll
                        for (int k = 0; k < NTAN; ++k) A.tangent_out[((int64_t)k * NQ + q) * ne + e] = tg[k];
                        const float q2 = tg[3] * tg[3] + tg[4] * tg[4] + tg[5] * tg[5];
                        nbad += (y && 1.0f - 3.0f * tg[0] - fmaxf(tg[1], 0.0f) * q2 < 0.05f) ? 1 : 0;
                    }
                }
            }
        }
        if (A.tau_out0) {
            A.tau_out0[e * NQ + q] = txx;
            A.tau_out1[e * NQ + q] = tyy;
            A.tau_out2[e * NQ + q] = txy;
            if (A.P_out) A.P_out[e * NQ + q] = Pmom;
        }
        if (!ITER && !LIN && A.phi_out) {
            // stress power less the elastic power (τ − τ_old) : τ / (2GΔt), the
            // out-of-plane deviator τzz = −(τxx + τyy) included (shear_heating.jl:77-88)
            const F tzz = -(txx + tyy), tzzo = -(txo + tyo);
            const F power = txx * gxx + tyy * gyy + txy * (gxy + gyx);
            const F elastic = (txx * (txx - txo) + tyy * (tyy - tyo) + tzz * (tzz - tzzo) +
                               F(2) * txy * (txy - txyo)) * m.inv2Gdt;
            A.phi_out[e * NQ + q] = power - elastic;
        }
        const F dOm = (F)kT.w[q] * det;
        const F Ptot = Pmom + Pnq;
        const F sxx = txx - Ptot, syy = tyy - Ptot;
        // reference-space stress: R_i[a] = Σ_j ∂N_a/∂ξ_j s_ji
        const F s0x = (i00 * sxx + i01 * txy) * dOm;
        const F s1x = (i10 * sxx + i11 * txy) * dOm;
        const F s0y = (i00 * txy + i01 * syy) * dOm;
        const F s1y = (i10 * txy + i11 * syy) * dOm;
        F bx = 0, by = 0;
        if (grav) {
            if (nodeph) {
                F r0 = 0, al = 0, be = 0;
#pragma unroll
                for (int a = 0; a < NV; ++a) {
                    const F Na = (F)kT.N[q][a];
                    r0 += Na * PT(PC_RHO0, pv[a]); al += Na * PT(PC_ALPHA, pv[a]); be += Na * PT(PC_BETA, pv[a]);
                }
                rho_c = r0; alpha_c = al; beta_c = be;
            }
            F Tq = 0;
            if (GEN && !LIN && A.thermal) {
#pragma unroll
                for (int i = 0; i < NP; ++i) Tq += (F)kT.NP1[q][i] * A.T[e * NP + i];
            }
            // the linearised density is ρ0 β δP (eos_density, momentum_residuals.jl:11-16)
            const F rho = LIN ? rho_c * (beta_c * Pq) : rho_c * ((F(1) - alpha_c * (Tq - A.Tref)) + beta_c * Pq);
            bx = rho * A.gx * dOm;
            by = rho * A.gy * dOm;
            bw[q] = rho * dOm;
        }
        if (F32) {
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const F d0 = (F)kT.dN[q][a][0], d1 = (F)kT.dN[q][a][1];
                Rx[a] = fmaf(d1, s1x, fmaf(d0, s0x, Rx[a]));
                Ry[a] = fmaf(d1, s1y, fmaf(d0, s0y, Ry[a]));
            }
        } else {
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const F d0 = (F)kT.dN[q][a][0], d1 = (F)kT.dN[q][a][1];
                Rx[a] += d0 * s0x + d1 * s1x;
                Ry[a] += d0 * s0y + d1 * s1y;
                if (grav) {
                    const F Na = (F)kT.N[q][a];
                    Rx[a] -= Na * bx;
                    Ry[a] -= Na * by;
                }
            }
        }
    }
    if (F32 && grav) {
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            F m = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) m = fmaf((F)kT.N[q][a], bw[q], m);
            Rx[a] = fmaf(-A.gx, m, Rx[a]);
            Ry[a] = fmaf(-A.gy, m, Ry[a]);
        }
    }

    if (PLASTIC && !ITER && A.yield_count) {
        const uint64_t mask = __ballot(yel);
        if (mask && (threadIdx.x & 63) == __ffsll((unsigned long long)__ballot(1)) - 1)
            atomicAdd(A.yield_count, (int32_t)__popcll(mask));
        A.yield_flag[e] = near;
        if (A.tangent_bad && nbad > 0) atomicAdd(A.tangent_bad, nbad);
    }

    // ---- shared nodes: element-slot residuals for the gather
    typedef typename Vec2<F>::T V2;
    if (TILE) {
        // sum each tile node's slots in LDS in ascending element order (cell
        // rows, cells, triangles); k_tile_update adds the perimeter nodes'
        // partial sums of the neighbouring tiles and updates every node
        __shared__ V2 s_re_own[TILE && !SHARE_TAU ? BLOCK * NSH : 1];
        V2* s_re = SHARE_TAU ? reinterpret_cast<V2*>(s_tau) : s_re_own;
        if (SHARE_TAU) __syncthreads();   // the block's τ_old stage is read out
#pragma unroll
        for (int a = 0; a < NSH; ++a) {
            V2 r; r.x = Rx[a]; r.y = Ry[a];
            s_re[threadIdx.x * NSH + a] = r;
        }
        __syncthreads();
        constexpr int TW = 32, TH = 2 * TCY, NTN = (TW + 1) * (TH + 1), NPER = 2 * (TW + 1) + 2 * (TH - 1);
        for (int k = threadIdx.x; k < NTN; k += BLOCK) {
            const int jx = k % (TW + 1), jy = k / (TW + 1);
            const int lx0 = max((jx - 1) >> 1, 0), lx1 = min(jx >> 1, 15);
            const int ly0 = max((jy - 1) >> 1, 0), ly1 = min(jy >> 1, TCY - 1);
            F rx = 0, ry = 0;
            for (int ly = ly0; ly <= ly1; ++ly)
                for (int lx = lx0; lx <= lx1; ++lx) {
                    // tiles start on even cells (16 columns, an even number of rows)
                    const int par = (lx + ly) & 1, pos = (jx - 2 * lx) + 3 * (jy - 2 * ly);
#pragma unroll
                    for (int tri = 0; tri < 2; ++tri) {
                        const int sl = kT7Slot[par][tri][pos];
                        if (sl >= 0) {
                            const V2 r = s_re[(ly * 32 + lx * 2 + tri) * NSH + sl];
                            rx += r.x;
                            ry += r.y;
                        }
                    }
                }
            if (jx == 0 || jx == TW || jy == 0 || jy == TH) {
                V2 r; r.x = rx; r.y = ry;
                A.bpart[tile * NPER + tile_perim(jx, jy, TW, TH)] = r;
                continue;
            }
            V2 r; r.x = rx; r.y = ry;
            A.Rn[(tty * TH + jy) * (int64_t)A.sx + (ttx * TW + jx)] = r;
        }
    } else {
#pragma unroll
        for (int a = 0; a < NSH; ++a) {
            V2 r; r.x = Rx[a]; r.y = Ry[a];
            A.Re[e * NSH + a] = r;
        }
    }
    // ---- private bubble node: Dirichlet, residual store, Chebyshev update
    {
        const int32_t b = n[NSH];
        const uint8_t fl = A.bcflag[b];
        if (LIN) { Rx[NSH] += A.rhs_vx[b]; Ry[NSH] += A.rhs_vy[b]; }
        if (!LIN && A.bf_x) { Rx[NSH] -= A.bf_x[b]; Ry[NSH] -= A.bf_y[b]; }
        const F rx = (fl & 1) ? F(0) : Rx[NSH];
        const F ry = (fl & 2) ? F(0) : Ry[NSH];
        if (A.Rv_out_x) { A.Rv_out_x[b] = rx; A.Rv_out_y[b] = ry; }
        if (A.Rv0_out_x) { A.Rv0_out_x[b] = rx; A.Rv0_out_y[b] = ry; }
        if (ITER && A.do_update) {
            const F cax = A.cheb ? A.cheb[0] : A.alpha_x, cay = A.cheb ? A.cheb[1] : A.alpha_y;
            const F cbx = A.cheb ? A.cheb[2] : A.beta_x, cby = A.cheb ? A.cheb[3] : A.beta_y;
            F nrx = cbx * A.rate_x[b] + rx / A.PCx[b];
            F nry = cby * A.rate_y[b] + ry / A.PCy[b];
            F nvx = vx[NSH] + (-cax) * nrx;
            F nvy = vy[NSH] + (-cay) * nry;
            if (fl & 1) { nrx = 0; nvx = A.bcval_x[b]; }
            if (fl & 2) { nry = 0; nvy = A.bcval_y[b]; }
            A.rate_x[b] = nrx; A.rate_y[b] = nry;
            A.vx[b] = nvx; A.vy[b] = nvy;
        }
    }
#undef PT
}

template <class F, int FL, int BLOCK>
__global__ __launch_bounds__(BLOCK) FEM_ELEM_ATTR void k_t7_element(ElemArgs<F> A) {
    t7_element_body<F, FL, BLOCK>(A, xcd_tile());
}

// The problem of a grouped launch that owns block b: the launch covers the
// blocks of nprob problems back to back, problem p from boff[p] (ascending,
// boff[0] = 0) to boff[p + 1]; -1 for the padding blocks past boff[nprob].
__device__ inline int group_of(int64_t b, const int32_t* __restrict__ boff, int32_t nprob) {
    if (b >= boff[nprob]) return -1;
    int lo = 0, hi = nprob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (boff[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One launch of the element kernel for many independent problems (batched
// solves): each problem keeps its own arguments (arrays, scalars, Chebyshev
// pointer) in args[p] and its own block range.
template <class F, int FL, int BLOCK>
__global__ __launch_bounds__(BLOCK) FEM_ELEM_ATTR void k_t7_element_group(const ElemArgs<F>* __restrict__ args,
                                                                          const int32_t* __restrict__ boff,
                                                                          int32_t nprob) {
    const int64_t b = xcd_tile();
    const int p = group_of(b, boff, nprob);
    if (p < 0) return;
    t7_element_body<F, FL, BLOCK>(args[p], b - boff[p]);
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
    const F* cheb;              // nullable: the four Chebyshev scalars in device memory instead
    int32_t do_update;
    const F* rhs_vx; const F* rhs_vy;   // nullable: nodal right-hand side added to the gathered sum
    const F* bf_x; const F* bf_y;       // nullable: nodal loads subtracted from the gathered sum
};

template <class F, bool STRUCT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_gather_update(GatherArgs<F> A) {
    (void)STRUCT;
    const int64_t i = xcd_tile() * BLOCK + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    F rx = 0, ry = 0;
    const int32_t k0 = A.off[i], k1 = A.off[i + 1];
    for (int32_t k = k0; k < k1; ++k) {
        const typename Vec2<F>::T r = A.Re[A.ent[k]];
        rx += r.x;
        ry += r.y;
    }
    if (A.bf_x) { rx -= A.bf_x[node]; ry -= A.bf_y[node]; }
    node_update<F>(node, rx, ry, A.rhs_vx, A.rhs_vy, A.bcflag, A.bcval_x, A.bcval_y, A.Rv_out_x, A.Rv_out_y,
                   A.Rv0_out_x, A.Rv0_out_y, A.do_update, A.cheb ? A.cheb[0] : A.alpha_x,
                   A.cheb ? A.cheb[1] : A.alpha_y, A.cheb ? A.cheb[2] : A.beta_x, A.cheb ? A.cheb[3] : A.beta_y,
                   A.rate_x, A.rate_y, A.PCx, A.PCy, A.vx, A.vy);
}

// Grid nodes after the EF_TILE element kernel, one thread per T6 grid node:
// an interior node of a tile reads its summed residual, a node on a tile line
// adds the partial sums of the (two or four) tiles holding it in ascending
// tile order; then the update of k_gather_update.
template <class F>
struct TileUpdArgs {
    GatherArgs<F> g;                  // node data (off, ent, Re, node_of unused)
    int32_t ntx, nty, TH, sx;         // tiles per row / column, grid rows per tile, 2 nx + 1
    int64_t nT6;                      // grid nodes
    const typename Vec2<F>::T* bpart; // [ntiles][perimeter nodes]
    const typename Vec2<F>::T* Rn;    // [nT6], interior nodes of the tiles
};

template <class F, int BLOCK>
__device__ __forceinline__ void tile_update_body(const TileUpdArgs<F>& P, const int64_t blk) {
    const int64_t node = blk * BLOCK + threadIdx.x;
    if (node >= P.nT6) return;
    constexpr int TW = 32;
    const int TH = P.TH, NPER = 2 * (TW + 1) + 2 * (TH - 1);
    const int64_t iy = node / P.sx, ix = node - iy * P.sx;
    F rx = 0, ry = 0;
    if (ix % TW != 0 && iy % TH != 0) {
        const typename Vec2<F>::T r = P.Rn[node];
        rx = r.x;
        ry = r.y;
    } else {
        const int64_t tx1 = ix / TW, ty1 = iy / TH;
        const int64_t tx0 = (ix % TW == 0 && tx1 > 0) ? tx1 - 1 : tx1;
        const int64_t ty0 = (iy % TH == 0 && ty1 > 0) ? ty1 - 1 : ty1;
        for (int64_t ty = ty0; ty <= ty1 && ty < P.nty; ++ty)
            for (int64_t tx = tx0; tx <= tx1 && tx < P.ntx; ++tx) {
                const typename Vec2<F>::T r =
                    P.bpart[(ty * P.ntx + tx) * NPER + tile_perim((int)(ix - tx * TW), (int)(iy - ty * TH), TW, TH)];
                rx += r.x;
                ry += r.y;
            }
    }
    const GatherArgs<F>& A = P.g;
    node_update<F>(node, rx, ry, A.rhs_vx, A.rhs_vy, A.bcflag, A.bcval_x, A.bcval_y, A.Rv_out_x, A.Rv_out_y,
                   A.Rv0_out_x, A.Rv0_out_y, A.do_update, A.cheb ? A.cheb[0] : A.alpha_x,
                   A.cheb ? A.cheb[1] : A.alpha_y, A.cheb ? A.cheb[2] : A.beta_x, A.cheb ? A.cheb[3] : A.beta_y,
                   A.rate_x, A.rate_y, A.PCx, A.PCy, A.vx, A.vy);
}

template <class F, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_tile_update(TileUpdArgs<F> P) {
    tile_update_body<F, BLOCK>(P, xcd_tile());
}

// grouped launch of k_tile_update (see k_t7_element_group)
template <class F, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_tile_update_group(const TileUpdArgs<F>* __restrict__ args,
                                                             const int32_t* __restrict__ boff, int32_t nprob) {
    const int64_t b = xcd_tile();
    const int p = group_of(b, boff, nprob);
    if (p < 0) return;
    tile_update_body<F, BLOCK>(args[p], b - boff[p]);
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
    int32_t cap;                 // PLASTIC: the tensile cap instead of Drucker-Prager
    const F* gamma;              // nullable: accepted plastic strain (cohesion softening)
    const F* eta_ip;             // nullable: viscosity at the points [nels][NQ]; the element
                                 // maximum replaces the phase maximum (momentum_residuals.jl:43)
    F* Ablk;                     // nullable: the whole velocity block, [14][14][nels] (row = c·7 + a)
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
    if (A.eta_ip) {
        etamax = A.eta_ip[e * NQ];
#pragma unroll
        for (int q = 1; q < NQ; ++q) etamax = fmax(etamax, A.eta_ip[e * NQ + q]);
    }
    // tangents at the quadrature points: C[q][t][m] = ∂τ_t/∂ε_m
    F Cq[NQ][3][3];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        F ax = 0, ay = 0, bx = 0, by = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const F d0 = (F)kT.dN[q][a][0], d1 = (F)kT.dN[q][a][1];
            ax += d0 * vx[a]; ay += d1 * vx[a]; bx += d0 * vy[a]; by += d1 * vy[a];
        }
        const F gxx = ax * iJ[0][0] + ay * iJ[1][0], gxy = ax * iJ[0][1] + ay * iJ[1][1];
        const F gyx = bx * iJ[0][0] + by * iJ[1][0], gyy = bx * iJ[0][1] + by * iJ[1][1];
        IpMat<F> m;
        F invG = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) invG += (F)kT.N[q][a] * PT(PC_INVG, pv[a]);
        m.etave = F(1) / (F(1) / etamax + invG / A.dt);
        m.inv2Gdt = invG / (F(2) * A.dt);
        if (PLASTIC) {
            F cphi = 0, C = 0, sphi = 0, spsi = 0, ereg = 0, Kb = 0;
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                const F Na = (F)kT.N[q][a];
                cphi += Na * PT(PC_COSPHI, pv[a]); C += Na * PT(PC_C, pv[a]);
                sphi += Na * PT(PC_SINPHI, pv[a]); spsi += Na * PT(PC_SINPSI, pv[a]);
                ereg += Na * PT(PC_EREG, pv[a]); Kb += Na * PT(PC_KB, pv[a]);
            }
            m.yield0 = cphi * C;
            m.sphi = sphi;
            m.invden = F(1) / ((m.etave + ereg) + ((Kb * A.dt) * (-spsi)) * (-sphi));
        }
        F txo = 0, tyo = 0, txyo = 0;
        if (ELASTIC) { txo = A.tau_old0[e * NQ + q]; tyo = A.tau_old1[e * NQ + q]; txyo = A.tau_old2[e * NQ + q]; }
        F Pq = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i) Pq += (F)kT.NP1[q][i] * P[i];
        if (A.Pf) {
#pragma unroll
            for (int i = 0; i < NP; ++i) Pq -= (F)kT.NP1[q][i] * A.Pf[e * NP + i];
        }
        D3<F> exx(gxx, F(1), F(0), F(0)), eyy(gyy, F(0), F(1), F(0)), exy((gxy + gyx) * F(0.5), F(0), F(0), F(1));
        D3<F> txx, tyy, txy;
        bool capped = false;
        if constexpr (PLASTIC && std::is_same<F, double>::value) {
            if (A.cap) {
                // the cap's tangent carries the corrected pressure, which also
                // depends on the strain rate: C holds ∂(τ − P_local I)/∂ε
                capped = true;
                D3<F> ttxx, ttyy, ttxy;
                stress2d<D3<F>, F, false>(exx, eyy, exy, m, txo, tyo, txyo, F(0), ttxx, ttyy, ttxy);
                auto at = [&](int col) -> double {
                    double acc = 0;
#pragma unroll
                    for (int a = 0; a < NV; ++a) acc += kT.N[q][a] * PT(col, pv[a]);
                    return acc;
                };
                double C = at(PC_C);
                if (A.gamma) C = fmin(fmax(C + at(PC_HC) * A.gamma[e * NQ + q], at(PC_CMIN)), C);
                CapPoint cp;
                cp.p_tr = Pq;
                cp.eve = m.etave;
                cp.KDt = at(PC_KB) * A.dt;
                cp.k = at(PC_SINPHI);
                cp.kq = at(PC_SINPSI);
                cp.c = C * at(PC_COSPHI);
                cp.pT = at(PC_PT);
                cp.ereg = at(PC_EREG);
                cp.g = cap_geo(cp.k, cp.kq, cp.c, cp.pT);
                D3<F> Pe;
                cap_update<D3<F>>(ttxx, ttyy, ttxy, cp, txx, tyy, txy, Pe);
                txx = txx - Pe;
                tyy = tyy - Pe;
            }
        }
        if (!capped) stress2d<D3<F>, F, PLASTIC>(exx, eyy, exy, m, txo, tyo, txyo, Pq, txx, tyy, txy);
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
            const F da0 = (F)kT.dN[q][a][0], da1 = (F)kT.dN[q][a][1];
            const F nax = da0 * iJ[0][0] + da1 * iJ[1][0];
            const F nay = da0 * iJ[0][1] + da1 * iJ[1][1];
            const F dOm = (F)kT.w[q] * det;
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
                const F d0 = (F)kT.dN[q][b][0], d1 = (F)kT.dN[q][b][1];
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
            const F ta0 = (F)kT.TB[j][a][0], ta1 = (F)kT.TB[j][a][1];
            const F fx = (ta0 * iJ[0][0] + ta1 * iJ[1][0]) * det * ps[j];
            const F fy = (ta0 * iJ[0][1] + ta1 * iJ[1][1]) * det * ps[j];
#pragma unroll
            for (int b = 0; b < NV; ++b) {
                const F t0 = (F)kT.TB[j][b][0], t1 = (F)kT.TB[j][b][1];
                const F bx = (t0 * iJ[0][0] + t1 * iJ[1][0]) * det;
                const F by = (t0 * iJ[0][1] + t1 * iJ[1][1]) * det;
                Jx[0][b] += fx * bx; Jx[1][b] += fx * by;
                Jy[0][b] += fy * bx; Jy[1][b] += fy * by;
            }
        }
        if (A.Ablk) {   // stored for the frozen adjoint operator (adjoint_operator.jl:412-436)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int b = 0; b < NV; ++b) {
                    A.Ablk[((int64_t)a * 2 * NV + c * NV + b) * ne + e] = Jx[c][b];
                    A.Ablk[((int64_t)(NV + a) * 2 * NV + c * NV + b) * ne + e] = Jy[c][b];
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
                                                           double dt, F* MP, F* gammaP, F* pscale,
                                                           const double* eta_ip) {
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
            eta += kT.N[q][a] * s_tab[PC_ETA * MAXPH + pv[a]];
            bq += kT.N[q][a] * (1.0 / s_tab[PC_BULK * MAXPH + pv[a]]);
        }
        if (eta_ip) eta = eta_ip[e * NQ + q];   // a viscosity field at the points (pressure_scaling.jl:185)
        const double gnum = gamma_fact * eta;
        const double gphy = bq == 0.0 ? gnum : dt / bq;
        const double gq = gphy * gnum / (gphy + gnum);
        const double dO = kT.w[q] * det;
        for (int i = 0; i < NP; ++i) {
            const double w = kT.NP1[q][i] * dO;
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
constexpr int RED_MAXK = 12;

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
static __global__ __launch_bounds__(RED_BLOCK) void k_reduce_final(int nparts, int K, int op, const double* partial,
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

// sets *flag when any entry is nonzero (a yes/no answer: atomic order is irrelevant)
template <class F>
__global__ void k_any_nonzero(int64_t n, const F* x, int32_t* flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && x[i] != F(0)) atomicOr(flag, 1);
}

// y = (To) (x * scale)
template <class To, class Ti>
__global__ void k_cast(int64_t n, const Ti* x, To* y, double scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (To)(x[i] * scale);
}
// y = (float)(a - b), b nullable
static __global__ void k_cast_sub(int64_t n, const double* a, const double* b, float* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)(b ? a[i] - b[i] : a[i]);
}
// x64 += scale * δ32 (defect-correction update of the FP64 solution)
static __global__ void k_add_correction(int64_t n, double* x, const float* d, double scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += scale * (double)d[i];
}

// folds a two-component correction into the FP64 state and clears it
static __global__ void k_fold2(int64_t n, double* x0, double* x1, float* d0, float* d1) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x0[i] += 1.0 * (double)d0[i];
        x1[i] += 1.0 * (double)d1[i];
        d0[i] = 0.0f;
        d1[i] = 0.0f;
    }
}

// the correction problem's right-hand sides: FP32 copies of R_v(x) (nv) and R_P(x) (np)
static __global__ void k_cast_rhs(int64_t nv, int64_t np, const double* Rx, const double* Ry, const double* RP,
                                  float* rx, float* ry, float* rp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) {
        rx[i] = (float)(Rx[i] * 1.0);
        ry[i] = (float)(Ry[i] * 1.0);
    }
    if (i < np) rp[i] = (float)(RP[i] * 1.0);
}

template <class T>
__global__ void k_fill(int64_t n, T* p, T v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace fem
