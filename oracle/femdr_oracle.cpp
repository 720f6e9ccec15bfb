// femdr_oracle.cpp -- CPU FP64 restatement of FEMTools.jl's matrix-free DR
// Stokes hot path. TEST INFRASTRUCTURE ONLY: it is the checker the HIP
// library is compared against; the product path never links or calls it.
//
// Reference citations are file:line into the FEMTools.jl repository.
// Element contributions are computed independently per element (OpenMP over
// elements) and summed into nodes by a gather over node->element lists in
// ascending element order, which reproduces a sequential scatter in element
// order bit-for-bit regardless of the thread count.
#include "femdr_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kEps = std::numeric_limits<double>::epsilon();

// Julia eps(x): spacing from |x| to the next larger float.
double julia_eps(double x) {
    double a = std::fabs(x);
    if (!std::isfinite(a)) return std::numeric_limits<double>::quiet_NaN();
    return std::nextafter(a, kInf) - a;
}

// ---------------------------------------------------------------------------
// Forward-mode dual numbers (stand-in for ForwardDiff.Dual used by
// src/stokes/assemblers/momentum_residuals.jl:781-787).
// ---------------------------------------------------------------------------
template <int N>
struct Dual {
    double v;
    double d[N];
    Dual() : v(0) { for (int i = 0; i < N; ++i) d[i] = 0; }
    Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0; }
};
template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v / b.v; double ib = 1.0 / (b.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) * ib;
    return r;
}
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { Dual<N> r = a; r.v += b; return r; }
template <int N> Dual<N> operator+(double a, const Dual<N>& b) { Dual<N> r = b; r.v += a; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { Dual<N> r = a; r.v -= b; return r; }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { Dual<N> r = -b; r.v += a; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) { Dual<N> r; r.v = a.v * b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r; }
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b) { Dual<N> r; r.v = a.v / b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b; return r; }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> Dual<N> dsqrt(const Dual<N>& a) { Dual<N> r; r.v = std::sqrt(a.v); double h = 0.5 / r.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h; return r; }
inline double dsqrt(double a) { return std::sqrt(a); }
template <int N> double val(const Dual<N>& a) { return a.v; }
inline double val(double a) { return a; }

// ---------------------------------------------------------------------------
// Reference elements
// ---------------------------------------------------------------------------
// T7 = T6 + cubic bubble, src/elements/shape_functions.jl:115-171.
void t7_basis(double xi, double eta, double N[7], double dN[7][2]) {
    const double L1 = 1 - xi - eta;
    const double b = 27.0 * (1 - xi - eta) * xi * eta;
    const double dbx = 27.0 * eta * (1 - 2 * xi - eta);
    const double dby = 27.0 * xi * (1 - xi - 2 * eta);
    N[0] = L1 * (2 * L1 - 1) + b / 9;
    N[1] = xi * (2 * xi - 1) + b / 9;
    N[2] = eta * (2 * eta - 1) + b / 9;
    N[3] = 4 * (1 - xi - eta) * xi - 4 * b / 9;
    N[4] = 4 * xi * eta - 4 * b / 9;
    N[5] = 4 * eta * (1 - xi - eta) - 4 * b / 9;
    N[6] = b;
    dN[0][0] = 1 - 4 * L1 + dbx / 9;            dN[0][1] = 1 - 4 * L1 + dby / 9;
    dN[1][0] = 4 * xi - 1 + dbx / 9;            dN[1][1] = dby / 9;
    dN[2][0] = dbx / 9;                         dN[2][1] = 4 * eta - 1 + dby / 9;
    dN[3][0] = 4 * (1 - 2 * xi - eta) - 4 * dbx / 9; dN[3][1] = -4 * xi - 4 * dby / 9;
    dN[4][0] = 4 * eta - 4 * dbx / 9;           dN[4][1] = 4 * xi - 4 * dby / 9;
    dN[5][0] = -4 * eta - 4 * dbx / 9;          dN[5][1] = 4 * (1 - xi - 2 * eta) - 4 * dby / 9;
    dN[6][0] = dbx;                             dN[6][1] = dby;
}

// 7-point degree-5 Dunavant rule, src/elements/integration_points.jl:188-202.
void t7_rule(double xi[7], double eta[7], double w[7]) {
    const double a1 = 0.101286507323456, a2 = 0.470142064105115;
    const double b1 = 1 - 2 * a1, b2 = 1 - 2 * a2;
    const double w1 = 0.062969590272414, w2 = 0.066197076394253, wc = 0.1125;
    const double X[7] = {a1, b1, a1, a2, b2, a2, 1.0 / 3};
    const double Y[7] = {a1, a1, b1, a2, a2, b2, 1.0 / 3};
    const double W[7] = {w1, w1, w1, w2, w2, w2, wc};
    for (int q = 0; q < 7; ++q) { xi[q] = X[q]; eta[q] = Y[q]; w[q] = W[q]; }
}

// P1 on the reference triangle, src/elements/shape_functions.jl:231-239.
void p1_basis(double xi, double eta, double N[3]) { N[0] = 1 - xi - eta; N[1] = xi; N[2] = eta; }

struct Tables {
    double Nq[7][7], NqP[7][3], dNdxi[7][7][2], w[7];
    Tables() {
        double xi[7], eta[7];
        t7_rule(xi, eta, w);
        for (int q = 0; q < 7; ++q) {
            t7_basis(xi[q], eta[q], Nq[q], dNdxi[q]);
            p1_basis(xi[q], eta[q], NqP[q]);
        }
    }
};
const Tables& tables() { static Tables t; return t; }

// ---------------------------------------------------------------------------
// Phase helpers, src/utils.jl:37-44 (interp2ip_phase), :70-74 (_gather_phase)
// ---------------------------------------------------------------------------
inline int phase_of(const int32_t* ph, int mode, int64_t e, int a, int nloc, int64_t global) {
    if (mode == ORC_PHASE_CELL) return ph[e] - 1;
    if (mode == ORC_PHASE_ELEMENT) return ph[e * nloc + a] - 1;
    return ph[global] - 1;
}
inline double interp_phase(const double* N, int n, const double* var, const int* ph) {
    double out = 0;
    for (int i = 0; i < n; ++i) out += N[i] * var[ph[i]];
    return out;
}

// ---------------------------------------------------------------------------
// Constitutive update at one quadrature point.
//   viscoelastic_coefficients_phase  src/stokes/assemblers/rheology.jl:35-42
//   deviatoric_stress (2-D)          rheology.jl:226-245
//   _deviatoric_stress_with_multiplier (DP) rheology.jl:272-334
// ---------------------------------------------------------------------------
struct Rheo {
    const orc_material* mat;
    const double* eta;   // per-phase viscosity actually used (mat->eta or element max)
    const orc_dp* dp;
    double dt;
    const orc_cap* cap = nullptr;     // tensile cap in place of dp
    const double* gamma = nullptr;    // the element's accepted plastic strain [nq] (softening)
    const double* eta_ip = nullptr;   // the element's viscosity at the points [nq] (_viscosity_at_q)
    double eta_pt = 0;                // ... at the current point (set per point from eta_ip)
};

template <class S, class SP = double>
double stress_ip(const S gvx[2], const S gvy[2], const double* Nv, int NV, const int* ph,
                 const Rheo& r, const double tau_o[3], SP Pq, S tau[3]) {
    S exx = gvx[0];
    S eyy = gvy[1];
    S exy = (gvx[1] + gvy[0]) / 2.0;
    S tr = (exx + eyy) / 3.0;
    double invG[ORC_MAX_PHASES];
    for (int p = 0; p < r.mat->nphases; ++p) invG[p] = 1.0 / r.mat->G[p];
    // a sampled viscosity is used as given, phase tuples are interpolated (rheology.jl:1-9)
    const double etaq = r.eta_ip ? r.eta_pt : interp_phase(Nv, NV, r.eta, ph);
    const double invGq = interp_phase(Nv, NV, invG, ph);
    const double etave = 1.0 / (1.0 / etaq + invGq / r.dt);
    const double inv2Gdt = invGq / (2 * r.dt);
    S txx = 2 * etave * ((exx - tr) + tau_o[0] * inv2Gdt);
    S tyy = 2 * etave * ((eyy - tr) + tau_o[1] * inv2Gdt);
    S txy = 2 * etave * (exy + tau_o[2] * inv2Gdt);
    double lambda = 0;
    if (r.dp) {
        const orc_dp* dp = r.dp;
        const double cphi = interp_phase(Nv, NV, dp->cos_phi, ph);
        const double sphi = interp_phase(Nv, NV, dp->sin_phi, ph);
        const double spsi = interp_phase(Nv, NV, dp->sin_psi, ph);
        const double C = interp_phase(Nv, NV, dp->C, ph);
        const double ereg = interp_phase(Nv, NV, dp->eta_reg, ph);
        const double Kb = interp_phase(Nv, NV, dp->Kb, ph);
        // second_invariant, rheology.jl:204-212
        S tzz = -txx - tyy;
        S tII = dsqrt((txx * txx + tyy * tyy + tzz * tzz) / 2.0 + txy * txy + kEps * kEps);
        S tIIs = tII + kEps * kEps;
        S F = tII - cphi * C - sphi * Pq;
        const double dFdP = -sphi, dQdP = -spsi;
        S dQxx = (2.0 * txx + tyy) / (2.0 * tIIs);
        S dQyy = (txx + 2.0 * tyy) / (2.0 * tIIs);
        S dQxy = txy / tIIs;
        if (val(F) > 0) {
            S lam = F / (etave + ereg + Kb * r.dt * dQdP * dFdP);
            lambda = val(lam);
            if (lambda > 0) {
                txx = txx - 2 * etave * lam * dQxx;
                tyy = tyy - 2 * etave * lam * dQyy;
                txy = txy - 2 * etave * lam * dQxy;
            }
        }
    }
    tau[0] = txx; tau[1] = tyy; tau[2] = txy;
    return lambda;
}

// ---------------------------------------------------------------------------
// Drucker-Prager tensile cap, src/stokes/assemblers/rheology.jl:510-787
// (Popov, Berlie & Kaus 2025, Eqs. 13-18, 42). Pressure is compression-
// positive; pT <= 0.
// ---------------------------------------------------------------------------
struct CapGeom { double a, b, p_y, R_y, p_d, tau_d, p_q; };

// cap_geometry (rheology.jl:546-557)
CapGeom cap_geometry(double k, double kq, double c, double pT) {
    CapGeom g;
    g.a = std::sqrt(1 + k * k);
    g.b = std::sqrt(1 + kq * kq);
    g.p_y = (g.a * pT + c) / (g.a - k);
    g.R_y = g.p_y - pT;
    g.p_d = g.p_y - g.R_y * k / g.a;
    g.tau_d = g.R_y / g.a;
    g.p_q = g.p_d + kq * g.tau_d;
    return g;
}

template <class S> S sq(const S& x) { return x * x; }

// cap_yield_function (rheology.jl:582-589): shear line or a·(R̂ − R_y)
template <class S>
S cap_yield(const S& s, const S& p, double k, double c, const CapGeom& g) {
    if (val(p) + k * val(s) >= g.p_y) return s - k * p - c;
    const double e = std::numeric_limits<double>::epsilon();
    return g.a * (dsqrt(s * s + sq(p - g.p_y) + e * e) - g.R_y);
}

// cap_invariants (rheology.jl:600-609): F, Aτ = (∂Q/∂s)/2, Ap = −∂Q/∂p
template <class S>
void cap_invariants(const S& s, const S& p, double k, double kq, double c, double pT, S& F, S& At, S& Ap) {
    const CapGeom g = cap_geometry(k, kq, c, pT);
    F = cap_yield(s, p, k, c, g);
    if (val(p) + kq * val(s) >= g.p_q) { At = S(0.5); Ap = S(kq); return; }
    const S Rq = dsqrt(s * s + sq(p - g.p_q));
    if (val(Rq) == 0) { At = S(0.0); Ap = S(0.0); return; }
    At = g.b * s / (2.0 * Rq);
    Ap = -g.b * (p - g.p_q) / Rq;
}

// the local system of cap_return_map and its Jacobian in x = (s, p, λ)
// (rheology.jl:611-615); the partials of F, Aτ, Ap by branch
struct CapSys {
    double s_tr, p_tr, eve, KDt, k, kq, c, pT, ereg;
    void residual(const double x[3], double r[3]) const {
        double F, At, Ap;
        cap_invariants<double>(x[0], x[1], k, kq, c, pT, F, At, Ap);
        r[0] = x[0] - s_tr + 2 * eve * x[2] * At;
        r[1] = x[1] - p_tr - KDt * x[2] * Ap;
        r[2] = F - ereg * x[2];
    }
    void jacobian(const double x[3], double J[3][3]) const {
        typedef Dual<2> D2;
        D2 s(x[0]), p(x[1]);
        s.d[0] = 1; p.d[1] = 1;
        D2 F, At, Ap;
        cap_invariants<D2>(s, p, k, kq, c, pT, F, At, Ap);
        const double lam = x[2];
        J[0][0] = 1 + 2 * eve * lam * At.d[0]; J[0][1] = 2 * eve * lam * At.d[1]; J[0][2] = 2 * eve * At.v;
        J[1][0] = -KDt * lam * Ap.d[0]; J[1][1] = 1 - KDt * lam * Ap.d[1]; J[1][2] = -KDt * Ap.v;
        J[2][0] = F.d[0]; J[2][1] = F.d[1]; J[2][2] = -ereg;
    }
};

// 3x3 solve by Gaussian elimination with partial pivoting; false if singular
bool solve3(double A[3][3], double b[3], double x[3]) {
    int piv[3] = {0, 1, 2};
    for (int col = 0; col < 3; ++col) {
        int best = col;
        for (int r = col + 1; r < 3; ++r)
            if (std::fabs(A[piv[r]][col]) > std::fabs(A[piv[best]][col])) best = r;
        std::swap(piv[col], piv[best]);
        const double d = A[piv[col]][col];
        if (d == 0 || !std::isfinite(d)) return false;
        for (int r = col + 1; r < 3; ++r) {
            const double f = A[piv[r]][col] / d;
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

struct CapRet {
    double s, p, lam, At, Ap;
    bool converged;
    double dx[3];   // ∂(s, p, λ)/∂s_trial at the solution (implicit function)
};

// cap_return_map (rheology.jl:652-690): Newton on the residual scaled by
// max(|s_trial|, |p_trial|, |F_trial|, eps), halving backtracking (24 tries)
// subject to s >= 0, λ >= 0, a finite residual and the Armijo decrease
// ‖r_new‖² <= (1 − α/10⁴)‖r‖², convergence at ‖r‖∞ <= 100 eps, then one full
// step; dx is the derivative that full step carries through dual numbers.
CapRet cap_return(double s_tr, double p_tr, double eve, double KDt, double k, double kq, double c, double pT,
                  double ereg, int maxiter = 40) {
    CapSys sys{s_tr, p_tr, eve, KDt, k, kq, c, pT, ereg};
    CapRet out;
    double F0, At0, Ap0;
    cap_invariants<double>(s_tr, p_tr, k, kq, c, pT, F0, At0, Ap0);
    out.s = s_tr; out.p = p_tr; out.lam = 0; out.At = At0; out.Ap = Ap0;
    out.dx[0] = 1; out.dx[1] = 0; out.dx[2] = 0;
    if (F0 <= 0) { out.converged = true; return out; }
    out.converged = false;
    if (!(std::isfinite(KDt) && KDt > 0 && std::isfinite(eve) && eve > 0)) return out;
    const double e = std::numeric_limits<double>::epsilon();
    const double scale = std::max(std::max(std::fabs(s_tr), std::fabs(p_tr)), std::max(std::fabs(F0), e));
    const double tol = 100 * e;
    double x[3] = {s_tr, p_tr, 0.0}, r[3];
    auto resid = [&](const double* y, double* rr) {
        sys.residual(y, rr);
        for (int i = 0; i < 3; ++i) rr[i] /= scale;
    };
    auto inf_norm = [](const double* v) { return std::max(std::fabs(v[0]), std::max(std::fabs(v[1]), std::fabs(v[2]))); };
    resid(x, r);
    bool conv = inf_norm(r) <= tol;
    for (int it = 0; it < maxiter && !conv; ++it) {
        double J[3][3], rhs[3] = {r[0], r[1], r[2]}, step[3];
        sys.jacobian(x, J);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] /= scale;
        if (!solve3(J, rhs, step)) return out;
        if (!(std::isfinite(step[0]) && std::isfinite(step[1]) && std::isfinite(step[2]))) return out;
        double alpha = 1.0;
        bool accepted = false;
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        for (int ls = 0; ls < 24; ++ls) {
            const double cand[3] = {x[0] - alpha * step[0], x[1] - alpha * step[1], x[2] - alpha * step[2]};
            if (cand[0] >= 0 && cand[2] >= 0) {
                double rn[3];
                resid(cand, rn);
                if (std::isfinite(rn[0]) && std::isfinite(rn[1]) && std::isfinite(rn[2]) &&
                    rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2] <= (1 - alpha / 10000) * r2) {
                    for (int i = 0; i < 3; ++i) { x[i] = cand[i]; r[i] = rn[i]; }
                    accepted = true;
                    break;
                }
            }
            alpha /= 2;
        }
        if (!accepted) return out;
        conv = inf_norm(r) <= tol;
    }
    if (!conv) return out;
    // the final full step (value unchanged to round-off); its derivative with
    // respect to s_trial is J⁻¹ e₁ (∂r/∂s_trial = −e₁)
    double J[3][3], Jc[3][3], rhs[3] = {r[0], r[1], r[2]}, step[3];
    sys.jacobian(x, J);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { J[i][j] /= scale; Jc[i][j] = J[i][j] * scale; }
    if (!solve3(J, rhs, step)) return out;
    for (int i = 0; i < 3; ++i) x[i] -= step[i];
    double e1[3] = {1, 0, 0};
    if (!solve3(Jc, e1, out.dx)) return out;
    double F, At, Ap;
    cap_invariants<double>(x[0], x[1], k, kq, c, pT, F, At, Ap);
    out.s = x[0]; out.p = x[1]; out.lam = x[2]; out.At = At; out.Ap = Ap;
    out.converged = true;
    return out;
}

// second_invariant with the eps floor (rheology.jl:206-212)
template <class S>
S second_invariant2(const S t[3]) {
    const double e = std::numeric_limits<double>::epsilon();
    const S tzz = -t[0] - t[1];
    return dsqrt((t[0] * t[0] + t[1] * t[1] + tzz * tzz) / 2.0 + t[2] * t[2] + e * e);
}

// cap_local_update (rheology.jl:708-717): radial deviatoric return and the
// corrected pressure; S carries the implicit derivative of the return with
// respect to the trial stress (P_trial is a constant of the velocity)
template <class S>
void cap_local_update(const S tt[3], double P_tr, double eve, double KDt, double k, double kq, double c, double pT,
                      double ereg, S tau[3], S& P, S* gdot = nullptr, S* thdot = nullptr) {
    const S s_tr = second_invariant2(tt);
    const CapRet ret = cap_return(val(s_tr), P_tr, eve, KDt, k, kq, c, pT, ereg);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (!ret.converged) {
        for (int i = 0; i < 3; ++i) tau[i] = S(nan);
        P = S(nan);
        if (gdot) *gdot = S(nan);
        if (thdot) *thdot = S(nan);
        return;
    }
    const S ds = s_tr - val(s_tr);   // the derivative part of s_trial
    const S s = ret.s + ret.dx[0] * ds, p = ret.p + ret.dx[1] * ds, lam = ret.lam + ret.dx[2] * ds;
    S F, At, Ap;
    if (ret.lam == 0) { At = S(ret.At); Ap = S(ret.Ap); }
    else cap_invariants<S>(s, p, k, kq, c, pT, F, At, Ap);
    const S g = lam * At, th = lam * Ap;
    for (int i = 0; i < 3; ++i) tau[i] = tt[i] - 2.0 * eve * g * tt[i] / s_tr;
    P = P_tr + KDt * th;
    if (gdot) *gdot = g;
    if (thdot) *thdot = th;
}

// _cap_ip_update (rheology.jl:722-741): VE trial stress, cap parameters and
// the softened cohesion interpolated to the point, the local update
template <class S>
void cap_stress_ip(const S gvx[2], const S gvy[2], const double* Nv, int NV, const int* ph, const Rheo& r,
                   const double tau_o[3], double Peq, const double* gamma, S tau[3], S& P_loc, S* gdot = nullptr,
                   S* thdot = nullptr) {
    Rheo rv = r;
    rv.dp = nullptr;
    S tt[3];
    stress_ip<S>(gvx, gvy, Nv, NV, ph, rv, tau_o, Peq, tt);
    double invG[ORC_MAX_PHASES];
    for (int p = 0; p < r.mat->nphases; ++p) invG[p] = 1.0 / r.mat->G[p];
    const double eve = 1.0 / (1.0 / interp_phase(Nv, NV, r.eta, ph) + interp_phase(Nv, NV, invG, ph) / r.dt);
    const orc_cap* cp = r.cap;
    double C = interp_phase(Nv, NV, cp->C, ph);
    if (gamma) {
        const double Cmin = interp_phase(Nv, NV, cp->C_min, ph), HC = interp_phase(Nv, NV, cp->H_C, ph);
        const double x = C + HC * *gamma;
        C = x > C ? C : (x < Cmin ? Cmin : x);   // clamp(C + H_C γ, C_min, C)
    }
    cap_local_update<S>(tt, Peq, eve, interp_phase(Nv, NV, cp->Kb, ph) * r.dt, interp_phase(Nv, NV, cp->sin_phi, ph),
                        interp_phase(Nv, NV, cp->sin_psi, ph), C * interp_phase(Nv, NV, cp->cos_phi, ph),
                        interp_phase(Nv, NV, cp->pT, ph), interp_phase(Nv, NV, cp->eta_reg, ph), tau, P_loc, gdot,
                        thdot);
}

// Element momentum residual, src/stokes/assemblers/momentum_residuals.jl:136-189
// (2-D with Pnum, T and gravity); accumulate_momentum :28-33; eos_density :11-16.
// Generic in NV/NP/nq; dNdx [nq][NV][2], Nq [nq][NV], NqP [nq][NP]. SP is
// double, or S to differentiate with respect to P (through the total
// pressure, the density and the Drucker-Prager return; the cap's return is
// evaluated at the value of P).
template <class S, class SP = double>
void momentum_element(int NV, int NP, int nq, const double* dNdx, const double* dOm,
                      const double* Nq, const double* NqP, const S* vx, const S* vy,
                      const SP* P, const S* Pnum /*nullable*/, const double* T,
                      const double* Pf /*nullable*/, const int* ph, const Rheo& r,
                      const double* tau_old /*nullable [3][nq]*/, S* Rx, S* Ry,
                      double* tau_out /*nullable [3][nq]*/, double* Pq_out /*nullable [nq]*/) {
    const orc_material* m = r.mat;
    double beta[ORC_MAX_PHASES];
    for (int p = 0; p < m->nphases; ++p) beta[p] = 1.0 / m->K[p];
    for (int a = 0; a < NV; ++a) { Rx[a] = S(0.0); Ry[a] = S(0.0); }
    for (int q = 0; q < nq; ++q) {
        const double* G = dNdx + (size_t)q * NV * 2;
        const double* Nv = Nq + (size_t)q * NV;
        const double* Np = NqP + (size_t)q * NP;
        // strain-rate gradients: grad v = dNdx' * vloc
        S gvx[2] = {S(0.0), S(0.0)}, gvy[2] = {S(0.0), S(0.0)};
        for (int a = 0; a < NV; ++a) {
            gvx[0] = gvx[0] + G[a * 2 + 0] * vx[a];
            gvx[1] = gvx[1] + G[a * 2 + 1] * vx[a];
            gvy[0] = gvy[0] + G[a * 2 + 0] * vy[a];
            gvy[1] = gvy[1] + G[a * 2 + 1] * vy[a];
        }
        double to[3] = {0, 0, 0};
        if (tau_old) { to[0] = tau_old[0 * nq + q]; to[1] = tau_old[1 * nq + q]; to[2] = tau_old[2 * nq + q]; }
        SP Pq = SP(0.0);
        for (int i = 0; i < NP; ++i) Pq = Pq + Np[i] * P[i];
        double Pfq = 0;
        if (Pf) for (int i = 0; i < NP; ++i) Pfq += Np[i] * Pf[i];
        const SP Peq = Pq - Pfq;
        S tau[3];
        S Pe = S(Peq);   // the cap corrects the pressure the momentum balance uses
        Rheo rq = r;
        if (r.eta_ip) rq.eta_pt = r.eta_ip[q];
        if (r.cap) cap_stress_ip<S>(gvx, gvy, Nv, NV, ph, rq, to, val(Peq), r.gamma ? r.gamma + q : nullptr, tau, Pe);
        else stress_ip<S, SP>(gvx, gvy, Nv, NV, ph, rq, to, Peq, tau);
        const S Pq_local = Pe + Pfq;
        if (tau_out) { tau_out[0 * nq + q] = val(tau[0]); tau_out[1 * nq + q] = val(tau[1]); tau_out[2 * nq + q] = val(tau[2]); }
        if (Pq_out) Pq_out[q] = val(Pq_local);
        S Ptot = Pq_local;
        if (Pnum) { S pn = S(0.0); for (int i = 0; i < NP; ++i) pn = pn + Np[i] * Pnum[i]; Ptot = Ptot + pn; }
        double Tq = 0;
        for (int i = 0; i < NP; ++i) Tq += Np[i] * T[i];
        const SP rho = interp_phase(Nv, NV, m->rho0, ph) *
                       (1 - interp_phase(Nv, NV, m->alpha, ph) * (Tq - m->Tref) +
                        interp_phase(Nv, NV, beta, ph) * Pq);
        for (int a = 0; a < NV; ++a) {
            Rx[a] = Rx[a] + ((G[a * 2 + 0] * (tau[0] - Ptot) + G[a * 2 + 1] * tau[2]) - Nv[a] * (rho * m->g[0])) * dOm[q];
            Ry[a] = Ry[a] + ((G[a * 2 + 0] * tau[2] + G[a * 2 + 1] * (tau[1] - Ptot)) - Nv[a] * (rho * m->g[1])) * dOm[q];
        }
    }
}

// Element continuity residual, src/stokes/assemblers/pressure_residual.jl:132-158;
// divergence :18-30.  bulk = etab or K per phase (pressure_residual.jl:1-5).
// SP is double, or S to differentiate with respect to P.
template <class S, class SP = double>
void pressure_element(int NV, int NP, int nq, const double* dNdx, const double* dOmP,
                      const double* NqP, const S* vx, const S* vy, const SP* P,
                      const double* P0, const double* T, const double* T0, const double* Q,
                      const int* phP, const double* alpha, const double* bulk, int nphases,
                      double dt, S* RP, const double* P0q = nullptr /* [nq]: P_old at the points */) {
    bool all_inf = true;
    for (int p = 0; p < nphases; ++p) all_inf = all_inf && std::isinf(bulk[p]);
    for (int i = 0; i < NP; ++i) RP[i] = S(0.0);
    for (int q = 0; q < nq; ++q) {
        const double* G = dNdx + (size_t)q * NV * 2;
        const double* Nv = NqP + (size_t)q * NP;
        double Qq = 0;
        if (Q) for (int i = 0; i < NP; ++i) Qq += Nv[i] * Q[i];
        const double etabq = all_inf ? kInf : interp_phase(Nv, NP, bulk, phP);
        const double aq = interp_phase(Nv, NP, alpha, phP);
        SP dP = SP(0.0);
        double dT = 0;
        if (P0q) {   // pressure_increment_at_ip with IntegrationPointPressure (pressure_residual.jl:90-92)
            for (int i = 0; i < NP; ++i) dP = dP + Nv[i] * P[i];
            dP = dP - P0q[q];
        } else {
            for (int i = 0; i < NP; ++i) dP = dP + Nv[i] * (P[i] - P0[i]);
        }
        for (int i = 0; i < NP; ++i) dT += Nv[i] * (T[i] - T0[i]);
        const SP dPdt = dP / (etabq * dt);
        const double dTdt = aq * dT / dt;
        S div = S(0.0);
        for (int a = 0; a < NV; ++a) div = div + G[a * 2 + 0] * vx[a];
        for (int a = 0; a < NV; ++a) div = div + G[a * 2 + 1] * vy[a];
        for (int i = 0; i < NP; ++i) RP[i] = RP[i] + Nv[i] * (((-div) - dPdt) + dTdt + Qq) * dOmP[q];
    }
}

// ---------------------------------------------------------------------------
// Mesh-level helpers
// ---------------------------------------------------------------------------
struct Csr {
    const int32_t* key = nullptr;
    int64_t nels = 0, nnodes = 0;
    std::vector<int64_t> off;
    std::vector<int32_t> ent;  // e*7 + a, ascending e per node (connectivity.jl:341-349)
};
Csr& node2elem(const orc_stokes* s) {
    static thread_local Csr c;
    if (c.key == s->el2n && c.nels == s->nels && c.nnodes == s->nnodes) return c;
    c.key = s->el2n; c.nels = s->nels; c.nnodes = s->nnodes;
    c.off.assign(s->nnodes + 1, 0);
    for (int64_t e = 0; e < s->nels; ++e)
        for (int a = 0; a < 7; ++a) c.off[s->el2n[e * 7 + a] + 1]++;
    for (int64_t n = 0; n < s->nnodes; ++n) c.off[n + 1] += c.off[n];
    c.ent.assign(c.off[s->nnodes], 0);
    std::vector<int64_t> fill(c.off.begin(), c.off.end() - 1);
    for (int64_t e = 0; e < s->nels; ++e)
        for (int a = 0; a < 7; ++a) c.ent[fill[s->el2n[e * 7 + a]]++] = (int32_t)(e * 7 + a);
    return c;
}

void gather(const orc_stokes* s, const std::vector<double>& Re, double* out) {
    const Csr& c = node2elem(s);
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < s->nnodes; ++n) {
        double acc = 0;
        for (int64_t k = c.off[n]; k < c.off[n + 1]; ++k) acc += Re[c.ent[k]];
        out[n] = acc;
    }
}

void elem_dNdx(const orc_stokes* s, int64_t e, double dNdx[7][7][2]) {
    const Tables& t = tables();
    for (int q = 0; q < 7; ++q) {
        const double* iJ = s->invJ + (e * 7 + q) * 4;
        for (int a = 0; a < 7; ++a) {
            dNdx[q][a][0] = t.dNdxi[q][a][0] * iJ[0] + t.dNdxi[q][a][1] * iJ[2];
            dNdx[q][a][1] = t.dNdxi[q][a][0] * iJ[1] + t.dNdxi[q][a][1] * iJ[3];
        }
    }
}

void elem_phases_v(const orc_stokes* s, int64_t e, int ph[7]) {
    for (int a = 0; a < 7; ++a) ph[a] = phase_of(s->phases_v, s->phase_mode_v, e, a, 7, s->el2n[e * 7 + a]);
}
void elem_phases_P(const orc_stokes* s, int64_t e, int ph[3]) {
    for (int a = 0; a < 3; ++a) ph[a] = phase_of(s->phases_P, s->phase_mode_P, e, a, 3, e * 3 + a);
}

const double* pressure_bulk(const orc_stokes* s, int finite_K) { return finite_K ? s->mat.K : s->mat.etab; }

void dirichlet(double* v, int64_t n, const int32_t* dofs, const double* vals) {
    for (int64_t i = 0; i < n; ++i) v[dofs[i]] = vals ? vals[i] : 0.0;
}

double nrm2(const double* x, int64_t n) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    return std::sqrt(s);
}

// _checked_λmax, src/utils.jl:152-158
bool checked_lmax(const double* J, const double* PC, int64_t n, const char* label, double* out) {
    double m = -kInf;
    bool nan = false;
    for (int64_t i = 0; i < n; ++i) {
        double r = J[i] / PC[i];
        if (std::isnan(r)) nan = true;
        m = std::max(m, r);
    }
    if (nan) m = std::numeric_limits<double>::quiet_NaN();
    *out = m;
    if (!(std::isfinite(m) && m > 0)) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s preconditioner produced invalid λmax = %g; check for zero or non-finite preconditioner entries", label, m);
        g_err = buf;
        return false;
    }
    return true;
}

// _stokes_λmin, src/stokes/solvers/DR.jl:32-37
double stokes_lmin(double step, const double* rate, const double* R, const double* R0, const double* PC, int64_t n) {
    double den = 0;
    for (int64_t i = 0; i < n; ++i) { double t = step * rate[i]; den += t * t; }
    if (den == 0) return 0;
    double num = 0;
    for (int64_t i = 0; i < n; ++i) num += (step * rate[i]) * ((R[i] - R0[i]) / PC[i]);
    return std::fabs(num) / den;
}

// _stokes_cheb, DR.jl:39-42
void stokes_cheb(double dtau, double lmin, double c_fact, double* a, double* b) {
    double c = 2 * std::sqrt(lmin) * c_fact;
    *a = 2 * dtau * dtau / (2 + c * dtau);
    *b = (2 - c * dtau) / (2 + c * dtau);
}

extern "C" {
// the two scalar helpers above, exported for their closed-form checks
double orc_stokes_lmin(double step, const double* rate, const double* R, const double* R0, const double* PC,
                       int64_t n) {
    return stokes_lmin(step, rate, R, R0, PC, n);
}
void orc_stokes_cheb(double dtau, double lmin, double c_fact, double* a, double* b) {
    stokes_cheb(dtau, lmin, c_fact, a, b);
}
}  // extern "C"

// The thermal side of a coupled solve (DR.jl:514-529, 604-661): the thermal
// state, its Chebyshev scalars and convergence, the shear-heating source.
struct ThermalHook {
    orc_thermal* t = nullptr;
    orc_stokes* s = nullptr;
    orc_stokes3* s3 = nullptr;       // a 3-D coupled solve (s is null)
    double dt = 1, Tref = 273, eps = 1e-6;
    int64_t nout = 1;
    bool shear = false;
    std::vector<double> phi;
    int64_t iter = 0;
    double alpha = 0, beta = 0, lmax = 0, nr0 = 0, err_T = kInf;
    bool converged = false;
};
void thermal_hook_begin(ThermalHook& h);
int thermal_hook_step(ThermalHook& h);

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_t7_basis(double xi, double eta, double* N, double* dN) {
    double n[7], d[7][2];
    t7_basis(xi, eta, n, d);
    for (int a = 0; a < 7; ++a) { N[a] = n[a]; dN[2 * a] = d[a][0]; dN[2 * a + 1] = d[a][1]; }
}
void orc_t7_rule(double* xi, double* eta, double* w) { t7_rule(xi, eta, w); }
void orc_t7_tables(double* Nq, double* NqP, double* dNdxi, double* w) {
    const Tables& t = tables();
    std::memcpy(Nq, t.Nq, sizeof t.Nq);
    std::memcpy(NqP, t.NqP, sizeof t.NqP);
    std::memcpy(dNdxi, t.dNdxi, sizeof t.dNdxi);
    std::memcpy(w, t.w, sizeof t.w);
}

void orc_t7_mesh_sizes(int32_t nx, int32_t ny, int64_t* nnodes, int64_t* nels) {
    *nnodes = (int64_t)(2 * nx + 1) * (2 * ny + 1) + 2LL * nx * ny;
    *nels = 2LL * nx * ny;
}

// generate_coordinates(T7): src/mesh/mesh.jl:412-440 (T6 grid from
// _tensor_grid_coordinates :352-361, LinRange interpolation, then centroids).
// generate_element2node(T7): src/mesh/connectivity.jl:111-140.
void orc_t7_mesh(int32_t nx, int32_t ny, double x0, double x1, double y0, double y1,
                 double* coords, int32_t* el2n) {
    const int64_t sx = 2 * nx + 1, sy = 2 * ny + 1;
    auto lerp = [](int64_t j, int64_t d, double a, double b) {
        double t = (double)j / (double)d;
        return (1 - t) * a + t * b;
    };
    for (int64_t iy = 0; iy < sy; ++iy)
        for (int64_t ix = 0; ix < sx; ++ix) {
            coords[(iy * sx + ix) * 2 + 0] = lerp(ix, sx - 1, x0, x1);
            coords[(iy * sx + ix) * 2 + 1] = lerp(iy, sy - 1, y0, y1);
        }
    const double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
    auto cent = [&](int64_t ex, int64_t ey, int a, int b, double* c) {
        c[0] = x0 + (3 * ex + a) * dx / 3;
        c[1] = y0 + (3 * ey + b) * dy / 3;
    };
    const int64_t nT6 = sx * sy;
    int64_t k = nT6, iel = 0;
    auto node = [&](int64_t ix, int64_t iy) { return (int32_t)(iy * sx + ix); };
    for (int64_t ey = 0; ey < ny; ++ey)
        for (int64_t ex = 0; ex < nx; ++ex) {
            const int64_t ix = 2 * ex, iy = 2 * ey;
            const int32_t BL = node(ix, iy), BR = node(ix + 2, iy), TR = node(ix + 2, iy + 2), TL = node(ix, iy + 2);
            const int32_t bot = node(ix + 1, iy), rgt = node(ix + 2, iy + 1), top = node(ix + 1, iy + 2), lft = node(ix, iy + 1);
            const int32_t ctr = node(ix + 1, iy + 1);
            const int32_t cA = (int32_t)(nT6 + iel), cB = (int32_t)(nT6 + iel + 1);
            int32_t* A = el2n + iel * 7;
            int32_t* B = el2n + (iel + 1) * 7;
            if ((ex + ey) % 2 == 0) {
                const int32_t a[7] = {BL, BR, TR, bot, rgt, ctr, cA};
                const int32_t b[7] = {BL, TR, TL, ctr, top, lft, cB};
                std::memcpy(A, a, sizeof a); std::memcpy(B, b, sizeof b);
                cent(ex, ey, 2, 1, coords + (k++) * 2);
                cent(ex, ey, 1, 2, coords + (k++) * 2);
            } else {
                const int32_t a[7] = {BL, BR, TL, bot, ctr, lft, cA};
                const int32_t b[7] = {BR, TR, TL, rgt, top, ctr, cB};
                std::memcpy(A, a, sizeof a); std::memcpy(B, b, sizeof b);
                cent(ex, ey, 1, 1, coords + (k++) * 2);
                cent(ex, ey, 2, 2, coords + (k++) * 2);
            }
            iel += 2;
        }
}

int64_t orc_box_boundary_nodes(int64_t nnodes, const double* coords, double x0, double x1,
                               double y0, double y1, int32_t* nodes) {
    int64_t k = 0;
    const double tol = 32 * kEps * std::max({std::fabs(x0), std::fabs(x1), std::fabs(y0), std::fabs(y1), 1.0});
    for (int64_t n = 0; n < nnodes; ++n) {
        double x = coords[2 * n], y = coords[2 * n + 1];
        if (std::fabs(x - x0) <= tol || std::fabs(x - x1) <= tol || std::fabs(y - y0) <= tol || std::fabs(y - y1) <= tol) {
            if (nodes) nodes[k] = (int32_t)n;
            ++k;
        }
    }
    return k;
}

// precompute_geometry_kernel!: src/mesh/mesh.jl:674-682 with _geometry_entry
// :550-553. Velocity geometry uses all 7 nodes and T7 gradients; pressure
// geometry uses the 3 corner nodes and P1 gradients (mixed_mesh.jl:297-328).
void orc_t7_geometry(int64_t nels, const double* coords, const int32_t* el2n,
                     double* invJ, double* dOmega, double* dOmegaP) {
    const Tables& t = tables();
    const double dP1[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nels; ++e) {
        for (int q = 0; q < 7; ++q) {
            double J[2][2] = {{0, 0}, {0, 0}};
            for (int a = 0; a < 7; ++a) {
                const double* c = coords + 2 * (int64_t)el2n[e * 7 + a];
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) J[i][j] += c[i] * t.dNdxi[q][a][j];
            }
            const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            double* iJ = invJ + (e * 7 + q) * 4;
            iJ[0] = J[1][1] / det; iJ[1] = -J[0][1] / det;
            iJ[2] = -J[1][0] / det; iJ[3] = J[0][0] / det;
            dOmega[e * 7 + q] = std::fabs(det) * t.w[q];
            double JP[2][2] = {{0, 0}, {0, 0}};
            for (int a = 0; a < 3; ++a) {
                const double* c = coords + 2 * (int64_t)el2n[e * 7 + a];
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) JP[i][j] += c[i] * dP1[a][j];
            }
            const double detP = JP[0][0] * JP[1][1] - JP[0][1] * JP[1][0];
            dOmegaP[e * 7 + q] = std::fabs(detP) * t.w[q];
        }
    }
}

void orc_pressure_element(int32_t NV, int32_t NP, int32_t nq, const double* dNdx, const double* dOm,
                          const double* dOmP, const double* NqP, const double* vx, const double* vy,
                          const double* P, const double* P0, const double* T, const double* T0,
                          const double* Q, const int32_t* phaseP, const double* alpha,
                          const double* bulk, double dt, double* RP_out) {
    (void)dOm;
    int ph[64];
    int np = 0;
    for (int i = 0; i < NP; ++i) { ph[i] = phaseP[i] - 1; np = std::max(np, ph[i] + 1); }
    pressure_element<double>(NV, NP, nq, dNdx, dOmP, NqP, vx, vy, P, P0, T, T0, Q, ph, alpha, bulk, np, dt, RP_out);
}

void orc_momentum_element(int32_t NV, int32_t NP, int32_t nq, const double* dNdx, const double* dOm,
                          const double* Nq, const double* NqP, const double* vx, const double* vy,
                          const double* P, const double* Pnum, const double* T, const int32_t* phase,
                          const orc_material* mat, const double* eta_override, const orc_dp* dp,
                          double dt, const double* tau_old, double* Rx, double* Ry, double* tau_out) {
    int ph[64];
    for (int i = 0; i < NV; ++i) ph[i] = phase[i] - 1;
    Rheo r{mat, eta_override ? eta_override : mat->eta, dp, dt};
    momentum_element<double>(NV, NP, nq, dNdx, dOm, Nq, NqP, vx, vy, P, Pnum, T, nullptr, ph, r,
                             tau_old, Rx, Ry, tau_out, nullptr);
}

double orc_stress_at_ip(const double* gradvx, const double* gradvy, const double* Nv, int32_t NV,
                        const int32_t* phase, const orc_material* mat, const orc_dp* dp, double dt,
                        const double* tau_old, double Pq, double* tau) {
    int ph[64];
    for (int i = 0; i < NV; ++i) ph[i] = phase[i] - 1;
    Rheo r{mat, mat->eta, dp, dt};
    return stress_ip<double>(gradvx, gradvy, Nv, NV, ph, r, tau_old, Pq, tau);
}

// assemble_viscosity_weighted_pressure_scaling!: pressure_scaling.jl:118-149,
// kernel :166-195, pressure_scale_at_ip :205-211, finalize :213-216.
void orc_pressure_scaling(orc_stokes* s, double gamma_fact, double dt, int32_t finite_K) {
    const Tables& t = tables();
    const double* bulk = pressure_bulk(s, finite_K);
    double invK[ORC_MAX_PHASES];
    for (int p = 0; p < s->mat.nphases; ++p) invK[p] = 1.0 / bulk[p];
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        int ph[7];
        elem_phases_v(s, e, ph);
        double MP[3] = {0, 0, 0}, gP[3] = {0, 0, 0};
        for (int q = 0; q < 7; ++q) {
            const double dO = s->dOmegaP[e * 7 + q];
            const double* eip = s->eta_scale_ip ? s->eta_scale_ip : s->eta_ip;
            const double etaq = eip ? eip[e * 7 + q] : interp_phase(t.Nq[q], 7, s->mat.eta, ph);
            const double gnum = gamma_fact * etaq;
            const double bq = interp_phase(t.Nq[q], 7, invK, ph);
            const double gphy = bq == 0 ? gnum : dt / bq;
            const double gq = gphy * gnum / (gphy + gnum);
            for (int a = 0; a < 3; ++a) {
                const double w = t.NqP[q][a] * dO;
                MP[a] += w;
                gP[a] += w * gq;
            }
        }
        for (int a = 0; a < 3; ++a) {
            s->MP[e * 3 + a] = MP[a];
            s->gammaP[e * 3 + a] = MP[a] == 0 ? 0.0 : gP[a] / MP[a];
        }
    }
}

// assemble_pressure_residual_kernel!: pressure_residual.jl:256-326.
void orc_pressure_residual(orc_stokes* s, double dt, int32_t finite_K, const double* P_old) {
    const Tables& t = tables();
    const double* bulk = pressure_bulk(s, finite_K);
    if (!P_old) P_old = s->P0;
    const double* Pq_old = P_old == s->P0 ? s->Pold_ip : nullptr;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        double dNdx[7][7][2];
        elem_dNdx(s, e, dNdx);
        double vx[7], vy[7];
        for (int a = 0; a < 7; ++a) { vx[a] = s->vx[s->el2n[e * 7 + a]]; vy[a] = s->vy[s->el2n[e * 7 + a]]; }
        int ph[3];
        elem_phases_P(s, e, ph);
        double RPe[3];
        pressure_element<double>(7, 3, 7, &dNdx[0][0][0], s->dOmegaP + e * 7, &t.NqP[0][0], vx, vy,
                                 s->P + e * 3, P_old + e * 3, s->T + e * 3, s->T0 + e * 3,
                                 s->Q ? s->Q + e * 3 : nullptr, ph, s->mat.alpha, bulk, s->mat.nphases, dt, RPe,
                                 Pq_old ? Pq_old + e * 7 : nullptr);
        for (int a = 0; a < 3; ++a) s->RP[e * 3 + a] = RPe[a];
    }
}

// assemble_momentum_residual_kernel!: momentum_residuals.jl:417-470, 497-528.
static void momentum_residual_impl(orc_stokes* s, double dt, const double* Pnum, double* tau_store[3]) {
    const Tables& t = tables();
    std::vector<double> Rex((size_t)s->nels * 7), Rey((size_t)s->nels * 7);
    Rheo r0{&s->mat, s->mat.eta, s->cap ? nullptr : s->dp, dt};
    r0.cap = s->cap;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        Rheo r = r0;
        if (s->cap && s->gamma_hist) r.gamma = s->gamma_hist + e * 7;
        if (s->eta_ip) r.eta_ip = s->eta_ip + e * 7;
        double dNdx[7][7][2];
        elem_dNdx(s, e, dNdx);
        double vx[7], vy[7];
        for (int a = 0; a < 7; ++a) { vx[a] = s->vx[s->el2n[e * 7 + a]]; vy[a] = s->vy[s->el2n[e * 7 + a]]; }
        int ph[7];
        elem_phases_v(s, e, ph);
        double to[21];
        bool has_old = s->tau_old[0] || s->tau_old[1] || s->tau_old[2];
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < 7; ++q) to[c * 7 + q] = s->tau_old[c] ? s->tau_old[c][e * 7 + q] : 0.0;
        double tout[21];
        momentum_element<double>(7, 3, 7, &dNdx[0][0][0], s->dOmega + e * 7, &t.Nq[0][0], &t.NqP[0][0], vx, vy,
                                 s->P + e * 3, Pnum ? Pnum + e * 3 : nullptr, s->T + e * 3,
                                 s->Pf ? s->Pf + e * 3 : nullptr, ph, r, has_old ? to : nullptr,
                                 &Rex[e * 7], &Rey[e * 7], tau_store ? tout : nullptr,
                                 tau_store && s->P_store ? s->P_store + e * 7 : nullptr);
        if (tau_store)
            for (int c = 0; c < 3; ++c)
                if (tau_store[c])
                    for (int q = 0; q < 7; ++q) tau_store[c][e * 7 + q] = tout[c * 7 + q];
    }
    gather(s, Rex, s->Rvx);
    gather(s, Rey, s->Rvy);
}

void orc_momentum_residual(orc_stokes* s, double dt, const double* Pnum) {
    momentum_residual_impl(s, dt, Pnum, nullptr);
    // _subtract_body_force! (DR.jl:1-7), before the Dirichlet rows (DR.jl:553, 686)
    double* R[2] = {s->Rvx, s->Rvy};
    for (int c = 0; c < 2; ++c)
        if (s->body_force[c])
            for (int64_t n = 0; n < s->nnodes; ++n) R[c][n] -= s->body_force[c][n];
}

// _set_pressure_source! and its area kernel (pressure_residual.jl:32-77)
int orc_set_pressure_source(orc_stokes* s, const double* Qq, double Q2D, int32_t has_Q2D) {
    const Tables& t = tables();
    if (!has_Q2D) {
        for (int64_t i = 0; i < s->nnodesP; ++i)
            if (!std::isfinite(Qq[i])) { g_err = "Qq must be finite"; return -1; }
        std::copy(Qq, Qq + s->nnodesP, s->Q);
        return 0;
    }
    if (!std::isfinite(Q2D)) { g_err = "Q2D must be finite"; return -1; }
    double area = 0;
    int64_t invalid = 0;
    for (int64_t e = 0; e < s->nels; ++e) {
        const double* Ql = Qq + e * 3;
        for (int i = 0; i < 3; ++i) invalid += !std::isfinite(Ql[i]) || Ql[i] < 0;
        for (int q = 0; q < 7; ++q) {
            double qv = 0;
            for (int i = 0; i < 3; ++i) qv += t.NqP[q][i] * Ql[i];
            invalid += !std::isfinite(qv) || qv < 0;
            area += qv * s->dOmegaP[e * 7 + q];
        }
    }
    if (invalid) { g_err = "Qq must be finite and nonnegative"; return -1; }
    if (!(std::isfinite(area) && area > 0)) {
        if (Q2D != 0) { g_err = "nonzero Q2D requires nonempty positive Qq support"; return -1; }
        std::fill(s->Q, s->Q + s->nnodesP, 0.0);
        return 0;
    }
    for (int64_t i = 0; i < s->nnodesP; ++i) s->Q[i] = Qq[i] * (Q2D / area);
    return 0;
}

// update_stokes_current_stress! (DR.jl:820-897): momentum assembly with
// stress storage, Pnum = nothing; Rv is overwritten as in the reference.
void orc_update_current_stress(orc_stokes* s, double dt) {
    momentum_residual_impl(s, dt, nullptr, s->tau_store);
}

}  // extern "C"

// One element of the augmented Jacobian: block[i][j][row][col] = ∂R_i[row]/∂v_j[col]
// with Pnum(v) = γP RP(v)/M_P inline (momentum_residuals.jl:929-970) and the
// element-maximum viscosity for every phase (momentum_residuals.jl:36-44).
static void augmented_element_block(const orc_stokes* s, int64_t e, double dt, const double* bulk,
                                    double block[2][2][7][7]) {
    const Tables& t = tables();
    typedef Dual<7> D7;
    double dNdx[7][7][2];
    elem_dNdx(s, e, dNdx);
    double vx[7], vy[7];
    for (int a = 0; a < 7; ++a) { vx[a] = s->vx[s->el2n[e * 7 + a]]; vy[a] = s->vy[s->el2n[e * 7 + a]]; }
    int ph[7], phP[3];
    elem_phases_v(s, e, ph);
    elem_phases_P(s, e, phP);
    // _element_max_phase_property: every phase gets the element max.
    double emax = s->mat.eta[ph[0]];
    for (int a = 1; a < 7; ++a) emax = std::max(emax, s->mat.eta[ph[a]]);
    if (s->eta_ip) {   // a sampled field: maximum over the element's points (momentum_residuals.jl:43)
        emax = s->eta_ip[e * 7];
        for (int q = 1; q < 7; ++q) emax = std::max(emax, s->eta_ip[e * 7 + q]);
    }
    double eta_pc[ORC_MAX_PHASES];
    for (int p = 0; p < s->mat.nphases; ++p) eta_pc[p] = emax;
    Rheo r{&s->mat, eta_pc, s->cap ? nullptr : s->dp, dt};
    r.cap = s->cap;
    if (s->cap && s->gamma_hist) r.gamma = s->gamma_hist + e * 7;
    double to[21];
    bool has_old = s->tau_old[0] || s->tau_old[1] || s->tau_old[2];
    for (int c = 0; c < 3; ++c)
        for (int q = 0; q < 7; ++q) to[c * 7 + q] = s->tau_old[c] ? s->tau_old[c][e * 7 + q] : 0.0;
    for (int j = 0; j < 2; ++j) {
        D7 ux[7], uy[7];
        for (int a = 0; a < 7; ++a) {
            ux[a] = D7(vx[a]); uy[a] = D7(vy[a]);
            if (j == 0) ux[a].d[a] = 1.0; else uy[a].d[a] = 1.0;
        }
        D7 RPe[3];
        pressure_element<D7>(7, 3, 7, &dNdx[0][0][0], s->dOmegaP + e * 7, &t.NqP[0][0], ux, uy,
                             s->P + e * 3, s->P0 + e * 3, s->T + e * 3, s->T0 + e * 3,
                             s->Q ? s->Q + e * 3 : nullptr, phP, s->mat.alpha, bulk, s->mat.nphases, dt, RPe);
        D7 Pn[3];
        for (int a = 0; a < 3; ++a) Pn[a] = s->gammaP[e * 3 + a] * RPe[a] / s->MP[e * 3 + a];
        D7 Rx[7], Ry[7];
        momentum_element<D7>(7, 3, 7, &dNdx[0][0][0], s->dOmega + e * 7, &t.Nq[0][0], &t.NqP[0][0], ux, uy,
                             s->P + e * 3, Pn, s->T + e * 3, s->Pf ? s->Pf + e * 3 : nullptr, ph, r,
                             has_old ? to : nullptr, Rx, Ry, nullptr, nullptr);
        for (int row = 0; row < 7; ++row)
            for (int col = 0; col < 7; ++col) {
                block[0][j][row][col] = Rx[row].d[col];
                block[1][j][row][col] = Ry[row].d[col];
            }
    }
}

extern "C" {

// assemble_augmented_momentum_jacobian_matrices_atomix!: momentum_residuals.jl
// :998-1033, kernel :1044-1066, problem :929-970, folding :1073-1117; the
// element viscosity is the element maximum (:36-44). Each (i, j) block is the
// forward-mode derivative of R_i with respect to v_j with Pnum(v) inline.
void orc_augmented_jacobian(orc_stokes* s, double dt, int32_t finite_K) {
    const double* bulk = pressure_bulk(s, finite_K);
    std::vector<double> Jx((size_t)s->nels * 7), Jy((size_t)s->nels * 7), Dx((size_t)s->nels * 7), Dy((size_t)s->nels * 7);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        double block[2][2][7][7];  // [i][j][row][col] = dR_i[row]/dv_j[col]
        augmented_element_block(s, e, dt, bulk, block);
        for (int row = 0; row < 7; ++row) {
            double sx = 0, sy = 0;
            for (int j = 0; j < 2; ++j) {
                double bx = 0, by = 0;
                for (int col = 0; col < 7; ++col) { bx += std::fabs(block[0][j][row][col]); by += std::fabs(block[1][j][row][col]); }
                sx += bx; sy += by;
            }
            Jx[e * 7 + row] = sx; Jy[e * 7 + row] = sy;
            Dx[e * 7 + row] = std::fabs(block[0][0][row][row]);
            Dy[e * 7 + row] = std::fabs(block[1][1][row][row]);
        }
    }
    gather(s, Jx, s->dRdvx);
    gather(s, Jy, s->dRdvy);
    gather(s, Dx, s->PCx);
    gather(s, Dy, s->PCy);
}

// update_stokes_velocity_kernel!: src/stokes/helpers.jl:31-41 with the
// negated step of DR.jl:690-695 (alpha here is the positive Chebyshev step).
void orc_dr_update(orc_stokes* s, const double* alpha, const double* beta) {
    double* rate[2] = {s->rate_x, s->rate_y};
    double* v[2] = {s->vx, s->vy};
    const double* R[2] = {s->Rvx, s->Rvy};
    const double* PC[2] = {s->PCx, s->PCy};
    for (int c = 0; c < 2; ++c) {
        const double a = -alpha[c], b = beta[c];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < s->nnodes; ++i) {
            double nr = b * rate[c][i] + R[c][i] / PC[c][i];
            rate[c][i] = nr;
            v[c][i] += a * nr;
        }
    }
}

// k bare inner DR iterations with fixed Chebyshev coefficients (DR.jl:665-697):
// pressure residual, Pnum, momentum residual, Dirichlet, update. Used to time
// the CPU side of the hot path.
void orc_dr_iterations(orc_stokes* s, double dt, int32_t finite_K, const double* alpha, const double* beta,
                       int64_t k) {
    double* rate[2] = {s->rate_x, s->rate_y};
    double* Rv[2] = {s->Rvx, s->Rvy};
    double* v[2] = {s->vx, s->vy};
    for (int64_t it = 0; it < k; ++it) {
        orc_pressure_residual(s, dt, finite_K, s->P0);
        for (int64_t i = 0; i < s->nnodesP; ++i) s->Pnum[i] = s->gammaP[i] * s->RP[i] / s->MP[i];
        orc_momentum_residual(s, dt, s->Pnum);
        for (int c = 0; c < 2; ++c) dirichlet(Rv[c], s->nbc[c], s->bc_dofs[c], nullptr);
        for (int c = 0; c < 2; ++c) dirichlet(rate[c], s->nbc[c], s->bc_dofs[c], nullptr);
        orc_dr_update(s, alpha, beta);
        for (int c = 0; c < 2; ++c) dirichlet(v[c], s->nbc[c], s->bc_dofs[c], s->bc_vals[c]);
    }
}

}  // extern "C"

// assemble_velocity_operator (adjoint_operator.jl:205-224): the 14×14 element
// blocks [nels][14][14], velocity DoFs stacked by component
static void velocity_blocks(const orc_stokes* s, double dt, const double* bulk, std::vector<double>& A) {
    A.resize((size_t)s->nels * 196);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        double block[2][2][7][7];
        augmented_element_block(s, e, dt, bulk, block);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int r = 0; r < 7; ++r)
                    for (int c = 0; c < 7; ++c) A[e * 196 + (i * 7 + r) * 14 + j * 7 + c] = block[i][j][r][c];
    }
}

// estimate_velocity_λmax (adjoint_operator.jl:240-281): x, y hold the start
// (the seed or the previous estimate's vector) and return the last iterate
static bool estimate_velocity_lmax(const orc_stokes* s, const std::vector<double>& A, std::vector<double> x[2],
                                   int max_it, double rtol, double* lam_out, int64_t* its) {
    const int64_t nv = s->nnodes, ne = s->nels;
    const double* PC[2] = {s->PCx, s->PCy};
    for (int c = 0; c < 2; ++c) dirichlet(x[c].data(), s->nbc[c], s->bc_dofs[c], nullptr);
    std::vector<double> z[2], a[2], Re[2];
    for (int c = 0; c < 2; ++c) { z[c].resize(nv); a[c].resize(nv); Re[c].resize((size_t)ne * 7); }
    double lam = 0;
    for (int it = 1; it <= max_it; ++it) {
        double sx = 0, sy = 0;
        for (int64_t i = 0; i < nv; ++i) sx += x[0][i] * x[0][i];
        for (int64_t i = 0; i < nv; ++i) sy += x[1][i] * x[1][i];
        const double n = std::sqrt(sx + sy);
        if (!(n > 0)) { g_err = "power iteration collapsed to the zero vector"; return false; }
        for (int c = 0; c < 2; ++c)
            for (int64_t i = 0; i < nv; ++i) {
                x[c][i] /= n;
                z[c][i] = x[c][i] / std::sqrt(PC[c][i]);
            }
        // velocity_operator_apply_kernel!: A · (x_loc; y_loc) per element, summed per node
        for (int64_t e = 0; e < ne; ++e) {
            double u[14];
            for (int k = 0; k < 7; ++k) { u[k] = z[0][s->el2n[e * 7 + k]]; u[7 + k] = z[1][s->el2n[e * 7 + k]]; }
            for (int r = 0; r < 14; ++r) {
                double acc = 0;
                for (int k = 0; k < 14; ++k) acc += A[e * 196 + r * 14 + k] * u[k];
                Re[r / 7][e * 7 + r % 7] = acc;
            }
        }
        gather(s, Re[0], a[0].data());
        gather(s, Re[1], a[1].data());
        for (int c = 0; c < 2; ++c) {
            for (int64_t i = 0; i < nv; ++i) a[c][i] /= std::sqrt(PC[c][i]);
            dirichlet(a[c].data(), s->nbc[c], s->bc_dofs[c], nullptr);
        }
        double ax = 0, ay = 0;
        for (int64_t i = 0; i < nv; ++i) ax += a[0][i] * a[0][i];
        for (int64_t i = 0; i < nv; ++i) ay += a[1][i] * a[1][i];
        const double ln = std::sqrt(ax + ay);
        x[0] = a[0];
        x[1] = a[1];
        if (it > 1 && std::fabs(ln - lam) <= rtol * ln) {
            *lam_out = ln;
            *its = it;
            return true;
        }
        lam = ln;
    }
    if (!(std::isfinite(lam) && lam > 0)) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "power iteration produced a non-positive λmax: %g", lam);
        g_err = buf;
        return false;
    }
    *lam_out = lam;
    *its = max_it;
    return true;
}

extern "C" {

// solve! (DR.jl:302-337) + _solve_stokes_dyrel! (DR.jl:348-802); th is the
// thermal side of solve_coupled! (null for solve!).
static int stokes_solve_impl(orc_stokes* s, const orc_params* p, orc_stats* out, ThermalHook* th) {
    std::memset(out, 0, sizeof *out);
    g_err.clear();
    const double dt = p->dt;
    const int finite_K = p->finite_K;
    const int64_t ncheck = p->check_interval;
    const int64_t nv = s->nnodes, np = s->nnodesP;
    const bool freeze = p->freeze_jacobian < 0 ? (s->dp == nullptr && s->cap == nullptr) : (p->freeze_jacobian != 0);
    if (!p->skip_scaling) orc_pressure_scaling(s, p->pressure_factor, dt, finite_K);

    double* rate[2] = {s->rate_x, s->rate_y};
    double* Rv[2] = {s->Rvx, s->Rvy};
    double* Rv0[2] = {s->Rv0x, s->Rv0y};
    double* dRdv[2] = {s->dRdvx, s->dRdvy};
    double* PC[2] = {s->PCx, s->PCy};
    double* v[2] = {s->vx, s->vy};
    for (int c = 0; c < 2; ++c) { std::fill(rate[c], rate[c] + nv, 0.0); std::fill(Rv0[c], Rv0[c] + nv, 0.0); }
    const char* labels[2] = {"stokes vx", "stokes vy"};

    // measure_λmax (DR.jl:375-378, 394, 425-426, 459-490): the velocity blocks
    // A of every Jacobian assembly and a power iteration on P^-1/2 A P^-1/2
    // (estimate_velocity_λmax, adjoint_operator.jl:240-281), warm-started from
    // the previous estimate's vector; λmax_v = min(Gershgorin, safety · measured)
    if (p->measure_lmax && s->eta_ip) {
        g_err = "quadrature viscosity requires measure_λmax=false";
        out->status = -1;
        return -1;
    }
    const int power_max = p->lmax_power_iterations > 0 ? p->lmax_power_iterations : 12;
    const double power_rtol = p->lmax_power_rtol > 0 ? p->lmax_power_rtol : 1e-2;
    double safety = p->lmax_safety > 0 ? p->lmax_safety : 1.1;
    if (s->dp || s->cap) safety = std::max(safety, 1.5);   // _stokes_spectral_safety (DR.jl:931-933)
    std::vector<double> px[2], Ablk;
    int64_t power_its = 0;
    auto measure = [&](double lmax[2]) -> bool {
        if (!p->measure_lmax) return true;
        velocity_blocks(s, dt, pressure_bulk(s, finite_K), Ablk);
        if (px[0].empty())
            for (int c = 0; c < 2; ++c) {
                px[c].resize(nv);
                for (int64_t i = 0; i < nv; ++i) px[c][i] = c == 0 ? std::sin((double)(i + 1) * 0.7) : std::cos((double)(i + 1) * 1.3);
            }
        double meas = 0;
        int64_t used = 0;
        if (!estimate_velocity_lmax(s, Ablk, px, power_max, power_rtol, &meas, &used)) return false;
        power_its += used;
        const double g = std::max(lmax[0], lmax[1]);
        lmax[0] = lmax[1] = std::min(g, safety * meas);
        return true;
    };

    orc_augmented_jacobian(s, dt, finite_K);
    double lmax[2];
    for (int c = 0; c < 2; ++c)
        if (!checked_lmax(dRdv[c], PC[c], nv, labels[c], &lmax[c])) { out->status = -6; return -6; }
    double lmax_g = std::max(lmax[0], lmax[1]);
    if (!measure(lmax)) { out->status = -1; return -1; }
    int64_t jac_assemblies = 1;
    double dtau[2], al[2], be[2];
    for (int c = 0; c < 2; ++c) {
        dtau[c] = 2 / std::sqrt(lmax[c]) * p->CFL_v;
        stokes_cheb(dtau[c], 0.0, p->c_fact, &al[c], &be[c]);
    }

    if (th) thermal_hook_begin(*th);
    auto thermal_ok = [&]() { return !th || th->converged; };

    double err_min = kInf, eps_tol = p->tolerance, err = 2 * eps_tol;
    double err_abs = kInf, err_rel = kInf, err_v0 = 1.0, err_P0 = 1.0, err_v00 = 1.0;
    double err_v = kInf, err_P = kInf;
    int64_t iter = 0, itPH_done = 0;
    bool converged = false;
    double rel_drop = p->rel_drop0;
    std::vector<double> tmp(np);

    auto err_pressure = [&]() {
        for (int64_t i = 0; i < np; ++i) tmp[i] = s->RP[i] / s->MP[i];
        return nrm2(tmp.data(), np) / std::sqrt((double)np);
    };
    auto err_velocity = [&]() {
        return std::max(nrm2(Rv[0], nv), nrm2(Rv[1], nv)) / (2 * std::sqrt((double)nv));
    };
    // collect_history (DR.jl:558-561, 722-727): (iter, err_v, per-component, 0, err_P)
    if (s->history) s->history_len = 0;
    auto history = [&](int64_t it, double ev) {
        if (!s->history || s->history_len >= s->history_cap) return;
        double* h = s->history + 6 * s->history_len++;
        h[0] = (double)it; h[1] = ev;
        h[2] = nrm2(Rv[0], nv) / (2 * std::sqrt((double)nv));
        h[3] = nrm2(Rv[1], nv) / (2 * std::sqrt((double)nv));
        h[4] = 0.0; h[5] = err_pressure();
    };

    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        orc_pressure_residual(s, dt, finite_K, s->P0);
        orc_momentum_residual(s, dt, nullptr);
        for (int c = 0; c < 2; ++c) dirichlet(Rv[c], s->nbc[c], s->bc_dofs[c], nullptr);
        err_P = err_pressure();
        err_v = err_velocity();
        history(iter, err_v);
        if (itPH == 1) { err_P0 = err_P + julia_eps(err_P); err_v0 = err_v + julia_eps(err_v); }
        else if (itPH == 2) { err_P0 = err_P + julia_eps(err_P); }
        const double err_v_rel = err_v / err_v0, err_P_rel = err_P / err_P0;
        err_abs = std::max(err_v, err_P);
        err_rel = std::max(err_v_rel, err_P_rel);
        err = std::min(err_abs, err_rel);
        if (std::isnan(err_v) || std::isnan(err_P) || std::isnan(err)) {
            g_err = "NaN detected in outer loop at PH=" + std::to_string(itPH); out->status = -4; return -4;
        }
        if (err > 1.0e10) { g_err = "Kaboom! Error > 1e10 in outer loop at PH=" + std::to_string(itPH); out->status = -5; return -5; }
        if (p->verbose)
            printf("itPH = %02lld iter = %06lld err = %.3e abs = %.3e rel = %.3e - norm[Rv=%.3e %.3e, Rp=%.3e %.3e]\n",
                   (long long)itPH, (long long)iter, err, err_abs, err_rel, err_v, err_v_rel, err_P, err_P_rel);
        if (err < eps_tol && thermal_ok()) { converged = true; break; }
        if (err > err_min * 1.05) rel_drop = std::max(rel_drop * 0.1, 1.0e-3);
        err_min = std::min(err_min, err);
        const double eps_vel = err * rel_drop;
        int64_t itPT = 0;
        while ((err > eps_vel || !thermal_ok()) && itPT <= p->iter_max) {
            ++itPT; ++iter;
            const bool do_check = iter % ncheck == 0;
            if (do_check) for (int c = 0; c < 2; ++c) std::copy(Rv[c], Rv[c] + nv, Rv0[c]);
            if (th) {
                const int rc = thermal_hook_step(*th);
                if (rc != 0) { out->status = rc; return rc; }
            }
            orc_pressure_residual(s, dt, finite_K, s->P0);
            for (int64_t i = 0; i < np; ++i) s->Pnum[i] = s->gammaP[i] * s->RP[i] / s->MP[i];
            orc_momentum_residual(s, dt, s->Pnum);
            for (int c = 0; c < 2; ++c) dirichlet(Rv[c], s->nbc[c], s->bc_dofs[c], nullptr);
            for (int c = 0; c < 2; ++c) dirichlet(rate[c], s->nbc[c], s->bc_dofs[c], nullptr);
            orc_dr_update(s, al, be);
            for (int c = 0; c < 2; ++c) dirichlet(v[c], s->nbc[c], s->bc_dofs[c], s->bc_vals[c]);
            if (do_check) {
                const double evi = err_velocity();
                if (iter == ncheck) err_v00 = evi + julia_eps(evi);
                err = evi <= eps_tol ? evi : std::max(evi / err_v00, evi);
                history(iter, evi);
                if (p->verbose > 1)
                    printf("  it = %lld, iter = %lld, err = %.3e abs = %.3e err_v00 = %.3e lmax = %.4e %.4e\n",
                           (long long)itPT, (long long)iter, err, evi, err_v00, lmax[0], lmax[1]);
                if (std::isnan(err)) { g_err = "NaN detected in inner loop PH=" + std::to_string(itPH) + " PT=" + std::to_string(itPT); out->status = -4; return -4; }
                if (err > 1.0e10) { g_err = "Kaboom! Error > 1e10 in inner loop PH=" + std::to_string(itPH) + " PT=" + std::to_string(itPT); out->status = -5; return -5; }
                double lmin[2];
                for (int c = 0; c < 2; ++c) lmin[c] = stokes_lmin(al[c], rate[c], Rv[c], Rv0[c], PC[c], nv);
                if (!freeze) {
                    ++jac_assemblies;
                    orc_augmented_jacobian(s, dt, finite_K);
                    for (int c = 0; c < 2; ++c)
                        if (!checked_lmax(dRdv[c], PC[c], nv, labels[c], &lmax[c])) { out->status = -6; return -6; }
                    lmax_g = std::max(lmax[0], lmax[1]);
                    if (!measure(lmax)) { out->status = -1; return -1; }
                }
                for (int c = 0; c < 2; ++c) {
                    dtau[c] = 2 / std::sqrt(lmax[c]) * p->CFL_v;
                    stokes_cheb(dtau[c], lmin[c], p->c_fact, &al[c], &be[c]);
                }
            }
            if (iter > p->max_iterations) break;
        }
        for (int64_t i = 0; i < np; ++i) s->P[i] += s->gammaP[i] * s->RP[i] / s->MP[i];
        if (iter > p->max_iterations) break;
    }
    out->itPH = itPH_done;
    out->iter = iter;
    out->err = err;
    out->err_abs = err_abs;
    out->err_rel = err_rel;
    out->err_v = err_v;
    out->err_P = err_P;
    out->converged = converged;
    out->reached_total_iter = iter > p->max_iterations;
    out->lambda_max = std::max(lmax[0], lmax[1]);
    out->lambda_max_gershgorin = lmax_g;
    out->jacobian_assemblies = jac_assemblies;
    out->lambda_max_iterations = power_its;
    out->status = 0;
    return 0;
}

int orc_stokes_solve(orc_stokes* s, const orc_params* p, orc_stats* out) { return stokes_solve_impl(s, p, out, nullptr); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Thermal diffusion
// ---------------------------------------------------------------------------
namespace {
// interp2ip_phase (src/utils.jl:37-44) over per-node phase ids (1-based)
double interp_nodes(const double* N, int nn, const double* prop, const int32_t* ph) {
    double s = 0;
    for (int i = 0; i < nn; ++i) s += N[i] * prop[ph[i] - 1];
    return s;
}
}  // namespace

extern "C" {

// integrate_residual, src/heat_diffusion/assembly/residual.jl:58-87. The
// Jacobian is the exact derivative of the same expression: with
// D = Δt/(ρ Cp) and ρ = ρ0 (1 - α (T - Tref) + β P), ∂D/∂T_j = D ρ0 α N_j / ρ.
// The quadrature-point source phi (nullable, [nq]) enters as Δt/(ρ Cp) Φ_q
// (residual.jl:75-79), in the Jacobian through ∂D/∂T_j.
static void thermal_integrate(int32_t nn, int32_t nq, const double* N, const double* dNdx, const double* dOm,
                              const double* T, const double* T0, const double* s, const int32_t* ph,
                              const orc_thermal_material* m, const double* P, double dt, double Tref,
                              const double* phi, double* Re, double* J, int dim = 2) {
    double beta[ORC_MAX_PHASES];
    for (int p = 0; p < m->nphases; ++p) beta[p] = 1.0 / m->K[p];   // β = map(inv, K) (residual.jl:62-64)
    for (int i = 0; i < nn; ++i) Re[i] = 0;
    if (J)
        for (int i = 0; i < nn * nn; ++i) J[i] = 0;
    for (int q = 0; q < nq; ++q) {
        const double* Nv = N + (int64_t)q * nn;
        const double* G = dNdx + (int64_t)q * nn * dim;
        double Tq = 0, Pq = 0, T0q = 0, sq = 0, gT[3] = {0, 0, 0};
        for (int i = 0; i < nn; ++i) {
            Tq += Nv[i] * T[i];
            Pq += Nv[i] * P[i];
            T0q += Nv[i] * T0[i];
            sq += Nv[i] * s[i];
            for (int d = 0; d < dim; ++d) gT[d] += G[dim * i + d] * T[i];
        }
        auto gdot = [&](int i, const double* v) {
            double r = 0;
            for (int d = 0; d < dim; ++d) r += G[dim * i + d] * v[d];
            return r;
        };
        const double kq = interp_nodes(Nv, nn, m->k, ph);
        const double aq = interp_nodes(Nv, nn, m->alpha, ph);
        const double bq = interp_nodes(Nv, nn, beta, ph);
        const double r0 = interp_nodes(Nv, nn, m->rho0, ph);
        const double rho = r0 * (1 - aq * (Tq - Tref) + bq * Pq);
        const double D = dt / (rho * interp_nodes(Nv, nn, m->Cp, ph));
        double rate = T0q - Tq + D * sq;
        if (phi) rate += D * phi[q];
        for (int i = 0; i < nn; ++i) {
            const double KTi = kq * gdot(i, gT);
            Re[i] += rate * Nv[i] * dOm[q] - D * KTi * dOm[q];
        }
        if (J) {
            const double c = D * r0 * aq / rho;
            for (int i = 0; i < nn; ++i) {
                const double KTi = kq * gdot(i, gT);
                for (int j = 0; j < nn; ++j) {
                    const double dD = c * Nv[j];
                    const double Kij = kq * gdot(i, G + dim * j);
                    double dr = -Nv[j] + dD * sq;
                    if (phi) dr += dD * phi[q];
                    J[i * nn + j] += (dr * Nv[i] - dD * KTi - D * Kij) * dOm[q];
                }
            }
        }
    }
}

void orc_thermal_integrate(int32_t nn, int32_t nq, const double* N, const double* dNdx, const double* dOm,
                           const double* T, const double* T0, const double* s, const int32_t* ph,
                           const orc_thermal_material* m, const double* P, double dt, double Tref,
                           double* Re, double* J) {
    thermal_integrate(nn, nq, N, dNdx, dOm, T, T0, s, ph, m, P, dt, Tref, nullptr, Re, J);
}

// assemble_diffusion_matrices_atomix! (heat residual.jl:90-121) via
// assemble_dr_matrices_atomix! (dynamic_relaxation/assembly.jl:43-60); element
// contributions are added in ascending element order.
static void thermal_residual3(orc_thermal* t, double dt, double Tref, int32_t compute_jacobian);
static void thermal_residual_q9(orc_thermal* t, double dt, double Tref, int32_t compute_jacobian);

void orc_thermal_residual(orc_thermal* t, double dt, double Tref, int32_t compute_jacobian) {
    if (t->dim == 3) {
        thermal_residual3(t, dt, Tref, compute_jacobian);
        return;
    }
    if (t->nloc == 9) {
        thermal_residual_q9(t, dt, Tref, compute_jacobian);
        return;
    }
    const Tables& tb = tables();
    for (int64_t n = 0; n < t->nnodes; ++n) t->R[n] = 0;
    if (compute_jacobian)
        for (int64_t n = 0; n < t->nnodes; ++n) t->dRdT[n] = t->PC[n] = 0;
    double N[7][7], dNdx[7][7][2], dOm[7], Tl[7], T0l[7], sl[7], Pl[7], Re[7], J[49];
    int32_t ph[7];
    for (int q = 0; q < 7; ++q)
        for (int a = 0; a < 7; ++a) N[q][a] = tb.Nq[q][a];
    for (int64_t e = 0; e < t->nels; ++e) {
        const int32_t* nd = t->el2n + e * 7;
        for (int q = 0; q < 7; ++q) {
            const double* iJ = t->invJ + (e * 7 + q) * 4;
            for (int a = 0; a < 7; ++a) {
                dNdx[q][a][0] = tb.dNdxi[q][a][0] * iJ[0] + tb.dNdxi[q][a][1] * iJ[2];
                dNdx[q][a][1] = tb.dNdxi[q][a][0] * iJ[1] + tb.dNdxi[q][a][1] * iJ[3];
            }
            dOm[q] = t->dOmega[e * 7 + q];
        }
        for (int a = 0; a < 7; ++a) {
            Tl[a] = t->T[nd[a]]; T0l[a] = t->T0[nd[a]]; sl[a] = t->source[nd[a]]; Pl[a] = t->P[nd[a]];
            ph[a] = t->phases[nd[a]];
        }
        thermal_integrate(7, 7, &N[0][0], &dNdx[0][0][0], dOm, Tl, T0l, sl, ph, &t->mat, Pl, dt, Tref,
                          t->source_ip ? t->source_ip + e * 7 : nullptr, Re, compute_jacobian ? J : nullptr);
        for (int a = 0; a < 7; ++a) t->R[nd[a]] += Re[a];
        if (compute_jacobian) {
            for (int a = 0; a < 7; ++a) {
                double rs = 0;
                for (int b = 0; b < 7; ++b) rs += std::fabs(J[a * 7 + b]);
                t->dRdT[nd[a]] += rs;
                t->PC[nd[a]] += std::fabs(J[a * 7 + a]);
            }
        }
    }
}

// solve_dynamic_relaxation!, src/dynamic_relaxation/solver.jl:61-122
int orc_thermal_solve(orc_thermal* t, const orc_thermal_params* p, orc_thermal_stats* out) {
    g_err.clear();
    const int64_t n = t->nnodes;
    double a_dr = 0, b_dr = 0, nr0 = 0, last_rel = std::numeric_limits<double>::quiet_NaN(), lmax = 0;
    for (int64_t it = 1; it <= p->max_iterations; ++it) {
        const bool do_jac = (it % p->check_interval == 0) || it == 1;
        if (do_jac)
            for (int64_t i = 0; i < n; ++i) t->R0[i] = t->R[i];
        orc_thermal_residual(t, p->dt, p->Tref, do_jac);
        dirichlet(t->R, t->nbc, t->bc_dofs, nullptr);
        dirichlet(t->rate, t->nbc, t->bc_dofs, nullptr);
        if (do_jac && !checked_lmax(t->dRdT, t->PC, n, "thermal diffusion", &lmax)) {
            out->status = -6;
            return -6;
        }
        for (int64_t i = 0; i < n; ++i) t->rate[i] = t->R[i] / t->PC[i] + b_dr * t->rate[i];
        for (int64_t i = 0; i < n; ++i) t->T[i] += a_dr * t->rate[i];
        dirichlet(t->T, t->nbc, t->bc_dofs, t->bc_vals);
        if (do_jac) {
            const double nr = nrm2(t->R, n);
            if (it == 1) nr0 = std::max(nr, julia_eps(nr));
            if (std::isnan(nr / nr0)) {
                g_err = "NaNs at PT iter " + std::to_string(it);
                out->status = -4;
                return -4;
            }
            const double dtau = 2 / std::sqrt(lmax) * t->CFL;
            double den = 0, num = 0;
            for (int64_t i = 0; i < n; ++i) {
                const double r = dtau * t->rate[i];
                den += r * r;
                num += r * ((t->R[i] - t->R0[i]) / t->PC[i]);
            }
            const double lmin = (it == 1 || den == 0) ? 0.0 : std::fabs(num) / den;
            const double c = 2 * std::sqrt(lmin) * t->c_fact;
            a_dr = 2 * dtau * dtau / (2 + c * dtau);
            b_dr = (2 - c * dtau) / (2 + c * dtau);
            last_rel = nr / nr0;
            if (p->verbose) printf("  PT %05lld  res = %6.2e\n", (long long)it, last_rel);
            if (last_rel < p->tolerance) {
                out->converged = 1;
                out->iterations = it;
                out->residual = last_rel;
                out->status = 0;
                return 0;
            }
        }
    }
    out->converged = 0;
    out->iterations = p->max_iterations;
    out->residual = last_rel;
    out->status = 0;
    return 0;
}

// assemble_shear_heating! (shear_heating.jl:18-92), plane strain: the stress
// power τ : ∇v less the elastic power, whose sum includes τzz = −(τxx + τyy)
// (shear_dissipation, shear_heating.jl:77-88).
void orc_shear_heating(const orc_stokes* s, double dt, double* phi) {
    const Tables& tb = tables();
    Rheo r0{&s->mat, s->mat.eta, s->cap ? nullptr : s->dp, dt};
    r0.cap = s->cap;
    double invG[ORC_MAX_PHASES];
    for (int p = 0; p < s->mat.nphases; ++p) invG[p] = 1.0 / s->mat.G[p];
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        Rheo r = r0;
        double dNdx[7][7][2];
        elem_dNdx(s, e, dNdx);
        int ph[7];
        elem_phases_v(s, e, ph);
        for (int q = 0; q < 7; ++q) {
            double gvx[2] = {0, 0}, gvy[2] = {0, 0};
            for (int a = 0; a < 7; ++a) {
                const int32_t g = s->el2n[e * 7 + a];
                gvx[0] += dNdx[q][a][0] * s->vx[g];
                gvx[1] += dNdx[q][a][1] * s->vx[g];
                gvy[0] += dNdx[q][a][0] * s->vy[g];
                gvy[1] += dNdx[q][a][1] * s->vy[g];
            }
            double to[3] = {0, 0, 0};
            for (int c = 0; c < 3; ++c)
                if (s->tau_old[c]) to[c] = s->tau_old[c][e * 7 + q];
            double Pq = 0, Pfq = 0;
            for (int i = 0; i < 3; ++i) {
                Pq += tb.NqP[q][i] * s->P[e * 3 + i];
                if (s->Pf) Pfq += tb.NqP[q][i] * s->Pf[e * 3 + i];
            }
            double tau[3], Pe;
            if (r.cap)
                cap_stress_ip<double>(gvx, gvy, tb.Nq[q], 7, ph, r, to, Pq - Pfq,
                                      s->gamma_hist ? s->gamma_hist + e * 7 + q : nullptr, tau, Pe);
            else
                stress_ip<double>(gvx, gvy, tb.Nq[q], 7, ph, r, to, Pq - Pfq, tau);
            const double inv2Gdt = interp_phase(tb.Nq[q], 7, invG, ph) / (2 * dt);
            const double tzz = -(tau[0] + tau[1]), tzzo = -(to[0] + to[1]);
            const double power = tau[0] * gvx[0] + tau[1] * gvy[1] + tau[2] * (gvx[1] + gvy[0]);
            const double elastic = (tau[0] * (tau[0] - to[0]) + tau[1] * (tau[1] - to[1]) + tzz * (tzz - tzzo) +
                                    2 * tau[2] * (tau[2] - to[2])) * inv2Gdt;
            phi[e * 7 + q] = power - elastic;
        }
    }
}

void orc_cap_geometry(double k, double kq, double c, double pT, double* out) {
    const CapGeom g = cap_geometry(k, kq, c, pT);
    out[0] = g.a; out[1] = g.b; out[2] = g.p_y; out[3] = g.R_y; out[4] = g.p_d; out[5] = g.tau_d; out[6] = g.p_q;
}

double orc_cap_yield_function(double tII, double P, double k, double kq, double c, double pT) {
    return cap_yield<double>(tII, P, k, c, cap_geometry(k, kq, c, pT));
}

void orc_cap_invariants(double s, double p, double k, double kq, double c, double pT, double* out) {
    cap_invariants<double>(s, p, k, kq, c, pT, out[0], out[1], out[2]);
}

int orc_cap_return_map(double s_trial, double p_trial, double eta_ve, double KDt, double k, double kq, double c,
                       double pT, double eta_reg, int32_t maxiter, double* out) {
    const CapRet r = cap_return(s_trial, p_trial, eta_ve, KDt, k, kq, c, pT, eta_reg, maxiter);
    out[0] = r.s; out[1] = r.p; out[2] = r.lam; out[3] = r.At; out[4] = r.Ap;
    return r.converged ? 1 : 0;
}

void orc_cap_local_update(const double* tau_trial, double P_trial, double eta_ve, double KDt, double k, double kq,
                          double c, double pT, double eta_reg, double* tau_out, double* out) {
    double P, g, th;
    cap_local_update<double>(tau_trial, P_trial, eta_ve, KDt, k, kq, c, pT, eta_reg, tau_out, P, &g, &th);
    out[0] = P; out[1] = g; out[2] = th;
}

// update_stokes_plastic_history_kernel! (momentum_residuals.jl:584-610) with
// the finiteness check of DR.jl:950-952
int orc_commit_plastic_history(orc_stokes* s, double dt) {
    if (!s->cap || !s->gamma_hist || !s->theta_hist) return 0;
    const Tables& tb = tables();
    Rheo r0{&s->mat, s->mat.eta, nullptr, dt};
    r0.cap = s->cap;
    bool bad = false;
    for (int64_t e = 0; e < s->nels; ++e) {
        double dNdx[7][7][2];
        elem_dNdx(s, e, dNdx);
        int ph[7];
        elem_phases_v(s, e, ph);
        for (int q = 0; q < 7; ++q) {
            double gvx[2] = {0, 0}, gvy[2] = {0, 0};
            for (int a = 0; a < 7; ++a) {
                const int32_t g = s->el2n[e * 7 + a];
                gvx[0] += dNdx[q][a][0] * s->vx[g];
                gvx[1] += dNdx[q][a][1] * s->vx[g];
                gvy[0] += dNdx[q][a][0] * s->vy[g];
                gvy[1] += dNdx[q][a][1] * s->vy[g];
            }
            double to[3] = {0, 0, 0};
            for (int c = 0; c < 3; ++c)
                if (s->tau_old[c]) to[c] = s->tau_old[c][e * 7 + q];
            double Pq = 0, Pfq = 0;
            for (int i = 0; i < 3; ++i) {
                Pq += tb.NqP[q][i] * s->P[e * 3 + i];
                if (s->Pf) Pfq += tb.NqP[q][i] * s->Pf[e * 3 + i];
            }
            double tau[3], Pe, gd, thd;
            const double gam = s->gamma_hist[e * 7 + q];
            cap_stress_ip<double>(gvx, gvy, tb.Nq[q], 7, ph, r0, to, Pq - Pfq, &gam, tau, Pe, &gd, &thd);
            s->gamma_hist[e * 7 + q] += dt * gd;
            s->theta_hist[e * 7 + q] += dt * thd;
            bad |= !std::isfinite(s->gamma_hist[e * 7 + q]) || !std::isfinite(s->theta_hist[e * 7 + q]);
        }
    }
    if (bad) {
        g_err = "non-finite plastic history: a tensile-cap local return map failed";
        return -1;
    }
    return 0;
}

// GeoParams.rotate_elastic_stress, 2-D (GeoParams 0.7.21, the version
// Project.toml pins; restated from its documented behaviour and the
// reference's tests, test/test_stress_rotation.jl:48-77): the stress turns by
// θ = ωΔt, τ' = R(θ) τ R(θ)ᵀ with R = [cos θ  −sin θ; sin θ  cos θ].
void orc_rotate_elastic_stress2(double omega, const double* tau, double dt, double* out) {
    const double th = omega * dt;
    const double c = std::cos(th), sn = std::sin(th);
    const double R[2][2] = {{c, -sn}, {sn, c}};
    const double T[2][2] = {{tau[0], tau[2]}, {tau[2], tau[1]}};
    double RT[2][2], M[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) RT[i][j] = R[i][0] * T[0][j] + R[i][1] * T[1][j];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) M[i][j] = RT[i][0] * R[j][0] + RT[i][1] * R[j][1];
    out[0] = M[0][0];
    out[1] = M[1][1];
    out[2] = M[0][1];
}

// rotate_stress! (src/stokes/stress_rotation.jl:28-61), 2-D: at every
// quadrature point ω = ½(∂vx/∂y − ∂vy/∂x) and tau_old = rotated tau
void orc_rotate_stress(const orc_stokes* s, double dt, const double* const* tau, double* const* tau_old) {
    for (int64_t e = 0; e < s->nels; ++e) {
        double dNdx[7][7][2];
        elem_dNdx(s, e, dNdx);
        for (int q = 0; q < 7; ++q) {
            double dvx_dy = 0, dvy_dx = 0;
            for (int a = 0; a < 7; ++a) {
                const int32_t g = s->el2n[e * 7 + a];
                dvx_dy += dNdx[q][a][1] * s->vx[g];
                dvy_dx += dNdx[q][a][0] * s->vy[g];
            }
            const double t[3] = {tau[0][e * 7 + q], tau[1][e * 7 + q], tau[2][e * 7 + q]};
            double r[3];
            orc_rotate_elastic_stress2((dvx_dy - dvy_dx) / 2, t, dt, r);
            for (int c = 0; c < 3; ++c) tau_old[c][e * 7 + q] = r[c];
        }
    }
}

}  // extern "C"

namespace {

// _transfer_temperature_kernel! (DR.jl:44-47): T_P[DoFsP[i]] = T[el2nP[i]];
// the P1 DoFs of element e sit at its three vertex nodes
void transfer_T(const orc_stokes* s, const double* T, double* TP) {
    for (int64_t e = 0; e < s->nels; ++e)
        for (int a = 0; a < 3; ++a) TP[e * 3 + a] = T[s->el2n[e * 7 + a]];
}

// Hex27: the four pressure DoFs sit at the cell centre and the +x, +y, +z
// face centres, local nodes 27, 23, 24, 26 (mixed_mesh.jl:413-417)
void transfer_T3(const orc_stokes3* s, const double* T, double* TP) {
    const int rows[4] = {26, 22, 23, 25};
    for (int64_t e = 0; e < s->nels; ++e)
        for (int a = 0; a < 4; ++a) TP[e * 4 + a] = T[s->el2n[e * 27 + rows[a]]];
}

void hook_transfer(ThermalHook& h) {
    if (h.s3) transfer_T3(h.s3, h.t->T, const_cast<double*>(h.s3->T));
    else transfer_T(h.s, h.t->T, h.s->T);
}

void hook_shear(ThermalHook& h) {
    if (h.s3) orc_shear_heating3(h.s3, h.dt, h.phi.data());
    else orc_shear_heating(h.s, h.dt, h.phi.data());
}

// DR.jl:524-530: zero ∂T∂τ and R0, initial dissipation
void thermal_hook_begin(ThermalHook& h) {
    orc_thermal* t = h.t;
    std::fill(t->rate, t->rate + t->nnodes, 0.0);
    std::fill(t->R0, t->R0 + t->nnodes, 0.0);
    if (h.shear) hook_shear(h);
}

// One thermal DR iteration of the coupled inner loop (DR.jl:604-661)
int thermal_hook_step(ThermalHook& h) {
    orc_thermal* t = h.t;
    const int64_t n = t->nnodes;
    ++h.iter;
    const bool chk = h.iter == 1 || h.iter % h.nout == 0;
    if (chk) std::copy(t->R, t->R + n, t->R0);
    orc_thermal_residual(t, h.dt, h.Tref, chk);
    dirichlet(t->R, t->nbc, t->bc_dofs, nullptr);
    dirichlet(t->rate, t->nbc, t->bc_dofs, nullptr);
    if (chk && !checked_lmax(t->dRdT, t->PC, n, "thermal diffusion", &h.lmax)) return -6;
    for (int64_t i = 0; i < n; ++i) t->rate[i] = t->R[i] / t->PC[i] + h.beta * t->rate[i];
    for (int64_t i = 0; i < n; ++i) t->T[i] += h.alpha * t->rate[i];
    dirichlet(t->T, t->nbc, t->bc_dofs, t->bc_vals);
    hook_transfer(h);
    if (chk) {
        const double nr = nrm2(t->R, n);
        if (h.iter == 1) h.nr0 = std::max(nr, julia_eps(nr));
        h.err_T = nr / h.nr0;
        if (!std::isfinite(h.err_T)) {
            g_err = "non-finite thermal residual at coupled iteration " + std::to_string(h.iter);
            return -4;
        }
        const double dtau = 2 / std::sqrt(h.lmax) * t->CFL;
        double den = 0, num = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double r = dtau * t->rate[i];
            den += r * r;
        }
        if (!(h.iter == 1 || den == 0))
            for (int64_t i = 0; i < n; ++i) num += (dtau * t->rate[i]) * ((t->R[i] - t->R0[i]) / t->PC[i]);
        const double lmin = (h.iter == 1 || den == 0) ? 0.0 : std::fabs(num) / den;
        stokes_cheb(dtau, lmin, t->c_fact, &h.alpha, &h.beta);
        h.converged = h.err_T < h.eps;
        // the dissipation is refreshed after the spectral estimate (DR.jl:659-661)
        if (h.shear) hook_shear(h);
    }
    return 0;
}

}  // namespace

extern "C" {

// solve_coupled! (DR.jl:81-131) with the coupled loop of _solve_stokes_dyrel!
int orc_coupled_solve(orc_thermal* t, orc_stokes* s, const orc_params* p, const orc_coupled_params* cp,
                      orc_stats* out, orc_coupled_stats* cst) {
    std::memset(out, 0, sizeof *out);
    std::memset(cst, 0, sizeof *cst);
    g_err.clear();
    if (t->nnodes != s->nnodes || (t->nloc != 0 && t->nloc != 7)) {
        g_err = "thermal mesh nodes must match the Stokes velocity-node layout";
        out->status = -2;
        return -2;
    }
    transfer_T(s, t->T, s->T);
    transfer_T(s, t->T0, s->T0);
    ThermalHook h;
    h.t = t;
    h.s = s;
    h.dt = p->dt;
    h.Tref = cp->Tref;
    h.eps = cp->tolerance_T;
    h.nout = p->check_interval;
    h.shear = cp->shear_heating != 0;
    const double* saved = t->source_ip;
    if (h.shear) {
        h.phi.assign((size_t)s->nels * 7, 0.0);
        t->source_ip = h.phi.data();
    }
    const int rc = stokes_solve_impl(s, p, out, &h);
    t->source_ip = saved;
    cst->err_T = h.err_T;
    cst->thermal_iterations = h.iter;
    cst->thermal_converged = h.converged;
    if (rc == 0) out->converged = out->converged && h.converged;
    return rc;
}

}  // extern "C"

// ===========================================================================
// 3-D: Hex27 velocity, P1-disc pressure with four modes per cell
// ===========================================================================
namespace {

// QuadraticElement{3, 27} node order (src/elements/shape_functions.jl:411-446)
const double kHex27Nodes[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1}, {0, -1, 1},  {1, 0, 1},  {0, 1, 1}, {-1, 0, 1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},  {-1, 1, 0},  {0, 0, -1},  {0, -1, 0}, {1, 0, 0}, {0, 1, 0},
    {-1, 0, 0},   {0, 0, 1},   {0, 0, 0}};

// one-dimensional quadratic Lagrange basis on -1, 0, +1 (shape_functions.jl:173-196)
double lq(double x, double node) { return node < 0 ? x * (x - 1) * 0.5 : (node == 0 ? 1 - x * x : x * (x + 1) * 0.5); }
double dlq(double x, double node) { return node < 0 ? x - 0.5 : (node == 0 ? -2 * x : x + 0.5); }

void hex27_basis(double xi, double eta, double zeta, double N[27], double dN[27][3]) {
    for (int a = 0; a < 27; ++a) {
        const double* n = kHex27Nodes[a];
        const double fx = lq(xi, n[0]), fy = lq(eta, n[1]), fz = lq(zeta, n[2]);
        N[a] = fx * fy * fz;
        dN[a][0] = dlq(xi, n[0]) * fy * fz;
        dN[a][1] = fx * dlq(eta, n[1]) * fz;
        dN[a][2] = fx * fy * dlq(zeta, n[2]);
    }
}

// 3x3x3 Gauss rule, ξ fastest (integration_points.jl:321-337)
void hex27_rule(double xi[27], double eta[27], double zeta[27], double w[27]) {
    const double a = std::sqrt(3.0 / 5.0);
    const double pts[3] = {-a, 0.0, a}, ws[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int i = 0; i < 27; ++i) {
        const int ix = i % 3, iy = (i / 3) % 3, iz = i / 9;
        xi[i] = pts[ix]; eta[i] = pts[iy]; zeta[i] = pts[iz];
        w[i] = ws[ix] * ws[iy] * ws[iz];
    }
}

// LinearElement{3, 4} modes (1 - ξ - η - ζ, ξ, η, ζ) evaluated on the cube
// (shape_functions.jl:288-298, mixed_mesh.jl:399-417)
void hexp1_basis(double xi, double eta, double zeta, double N[4]) {
    N[0] = 1 - xi - eta - zeta; N[1] = xi; N[2] = eta; N[3] = zeta;
}

struct Tables3 {
    double Nq[27][27], NqP[27][4], dNdxi[27][27][3], w[27];
    Tables3() {
        double xi[27], eta[27], zeta[27];
        hex27_rule(xi, eta, zeta, w);
        for (int q = 0; q < 27; ++q) {
            hex27_basis(xi[q], eta[q], zeta[q], Nq[q], dNdxi[q]);
            hexp1_basis(xi[q], eta[q], zeta[q], NqP[q]);
        }
    }
};
const Tables3& tables3() { static Tables3 t; return t; }

// 3-D deviatoric stress with the Drucker-Prager return at one point
// (rheology.jl:413-472); gv[i][j] = ∂v_i/∂x_j; tau = (xx, yy, zz, xy, xz, yz).
template <class S>
double stress_ip3(const S gv[3][3], const double* Nv, int NV, const int* ph, const Rheo& r, const double tau_o[6],
                  double Pq, S tau[6]) {
    S tr = (gv[0][0] + gv[1][1] + gv[2][2]) / 3.0;
    S eps[6] = {gv[0][0] - tr, gv[1][1] - tr, gv[2][2] - tr, (gv[0][1] + gv[1][0]) / 2.0,
                (gv[0][2] + gv[2][0]) / 2.0, (gv[1][2] + gv[2][1]) / 2.0};
    double invG[ORC_MAX_PHASES];
    for (int p = 0; p < r.mat->nphases; ++p) invG[p] = 1.0 / r.mat->G[p];
    const double etaq = interp_phase(Nv, NV, r.eta, ph);
    const double invGq = interp_phase(Nv, NV, invG, ph);
    const double etave = 1.0 / (1.0 / etaq + invGq / r.dt);
    const double inv2Gdt = invGq / (2 * r.dt);
    for (int k = 0; k < 6; ++k) tau[k] = 2 * etave * (eps[k] + tau_o[k] * inv2Gdt);
    double lambda = 0;
    if (r.dp) {
        const orc_dp* dp = r.dp;
        const double cphi = interp_phase(Nv, NV, dp->cos_phi, ph);
        const double sphi = interp_phase(Nv, NV, dp->sin_phi, ph);
        const double spsi = interp_phase(Nv, NV, dp->sin_psi, ph);
        const double C = interp_phase(Nv, NV, dp->C, ph);
        const double ereg = interp_phase(Nv, NV, dp->eta_reg, ph);
        const double Kb = interp_phase(Nv, NV, dp->Kb, ph);
        // second_invariant, rheology.jl:395-400
        S tII = dsqrt((tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]) / 2.0 + tau[3] * tau[3] +
                      tau[4] * tau[4] + tau[5] * tau[5] + kEps * kEps);
        S tIIs = tII + kEps * kEps;
        S F = tII - cphi * C - sphi * Pq;
        const double dFdP = -sphi, dQdP = -spsi;
        S dQ[6] = {tau[0] / (2.0 * tIIs), tau[1] / (2.0 * tIIs), tau[2] / (2.0 * tIIs),
                   tau[3] / tIIs, tau[4] / tIIs, tau[5] / tIIs};
        if (val(F) > 0) {
            S lam = F / (etave + ereg + Kb * r.dt * dQdP * dFdP);
            lambda = val(lam);
            if (lambda > 0)
                for (int k = 0; k < 6; ++k) tau[k] = tau[k] - 2 * etave * lam * dQ[k];
        }
    }
    return lambda;
}

// 3-D element momentum residual (momentum_residuals.jl:209-258,
// accumulate_momentum :28-33, eos_density :11-16). dNdx [nq][NV][3],
// tau_old [6][nq]; T (nullable): the element's temperature at its pressure
// DoFs, interpolated with the pressure basis into the density (T = 0 without)
template <class S>
void momentum_element3(int NV, int NP, int nq, const double* dNdx, const double* dOm, const double* Nq,
                       const double* NqP, const S* const v[3], const double* P, const S* Pnum, const int* ph,
                       const Rheo& r, const double g[3], const double* tau_old, S* const R[3], double* tau_out,
                       const double* T = nullptr) {
    const orc_material* m = r.mat;
    double beta[ORC_MAX_PHASES];
    for (int p = 0; p < m->nphases; ++p) beta[p] = 1.0 / m->K[p];
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < NV; ++a) R[c][a] = S(0.0);
    for (int q = 0; q < nq; ++q) {
        const double* G = dNdx + (size_t)q * NV * 3;
        const double* Nv = Nq + (size_t)q * NV;
        const double* Np = NqP + (size_t)q * NP;
        S gv[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                S acc = S(0.0);
                for (int a = 0; a < NV; ++a) acc = acc + G[a * 3 + j] * v[i][a];
                gv[i][j] = acc;
            }
        double to[6] = {0, 0, 0, 0, 0, 0};
        if (tau_old)
            for (int k = 0; k < 6; ++k) to[k] = tau_old[k * nq + q];
        double Pq = 0;
        for (int i = 0; i < NP; ++i) Pq += Np[i] * P[i];
        S tau[6];
        stress_ip3<S>(gv, Nv, NV, ph, r, to, Pq, tau);
        if (tau_out)
            for (int k = 0; k < 6; ++k) tau_out[k * nq + q] = val(tau[k]);
        S Ptot = S(Pq);
        if (Pnum) {
            S pn = S(0.0);
            for (int i = 0; i < NP; ++i) pn = pn + Np[i] * Pnum[i];
            Ptot = Ptot + pn;
        }
        double Tq = 0.0;
        if (T)
            for (int i = 0; i < NP; ++i) Tq += Np[i] * T[i];
        const double rho = interp_phase(Nv, NV, m->rho0, ph) *
                           (1 - interp_phase(Nv, NV, m->alpha, ph) * (Tq - m->Tref) + interp_phase(Nv, NV, beta, ph) * Pq);
        // τ as a symmetric matrix; column i holds τ_ji
        const S* tm[3][3] = {{&tau[0], &tau[3], &tau[4]}, {&tau[3], &tau[1], &tau[5]}, {&tau[4], &tau[5], &tau[2]}};
        for (int i = 0; i < 3; ++i)
            for (int a = 0; a < NV; ++a) {
                S acc = S(0.0);
                for (int j = 0; j < 3; ++j) {
                    S sji = *tm[j][i];
                    if (i == j) sji = sji - Ptot;
                    acc = acc + G[a * 3 + j] * sji;
                }
                R[i][a] = R[i][a] + (acc - Nv[a] * (rho * g[i])) * dOm[q];
            }
    }
}

// 3-D element continuity residual (pressure_residual.jl:132-158), no Q; with
// T (the element's temperature at its pressure DoFs, nullable) the thermal term
// α_q N^P·(T − T0)/Δt, α interpolated like the bulk modulus
template <class S>
void pressure_element3(int NV, int NP, int nq, const double* dNdx, const double* dOmP, const double* NqP,
                       const S* const v[3], const double* P, const double* P0, const int* phP, const double* bulk,
                       int nphases, double dt, S* RP, const double* alpha = nullptr, const double* T = nullptr,
                       const double* T0 = nullptr) {
    bool all_inf = true;
    for (int p = 0; p < nphases; ++p) all_inf = all_inf && std::isinf(bulk[p]);
    for (int i = 0; i < NP; ++i) RP[i] = S(0.0);
    for (int q = 0; q < nq; ++q) {
        const double* G = dNdx + (size_t)q * NV * 3;
        const double* Nv = NqP + (size_t)q * NP;
        const double etabq = all_inf ? kInf : interp_phase(Nv, NP, bulk, phP);
        double dP = 0;
        for (int i = 0; i < NP; ++i) dP += Nv[i] * (P[i] - P0[i]);
        const double dPdt = dP / (etabq * dt);
        double dTdt = 0.0;
        if (T) {
            double dT = 0;
            for (int i = 0; i < NP; ++i) dT += Nv[i] * (T[i] - T0[i]);
            dTdt = interp_phase(Nv, NP, alpha, phP) * dT / dt;
        }
        S div = S(0.0);
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < NV; ++a) div = div + G[a * 3 + c] * v[c][a];
        for (int i = 0; i < NP; ++i) RP[i] = RP[i] + Nv[i] * (((-div) - dPdt) + dTdt + 0.0) * dOmP[q];
    }
}

void elem_dNdx3(const orc_stokes3* s, int64_t e, double dNdx[27][27][3]) {
    const Tables3& t = tables3();
    for (int q = 0; q < 27; ++q) {
        const double* iJ = s->invJ + (e * 27 + q) * 9;
        for (int a = 0; a < 27; ++a)
            for (int k = 0; k < 3; ++k)
                dNdx[q][a][k] = t.dNdxi[q][a][0] * iJ[0 * 3 + k] + t.dNdxi[q][a][1] * iJ[1 * 3 + k] +
                                t.dNdxi[q][a][2] * iJ[2 * 3 + k];
    }
}

const double* pressure_bulk3(const orc_stokes3* s, int finite_K) { return finite_K ? s->mat.K : s->mat.etab; }

void gather3(const orc_stokes3* s, const std::vector<double>& Re, double* out) {
    // element contributions in ascending element order
    std::fill(out, out + s->nnodes, 0.0);
    for (int64_t e = 0; e < s->nels; ++e)
        for (int a = 0; a < 27; ++a) out[s->el2n[e * 27 + a]] += Re[e * 27 + a];
}

}  // namespace

extern "C" {

void orc_hex27_basis(double xi, double eta, double zeta, double* N, double* dN) {
    hex27_basis(xi, eta, zeta, N, reinterpret_cast<double(*)[3]>(dN));
}
void orc_hex27_rule(double* xi, double* eta, double* zeta, double* w) { hex27_rule(xi, eta, zeta, w); }
void orc_hex27_tables(double* Nq, double* NqP, double* dNdxi, double* w) {
    const Tables3& t = tables3();
    std::memcpy(Nq, t.Nq, sizeof t.Nq);
    std::memcpy(NqP, t.NqP, sizeof t.NqP);
    std::memcpy(dNdxi, t.dNdxi, sizeof t.dNdxi);
    std::memcpy(w, t.w, sizeof t.w);
}

void orc_hex27_mesh_sizes(int32_t nx, int32_t ny, int32_t nz, int64_t* nnodes, int64_t* nels) {
    *nnodes = (int64_t)(2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1);
    *nels = (int64_t)nx * ny * nz;
}

// generate_coordinates (mesh.jl:340-398) and generate_element2node
// (connectivity.jl:273-322) for QuadraticElement{3, 27}
void orc_hex27_mesh(int32_t nx, int32_t ny, int32_t nz, double x0, double x1, double y0, double y1, double z0,
                    double z1, double* coords, int32_t* el2n) {
    const int64_t sx = 2 * nx + 1, sy = 2 * ny + 1, sz = 2 * nz + 1;
    for (int64_t iz = 0; iz < sz; ++iz)
        for (int64_t iy = 0; iy < sy; ++iy)
            for (int64_t ix = 0; ix < sx; ++ix) {
                const int64_t n = (iz * sy + iy) * sx + ix;
                coords[3 * n + 0] = x0 + (x1 - x0) * (double)ix / (double)(sx - 1);
                coords[3 * n + 1] = y0 + (y1 - y0) * (double)iy / (double)(sy - 1);
                coords[3 * n + 2] = z0 + (z1 - z0) * (double)iz / (double)(sz - 1);
            }
    auto node = [&](int64_t ix, int64_t iy, int64_t iz) { return (int32_t)(iz * sx * sy + iy * sx + ix); };
    int64_t e = 0;
    for (int64_t ez = 0; ez < nz; ++ez)
        for (int64_t ey = 0; ey < ny; ++ey)
            for (int64_t ex = 0; ex < nx; ++ex, ++e) {
                const int64_t ix = 2 * ex, iy = 2 * ey, iz = 2 * ez;
                for (int a = 0; a < 27; ++a) {
                    const double* r = kHex27Nodes[a];
                    el2n[e * 27 + a] = node(ix + 1 + (int64_t)r[0], iy + 1 + (int64_t)r[1], iz + 1 + (int64_t)r[2]);
                }
            }
}

// precompute_geometry (mesh.jl:550-553, 674-682): J = Σ_a x_a ⊗ ∂N_a/∂ξ at each
// point, stored as J^-1 (row-major, [ξ-index][x-index]) and |det J| w
void orc_hex27_geometry(int64_t nels, const double* coords, const int32_t* el2n, double* invJ, double* dOmega) {
    const Tables3& t = tables3();
    for (int64_t e = 0; e < nels; ++e)
        for (int q = 0; q < 27; ++q) {
            double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // J[i][j] = ∂x_i/∂ξ_j
            for (int a = 0; a < 27; ++a) {
                const double* x = coords + 3 * (int64_t)el2n[e * 27 + a];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] += x[i] * t.dNdxi[q][a][j];
            }
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            double inv[3][3];   // (J^-1)[j][i]: ∂ξ_j/∂x_i
            inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
            inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
            inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
            inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) invJ[(e * 27 + q) * 9 + j * 3 + i] = inv[j][i];
            dOmega[e * 27 + q] = std::fabs(det) * t.w[q];
        }
}

void orc_momentum_element3(int32_t NV, int32_t NP, int32_t nq, const double* dNdx, const double* dOm, const double* Nq,
                           const double* NqP, const double* vx, const double* vy, const double* vz, const double* P,
                           const double* Pnum, const int32_t* phase, const orc_material* mat, const double* g,
                           const orc_dp* dp, double dt, const double* tau_old, double* Rx, double* Ry, double* Rz,
                           double* tau_out) {
    std::vector<int> ph(NV);
    for (int a = 0; a < NV; ++a) ph[a] = phase[a] - 1;
    Rheo r{mat, mat->eta, dp, dt};
    const double* v[3] = {vx, vy, vz};
    double* R[3] = {Rx, Ry, Rz};
    momentum_element3<double>(NV, NP, nq, dNdx, dOm, Nq, NqP, v, P, Pnum, ph.data(), r, g, tau_old, R, tau_out);
}

// assemble_viscosity_weighted_pressure_scaling! with the Hex27 Jacobi weight
// N_a² dΩ (pressure_scaling.jl:166-216)
void orc_pressure_scaling3(orc_stokes3* s, double gamma_fact, double dt, int32_t finite_K) {
    const Tables3& t = tables3();
    const double* bulk = pressure_bulk3(s, finite_K);
    double invK[ORC_MAX_PHASES];
    for (int p = 0; p < s->mat.nphases; ++p) invK[p] = 1.0 / bulk[p];
    for (int64_t e = 0; e < s->nels; ++e) {
        int ph[27];
        for (int a = 0; a < 27; ++a) ph[a] = s->phases[e] - 1;
        double MP[4] = {0, 0, 0, 0}, gP[4] = {0, 0, 0, 0};
        for (int q = 0; q < 27; ++q) {
            const double dO = s->dOmega[e * 27 + q];
            const double etaq = interp_phase(t.Nq[q], 27, s->mat.eta, ph);
            const double gnum = gamma_fact * etaq;
            const double bq = interp_phase(t.Nq[q], 27, invK, ph);
            const double gphy = bq == 0 ? gnum : dt / bq;
            const double gq = gphy * gnum / (gphy + gnum);
            for (int a = 0; a < 4; ++a) {
                const double w = t.NqP[q][a] * t.NqP[q][a] * dO;
                MP[a] += w;
                gP[a] += w * gq;
            }
        }
        for (int a = 0; a < 4; ++a) {
            s->MP[e * 4 + a] = MP[a];
            s->gammaP[e * 4 + a] = MP[a] == 0 ? 0.0 : gP[a] / MP[a];
        }
    }
}

void orc_pressure_residual3(orc_stokes3* s, double dt, int32_t finite_K) {
    const Tables3& t = tables3();
    const double* bulk = pressure_bulk3(s, finite_K);
    for (int64_t e = 0; e < s->nels; ++e) {
        static thread_local double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        double vl[3][27];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 27; ++a) vl[c][a] = s->v[c][s->el2n[e * 27 + a]];
        const double* v[3] = {vl[0], vl[1], vl[2]};
        int ph[4];
        for (int a = 0; a < 4; ++a) ph[a] = s->phases[e] - 1;
        double RPe[4];
        pressure_element3<double>(27, 4, 27, &dNdx[0][0][0], s->dOmega + e * 27, &t.NqP[0][0], v, s->P + e * 4,
                                  s->P0 + e * 4, ph, bulk, s->mat.nphases, dt, RPe, s->mat.alpha,
                                  s->T ? s->T + e * 4 : nullptr, s->T ? s->T0 + e * 4 : nullptr);
        for (int a = 0; a < 4; ++a) s->RP[e * 4 + a] = RPe[a];
    }
}

static void momentum_residual3_impl(orc_stokes3* s, double dt, const double* Pnum, double* const* tau_store) {
    const Tables3& t = tables3();
    std::vector<double> Re[3];
    for (int c = 0; c < 3; ++c) Re[c].assign((size_t)s->nels * 27, 0.0);
    Rheo r{&s->mat, s->mat.eta, s->dp, dt};
    const bool has_old = s->tau_old[0] != nullptr;
    for (int64_t e = 0; e < s->nels; ++e) {
        static thread_local double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        double vl[3][27];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 27; ++a) vl[c][a] = s->v[c][s->el2n[e * 27 + a]];
        const double* v[3] = {vl[0], vl[1], vl[2]};
        int ph[27];
        for (int a = 0; a < 27; ++a) ph[a] = s->phases[e] - 1;
        double to[6 * 27], tout[6 * 27];
        if (has_old)
            for (int k = 0; k < 6; ++k)
                for (int q = 0; q < 27; ++q) to[k * 27 + q] = s->tau_old[k][e * 27 + q];
        double Rl[3][27];
        double* R[3] = {Rl[0], Rl[1], Rl[2]};
        momentum_element3<double>(27, 4, 27, &dNdx[0][0][0], s->dOmega + e * 27, &t.Nq[0][0], &t.NqP[0][0], v,
                                  s->P + e * 4, Pnum ? Pnum + e * 4 : nullptr, ph, r, s->g, has_old ? to : nullptr, R,
                                  tau_store ? tout : nullptr, s->T ? s->T + e * 4 : nullptr);
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 27; ++a) Re[c][e * 27 + a] = Rl[c][a];
        if (tau_store)
            for (int k = 0; k < 6; ++k)
                for (int q = 0; q < 27; ++q) tau_store[k][e * 27 + q] = tout[k * 27 + q];
    }
    for (int c = 0; c < 3; ++c) gather3(s, Re[c], s->Rv[c]);
}

void orc_momentum_residual3(orc_stokes3* s, double dt, const double* Pnum) { momentum_residual3_impl(s, dt, Pnum, nullptr); }
void orc_update_current_stress3(orc_stokes3* s, double dt) { momentum_residual3_impl(s, dt, nullptr, s->tau_store); }

// GeoParams.rotate_elastic_stress, 3-D (restated; test/test_stress_rotation.jl
// :112-160): ω = curl v; the stress co-rotates by the rotation vector ωΔt/2,
// R = I + sin θ K + (1 − cos θ) K² (Rodrigues; K the cross-product matrix of
// the unit axis, θ = |ω|Δt/2), τ' = R τ Rᵀ; no rotation when ω = 0. tau in
// assembler order (xx, yy, zz, xy, xz, yz).
void orc_rotate_elastic_stress3(const double* omega, const double* tau, double dt, double* out) {
    const double r[3] = {omega[0] * dt / 2, omega[1] * dt / 2, omega[2] * dt / 2};
    const double th = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (th == 0) {
        for (int c = 0; c < 6; ++c) out[c] = tau[c];
        return;
    }
    const double n[3] = {r[0] / th, r[1] / th, r[2] / th};
    const double K[3][3] = {{0, -n[2], n[1]}, {n[2], 0, -n[0]}, {-n[1], n[0], 0}};
    double K2[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K2[i][j] = K[i][0] * K[0][j] + K[i][1] * K[1][j] + K[i][2] * K[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = (i == j ? 1.0 : 0.0) + std::sin(th) * K[i][j] + (1 - std::cos(th)) * K2[i][j];
    const double T[3][3] = {{tau[0], tau[3], tau[4]}, {tau[3], tau[1], tau[5]}, {tau[4], tau[5], tau[2]}};
    double RT[3][3], M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) RT[i][j] = R[i][0] * T[0][j] + R[i][1] * T[1][j] + R[i][2] * T[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = RT[i][0] * R[j][0] + RT[i][1] * R[j][1] + RT[i][2] * R[j][2];
    out[0] = M[0][0]; out[1] = M[1][1]; out[2] = M[2][2];
    out[3] = M[0][1]; out[4] = M[0][2]; out[5] = M[1][2];
}

// rotate_stress! (stress_rotation.jl:28-61), 3-D
void orc_rotate_stress3(const orc_stokes3* s, double dt, const double* const* tau, double* const* tau_old) {
    for (int64_t e = 0; e < s->nels; ++e) {
        static thread_local double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        for (int q = 0; q < 27; ++q) {
            double G[3][3] = {};   // G[c][k] = ∂v_c/∂x_k
            for (int a = 0; a < 27; ++a) {
                const int32_t g = s->el2n[e * 27 + a];
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < 3; ++k) G[c][k] += dNdx[q][a][k] * s->v[c][g];
            }
            const double w[3] = {G[2][1] - G[1][2], G[0][2] - G[2][0], G[1][0] - G[0][1]};
            double t[6], o[6];
            for (int c = 0; c < 6; ++c) t[c] = tau[c][e * 27 + q];
            orc_rotate_elastic_stress3(w, t, dt, o);
            for (int c = 0; c < 6; ++c) tau_old[c][e * 27 + q] = o[c];
        }
    }
}

// augmented Jacobian diagnostics in 3-D: as orc_augmented_jacobian, one
// forward-mode pass per velocity direction with 27 partials
void orc_augmented_jacobian3(orc_stokes3* s, double dt, int32_t finite_K) {
    const Tables3& t = tables3();
    const double* bulk = pressure_bulk3(s, finite_K);
    std::vector<double> Jr[3], Dg[3];
    for (int c = 0; c < 3; ++c) { Jr[c].assign((size_t)s->nels * 27, 0.0); Dg[c].assign((size_t)s->nels * 27, 0.0); }
    typedef Dual<27> D27;
    const bool has_old = s->tau_old[0] != nullptr;
    for (int64_t e = 0; e < s->nels; ++e) {
        static thread_local double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        int ph[27], phP[4];
        for (int a = 0; a < 27; ++a) ph[a] = s->phases[e] - 1;
        for (int a = 0; a < 4; ++a) phP[a] = s->phases[e] - 1;
        double eta_pc[ORC_MAX_PHASES];
        for (int p = 0; p < s->mat.nphases; ++p) eta_pc[p] = s->mat.eta[ph[0]];   // cell phase: element max
        Rheo r{&s->mat, eta_pc, s->dp, dt};
        double to[6 * 27];
        if (has_old)
            for (int k = 0; k < 6; ++k)
                for (int q = 0; q < 27; ++q) to[k * 27 + q] = s->tau_old[k][e * 27 + q];
        static thread_local double block[3][3][27][27];   // [i][j][row][col]
        for (int j = 0; j < 3; ++j) {
            static thread_local D27 u[3][27];
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 27; ++a) {
                    u[c][a] = D27(s->v[c][s->el2n[e * 27 + a]]);
                    if (c == j) u[c][a].d[a] = 1.0;
                }
            const D27* uv[3] = {u[0], u[1], u[2]};
            D27 RPe[4];
            pressure_element3<D27>(27, 4, 27, &dNdx[0][0][0], s->dOmega + e * 27, &t.NqP[0][0], uv, s->P + e * 4,
                                   s->P0 + e * 4, phP, bulk, s->mat.nphases, dt, RPe);
            D27 Pn[4];
            for (int a = 0; a < 4; ++a) Pn[a] = s->gammaP[e * 4 + a] * RPe[a] / s->MP[e * 4 + a];
            static thread_local D27 Rl[3][27];
            D27* R[3] = {Rl[0], Rl[1], Rl[2]};
            momentum_element3<D27>(27, 4, 27, &dNdx[0][0][0], s->dOmega + e * 27, &t.Nq[0][0], &t.NqP[0][0], uv,
                                   s->P + e * 4, Pn, ph, r, s->g, has_old ? to : nullptr, R, nullptr);
            for (int i = 0; i < 3; ++i)
                for (int row = 0; row < 27; ++row)
                    for (int col = 0; col < 27; ++col) block[i][j][row][col] = Rl[i][row].d[col];
        }
        for (int i = 0; i < 3; ++i)
            for (int row = 0; row < 27; ++row) {
                double sum = 0;
                for (int j = 0; j < 3; ++j)
                    for (int col = 0; col < 27; ++col) sum += std::fabs(block[i][j][row][col]);
                Jr[i][e * 27 + row] = sum;
                Dg[i][e * 27 + row] = std::fabs(block[i][i][row][row]);
            }
    }
    for (int c = 0; c < 3; ++c) {
        gather3(s, Jr[c], s->dRdv[c]);
        gather3(s, Dg[c], s->PC[c]);
    }
}

// solve! (DR.jl:302-337) + _solve_stokes_dyrel! (DR.jl:348-802), D = 3; with
// th, the coupled loop of solve_coupled! (DR.jl:514-661)
static int stokes_solve3_impl(orc_stokes3* s, const orc_params* p, orc_stats* out, ThermalHook* th) {
    std::memset(out, 0, sizeof *out);
    g_err.clear();
    const double dt = p->dt;
    const int finite_K = p->finite_K;
    const int64_t ncheck = p->check_interval;
    const int64_t nv = s->nnodes, np = s->nnodesP;
    const bool freeze = p->freeze_jacobian < 0 ? (s->dp == nullptr) : (p->freeze_jacobian != 0);
    if (!p->skip_scaling) orc_pressure_scaling3(s, p->pressure_factor, dt, finite_K);
    for (int c = 0; c < 3; ++c) {
        std::fill(s->rate[c], s->rate[c] + nv, 0.0);
        std::fill(s->Rv0[c], s->Rv0[c] + nv, 0.0);
    }
    const char* labels[3] = {"stokes vx", "stokes vy", "stokes vz"};
    orc_augmented_jacobian3(s, dt, finite_K);
    double lmax[3];
    for (int c = 0; c < 3; ++c)
        if (!checked_lmax(s->dRdv[c], s->PC[c], nv, labels[c], &lmax[c])) { out->status = -6; return -6; }
    double lmax_g = std::max(lmax[0], std::max(lmax[1], lmax[2]));
    int64_t jac_assemblies = 1;
    double dtau[3], al[3], be[3];
    for (int c = 0; c < 3; ++c) {
        dtau[c] = 2 / std::sqrt(lmax[c]) * p->CFL_v;
        stokes_cheb(dtau[c], 0.0, p->c_fact, &al[c], &be[c]);
    }
    if (th) thermal_hook_begin(*th);
    auto thermal_ok = [&]() { return !th || th->converged; };
    double err_min = kInf, eps_tol = p->tolerance, err = 2 * eps_tol;
    double err_abs = kInf, err_rel = kInf, err_v0 = 1.0, err_P0 = 1.0, err_v00 = 1.0;
    double err_v = kInf, err_P = kInf;
    int64_t iter = 0, itPH_done = 0;
    bool converged = false;
    double rel_drop = p->rel_drop0;
    std::vector<double> tmp(np);
    auto err_pressure = [&]() {
        for (int64_t i = 0; i < np; ++i) tmp[i] = s->RP[i] / s->MP[i];
        return nrm2(tmp.data(), np) / std::sqrt((double)np);
    };
    auto err_velocity = [&]() {
        return std::max(nrm2(s->Rv[0], nv), std::max(nrm2(s->Rv[1], nv), nrm2(s->Rv[2], nv))) /
               (2 * std::sqrt((double)nv));
    };
    auto update = [&]() {
        for (int c = 0; c < 3; ++c)
            for (int64_t i = 0; i < nv; ++i) {
                s->rate[c][i] = be[c] * s->rate[c][i] + s->Rv[c][i] / s->PC[c][i];
                s->v[c][i] += -al[c] * s->rate[c][i];
            }
    };
    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        orc_pressure_residual3(s, dt, finite_K);
        orc_momentum_residual3(s, dt, nullptr);
        for (int c = 0; c < 3; ++c) dirichlet(s->Rv[c], s->nbc[c], s->bc_dofs[c], nullptr);
        err_P = err_pressure();
        err_v = err_velocity();
        if (itPH == 1) { err_P0 = err_P + julia_eps(err_P); err_v0 = err_v + julia_eps(err_v); }
        else if (itPH == 2) { err_P0 = err_P + julia_eps(err_P); }
        err_abs = std::max(err_v, err_P);
        err_rel = std::max(err_v / err_v0, err_P / err_P0);
        err = std::min(err_abs, err_rel);
        if (std::isnan(err_v) || std::isnan(err_P) || std::isnan(err)) {
            g_err = "NaN detected in outer loop at PH=" + std::to_string(itPH); out->status = -4; return -4;
        }
        if (err > 1.0e10) { g_err = "Kaboom! Error > 1e10 in outer loop at PH=" + std::to_string(itPH); out->status = -5; return -5; }
        if (p->verbose)
            printf("itPH = %02lld iter = %06lld err = %.3e abs = %.3e rel = %.3e\n", (long long)itPH, (long long)iter,
                   err, err_abs, err_rel);
        if (err < eps_tol && thermal_ok()) { converged = true; break; }
        if (err > err_min * 1.05) rel_drop = std::max(rel_drop * 0.1, 1.0e-3);
        err_min = std::min(err_min, err);
        const double eps_vel = err * rel_drop;
        int64_t itPT = 0;
        while ((err > eps_vel || !thermal_ok()) && itPT <= p->iter_max) {
            ++itPT; ++iter;
            const bool do_check = iter % ncheck == 0;
            if (do_check) for (int c = 0; c < 3; ++c) std::copy(s->Rv[c], s->Rv[c] + nv, s->Rv0[c]);
            if (th) {
                const int rc = thermal_hook_step(*th);
                if (rc != 0) { out->status = rc; return rc; }
            }
            orc_pressure_residual3(s, dt, finite_K);
            for (int64_t i = 0; i < np; ++i) s->Pnum[i] = s->gammaP[i] * s->RP[i] / s->MP[i];
            orc_momentum_residual3(s, dt, s->Pnum);
            for (int c = 0; c < 3; ++c) dirichlet(s->Rv[c], s->nbc[c], s->bc_dofs[c], nullptr);
            for (int c = 0; c < 3; ++c) dirichlet(s->rate[c], s->nbc[c], s->bc_dofs[c], nullptr);
            update();
            for (int c = 0; c < 3; ++c) dirichlet(s->v[c], s->nbc[c], s->bc_dofs[c], s->bc_vals[c]);
            if (do_check) {
                const double evi = err_velocity();
                if (iter == ncheck) err_v00 = evi + julia_eps(evi);
                err = evi <= eps_tol ? evi : std::max(evi / err_v00, evi);
                if (std::isnan(err)) { g_err = "NaN detected in inner loop PH=" + std::to_string(itPH) + " PT=" + std::to_string(itPT); out->status = -4; return -4; }
                if (err > 1.0e10) { g_err = "Kaboom! Error > 1e10 in inner loop PH=" + std::to_string(itPH) + " PT=" + std::to_string(itPT); out->status = -5; return -5; }
                double lmin[3];
                for (int c = 0; c < 3; ++c) lmin[c] = stokes_lmin(al[c], s->rate[c], s->Rv[c], s->Rv0[c], s->PC[c], nv);
                if (!freeze) {
                    ++jac_assemblies;
                    orc_augmented_jacobian3(s, dt, finite_K);
                    for (int c = 0; c < 3; ++c)
                        if (!checked_lmax(s->dRdv[c], s->PC[c], nv, labels[c], &lmax[c])) { out->status = -6; return -6; }
                    lmax_g = std::max(lmax[0], std::max(lmax[1], lmax[2]));
                }
                for (int c = 0; c < 3; ++c) {
                    dtau[c] = 2 / std::sqrt(lmax[c]) * p->CFL_v;
                    stokes_cheb(dtau[c], lmin[c], p->c_fact, &al[c], &be[c]);
                }
            }
            if (iter > p->max_iterations) break;
        }
        for (int64_t i = 0; i < np; ++i) s->P[i] += s->gammaP[i] * s->RP[i] / s->MP[i];
        if (iter > p->max_iterations) break;
    }
    out->itPH = itPH_done;
    out->iter = iter;
    out->err = err;
    out->err_abs = err_abs;
    out->err_rel = err_rel;
    out->err_v = err_v;
    out->err_P = err_P;
    out->converged = converged;
    out->reached_total_iter = iter > p->max_iterations;
    out->lambda_max = std::max(lmax[0], std::max(lmax[1], lmax[2]));
    out->lambda_max_gershgorin = lmax_g;
    out->jacobian_assemblies = jac_assemblies;
    out->status = 0;
    return 0;
}

int orc_stokes_solve3(orc_stokes3* s, const orc_params* p, orc_stats* out) { return stokes_solve3_impl(s, p, out, nullptr); }

// assemble_shear_heating! (shear_heating.jl:18-103), 3-D: the stress power
// τ : ∇v (shear terms twice) less the elastic power τ : (τ − τ_old)/(2GΔt)
void orc_shear_heating3(const orc_stokes3* s, double dt, double* phi) {
    const Tables3& t = tables3();
    Rheo r{&s->mat, s->mat.eta, s->dp, dt};
    double invG[ORC_MAX_PHASES];
    for (int p = 0; p < s->mat.nphases; ++p) invG[p] = 1.0 / s->mat.G[p];
    const bool has_old = s->tau_old[0] != nullptr;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        int ph[27];
        for (int a = 0; a < 27; ++a) ph[a] = s->phases[e] - 1;
        for (int q = 0; q < 27; ++q) {
            double gv[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double acc = 0;
                    for (int a = 0; a < 27; ++a) acc += dNdx[q][a][j] * s->v[i][s->el2n[e * 27 + a]];
                    gv[i][j] = acc;
                }
            double to[6] = {0, 0, 0, 0, 0, 0};
            if (has_old)
                for (int k = 0; k < 6; ++k) to[k] = s->tau_old[k][e * 27 + q];
            double Pq = 0;
            for (int i = 0; i < 4; ++i) Pq += t.NqP[q][i] * s->P[e * 4 + i];
            double tau[6];
            stress_ip3<double>(gv, t.Nq[q], 27, ph, r, to, Pq, tau);
            const double inv2Gdt = interp_phase(t.Nq[q], 27, invG, ph) / (2 * dt);
            const double power = tau[0] * gv[0][0] + tau[1] * gv[1][1] + tau[2] * gv[2][2] +
                                 tau[3] * (gv[0][1] + gv[1][0]) + tau[4] * (gv[0][2] + gv[2][0]) +
                                 tau[5] * (gv[1][2] + gv[2][1]);
            const double elastic = (tau[0] * (tau[0] - to[0]) + tau[1] * (tau[1] - to[1]) + tau[2] * (tau[2] - to[2]) +
                                    2 * (tau[3] * (tau[3] - to[3]) + tau[4] * (tau[4] - to[4]) +
                                         tau[5] * (tau[5] - to[5]))) * inv2Gdt;
            phi[e * 27 + q] = power - elastic;
        }
    }
}

// solve_coupled! on a Hex27 state (thermal mesh: the same Hex27 nodes)
int orc_coupled_solve3(orc_thermal* t, orc_stokes3* s, const orc_params* p, const orc_coupled_params* cp,
                       orc_stats* out, orc_coupled_stats* cst) {
    std::memset(out, 0, sizeof *out);
    std::memset(cst, 0, sizeof *cst);
    g_err.clear();
    if (t->nnodes != s->nnodes || !s->T || !s->T0) {
        g_err = t->nnodes != s->nnodes ? "thermal mesh nodes must match the Stokes velocity-node layout"
                                       : "the 3-D state needs its T and T0 fields for a coupled solve";
        out->status = -2;
        return -2;
    }
    transfer_T3(s, t->T, const_cast<double*>(s->T));
    transfer_T3(s, t->T0, const_cast<double*>(s->T0));
    ThermalHook h;
    h.t = t;
    h.s3 = s;
    h.dt = p->dt;
    h.Tref = cp->Tref;
    h.eps = cp->tolerance_T;
    h.nout = p->check_interval;
    h.shear = cp->shear_heating != 0;
    const double* saved = t->source_ip;
    if (h.shear) {
        h.phi.assign((size_t)s->nels * 27, 0.0);
        t->source_ip = h.phi.data();
    }
    const int rc = stokes_solve3_impl(s, p, out, &h);
    t->source_ip = saved;
    cst->err_T = h.err_T;
    cst->thermal_iterations = h.iter;
    cst->thermal_converged = h.converged;
    if (rc == 0) out->converged = out->converged && h.converged;
    return rc;
}

}  // extern "C"

// ===========================================================================
// CellPressureStokesDR: the viscous 3-D Hex27 solver with the four cell-local
// pressure modes (1, ξ, η, ζ) (stokes_types.jl:545-613, DR.jl:960-1057)
// ===========================================================================
namespace {

// pressure_mode_table (momentum_residuals.jl:618-621): the modes at the points
struct Modes3 {
    double m[27][4];
    Modes3() {
        double xi[27], eta[27], zeta[27], w[27];
        hex27_rule(xi, eta, zeta, w);
        for (int q = 0; q < 27; ++q) { m[q][0] = 1; m[q][1] = xi[q]; m[q][2] = eta[q]; m[q][3] = zeta[q]; }
    }
};
const Modes3& modes3() { static Modes3 t; return t; }

}  // namespace

extern "C" {

// stokes_pressure_residual_3d_kernel! (pressure_residual.jl:337-366):
// RP[4e + m] = −Σ_q mode_m(q) (∇·v)_q dΩ_q
void orc_cp3_pressure_residual(const orc_stokes3* s, double* RP) {
    const Modes3& md = modes3();
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        static thread_local double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        double r[4] = {0, 0, 0, 0};
        for (int q = 0; q < 27; ++q) {
            double div = 0;
            for (int a = 0; a < 27; ++a) {
                const int32_t g = s->el2n[e * 27 + a];
                div += dNdx[q][a][0] * s->v[0][g] + dNdx[q][a][1] * s->v[1][g] + dNdx[q][a][2] * s->v[2][g];
            }
            const double dOm = s->dOmega[e * 27 + q];
            for (int m = 0; m < 4; ++m) r[m] -= md.m[q][m] * (div * dOm);
        }
        for (int m = 0; m < 4; ++m) RP[e * 4 + m] = r[m];
    }
}

// stokes_momentum_residual_3d_kernel! (momentum_residuals.jl:645-691): the
// viscous residual ∫ ∇N·(2η ε̇′ − P I) − N ρ0 g with P from the modes; the
// element contributions added in ascending element order
void orc_cp3_momentum_residual(const orc_stokes3* s, double* const* R) {
    const Tables3& t = tables3();
    const Modes3& md = modes3();
    std::vector<double> Re((size_t)s->nels * 27 * 3);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        static thread_local double dNdx[27][27][3];
        elem_dNdx3(s, e, dNdx);
        const int p = s->phases[e] - 1;
        const double eta = s->mat.eta[p], rho = s->mat.rho0[p];
        double* re = &Re[(size_t)e * 81];
        for (int i = 0; i < 81; ++i) re[i] = 0;
        for (int q = 0; q < 27; ++q) {
            double G[3][3] = {};
            for (int a = 0; a < 27; ++a) {
                const int32_t g = s->el2n[e * 27 + a];
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < 3; ++k) G[c][k] += dNdx[q][a][k] * s->v[c][g];
            }
            const double tr = (G[0][0] + G[1][1] + G[2][2]) / 3.0;
            double Pq = 0;
            for (int m = 0; m < 4; ++m) Pq += md.m[q][m] * s->P[e * 4 + m];
            double sig[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) sig[i][j] = eta * (G[i][j] + G[j][i]) - (i == j ? 2 * eta * tr + Pq : 0.0);
            const double dOm = s->dOmega[e * 27 + q];
            for (int a = 0; a < 27; ++a)
                for (int i = 0; i < 3; ++i) {
                    double acc = 0;
                    for (int j = 0; j < 3; ++j) acc += dNdx[q][a][j] * sig[i][j];
                    re[a * 3 + i] += (acc - t.Nq[q][a] * rho * s->g[i]) * dOm;
                }
        }
    }
    for (int c = 0; c < 3; ++c) std::fill(R[c], R[c] + s->nnodes, 0.0);
    for (int64_t e = 0; e < s->nels; ++e)
        for (int a = 0; a < 27; ++a)
            for (int c = 0; c < 3; ++c) R[c][s->el2n[e * 27 + a]] += Re[(size_t)e * 81 + a * 3 + c];
}

// stokes_preconditioner_3d_kernel! (momentum_residuals.jl:724-745): Jacobi
// diagonal η (|∇N_a|² + (∂N_a/∂x_c)²/3) dΩ and the lumped mode mass Σ mode² dΩ
void orc_cp3_preconditioner(const orc_stokes3* s, double* const* diag, double* M) {
    const Modes3& md = modes3();
    for (int c = 0; c < 3; ++c) std::fill(diag[c], diag[c] + s->nnodes, 0.0);
    static double dNdx[27][27][3];
    for (int64_t e = 0; e < s->nels; ++e) {
        elem_dNdx3(s, e, dNdx);
        const double eta = s->mat.eta[s->phases[e] - 1];
        double m4[4] = {0, 0, 0, 0};
        for (int q = 0; q < 27; ++q) {
            const double dOm = s->dOmega[e * 27 + q];
            for (int a = 0; a < 27; ++a) {
                const double* gr = dNdx[q][a];
                const double n2 = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
                for (int c = 0; c < 3; ++c) diag[c][s->el2n[e * 27 + a]] += eta * (n2 + gr[c] * gr[c] / 3) * dOm;
            }
            for (int m = 0; m < 4; ++m) m4[m] += md.m[q][m] * md.m[q][m] * dOm;
        }
        for (int m = 0; m < 4; ++m) M[e * 4 + m] = m4[m];
    }
}

// _relax_cell_pressure_stokes! (DR.jl:997-1057): Arrow-Hurwicz pressure step
// with the mean of the constant mode removed, Jacobi velocity step,
// Dirichlet values before the first residual and after every update, the
// absolute test max(max_c ‖Rv_c‖, ‖RP‖) < tolerance every check_interval
// iterations and at the last. s->PC holds the diagonal, s->MP the mode mass.
int orc_cp3_solve(orc_stokes3* s, const orc_cp3_params* p, orc_cp3_stats* out, double* history,
                  int64_t history_cap, int64_t* history_len) {
    std::memset(out, 0, sizeof *out);
    g_err.clear();
    const int64_t nv = s->nnodes, np = 4 * s->nels;
    orc_cp3_preconditioner(s, s->PC, s->MP);
    for (int c = 0; c < 3; ++c) dirichlet(s->v[c], s->nbc[c], s->bc_dofs[c], s->bc_vals[c]);
    double sumM = 0;
    for (int64_t e = 0; e < s->nels; ++e) sumM += s->MP[e * 4];
    double err_v = kInf, err_P = kInf, err = kInf;
    int64_t iter = 0, nh = 0;
    for (int64_t it = 1; it <= p->max_iterations; ++it) {
        iter = it;
        orc_cp3_pressure_residual(s, s->RP);
        for (int64_t i = 0; i < np; ++i) s->P[i] += p->pressure_step * s->RP[i] / s->MP[i];
        double pm = 0;
        for (int64_t e = 0; e < s->nels; ++e) pm += s->P[e * 4] * s->MP[e * 4];
        pm /= sumM;
        for (int64_t e = 0; e < s->nels; ++e) s->P[e * 4] -= pm;
        orc_cp3_momentum_residual(s, s->Rv);
        for (int c = 0; c < 3; ++c)
            if (s->body_force[c])
                for (int64_t i = 0; i < nv; ++i) s->Rv[c][i] -= s->body_force[c][i];
        for (int c = 0; c < 3; ++c) {
            dirichlet(s->Rv[c], s->nbc[c], s->bc_dofs[c], nullptr);
            for (int64_t i = 0; i < nv; ++i) s->v[c][i] -= p->velocity_step * s->Rv[c][i] / s->PC[c][i];
            dirichlet(s->v[c], s->nbc[c], s->bc_dofs[c], s->bc_vals[c]);
        }
        if (!(it % p->check_interval == 0 || it == p->max_iterations)) continue;
        err_v = std::max(nrm2(s->Rv[0], nv), std::max(nrm2(s->Rv[1], nv), nrm2(s->Rv[2], nv)));
        err_P = nrm2(s->RP, np);
        err = std::max(err_v, err_P);
        if (!std::isfinite(err)) {
            g_err = "non-finite residual in 3-D Stokes solve";
            out->status = -4;
            return -4;
        }
        if (history && nh < history_cap) {
            history[nh * 3] = (double)it; history[nh * 3 + 1] = err_v; history[nh * 3 + 2] = err_P;
            ++nh;
        }
        if (p->verbose) printf("iter = %06lld err = %.3e - norm[Rv=%.3e, Rp=%.3e]\n", (long long)it, err, err_v, err_P);
        if (err < p->tolerance) break;
    }
    if (history_len) *history_len = nh;
    out->converged = err < p->tolerance;
    out->iterations = iter;
    out->residual = err;
    out->err_v = err_v;
    out->err_P = err_P;
    out->reached_total_iter = iter >= p->max_iterations && !out->converged;
    out->status = 0;
    return 0;
}

}  // extern "C"

// The thermal residual on a Hex27 mesh (the 27-node QuadraticElement{3,27} with
// the 3x3x3 Gauss rule, heat residual.jl:58-87 and the assembler of
// residual.jl:90-121): t->invJ holds [nels][27][9] (∂ξ/∂x row-major per
// point) and t->dOmega [nels][27], as orc_hex27_geometry writes them.
static void thermal_residual3(orc_thermal* t, double dt, double Tref, int32_t compute_jacobian) {
    const Tables3& tb = tables3();
    constexpr int NN = 27, NQ3 = 27;
    for (int64_t n = 0; n < t->nnodes; ++n) t->R[n] = 0;
    if (compute_jacobian)
        for (int64_t n = 0; n < t->nnodes; ++n) t->dRdT[n] = t->PC[n] = 0;
    static thread_local std::vector<double> J(NN * NN);
    double N[NQ3][NN], dNdx[NQ3][NN][3], dOm[NQ3], Tl[NN], T0l[NN], sl[NN], Pl[NN], Re[NN];
    int32_t ph[NN];
    for (int q = 0; q < NQ3; ++q)
        for (int a = 0; a < NN; ++a) N[q][a] = tb.Nq[q][a];
    for (int64_t e = 0; e < t->nels; ++e) {
        const int32_t* nd = t->el2n + e * NN;
        for (int q = 0; q < NQ3; ++q) {
            const double* iJ = t->invJ + (e * NQ3 + q) * 9;
            for (int a = 0; a < NN; ++a)
                for (int d = 0; d < 3; ++d) {
                    double v = 0;
                    for (int k = 0; k < 3; ++k) v += tb.dNdxi[q][a][k] * iJ[k * 3 + d];
                    dNdx[q][a][d] = v;
                }
            dOm[q] = t->dOmega[e * NQ3 + q];
        }
        for (int a = 0; a < NN; ++a) {
            Tl[a] = t->T[nd[a]]; T0l[a] = t->T0[nd[a]]; sl[a] = t->source[nd[a]]; Pl[a] = t->P[nd[a]];
            ph[a] = t->phases[nd[a]];
        }
        thermal_integrate(NN, NQ3, &N[0][0], &dNdx[0][0][0], dOm, Tl, T0l, sl, ph, &t->mat, Pl, dt, Tref,
                          t->source_ip ? t->source_ip + e * NQ3 : nullptr, Re, compute_jacobian ? J.data() : nullptr,
                          3);
        for (int a = 0; a < NN; ++a) t->R[nd[a]] += Re[a];
        if (compute_jacobian) {
            for (int a = 0; a < NN; ++a) {
                double rs = 0;
                for (int b = 0; b < NN; ++b) rs += std::fabs(J[a * NN + b]);
                t->dRdT[nd[a]] += rs;
                t->PC[nd[a]] += std::fabs(J[a * NN + a]);
            }
        }
    }
}

// ===========================================================================
// 2-D: ThermalDiffusionDR on Q9 (QuadraticElement{2, 9}) meshes
// ===========================================================================
namespace {

// QuadraticElement{2, 9} local node order (shape_functions.jl:259-283)
const double kQ9Nodes[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}, {0, 0}};
// grid offsets of the local nodes within a cell (connectivity.jl:193-227)
const int kQ9Grid[9][2] = {{0, 0}, {2, 0}, {2, 2}, {0, 2}, {1, 0}, {2, 1}, {1, 2}, {0, 1}, {1, 1}};

// tensor product of the 1-D quadratic basis (_tensor_product_shape_functions)
void q9_basis(double xi, double eta, double N[9], double dN[9][2]) {
    for (int a = 0; a < 9; ++a) {
        const double fx = lq(xi, kQ9Nodes[a][0]), fy = lq(eta, kQ9Nodes[a][1]);
        N[a] = fx * fy;
        dN[a][0] = dlq(xi, kQ9Nodes[a][0]) * fy;
        dN[a][1] = fx * dlq(eta, kQ9Nodes[a][1]);
    }
}

// IntegrationPoints(::QuadraticElement{2, 9}) (integration_points.jl:218-241):
// the weights are the literals 25/81, 40/81, 64/81, not products of 1-D weights
void q9_rule(double xi[9], double eta[9], double w[9]) {
    const double a = std::sqrt(3.0 / 5.0);
    const double pts[3] = {-a, 0.0, a};
    const double ws[9] = {25.0 / 81, 40.0 / 81, 25.0 / 81, 40.0 / 81, 64.0 / 81, 40.0 / 81, 25.0 / 81, 40.0 / 81,
                          25.0 / 81};
    for (int q = 0; q < 9; ++q) {
        xi[q] = pts[q % 3];
        eta[q] = pts[q / 3];
        w[q] = ws[q];
    }
}

struct TablesQ9 {
    double Nq[9][9], dNdxi[9][9][2], w[9];
    TablesQ9() {
        double xi[9], eta[9];
        q9_rule(xi, eta, w);
        for (int q = 0; q < 9; ++q) q9_basis(xi[q], eta[q], Nq[q], dNdxi[q]);
    }
};
const TablesQ9& tables_q9() { static TablesQ9 t; return t; }

}  // namespace

extern "C" {

void orc_q9_basis(double xi, double eta, double* N, double* dN) {
    q9_basis(xi, eta, N, reinterpret_cast<double(*)[2]>(dN));
}
void orc_q9_rule(double* xi, double* eta, double* w) { q9_rule(xi, eta, w); }

void orc_q9_mesh_sizes(int32_t nx, int32_t ny, int64_t* nnodes, int64_t* nels) {
    *nnodes = (int64_t)(2 * nx + 1) * (2 * ny + 1);
    *nels = (int64_t)nx * ny;
}

// _tensor_grid_coordinates with LinRange(x0, x1, 2nx + 1): (1 - t) x0 + t x1
// with t = i/(2nx); the element loop has ex fastest
void orc_q9_mesh(int32_t nx, int32_t ny, double x0, double x1, double y0, double y1, double* coords,
                 int32_t* el2n) {
    const int64_t sx = 2 * (int64_t)nx + 1, sy = 2 * (int64_t)ny + 1;
    for (int64_t iy = 0; iy < sy; ++iy)
        for (int64_t ix = 0; ix < sx; ++ix) {
            const int64_t n = iy * sx + ix;
            const double tx = (double)ix / (double)(sx - 1), ty = (double)iy / (double)(sy - 1);
            coords[2 * n + 0] = (1 - tx) * x0 + tx * x1;
            coords[2 * n + 1] = (1 - ty) * y0 + ty * y1;
        }
    int64_t e = 0;
    for (int64_t ey = 0; ey < ny; ++ey)
        for (int64_t ex = 0; ex < nx; ++ex, ++e)
            for (int a = 0; a < 9; ++a)
                el2n[e * 9 + a] = (int32_t)((2 * ey + kQ9Grid[a][1]) * sx + 2 * ex + kQ9Grid[a][0]);
}

void orc_q9_geometry(int64_t nels, const double* coords, const int32_t* el2n, double* invJ, double* dOmega) {
    const TablesQ9& t = tables_q9();
    for (int64_t e = 0; e < nels; ++e)
        for (int q = 0; q < 9; ++q) {
            double J[2][2] = {{0, 0}, {0, 0}};   // J[i][j] = ∂x_i/∂ξ_j
            for (int a = 0; a < 9; ++a) {
                const double* x = coords + 2 * (int64_t)el2n[e * 9 + a];
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) J[i][j] += x[i] * t.dNdxi[q][a][j];
            }
            const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            double* iJ = invJ + (e * 9 + q) * 4;   // (J^-1)[j][i] = ∂ξ_j/∂x_i
            iJ[0] = J[1][1] / det;
            iJ[1] = -J[0][1] / det;
            iJ[2] = -J[1][0] / det;
            iJ[3] = J[0][0] / det;
            dOmega[e * 9 + q] = std::fabs(det) * t.w[q];
        }
}

}  // extern "C"

// assemble_diffusion_matrices_atomix! on a Q9 mesh: integrate_residual
// (residual.jl:58-87) at the nine Gauss points with the element's precomputed
// J^-1 and |det J| w, nodal contributions added in ascending element order
static void thermal_residual_q9(orc_thermal* t, double dt, double Tref, int32_t compute_jacobian) {
    const TablesQ9& tb = tables_q9();
    constexpr int NN = 9, NQ9 = 9;
    for (int64_t n = 0; n < t->nnodes; ++n) t->R[n] = 0;
    if (compute_jacobian)
        for (int64_t n = 0; n < t->nnodes; ++n) t->dRdT[n] = t->PC[n] = 0;
    double N[NQ9][NN], dNdx[NQ9][NN][2], dOm[NQ9], Tl[NN], T0l[NN], sl[NN], Pl[NN], Re[NN], J[NN * NN];
    int32_t ph[NN];
    for (int q = 0; q < NQ9; ++q)
        for (int a = 0; a < NN; ++a) N[q][a] = tb.Nq[q][a];
    for (int64_t e = 0; e < t->nels; ++e) {
        const int32_t* nd = t->el2n + e * NN;
        for (int q = 0; q < NQ9; ++q) {
            const double* iJ = t->invJ + (e * NQ9 + q) * 4;
            for (int a = 0; a < NN; ++a) {
                dNdx[q][a][0] = tb.dNdxi[q][a][0] * iJ[0] + tb.dNdxi[q][a][1] * iJ[2];
                dNdx[q][a][1] = tb.dNdxi[q][a][0] * iJ[1] + tb.dNdxi[q][a][1] * iJ[3];
            }
            dOm[q] = t->dOmega[e * NQ9 + q];
        }
        for (int a = 0; a < NN; ++a) {
            Tl[a] = t->T[nd[a]]; T0l[a] = t->T0[nd[a]]; sl[a] = t->source[nd[a]]; Pl[a] = t->P[nd[a]];
            ph[a] = t->phases[nd[a]];
        }
        thermal_integrate(NN, NQ9, &N[0][0], &dNdx[0][0][0], dOm, Tl, T0l, sl, ph, &t->mat, Pl, dt, Tref,
                          t->source_ip ? t->source_ip + e * NQ9 : nullptr, Re, compute_jacobian ? J : nullptr);
        for (int a = 0; a < NN; ++a) t->R[nd[a]] += Re[a];
        if (compute_jacobian) {
            for (int a = 0; a < NN; ++a) {
                double rs = 0;
                for (int b = 0; b < NN; ++b) rs += std::fabs(J[a * NN + b]);
                t->dRdT[nd[a]] += rs;
                t->PC[nd[a]] += std::fabs(J[a * NN + a]);
            }
        }
    }
}

// ===========================================================================
// 2-D: the frozen-operator adjoint of StokesDR (operator = :blocks)
// ===========================================================================
namespace {

constexpr int ANU = 14, ANP = 3;

// _symmetric_tangent (adjoint_operator.jl:172-182): no yield model and K = Inf
bool adjoint_symmetric(const orc_stokes* s) {
    if (s->dp || s->cap) return false;
    for (int p = 0; p < s->mat.nphases; ++p)
        if (!std::isinf(s->mat.K[p])) return false;
    return true;
}
// D and W are kept only where some phase has a finite bulk viscosity (:455-458)
bool adjoint_stores(const orc_stokes* s) {
    for (int p = 0; p < s->mat.nphases; ++p)
        if (std::isfinite(s->mat.etab[p])) return true;
    return false;
}

// element_adjoint_operator_blocks (adjoint_operator.jl:318-373): A is the
// augmented Jacobian (the element-maximum viscosity, Pnum inline with ηb);
// B, C, D and ∂Rv/∂Pnum are forward-mode derivatives of the element residuals
void adjoint_element(const orc_stokes* s, int64_t e, double dt, double A[ANU][ANU], double B[ANU][ANP],
                     double C[ANP][ANU], double D[ANP][ANP], double W[ANU][ANP]) {
    const Tables& t = tables();
    double block[2][2][7][7];
    augmented_element_block(s, e, dt, s->mat.etab, block);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 7; ++r)
                for (int c = 0; c < 7; ++c) A[i * 7 + r][j * 7 + c] = block[i][j][r][c];
    double dNdx[7][7][2];
    elem_dNdx(s, e, dNdx);
    int ph[7], phP[3];
    elem_phases_v(s, e, ph);
    elem_phases_P(s, e, phP);
    double vx[7], vy[7];
    for (int a = 0; a < 7; ++a) { vx[a] = s->vx[s->el2n[e * 7 + a]]; vy[a] = s->vy[s->el2n[e * 7 + a]]; }
    double to[21];
    const bool has_old = s->tau_old[0] || s->tau_old[1] || s->tau_old[2];
    for (int c = 0; c < 3; ++c)
        for (int q = 0; q < 7; ++q) to[c * 7 + q] = s->tau_old[c] ? s->tau_old[c][e * 7 + q] : 0.0;
    // B and W differentiate the residual with the state's per-phase viscosity
    // and yield model (the momentum closure of adjoint_operator.jl:338-345)
    Rheo r{&s->mat, s->mat.eta, s->dp, dt};
    r.eta_ip = s->eta_ip ? s->eta_ip + e * 7 : nullptr;
    const double* Pf = s->Pf ? s->Pf + e * 3 : nullptr;
    typedef Dual<3> D3;
    D3 ux[7], uy[7];
    for (int a = 0; a < 7; ++a) { ux[a] = D3(vx[a]); uy[a] = D3(vy[a]); }
    {   // B = ∂Rv/∂P with Pnum = 0 held as an independent variable
        D3 Pd[3], Pn[3];
        for (int i = 0; i < 3; ++i) { Pd[i] = D3(s->P[e * 3 + i]); Pd[i].d[i] = 1.0; }
        D3 Rx[7], Ry[7];
        momentum_element<D3, D3>(7, 3, 7, &dNdx[0][0][0], s->dOmega + e * 7, &t.Nq[0][0], &t.NqP[0][0], ux, uy, Pd,
                                 Pn, s->T + e * 3, Pf, ph, r, has_old ? to : nullptr, Rx, Ry, nullptr, nullptr);
        for (int a = 0; a < 7; ++a)
            for (int i = 0; i < 3; ++i) { B[a][i] = Rx[a].d[i]; B[7 + a][i] = Ry[a].d[i]; }
    }
    {   // W = ∂Rv/∂Pnum column-scaled by Γ = γP/M_P
        D3 Pn[3];
        for (int i = 0; i < 3; ++i) Pn[i].d[i] = 1.0;
        D3 Rx[7], Ry[7];
        momentum_element<D3>(7, 3, 7, &dNdx[0][0][0], s->dOmega + e * 7, &t.Nq[0][0], &t.NqP[0][0], ux, uy,
                             s->P + e * 3, Pn, s->T + e * 3, Pf, ph, r, has_old ? to : nullptr, Rx, Ry, nullptr,
                             nullptr);
        for (int i = 0; i < 3; ++i) {
            const double G = s->gammaP[e * 3 + i] / s->MP[e * 3 + i];
            for (int a = 0; a < 7; ++a) { W[a][i] = Rx[a].d[i] * G; W[7 + a][i] = Ry[a].d[i] * G; }
        }
    }
    {   // C = ∂RP/∂v
        typedef Dual<ANU> DU;
        DU wx[7], wy[7];
        for (int a = 0; a < 7; ++a) {
            wx[a] = DU(vx[a]); wx[a].d[a] = 1.0;
            wy[a] = DU(vy[a]); wy[a].d[7 + a] = 1.0;
        }
        DU RP[3];
        pressure_element<DU>(7, 3, 7, &dNdx[0][0][0], s->dOmegaP + e * 7, &t.NqP[0][0], wx, wy, s->P + e * 3,
                             s->P0 + e * 3, s->T + e * 3, s->T0 + e * 3, nullptr, phP, s->mat.alpha, s->mat.etab,
                             s->mat.nphases, dt, RP);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < ANU; ++k) C[i][k] = RP[i].d[k];
    }
    {   // D = ∂RP/∂P: the storage term −∫ N_i N_j /(ηb Δt)
        D3 Pd[3];
        for (int i = 0; i < 3; ++i) { Pd[i] = D3(s->P[e * 3 + i]); Pd[i].d[i] = 1.0; }
        D3 RP[3];
        pressure_element<D3, D3>(7, 3, 7, &dNdx[0][0][0], s->dOmegaP + e * 7, &t.NqP[0][0], ux, uy, Pd,
                                 s->P0 + e * 3, s->T + e * 3, s->T0 + e * 3, nullptr, phP, s->mat.alpha,
                                 s->mat.etab, s->mat.nphases, dt, RP);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[i][j] = RP[i].d[j];
    }
}

// the velocity part of one element's apply: Aᵀλ (the packed upper triangle when
// symmetric, _symmetric_matvec) plus the pressure coupling (B when symmetric,
// else Cᵀ), and its pressure part Bᵀλv + Dᵀ(Wᵀλv + λP)
void adjoint_element_apply(const double* A, const double* B, const double* C, const double* D, const double* W,
                           bool sym, bool stores, const double lv[ANU], const double lp[ANP], double resv[ANU],
                           double resp[ANP]) {
    for (int r = 0; r < ANU; ++r) {
        double a = 0, b = 0;
        for (int k = 0; k < ANU; ++k)
            a += (sym ? A[std::min(r, k) * ANU + std::max(r, k)] : A[k * ANU + r]) * lv[k];
        for (int i = 0; i < ANP; ++i) b += (sym ? B[r * ANP + i] : C[i * ANU + r]) * lp[i];
        resv[r] = a + b;
    }
    for (int i = 0; i < ANP; ++i) {
        double b = 0;
        for (int k = 0; k < ANU; ++k) b += B[k * ANP + i] * lv[k];
        double st = 0;
        if (stores) {
            double w[ANP];
            for (int j = 0; j < ANP; ++j) {
                double acc = 0;
                for (int k = 0; k < ANU; ++k) acc += W[k * ANP + j] * lv[k];
                w[j] = acc + lp[j];
            }
            for (int j = 0; j < ANP; ++j) st += D[j * ANP + i] * w[j];
        }
        resp[i] = b + st;
    }
}

void adjoint_apply_impl(const orc_stokes* s, const double* A, const double* B, const double* C, const double* D,
                        const double* W, bool sym, bool stores, const double* lx, const double* ly, const double* lP,
                        double* dvx, double* dvy, double* dP) {
    std::vector<double> Rx((size_t)s->nels * 7), Ry((size_t)s->nels * 7);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        double lv[ANU], lp[ANP], rv[ANU], rp[ANP];
        for (int a = 0; a < 7; ++a) { lv[a] = lx[s->el2n[e * 7 + a]]; lv[7 + a] = ly[s->el2n[e * 7 + a]]; }
        for (int i = 0; i < ANP; ++i) lp[i] = lP[e * 3 + i];
        adjoint_element_apply(A + e * ANU * ANU, B + e * ANU * ANP, C ? C + e * ANP * ANU : nullptr,
                              D ? D + e * ANP * ANP : nullptr, W ? W + e * ANU * ANP : nullptr, sym, stores, lv, lp,
                              rv, rp);
        for (int a = 0; a < 7; ++a) { Rx[e * 7 + a] = rv[a]; Ry[e * 7 + a] = rv[7 + a]; }
        for (int i = 0; i < ANP; ++i) dP[e * 3 + i] = rp[i];   // discontinuous pressure: unshared
    }
    gather(s, Rx, dvx);
    gather(s, Ry, dvy);
}

double power_seed(int64_t i, int d) { return d == 1 ? std::sin(0.7 * i) : d == 2 ? std::cos(1.3 * i) : std::sin(1.9 * i); }

}  // namespace

extern "C" {

int orc_adjoint_blocks(orc_stokes* s, double dt, double* A, double* B, double* C, double* D, double* W,
                       double* defects) {
    g_err.clear();
    if (s->cap) {
        g_err = "the frozen-operator adjoint covers models without a yield model or with the Drucker-Prager model";
        return -1;
    }
    std::vector<double> Jx((size_t)s->nels * 7), Jy((size_t)s->nels * 7), Dx((size_t)s->nels * 7),
        Dy((size_t)s->nels * 7), dA(s->nels), dC(s->nels);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < s->nels; ++e) {
        double Ae[ANU][ANU], Be[ANU][ANP], Ce[ANP][ANU], De[ANP][ANP], We[ANU][ANP];
        adjoint_element(s, e, dt, Ae, Be, Ce, De, We);
        std::memcpy(A + e * ANU * ANU, Ae, sizeof Ae);
        std::memcpy(B + e * ANU * ANP, Be, sizeof Be);
        std::memcpy(C + e * ANP * ANU, Ce, sizeof Ce);
        std::memcpy(D + e * ANP * ANP, De, sizeof De);
        std::memcpy(W + e * ANU * ANP, We, sizeof We);
        double nA = 0, nAs = 0, nC = 0, nCs = 0;
        for (int i = 0; i < ANU; ++i)
            for (int j = 0; j < ANU; ++j) { nA += Ae[i][j] * Ae[i][j]; nAs += sq(Ae[i][j] - Ae[j][i]); }
        for (int i = 0; i < ANP; ++i)
            for (int k = 0; k < ANU; ++k) { nC += Ce[i][k] * Ce[i][k]; nCs += sq(Ce[i][k] - Be[k][i]); }
        dA[e] = nA == 0 ? 0.0 : std::sqrt(nAs) / std::sqrt(nA);
        dC[e] = nC == 0 ? 0.0 : std::sqrt(nCs) / std::sqrt(nC);
        // _scatter_velocity_diagnostics! (adjoint_operator.jl:390-403)
        for (int a = 0; a < 7; ++a) {
            double sx = 0, sy = 0;
            for (int k = 0; k < ANU; ++k) { sx += std::fabs(Ae[a][k]); sy += std::fabs(Ae[7 + a][k]); }
            Jx[e * 7 + a] = sx; Jy[e * 7 + a] = sy;
            Dx[e * 7 + a] = std::fabs(Ae[a][a]);
            Dy[e * 7 + a] = std::fabs(Ae[7 + a][7 + a]);
        }
    }
    gather(s, Jx, s->dRdvx);
    gather(s, Jy, s->dRdvy);
    gather(s, Dx, s->PCx);
    gather(s, Dy, s->PCy);
    if (defects) {
        defects[0] = defects[1] = 0;
        for (int64_t e = 0; e < s->nels; ++e) {
            defects[0] = std::max(defects[0], dA[e]);
            defects[1] = std::max(defects[1], dC[e]);
        }
    }
    return 0;
}

void orc_adjoint_apply(const orc_stokes* s, const double* A, const double* B, const double* C, const double* D,
                       const double* W, int32_t symmetric, int32_t stores, const double* lx, const double* ly,
                       const double* lP, double* dvx, double* dvy, double* dP) {
    adjoint_apply_impl(s, A, B, C, D, W, symmetric != 0, stores != 0, lx, ly, lP, dvx, dvy, dP);
}

int orc_adjoint_solve(orc_stokes* s, const orc_adjoint_params* p, const double* obj_x, const double* obj_y,
                      double* lam_x, double* lam_y, double* lam_P, orc_adjoint_stats* st) {
    std::memset(st, 0, sizeof *st);
    g_err.clear();
    const int64_t nv = s->nnodes, np = s->nnodesP, ne = s->nels;
    bool assembled = false;
    for (int64_t i = 0; i < np && !assembled; ++i) assembled = s->gammaP[i] != 0;
    if (!assembled) {
        g_err = "dr.γP is unassembled; run the forward solve! on this state before solve_adjoint!";
        st->status = -1;
        return -1;
    }
    const bool sym = adjoint_symmetric(s), stores = adjoint_stores(s);
    std::vector<double> A((size_t)ne * ANU * ANU), B((size_t)ne * ANU * ANP), C((size_t)ne * ANP * ANU),
        D((size_t)ne * ANP * ANP), W((size_t)ne * ANU * ANP);
    double defects[2];
    if (orc_adjoint_blocks(s, p->dt, A.data(), B.data(), C.data(), D.data(), W.data(), defects) != 0) {
        st->status = -1;
        return -1;
    }
    // _assert_frozen_symmetry (adjoint_operator.jl:149-160)
    const double tol_sym = std::sqrt(std::numeric_limits<double>::epsilon());
    if (sym && (defects[0] > tol_sym || defects[1] > tol_sym)) {
        char buf[320];
        std::snprintf(buf, sizeof buf,
                      "the frozen adjoint operator assumes the element tangent is symmetric, but the assembled blocks "
                      "disagree: max ‖A - Aᵀ‖/‖A‖ = %g, max ‖C - Bᵀ‖/‖C‖ = %g over %lld elements",
                      defects[0], defects[1], (long long)ne);
        g_err = buf;
        st->status = -1;
        return -1;
    }
    auto apply = [&](const double* lx, const double* ly, const double* lP, double* dvx, double* dvy, double* dP) {
        adjoint_apply_impl(s, A.data(), B.data(), sym ? nullptr : C.data(), stores ? D.data() : nullptr,
                           stores ? W.data() : nullptr, sym, stores, lx, ly, lP, dvx, dvy, dP);
    };
    double* PC[2] = {s->PCx, s->PCy};
    double lmax_c[2];
    const char* labels[2] = {"adjoint vx", "adjoint vy"};
    for (int c = 0; c < 2; ++c)
        if (!checked_lmax(c == 0 ? s->dRdvx : s->dRdvy, PC[c], nv, labels[c], &lmax_c[c])) {
            st->status = -6;
            return -6;
        }
    st->lmax_gershgorin = std::max(lmax_c[0], lmax_c[1]);
    double lmax_v[2] = {lmax_c[0], lmax_c[1]};
    std::vector<double> x[2], z[2], y[2], zP(np, 0.0), dPs(np);
    auto project = [&](std::vector<double>* u) {
        for (int c = 0; c < 2; ++c) dirichlet(u[c].data(), s->nbc[c], s->bc_dofs[c], nullptr);
    };
    auto stacked_norm = [&](std::vector<double>* u) {
        double a = 0, b = 0;
        for (int64_t i = 0; i < nv; ++i) a += u[0][i] * u[0][i];
        for (int64_t i = 0; i < nv; ++i) b += u[1][i] * u[1][i];
        return std::sqrt(a + b);
    };
    if (p->measure_lmax) {
        // estimate_adjoint_λmax (adjoint_operator.jl:564-607), 100 iterations, rtol 1e-3
        for (int c = 0; c < 2; ++c) {
            x[c].resize(nv); z[c].resize(nv); y[c].resize(nv);
            for (int64_t i = 0; i < nv; ++i) x[c][i] = power_seed(i + 1, c + 1);
        }
        project(x);
        double lam = 0;
        int64_t its = 0;
        for (int64_t it = 1; it <= 100; ++it) {
            its = it;
            const double nx = stacked_norm(x);
            if (!(nx > 0)) {
                g_err = "power iteration collapsed to the zero vector";
                st->status = -1;
                return -1;
            }
            for (int c = 0; c < 2; ++c)
                for (int64_t i = 0; i < nv; ++i) x[c][i] /= nx;
            for (int c = 0; c < 2; ++c)
                for (int64_t i = 0; i < nv; ++i) z[c][i] = x[c][i] / std::sqrt(PC[c][i]);
            apply(z[0].data(), z[1].data(), zP.data(), y[0].data(), y[1].data(), dPs.data());
            for (int c = 0; c < 2; ++c)
                for (int64_t i = 0; i < nv; ++i) y[c][i] /= std::sqrt(PC[c][i]);
            project(y);
            const double ln = stacked_norm(y);
            x[0] = y[0];
            x[1] = y[1];
            const bool conv = it > 1 && std::fabs(ln - lam) <= 1e-3 * ln;
            lam = ln;
            if (conv) break;
        }
        if (!(std::isfinite(lam) && lam > 0)) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "power iteration produced a non-positive λmax: %g", lam);
            g_err = buf;
            st->status = -1;
            return -1;
        }
        lmax_v[0] = lmax_v[1] = lam;
        st->lmax_iterations = its;
    }
    st->lmax = lmax_v[0];
    double dtau[2], al[2], be[2];
    for (int c = 0; c < 2; ++c) {
        dtau[c] = 2 / std::sqrt(lmax_v[c]) * p->CFL_v;
        stokes_cheb(dtau[c], 0.0, p->c_fact, &al[c], &be[c]);
    }
    double* lam[2] = {lam_x, lam_y};
    const double* obj[2] = {obj_x, obj_y};
    std::vector<double> R[2], R0[2], rate[2], dv[2], RP(np);
    for (int c = 0; c < 2; ++c) {
        R[c].assign(nv, 0.0); R0[c].assign(nv, 0.0); rate[c].assign(nv, 0.0); dv[c].assign(nv, 0.0);
    }
    auto assemble = [&]() {
        apply(lam_x, lam_y, lam_P, dv[0].data(), dv[1].data(), RP.data());
        for (int c = 0; c < 2; ++c) {
            for (int64_t i = 0; i < nv; ++i) R[c][i] = obj[c][i] + dv[c][i];
            dirichlet(R[c].data(), s->nbc[c], s->bc_dofs[c], nullptr);
        }
    };
    auto velocity_error = [&]() {
        return std::max(nrm2(R[0].data(), nv), nrm2(R[1].data(), nv)) / std::sqrt((double)nv);
    };
    int64_t iter = 0, itPH_done = 0;
    double err = kInf, err_v0 = 1, err_P0 = 1, err_v = kInf, err_P = kInf;
    bool converged = false;
    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        assemble();
        err_v = velocity_error();
        err_P = nrm2(RP.data(), np) / std::sqrt((double)np);
        if (itPH == 1) {
            err_v0 = err_v + julia_eps(err_v);
            err_P0 = err_P + julia_eps(err_P);
        }
        err = std::max(std::min(err_v, err_v / err_v0), std::min(err_P, err_P / err_P0));
        if (p->verbose)
            printf("adj PH=%03lld iter=%06lld err=%.3e Rv=%.3e RP=%.3e\n", (long long)itPH, (long long)iter, err,
                   err_v, err_P);
        if (err < p->tolerance) {
            converged = true;
            break;
        }
        const double target = std::max(err_v * p->rel_drop, p->tolerance);
        int64_t inner = 0;
        while (err_v > target && inner < p->iter_max && iter < p->max_iterations) {
            ++inner;
            ++iter;
            for (int c = 0; c < 2; ++c) R0[c] = R[c];
            for (int c = 0; c < 2; ++c) {
                const double a = -al[c], b = be[c];
                for (int64_t i = 0; i < nv; ++i) {
                    const double nr = b * rate[c][i] + R[c][i] / PC[c][i];
                    rate[c][i] = nr;
                    lam[c][i] += a * nr;
                }
            }
            for (int c = 0; c < 2; ++c) {
                dirichlet(lam[c], s->nbc[c], s->bc_dofs[c], nullptr);
                dirichlet(rate[c].data(), s->nbc[c], s->bc_dofs[c], nullptr);
            }
            assemble();
            if (iter % p->check_interval == 0) {
                err_v = velocity_error();
                for (int c = 0; c < 2; ++c) {
                    const double lmin = stokes_lmin(al[c], rate[c].data(), R[c].data(), R0[c].data(), PC[c], nv);
                    stokes_cheb(dtau[c], lmin, p->c_fact, &al[c], &be[c]);
                }
            }
        }
        for (int64_t i = 0; i < np; ++i) lam_P[i] += s->gammaP[i] * RP[i] / s->MP[i];   // Arrow-Hurwicz
        if (iter >= p->max_iterations) break;
    }
    if (!converged) {
        assemble();
        err_v = velocity_error();
        err_P = nrm2(RP.data(), np) / std::sqrt((double)np);
        err = std::max(std::min(err_v, err_v / err_v0), std::min(err_P, err_P / err_P0));
    }
    st->itPH = itPH_done;
    st->iter = iter;
    st->err = err;
    st->err_v = err_v;
    st->err_P = err_P;
    st->converged = converged || err < p->tolerance;
    st->status = 0;
    return 0;
}

}  // extern "C"
