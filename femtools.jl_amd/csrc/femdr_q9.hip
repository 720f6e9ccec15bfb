// femdr_q9.hip -- ThermalDiffusionDR on structured Q9 (QuadraticElement{2, 9})
// meshes: the mesh constructor and the thermal element kernel behind
// fem_thermal_*; the pseudo-transient loop and the node gather are those of
// femdr_thermal.hip.
//
// Work layout as in the Hex27 thermal kernel, scaled to 2-D: 16 lanes serve
// one element (9 of them active), 16 elements per 256-thread block. Lanes
// load the nine nodal values (lane = local node), evaluate temperature,
// gradient and material at the nine Gauss points (lane = point), then form
// the nodal residuals and, on check iterations, the Jacobian row sums
// (lane = node). A structured box has one affine map: ∂ξ/∂x = diag(2/hx,
// 2/hy), |det J| = hx hy / 4. The eight shared local nodes write element
// slots e*8 + a, summed per node in ascending element order by the gather;
// the centre node (local node 9) is private and is finished by its element.
#include "femdr_host.hpp"

namespace {

constexpr int QNV = 9, QNQ = 9, QNSH = 8, QLANES = 16, QBLOCK = 256, QEPB = QBLOCK / QLANES;
constexpr int QCENTRE = 8;

// QuadraticElement{2, 9} local nodes as 1-D node indices per axis (0: -1,
// 1: 0, 2: +1; shape_functions.jl:259-283), which are also the refined-grid
// offsets of the nodes within their cell (connectivity.jl:193-227)
constexpr int8_t kQ9R[QNV][2] = {{0, 0}, {2, 0}, {2, 2}, {0, 2}, {1, 0}, {2, 1}, {1, 2}, {0, 1}, {1, 1}};

// IntegrationPoints(::QuadraticElement{2, 9}) weights, ξ fastest
// (integration_points.jl:218-241): the literals n/81, as the rule states them
constexpr double kQ9W[QNQ] = {25.0 / 81, 40.0 / 81, 25.0 / 81, 40.0 / 81, 64.0 / 81,
                              40.0 / 81, 25.0 / 81, 40.0 / 81, 25.0 / 81};

struct Q9Geo {
    int64_t nels;
    int32_t nx;
    int64_t sx;                      // refined-grid stride 2nx + 1
    double jac[2], det;              // ∂ξ/∂x per axis, |det J|
};

Q9Geo q9_geo(const fem_mesh* m) {
    Q9Geo g;
    g.nels = m->nels;
    g.nx = m->nx;
    g.sx = 2 * (int64_t)m->nx + 1;
    const double hx = (m->box[1] - m->box[0]) / m->nx, hy = (m->box[3] - m->box[2]) / m->ny;
    g.jac[0] = 2 / hx;
    g.jac[1] = 2 / hy;
    g.det = (hx / 2) * (hy / 2);
    return g;
}

__device__ inline int64_t q9_node(const Q9Geo& g, int64_t e, int a) {
    const int64_t ey = e / g.nx, ex = e - ey * g.nx;
    return (2 * ey + kQ9R[a][1]) * g.sx + 2 * ex + kQ9R[a][0];
}

// integrate_residual (heat_diffusion/assembly/residual.jl:58-87) for one Q9
// element, and on check iterations the row sums Σ_j |J_ij| and |J_ii| of its
// exact Jacobian (residual.jl:30-47 takes it with ForwardDiff): with
// D = Δt/(ρ Cp), ρ = ρ0 (1 - α (T - Tref) + β P), ∂D/∂T_j = D ρ0 α N_j / ρ.
// Nodal sums at a point run over the local nodes in element order.
template <bool JAC>
__global__ __launch_bounds__(QBLOCK) void k_thermal_q9(Q9Geo g, ThermBoxArgs A) {
    const int es = threadIdx.x / QLANES, l = threadIdx.x % QLANES;
    const int64_t e = xcd_tile() * QEPB + es;
    const bool ev = e < g.nels;
    __shared__ double sT[QEPB][QNV], sT0[QEPB][QNV], sS[QEPB][QNV], sP[QEPB][QNV];
    __shared__ double sK[QEPB][QNV], sCp[QEPB][QNV], sR0[QEPB][QNV], sAl[QEPB][QNV], sBe[QEPB][QNV];
    __shared__ double qRate[QEPB][QNQ], qD[QEPB][QNQ], qC[QEPB][QNQ], qK[QEPB][QNQ], qS[QEPB][QNQ];
    __shared__ double qW[QEPB][QNQ], qG[QEPB][QNQ][2];
    if (ev && l < QNV) {
        const int64_t n = q9_node(g, e, l);
        sT[es][l] = A.T[n]; sT0[es][l] = A.T0[n]; sS[es][l] = A.src[n]; sP[es][l] = A.P[n];
        const int p = A.phases[n] - 1;
        sK[es][l] = A.tab[TC_K * MAXPH + p]; sCp[es][l] = A.tab[TC_CP * MAXPH + p];
        sR0[es][l] = A.tab[TC_RHO0 * MAXPH + p]; sAl[es][l] = A.tab[TC_ALPHA * MAXPH + p];
        sBe[es][l] = A.tab[TC_BETA * MAXPH + p];
    }
    __syncthreads();
    if (ev && l < QNQ) {
        const int q = l;
        const double x0 = gauss_x(q % 3), x1 = gauss_x(q / 3);
        double Tq = 0, T0q = 0, sq = 0, Pq = 0, kq = 0, cq = 0, rq = 0, aq = 0, bq = 0, g0 = 0, g1 = 0;
        for (int a = 0; a < QNV; ++a) {
            const int i = kQ9R[a][0], j = kQ9R[a][1];
            const double L0 = lq(x0, i), L1 = lq(x1, j);
            const double N = L0 * L1;
            const double v = sT[es][a];
            Tq += N * v; T0q += N * sT0[es][a]; sq += N * sS[es][a]; Pq += N * sP[es][a];
            kq += N * sK[es][a]; cq += N * sCp[es][a]; rq += N * sR0[es][a]; aq += N * sAl[es][a];
            bq += N * sBe[es][a];
            g0 += dlq(x0, i) * L1 * v;
            g1 += L0 * dlq(x1, j) * v;
        }
        const double rho = rq * (1 - aq * (Tq - A.Tref) + bq * Pq);
        const double D = A.dt / (rho * cq);
        const double phq = A.phi ? A.phi[e * QNQ + q] : 0.0;
        double rate = T0q - Tq + D * sq;
        if (A.phi) rate += D * phq;   // Δt/(ρ Cp) Φ_q (residual.jl:75-79)
        qRate[es][q] = rate;
        qD[es][q] = D;
        qC[es][q] = D * rq * aq / rho;
        qK[es][q] = kq;
        qS[es][q] = sq + phq;
        qW[es][q] = g.det * kQ9W[q];
        qG[es][q][0] = g0 * g.jac[0];
        qG[es][q][1] = g1 * g.jac[1];
    }
    __syncthreads();
    if (!ev || l >= QNV) return;
    const int a = l, ai = kQ9R[a][0], aj = kQ9R[a][1];
    double r = 0;
    for (int q = 0; q < QNQ; ++q) {
        const double x0 = gauss_x(q % 3), x1 = gauss_x(q / 3);
        const double L0 = lq(x0, ai), L1 = lq(x1, aj);
        const double N = L0 * L1;
        const double ga0 = dlq(x0, ai) * L1 * g.jac[0], ga1 = L0 * dlq(x1, aj) * g.jac[1];
        const double KTi = qK[es][q] * (ga0 * qG[es][q][0] + ga1 * qG[es][q][1]);
        r += (qRate[es][q] * N - qD[es][q] * KTi) * qW[es][q];
    }
    double rs = 0, dg = 0;
    if (JAC) {
        for (int b = 0; b < QNV; ++b) {
            const int bi = kQ9R[b][0], bj = kQ9R[b][1];
            double J = 0;
            for (int q = 0; q < QNQ; ++q) {
                const double x0 = gauss_x(q % 3), x1 = gauss_x(q / 3);
                const double L0 = lq(x0, ai), L1 = lq(x1, aj), M0 = lq(x0, bi), M1 = lq(x1, bj);
                const double Na = L0 * L1, Nb = M0 * M1;
                const double ga0 = dlq(x0, ai) * L1 * g.jac[0], ga1 = L0 * dlq(x1, aj) * g.jac[1];
                const double gb0 = dlq(x0, bi) * M1 * g.jac[0], gb1 = M0 * dlq(x1, bj) * g.jac[1];
                const double k = qK[es][q];
                const double KTi = k * (ga0 * qG[es][q][0] + ga1 * qG[es][q][1]);
                const double Kab = k * (ga0 * gb0 + ga1 * gb1);
                const double dD = qC[es][q] * Nb;
                J += ((-Nb + dD * qS[es][q]) * Na - dD * KTi - qD[es][q] * Kab) * qW[es][q];
            }
            rs += fabs(J);
            if (b == a) dg = fabs(J);
        }
    }
    if (a != QCENTRE) {
        const int64_t sl = e * QNSH + a;
        A.Rs[sl] = r;
        if (JAC) A.Js[sl] = make_double2(rs, dg);
        return;
    }
    // the private centre node
    const int64_t b = q9_node(g, e, a);
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
        A.Tw[b] = fixed ? A.bcval[b] : sT[es][a] + A.alpha * nr;
    }
}

}  // namespace

int q9_thermal_element(fem_ctx* c, const fem_mesh* m, const ThermBoxArgs& a, bool jac) {
    const Q9Geo g = q9_geo(m);
    const dim3 grid(grid_xcd(m->nels, QEPB));
    if (jac) hipLaunchKernelGGL(k_thermal_q9<true>, grid, dim3(QBLOCK), 0, c->stream, g, a);
    else hipLaunchKernelGGL(k_thermal_q9<false>, grid, dim3(QBLOCK), 0, c->stream, g, a);
    return check_launch(c, "k_thermal_q9");
}

extern "C" int fem_mesh_q9_create(fem_ctx* c, int32_t nx, int32_t ny, double x0, double x1, double y0, double y1,
                                  fem_mesh** out) {
    if (!c || !out) return FEM_EARG;
    *out = nullptr;
    if (nx < 1 || ny < 1) return set_err(c, FEM_EARG, "element counts must be positive, got (%d, %d)", nx, ny);
    if (!(x1 > x0) || !(y1 > y0)) return set_err(c, FEM_EARG, "empty domain");
    const int64_t sx = 2 * (int64_t)nx + 1, sy = 2 * (int64_t)ny + 1;
    const int64_t nn = sx * sy, ne = (int64_t)nx * ny;
    if (nn > std::numeric_limits<int32_t>::max() || ne * QNSH > std::numeric_limits<int32_t>::max())
        return set_err(c, FEM_EDIM, "mesh too large for 32-bit node and element-slot indices");
    HIPCHK(c, hipSetDevice(c->device));
    fem_mesh* m = new fem_mesh();
    m->ctx = c;
    m->dim = 2;
    m->nloc = QNV;
    m->structured = true;
    m->nx = nx;
    m->ny = ny;
    const double bx[4] = {x0, x1, y0, y1};
    std::memcpy(m->box, bx, sizeof bx);
    m->nnodes = nn;
    m->nels = ne;
    m->nnodesP = 0;   // a thermal mesh: no pressure space
    // _tensor_grid_coordinates (mesh.jl:340-398): LinRange(x0, x1, 2nx + 1),
    // (1 - t) x0 + t x1 with t = i / 2nx, x fastest
    m->h_coords.resize((size_t)nn * 2);
    for (int64_t iy = 0; iy < sy; ++iy)
        for (int64_t ix = 0; ix < sx; ++ix) {
            const int64_t k = iy * sx + ix;
            const double tx = (double)ix / (double)(sx - 1), ty = (double)iy / (double)(sy - 1);
            m->h_coords[2 * k + 0] = (1 - tx) * x0 + tx * x1;
            m->h_coords[2 * k + 1] = (1 - ty) * y0 + ty * y1;
            if (ix == 0 || iy == 0 || ix == sx - 1 || iy == sy - 1) m->h_boundary.push_back((int32_t)k);
        }
    // generate_element2node (connectivity.jl:193-227), elements x fastest
    m->h_el2n.resize((size_t)ne * QNV);
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t ex = e % nx, ey = e / nx;
        for (int a = 0; a < QNV; ++a)
            m->h_el2n[e * QNV + a] = (int32_t)((2 * ey + kQ9R[a][1]) * sx + 2 * ex + kQ9R[a][0]);
    }
    // node -> element-slot lists of the eight shared local nodes in ascending
    // element order (generate_node2element, connectivity.jl:341-349); cell
    // centres (odd, odd grid indices) are private to their element
    std::vector<int64_t> sid(nn, -1);
    std::vector<int32_t> node_of;
    for (int64_t k = 0; k < nn; ++k) {
        const int64_t ix = k % sx, iy = k / sx;
        if ((ix & 1) && (iy & 1)) continue;
        sid[k] = (int64_t)node_of.size();
        node_of.push_back((int32_t)k);
    }
    m->ns = (int64_t)node_of.size();
    std::vector<int32_t> off(m->ns + 1, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < QNSH; ++a) off[sid[m->h_el2n[e * QNV + a]] + 1]++;
    for (int64_t i = 0; i < m->ns; ++i) off[i + 1] += off[i];
    std::vector<int32_t> ent(off[m->ns]);
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < QNSH; ++a) ent[fill[sid[m->h_el2n[e * QNV + a]]]++] = (int32_t)(e * QNSH + a);
    int r = FEM_OK;
    if ((r = dalloc_t(c, &m->d_off, off.size())) || (r = dalloc_t(c, &m->d_ent, ent.size())) ||
        (r = dalloc_t(c, &m->d_node_of, node_of.size()))) {
        fem_mesh_destroy(m);
        return r;
    }
    hipError_t he = hipMemcpy(m->d_off, off.data(), off.size() * 4, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(m->d_ent, ent.data(), ent.size() * 4, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(m->d_node_of, node_of.data(), node_of.size() * 4, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        fem_mesh_destroy(m);
        return set_err(c, FEM_EHIP, "hipMemcpy of the Q9 node lists failed: %s", hipGetErrorString(he));
    }
    *out = m;
    return FEM_OK;
}
