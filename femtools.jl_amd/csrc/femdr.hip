// femdr.hip -- host side of libfemdr.so: the C ABI of include/femdr.h, mesh
// construction, state, and the Powell-Hestenes/DYREL control loop of
// src/stokes/solvers/DR.jl:348-802 driving the kernels of femdr_kernels.hpp.
//
// The device never waits on the host between convergence checks: the
// ncheck-1 pseudo-iterations between two checks are launched back to back on
// the context stream; the host synchronises only to read the few reduced
// scalars of a check (norms, λmin sums, λmax) and of an outer iteration.
#include "femdr_host.hpp"

#include <map>
#include <mutex>

#include <cstdlib>
#include <functional>


// ---------------------------------------------------------------------------
// reference-element tables (shape_functions.jl:115-171, 231-239;
// integration_points.jl:188-202)
// ---------------------------------------------------------------------------
static void t7_basis(double xi, double eta, double N[7], double dN[7][2]) {
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

template <class T>
static int fill_t(fem_ctx* c, T* p, T v, int64_t n) {
    if (!c || n < 0) return FEM_EARG;
    if (n == 0) return FEM_OK;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, p, v);
    return check_launch(c, "k_fill");
}

// ---------------------------------------------------------------------------
// context and buffers
// ---------------------------------------------------------------------------
// Device scratch of fem_norms_noalloc: the block partials and the three sums
// of one norms evaluation. Each GPU holds its own copy for the life of the
// process; its device address is resolved once per device at context creation
// (with that device current) and only that address is used from the host.
constexpr int64_t NORMS_SCRATCH_BLOCKS = 2048;
static __device__ __align__(256) double scratch[NORMS_SCRATCH_BLOCKS * 3 + 8];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device `dev`
static hipError_t resolve_scratch(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (g_scratch_ptr[dev]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[dev] = static_cast<double*>(p);
    return e;
}

extern "C" {

int fem_ctx_create(int device, fem_ctx** out) {
    if (!out) return FEM_EARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FEM_EHIP;
    if (device < 0 || device >= n) return FEM_EARG;
    fem_ctx* c = new fem_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return FEM_EHIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return FEM_EHIP; }
    if (dalloc_t(c, &c->red_partial, RED_GRID * RED_MAXK) || dalloc_t(c, &c->red_out, RED_MAXK) ||
        hipHostMalloc((void**)&c->h_red, RED_MAXK * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->h_cheb, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        resolve_scratch(device) != hipSuccess) {
        delete c;
        return FEM_EHIP;
    }
    *out = c;
    return FEM_OK;
}

int fem_ctx_set_tuning(fem_ctx* c, const fem_tuning* t) {
    if (!c || !t) return FEM_EARG;
    if ((t->tiles != 0 && t->tiles != 1) || (t->graphs != 0 && t->graphs != 1))
        return set_err(c, FEM_EARG, "tiles and graphs take 0 or 1");
    if (t->tile_rows != 4 && t->tile_rows != 8 && t->tile_rows != 16)
        return set_err(c, FEM_EARG, "tile_rows must be 4, 8 or 16, got %d", t->tile_rows);
    if (t->fold_interval < 0) return set_err(c, FEM_EARG, "fold_interval must be >= 0, got %d", t->fold_interval);
    c->tune = *t;
    return FEM_OK;
}

int fem_ctx_get_tuning(fem_ctx* c, fem_tuning* t) {
    if (!c || !t) return FEM_EARG;
    *t = c->tune;
    return FEM_OK;
}

int fem_ctx_destroy(fem_ctx* c) {
    if (!c) return FEM_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->tev) hipEventDestroy(e);
    for (void* p : c->bufs) hipFree(p);
    hipFree(c->red_partial);
    hipFree(c->red_out);
    hipHostFree(c->h_red);
    hipHostFree(c->h_cheb);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipStreamDestroy(c->stream);
    delete c;
    return FEM_OK;
}

int fem_last_error(const fem_ctx* c, char* buf, size_t len) {
    if (!c || !buf || len == 0) return FEM_EARG;
    snprintf(buf, len, "%s", c->err.c_str());
    return FEM_OK;
}

int fem_sync(fem_ctx* c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}

int fem_buf_alloc(fem_ctx* c, size_t bytes, void** dptr) {
    if (!c || !dptr) return FEM_EARG;
    CHK(dalloc(c, dptr, bytes));
    HIPCHK(c, hipMemsetAsync(*dptr, 0, bytes ? bytes : 16, c->stream));
    c->bufs.insert(*dptr);
    return FEM_OK;
}
int fem_buf_free(fem_ctx* c, void* dptr) {
    if (!c) return FEM_EARG;
    if (!dptr) return FEM_OK;
    if (!c->bufs.count(dptr)) return set_err(c, FEM_EARG, "fem_buf_free: pointer not owned by this context");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->bufs.erase(dptr);
    HIPCHK(c, hipFree(dptr));
    return FEM_OK;
}
int fem_copy_h2d(fem_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return FEM_EARG;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}
int fem_copy_d2h(fem_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return FEM_EARG;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}
int fem_copy_d2d(fem_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return FEM_EARG;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return FEM_OK;
}
int fem_fill_f64(fem_ctx* c, double* p, double v, int64_t n) { return fill_t(c, p, v, n); }
int fem_fill_f32(fem_ctx* c, float* p, float v, int64_t n) { return fill_t(c, p, v, n); }
int fem_fill_i32(fem_ctx* c, int32_t* p, int32_t v, int64_t n) { return fill_t(c, p, v, n); }

}  // extern "C"

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------
static void gen_t7(int32_t nx, int32_t ny, double x0, double x1, double y0, double y1, std::vector<double>& coords,
                   std::vector<int32_t>& el2n) {
    // generate_coordinates(T7) mesh.jl:412-440; generate_element2node(T7)
    // connectivity.jl:111-140 (alternating diagonals, centroids appended).
    const int64_t sx = 2 * (int64_t)nx + 1, sy = 2 * (int64_t)ny + 1;
    const int64_t nT6 = sx * sy, nels = 2LL * nx * ny;
    coords.assign((size_t)(nT6 + nels) * 2, 0.0);
    el2n.assign((size_t)nels * 7, 0);
    auto lerp = [](int64_t j, int64_t d, double a, double b) {
        const double t = (double)j / (double)d;
        return (1 - t) * a + t * b;
    };
    for (int64_t iy = 0; iy < sy; ++iy)
        for (int64_t ix = 0; ix < sx; ++ix) {
            coords[(iy * sx + ix) * 2] = lerp(ix, sx - 1, x0, x1);
            coords[(iy * sx + ix) * 2 + 1] = lerp(iy, sy - 1, y0, y1);
        }
    const double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
    int64_t k = nT6, e = 0;
    auto node = [&](int64_t ix, int64_t iy) { return (int32_t)(iy * sx + ix); };
    auto cent = [&](int64_t ex, int64_t ey, int a, int b) {
        coords[k * 2] = x0 + (3 * ex + a) * dx / 3;
        coords[k * 2 + 1] = y0 + (3 * ey + b) * dy / 3;
        ++k;
    };
    for (int64_t ey = 0; ey < ny; ++ey)
        for (int64_t ex = 0; ex < nx; ++ex) {
            const int64_t ix = 2 * ex, iy = 2 * ey;
            const int32_t BL = node(ix, iy), BR = node(ix + 2, iy), TR = node(ix + 2, iy + 2), TL = node(ix, iy + 2);
            const int32_t bot = node(ix + 1, iy), rgt = node(ix + 2, iy + 1), top = node(ix + 1, iy + 2),
                          lft = node(ix, iy + 1), ctr = node(ix + 1, iy + 1);
            const int32_t cA = (int32_t)(nT6 + e), cB = (int32_t)(nT6 + e + 1);
            int32_t* A = &el2n[e * 7];
            int32_t* B = &el2n[(e + 1) * 7];
            if ((ex + ey) % 2 == 0) {
                const int32_t a[7] = {BL, BR, TR, bot, rgt, ctr, cA}, b[7] = {BL, TR, TL, ctr, top, lft, cB};
                std::memcpy(A, a, sizeof a); std::memcpy(B, b, sizeof b);
                cent(ex, ey, 2, 1); cent(ex, ey, 1, 2);
            } else {
                const int32_t a[7] = {BL, BR, TL, bot, ctr, lft, cA}, b[7] = {BR, TR, TL, rgt, top, ctr, cB};
                std::memcpy(A, a, sizeof a); std::memcpy(B, b, sizeof b);
                cent(ex, ey, 1, 1); cent(ex, ey, 2, 2);
            }
            e += 2;
        }
}

static int mesh_finish(fem_ctx* c, fem_mesh* m) {
    const int64_t ne = m->nels, nn = m->nnodes;
    m->nnodesP = 3 * ne;
    // bubble (local node 7) must be private to its element
    std::vector<int32_t> val(nn, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < 7; ++a) {
            int32_t n = m->h_el2n[e * 7 + a];
            if (n < 0 || n >= nn) return set_err(c, FEM_EARG, "el2n entry %d out of range at element %lld", n + 1, (long long)e + 1);
            val[n]++;
        }
    std::vector<uint8_t> is_bubble(nn, 0);
    for (int64_t e = 0; e < ne; ++e) {
        int32_t b = m->h_el2n[e * 7 + 6];
        if (val[b] != 1) return set_err(c, FEM_EARG, "T7 centroid node %d of element %lld is shared", b + 1, (long long)e + 1);
        is_bubble[b] = 1;
    }
    // shared-node list + node->element-slot lists in ascending element order
    // (generate_node2element, connectivity.jl:341-349)
    std::vector<int32_t> node_of;
    std::vector<int64_t> sid(nn, -1);
    bool identity = true;
    for (int64_t n = 0; n < nn; ++n)
        if (!is_bubble[n] && val[n] > 0) {
            sid[n] = (int64_t)node_of.size();
            if (sid[n] != n) identity = false;
            node_of.push_back((int32_t)n);
        }
    m->ns = (int64_t)node_of.size();
    std::vector<int32_t> off(m->ns + 1, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < NSH; ++a) off[sid[m->h_el2n[e * 7 + a]] + 1]++;
    for (int64_t i = 0; i < m->ns; ++i) off[i + 1] += off[i];
    std::vector<int32_t> ent(off[m->ns]);
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < NSH; ++a) ent[fill[sid[m->h_el2n[e * 7 + a]]]++] = (int32_t)(e * NSH + a);
    if ((int64_t)ne * NSH > std::numeric_limits<int32_t>::max())
        return set_err(c, FEM_EDIM, "mesh too large for 32-bit element-slot indices");
    // SoA connectivity and per-element geometry (affine T7: J from the corners)
    std::vector<int32_t> soa((size_t)7 * ne);
    std::vector<double> geo((size_t)5 * ne);
    for (int64_t e = 0; e < ne; ++e) {
        for (int a = 0; a < 7; ++a) soa[a * ne + e] = m->h_el2n[e * 7 + a];
        const double* p0 = &m->h_coords[2 * (int64_t)m->h_el2n[e * 7 + 0]];
        const double* p1 = &m->h_coords[2 * (int64_t)m->h_el2n[e * 7 + 1]];
        const double* p2 = &m->h_coords[2 * (int64_t)m->h_el2n[e * 7 + 2]];
        const double J00 = p1[0] - p0[0], J01 = p2[0] - p0[0], J10 = p1[1] - p0[1], J11 = p2[1] - p0[1];
        const double det = J00 * J11 - J01 * J10;
        if (!(std::fabs(det) > 0)) return set_err(c, FEM_EARG, "degenerate element %lld", (long long)e + 1);
        geo[0 * ne + e] = J11 / det;
        geo[1 * ne + e] = -J01 / det;
        geo[2 * ne + e] = -J10 / det;
        geo[3 * ne + e] = J00 / det;
        geo[4 * ne + e] = std::fabs(det);
    }
    CHK(dalloc_t(c, &m->d_el2n, soa.size()));
    CHK(dalloc_t(c, &m->d_geo, geo.size()));
    CHK(dalloc_t(c, &m->d_off, off.size()));
    CHK(dalloc_t(c, &m->d_ent, ent.size()));
    HIPCHK(c, hipMemcpy(m->d_el2n, soa.data(), soa.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_geo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_ent, ent.data(), ent.size() * 4, hipMemcpyHostToDevice));
    if (!identity) {
        CHK(dalloc_t(c, &m->d_node_of, node_of.size()));
        HIPCHK(c, hipMemcpy(m->d_node_of, node_of.data(), node_of.size() * 4, hipMemcpyHostToDevice));
    }
    return FEM_OK;
}

extern "C" {

int fem_mesh_t7_create(fem_ctx* c, int32_t nx, int32_t ny, double x0, double x1, double y0, double y1,
                       fem_mesh** out) {
    if (!c || !out) return FEM_EARG;
    *out = nullptr;
    if (nx < 1 || ny < 1) return set_err(c, FEM_EARG, "element counts must be positive, got (%d, %d)", nx, ny);
    if (!(x1 > x0) || !(y1 > y0)) return set_err(c, FEM_EARG, "empty domain");
    HIPCHK(c, hipSetDevice(c->device));
    fem_mesh* m = new fem_mesh();
    m->ctx = c;
    m->structured = true;
    m->box[0] = x0; m->box[1] = x1; m->box[2] = y0; m->box[3] = y1;
    m->nx = nx; m->ny = ny;
    gen_t7(nx, ny, x0, x1, y0, y1, m->h_coords, m->h_el2n);
    m->nnodes = (int64_t)m->h_coords.size() / 2;
    m->nels = (int64_t)m->h_el2n.size() / 7;
    // Γnodes: nodes on the box boundary (mesh.jl:79-80); only T6 grid nodes
    const int64_t sx = 2 * (int64_t)nx + 1, sy = 2 * (int64_t)ny + 1;
    for (int64_t iy = 0; iy < sy; ++iy)
        for (int64_t ix = 0; ix < sx; ++ix)
            if (ix == 0 || iy == 0 || ix == sx - 1 || iy == sy - 1) m->h_boundary.push_back((int32_t)(iy * sx + ix));
    int r = mesh_finish(c, m);
    if (r != FEM_OK) { fem_mesh_destroy(m); return r; }
    *out = m;
    return FEM_OK;
}

int fem_mesh_t7_upload(fem_ctx* c, const double* host_coords, int64_t nnodes, const int32_t* host_el2n_1based,
                       int64_t nels, fem_mesh** out) {
    if (!c || !out || !host_coords || !host_el2n_1based) return FEM_EARG;
    *out = nullptr;
    if (nnodes <= 0 || nels <= 0) return set_err(c, FEM_EARG, "empty mesh");
    HIPCHK(c, hipSetDevice(c->device));
    fem_mesh* m = new fem_mesh();
    m->ctx = c;
    m->nnodes = nnodes;
    m->nels = nels;
    m->h_coords.assign(host_coords, host_coords + 2 * nnodes);
    m->h_el2n.resize((size_t)nels * 7);
    for (int64_t i = 0; i < nels * 7; ++i) m->h_el2n[i] = host_el2n_1based[i] - 1;
    int r = mesh_finish(c, m);
    if (r != FEM_OK) { fem_mesh_destroy(m); return r; }
    *out = m;
    return FEM_OK;
}

int fem_mesh_get_info(const fem_mesh* m, fem_mesh_info* o) {
    if (!m || !o) return FEM_EARG;
    o->dim = m->dim; o->nv_local = m->nloc; o->np_local = m->dim == 3 ? 4 : (m->nloc == 9 ? 0 : 3);
    o->nq = m->nloc;
    o->nnodes = m->nnodes; o->nels = m->nels; o->nnodesP = m->nnodesP; o->n_shared = m->ns;
    o->n_boundary = (int64_t)m->h_boundary.size();
    return FEM_OK;
}
int fem_mesh_get_coords(const fem_mesh* m, double* out) {
    if (!m || !out) return FEM_EARG;
    std::memcpy(out, m->h_coords.data(), m->h_coords.size() * sizeof(double));
    return FEM_OK;
}
int fem_mesh_get_el2n(const fem_mesh* m, int32_t* out) {
    if (!m || !out) return FEM_EARG;
    for (size_t i = 0; i < m->h_el2n.size(); ++i) out[i] = m->h_el2n[i] + 1;
    return FEM_OK;
}
int fem_mesh_get_boundary_nodes(const fem_mesh* m, int32_t* out) {
    if (!m || !out) return FEM_EARG;
    for (size_t i = 0; i < m->h_boundary.size(); ++i) out[i] = m->h_boundary[i] + 1;
    return FEM_OK;
}
int fem_mesh_destroy(fem_mesh* m) {
    if (!m) return FEM_OK;
    hipFree(m->d_el2n); hipFree(m->d_geo); hipFree(m->d_off); hipFree(m->d_ent); hipFree(m->d_node_of);
    delete m;
    return FEM_OK;
}

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------
int fem_stokes_create(fem_ctx* c, fem_mesh* m, const fem_material* mat, const fem_state_opts* opts,
                      fem_stokes** out) {
    if (!c || !m || !mat || !out) return FEM_EARG;
    *out = nullptr;
    if (mat->nphases < 1 || mat->nphases > FEM_MAX_PHASES)
        return set_err(c, FEM_EARG, "nphases must be in 1..%d", FEM_MAX_PHASES);
    if (mat->ndim != m->dim)
        return set_err(c, FEM_EDIM, "material gravity must have %d components for a %d-D mesh", m->dim, m->dim);
    if (m->nloc == 9)
        return set_err(c, FEM_EARG, "Stokes states take T7 or Hex27 meshes; a Q9 mesh carries thermal states only");
    HIPCHK(c, hipSetDevice(c->device));
    fem_stokes* s = new fem_stokes();
    s->ctx = c;
    s->mesh = m;
    s->dim = m->dim;
    s->mat = *mat;
    s->opts = opts ? *opts : fem_state_opts{0.98, 0.98, 0.9, 1e-6};
    const int64_t nv = m->nnodes, np = m->nnodesP, nip = m->nels * m->nloc;
    for (int k = 0; k < FEM_FIELD_COUNT; ++k) {
        int64_t n;
        if (k <= FEM_FIELD_PC_V_Y) n = nv;
        else if (k <= FEM_FIELD_TAU_OLD_XY) n = nip;
        else if (k <= FEM_FIELD_PNUM) n = np;
        else if (k <= FEM_FIELD_PC_V_Z) n = m->dim == 3 ? nv : 0;
        else if (k <= FEM_FIELD_TAU_OLD_YZ) n = m->dim == 3 ? nip : 0;
        else n = m->dim == 2 ? nip : 0;
        s->len[k] = n;
        // +16 entries: 16-byte global->LDS loads of the last block may read
        // 8 bytes past the final element of an integration-point array
        int r = dalloc_t(c, &s->f[k], n + 16);
        if (r != FEM_OK) { fem_stokes_destroy(s); return r; }
        HIPCHK(c, hipMemset(s->f[k], 0, (n + 16) * sizeof(double)));
    }
    int r = FEM_OK;
    if ((r = dalloc_t(c, &s->d_phase_v, 7 * m->nels)) || (r = dalloc_t(c, &s->d_phase_P, 3 * m->nels)) ||
        (r = dalloc_t(c, &s->d_pscale, np)) || (r = dalloc_t(c, &s->d_Re, NSH * m->nels)) ||
        (r = dalloc_t(c, &s->d_Jrow, NSH * m->nels)) || (r = dalloc_t(c, &s->d_Jdiag, NSH * m->nels)) ||
        (r = dalloc_t(c, &s->d_bcflag, nv)) || (r = dalloc_t(c, &s->d_bcval_x, nv)) ||
        (r = dalloc_t(c, &s->d_bcval_y, nv)) || (r = dalloc_t(c, &s->d_ptab, PC_NCOL * MAXPH)) ||
        (r = dalloc_t(c, &s->d_flag, 4)) || (r = dalloc_t(c, &s->d_cheb, 8))) {
        fem_stokes_destroy(s);
        return r;
    }
    if (m->dim == 3 && ((r = dalloc_t(c, &s->d_bcval_z, nv)) || (r = dalloc_t(c, &s->d_slot3, 26 * 3 * m->nels)) ||
                        (r = dalloc_t(c, &s->d_jslot3, 26 * 6 * m->nels)))) {
        fem_stokes_destroy(s);
        return r;
    }
    HIPCHK(c, hipMemset(s->d_phase_v, 0, 7 * m->nels * 4));
    HIPCHK(c, hipMemset(s->d_phase_P, 0, 3 * m->nels * 4));
    HIPCHK(c, hipMemset(s->d_bcflag, 0, nv));
    HIPCHK(c, hipMemset(s->d_pscale, 0, np * 8));
    *out = s;
    return FEM_OK;
}

int fem_stokes_destroy(fem_stokes* s) {
    if (!s) return FEM_OK;
    hipStreamSynchronize(s->ctx->stream);
    for (int k = 0; k < FEM_FIELD_COUNT; ++k) hipFree(s->f[k]);
    hipFree(s->d_phase_v); hipFree(s->d_phase_P); hipFree(s->d_pscale); hipFree(s->d_Re);
    hipFree(s->d_bpart64); hipFree(s->d_Rn64);
    hipFree(s->d_Jrow); hipFree(s->d_Jdiag); hipFree(s->d_bcflag); hipFree(s->d_bcval_x);
    hipFree(s->d_bcval_y); hipFree(s->d_bc_stage); hipFree(s->d_bc_stage_v); hipFree(s->d_ptab);
    hipFree(s->d_flag);
    hipFree(s->d_cheb);
    hipFree(s->d_yflag);
    hipFree(s->d_bcval_z); hipFree(s->d_slot3); hipFree(s->d_jslot3);
    if (s->w32) {
        float** p = (float**)s->w32;
        for (size_t k = 0; k < sizeof(fem_stokes::W32) / sizeof(float*); ++k) hipFree(p[k]);
        delete s->w32;
    }
    if (s->w3) {
        float** p = (float**)s->w3;
        for (size_t k = 0; k < sizeof(fem_stokes::W3) / sizeof(float*); ++k) hipFree(p[k]);
        delete s->w3;
    }
    delete s;
    return FEM_OK;
}

int fem_stokes_field(fem_stokes* s, int32_t field, void** dptr, int64_t* len, int32_t* dtype) {
    if (!s || field < 0 || field >= FEM_FIELD_COUNT) return FEM_EARG;
    if (dptr) *dptr = s->f[field];
    if (len) *len = s->len[field];
    if (dtype) *dtype = 0;
    return FEM_OK;
}

int fem_stokes_set_phases(fem_stokes* s, int32_t which, int32_t layout, const int32_t* h, int64_t n) {
    if (!s || !h || which < 0 || which > 1) return FEM_EARG;
    fem_ctx* c = s->ctx;
    const fem_mesh* m = s->mesh;
    const int64_t ne = m->nels;
    if (m->dim == 3 && layout != FEM_PHASE_CELL)
        return set_err(c, FEM_EARG, "the 3-D path takes one phase per cell (a 1 x nels phase matrix)");
    const int nloc = which == 0 ? 7 : 3;
    std::vector<int32_t> loc((size_t)nloc * ne);   // [nloc][ne] 0-based
    if (layout == FEM_PHASE_CELL) {
        if (n != ne) return set_err(c, FEM_EDIM, "cell phases must have nels = %lld entries", (long long)ne);
        for (int64_t e = 0; e < ne; ++e)
            for (int a = 0; a < nloc; ++a) loc[a * ne + e] = h[e] - 1;
    } else if (layout == FEM_PHASE_ELEMENT) {
        if (n != (int64_t)nloc * ne)
            return set_err(c, FEM_EDIM, "element phases must be %d x nels", nloc);
        for (int64_t e = 0; e < ne; ++e)
            for (int a = 0; a < nloc; ++a) loc[a * ne + e] = h[e * nloc + a] - 1;
    } else if (layout == FEM_PHASE_NODAL) {
        const int64_t need = which == 0 ? m->nnodes : m->nnodesP;
        if (n != need) return set_err(c, FEM_EDIM, "nodal phases must have %lld entries", (long long)need);
        for (int64_t e = 0; e < ne; ++e)
            for (int a = 0; a < nloc; ++a)
                loc[a * ne + e] = (which == 0 ? h[m->h_el2n[e * 7 + a]] : h[e * 3 + a]) - 1;
    } else {
        return set_err(c, FEM_EARG, "unknown phase layout %d", layout);
    }
    for (int32_t p : loc)
        if (p < 0 || p >= s->mat.nphases) return set_err(c, FEM_EARG, "phase id %d out of range 1..%d", p + 1, s->mat.nphases);
    // element-constant phases use the cell fast path
    bool cell = true;
    for (int64_t e = 0; e < ne && cell; ++e)
        for (int a = 1; a < nloc; ++a)
            if (loc[a * ne + e] != loc[e]) { cell = false; break; }
    int32_t* d = which == 0 ? s->d_phase_v : s->d_phase_P;
    HIPCHK(c, hipMemcpy(d, loc.data(), (size_t)nloc * ne * 4, hipMemcpyHostToDevice));
    (which == 0 ? s->layout_v : s->layout_P) = cell ? FEM_PHASE_CELL : FEM_PHASE_ELEMENT;
    // compare the cell rows of both phase arrays (device copies)
    std::vector<int32_t> pv(ne), pp(ne);
    HIPCHK(c, hipMemcpy(pv.data(), s->d_phase_v, ne * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(pp.data(), s->d_phase_P, ne * 4, hipMemcpyDeviceToHost));
    s->phase_same = pv == pp;
    return FEM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// per-phase table, boundary conditions, kernel launches
// ---------------------------------------------------------------------------
int upload_ptab(fem_stokes* s, const fem_drucker_prager* dp, double dt, int finite_K, bool* bulk_inf,
                       bool* thermal, bool* eos_p, bool* elastic, const fem_drucker_prager_cap* cap) {
    fem_ctx* c = s->ctx;
    const fem_material& m = s->mat;
    if (dp && dp->nphases != m.nphases)
        return set_err(c, FEM_EDIM, "DruckerPrager has %d phases, material has %d", dp->nphases, m.nphases);
    if (cap && cap->nphases != m.nphases)
        return set_err(c, FEM_EDIM, "DruckerPragerCap has %d phases, material has %d", cap->nphases, m.nphases);
    if (cap)
        for (int p = 0; p < m.nphases; ++p)
            if (!(cap->Kb[p] > 0 && std::isfinite(cap->Kb[p])))
                return set_err(c, FEM_EARG, "phase %d: dilatant plasticity requires a positive finite bulk modulus, got %g",
                               p + 1, cap->Kb[p]);
    std::vector<double> t((size_t)PC_NCOL * MAXPH, 0.0);
    *bulk_inf = true; *thermal = false; *eos_p = false; *elastic = false;
    for (int p = 0; p < m.nphases; ++p) {
        const double invG = 1.0 / m.G[p];
        const double bulk = finite_K ? m.K[p] : m.etab[p];
        t[PC_ETA * MAXPH + p] = m.eta[p];
        t[PC_INVG * MAXPH + p] = invG;
        t[PC_RHO0 * MAXPH + p] = m.rho0[p];
        t[PC_ALPHA * MAXPH + p] = m.alpha[p];
        t[PC_BETA * MAXPH + p] = 1.0 / m.K[p];
        t[PC_BULK * MAXPH + p] = bulk;
        t[PC_ETAVE * MAXPH + p] = 1.0 / (1.0 / m.eta[p] + invG / dt);
        t[PC_INV2GDT * MAXPH + p] = invG / (2 * dt);
        if (dp) {
            const double etave = t[PC_ETAVE * MAXPH + p];
            t[PC_YIELD0 * MAXPH + p] = dp->cos_phi[p] * dp->C[p];
            t[PC_INVDEN * MAXPH + p] = 1.0 / ((etave + dp->eta_reg[p]) + ((dp->Kb[p] * dt) * (-dp->sin_psi[p])) * (-dp->sin_phi[p]));
            t[PC_COSPHI * MAXPH + p] = dp->cos_phi[p];
            t[PC_SINPHI * MAXPH + p] = dp->sin_phi[p];
            t[PC_SINPSI * MAXPH + p] = dp->sin_psi[p];
            t[PC_C * MAXPH + p] = dp->C[p];
            t[PC_EREG * MAXPH + p] = dp->eta_reg[p];
            t[PC_KB * MAXPH + p] = dp->Kb[p];
        }
        if (cap) {
            t[PC_COSPHI * MAXPH + p] = cap->cos_phi[p];
            t[PC_SINPHI * MAXPH + p] = cap->sin_phi[p];
            t[PC_SINPSI * MAXPH + p] = cap->sin_psi[p];
            t[PC_C * MAXPH + p] = cap->C[p];
            t[PC_EREG * MAXPH + p] = cap->eta_reg[p];
            t[PC_KB * MAXPH + p] = cap->Kb[p];
            t[PC_PT * MAXPH + p] = cap->pT[p];
            t[PC_CMIN * MAXPH + p] = cap->C_min[p];
            t[PC_HC * MAXPH + p] = cap->H_C[p];
        }
        if (!std::isinf(bulk)) *bulk_inf = false;
        if (m.alpha[p] != 0) *thermal = true;
        if (!std::isinf(m.K[p])) *eos_p = true;
        if (invG != 0) *elastic = true;
    }
    HIPCHK(c, hipMemcpyAsync(s->d_ptab, t.data(), t.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}

__global__ void k_set_bc(int64_t n, const int32_t* dofs, const double* vals, uint8_t bit, uint8_t* flag,
                         double* bcval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        flag[dofs[i]] |= bit;
        bcval[dofs[i]] = vals[i];
    }
}

int upload_bcs(fem_stokes* s, const fem_bc* bcs) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes;
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, nv, c->stream));
    double* bcval[3] = {s->d_bcval_x, s->d_bcval_y, s->d_bcval_z};
    for (int comp = 0; comp < s->dim; ++comp) {
        const fem_bc& b = bcs[comp];
        if (b.n == 0) continue;
        if (b.n < 0 || !b.host_dofs_1based || !b.host_vals) return set_err(c, FEM_EARG, "invalid boundary condition");
        // a serial v[dofs[i]] = vals[i] lets a later duplicate win; keep that
        std::vector<int64_t> last(nv, -1);
        for (int64_t i = 0; i < b.n; ++i) {
            const int32_t dof = b.host_dofs_1based[i] - 1;
            if (dof < 0 || dof >= nv) return set_err(c, FEM_EARG, "Dirichlet DoF %d out of range", dof + 1);
            last[dof] = i;
        }
        std::vector<int32_t> d;
        std::vector<double> v;
        for (int64_t i = 0; i < b.n; ++i) {
            const int32_t dof = b.host_dofs_1based[i] - 1;
            if (last[dof] == i) { d.push_back(dof); v.push_back(b.host_vals[i]); }
        }
        const int64_t nb = (int64_t)d.size();
        if (nb > s->bc_stage_cap) {
            hipFree(s->d_bc_stage);
            hipFree(s->d_bc_stage_v);
            s->d_bc_stage = nullptr;
            s->d_bc_stage_v = nullptr;
            CHK(dalloc_t(c, &s->d_bc_stage, nb));
            CHK(dalloc_t(c, &s->d_bc_stage_v, nb));
            s->bc_stage_cap = nb;
        }
        HIPCHK(c, hipMemcpyAsync(s->d_bc_stage, d.data(), nb * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(s->d_bc_stage_v, v.data(), nb * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_set_bc, dim3(grid_for(nb, 256)), dim3(256), 0, c->stream, nb, s->d_bc_stage,
                           s->d_bc_stage_v, (uint8_t)(1 << comp), s->d_bcflag, bcval[comp]);
        CHK(check_launch(c, "k_set_bc"));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return FEM_OK;
}

struct LaunchOpts {
    bool plastic = false, elastic = false, bulk_inf = true, thermal = false, eos_p = false;
    double* Ablk = nullptr;   // the Jacobian launch also stores the velocity blocks
    bool use_pnum = false, do_update = false;
    bool store_RP = false, store_Pnum = false, store_Rv = false, store_Rv0 = false, store_tau = false;
    double dt = 1.0;
    double alpha[2] = {0, 0}, beta[2] = {0, 0};
    const void* cheb = nullptr;       // nullable: the Chebyshev scalars (element type) in device memory
    int32_t* yield_count = nullptr;   // nullable: count of elements with a yielding point
    uint8_t* yield_flag = nullptr;    // with yield_count: near-yield flag per element
    double* phi_out = nullptr;        // nullable (FP64 residual launches): shear dissipation [nels][7]
    float* tangent_out = nullptr;     // nullable (FP64 plastic residual launches): return linearisations
    int32_t* tangent_bad = nullptr;   // with tangent_out: count of points whose linearisation is not positive
    bool cap = false;                 // the tensile cap is the yield model (FP64, EF_GEN launches)
    bool commit = false;              // a cap launch that commits the plastic history
};

constexpr int EBLOCK = 128;
constexpr int GBLOCK = 256;
// EF_TILE blocks: 16 x rows cells, 32 threads per cell row (fem_tuning.tile_rows)

// Inner iterations of structured 2-D meshes whose cell counts are multiples
// of the 16 x rows tile run the EF_TILE element kernel (FP32 corrections at
// 4, 8 or 16 rows; FP64 at 4 or 8: the FP64 τ_old and slot stages of a
// 16-row tile would leave one block per CU)
static int tile_rows(const fem_mesh* m) { return m->ctx->tune.tile_rows; }
static bool tile_shape_ok(const fem_mesh* m, int rows) {
    return m->dim == 2 && m->structured && m->nx > 0 && m->nx % 16 == 0 && m->ny % rows == 0;
}
static bool tile_ok(const fem_mesh* m) { return m->ctx->tune.tiles && tile_shape_ok(m, tile_rows(m)); }
static int64_t tile_count(const fem_mesh* m, int rows) { return (int64_t)(m->nx / 16) * (m->ny / rows); }
static int tile_perimeter_nodes(int rows) { return 2 * (32 + 1) + 2 * (2 * rows - 1); }
static bool tile64_ok(const fem_mesh* m) { return tile_ok(m) && tile_rows(m) <= 8; }

// One set of DR arrays in precision F. The FP64 solve runs on the state's own
// fields; a correction solve of the mixed-precision path runs on FP32 work
// arrays holding the correction (δv, δP) of the FP64 solution, with the FP64
// residual R(x) as right-hand side.
template <class F>
struct DRSet {
    F *vx = nullptr, *vy = nullptr, *rate_x = nullptr, *rate_y = nullptr;
    F *Rv_x = nullptr, *Rv_y = nullptr, *Rv0_x = nullptr, *Rv0_y = nullptr, *PC_x = nullptr, *PC_y = nullptr;
    F *P = nullptr, *Pold = nullptr, *RP = nullptr, *Pnum = nullptr, *gammaP = nullptr, *MP = nullptr,
      *pscale = nullptr;
    F *tau_old[3] = {nullptr, nullptr, nullptr}, *tau[3] = {nullptr, nullptr, nullptr};
    F *T = nullptr, *T0 = nullptr, *Q = nullptr, *Pf = nullptr;
    F *bcval_x = nullptr, *bcval_y = nullptr;
    const double* ptab = nullptr;   // per-phase table (read in FP64, converted at use)
    F* geo = nullptr;
    typename Vec2<F>::T* Re = nullptr;
    const F *rhs_P = nullptr, *rhs_vx = nullptr, *rhs_vy = nullptr;
    const double *xvx = nullptr, *xvy = nullptr, *xP = nullptr, *xPf = nullptr;
    const uint8_t* hot = nullptr;   // near-yield flags of the elements (EF_DELTA)
    const float* tangent = nullptr; // return linearisations about x (EF_DELTA|EF_TANGENT)
    typename Vec2<F>::T* bpart = nullptr;   // EF_TILE perimeter partial sums
    typename Vec2<F>::T* Rn = nullptr;      // EF_TILE interior-node sums
    bool tile = false;    // inner iterations run the EF_TILE element kernel
    bool lin = false;     // correction operator (EF_LIN)
    bool delta = false;   // ... with the return map about the state x (EF_DELTA)
};

static DRSet<double> set64(fem_stokes* s) {
    DRSet<double> d;
    double** f = s->f;
    d.vx = f[FEM_FIELD_V_X]; d.vy = f[FEM_FIELD_V_Y];
    d.rate_x = f[FEM_FIELD_DVDTAU_X]; d.rate_y = f[FEM_FIELD_DVDTAU_Y];
    d.Rv_x = f[FEM_FIELD_RV_X]; d.Rv_y = f[FEM_FIELD_RV_Y];
    d.Rv0_x = f[FEM_FIELD_RV0_X]; d.Rv0_y = f[FEM_FIELD_RV0_Y];
    d.PC_x = f[FEM_FIELD_PC_V_X]; d.PC_y = f[FEM_FIELD_PC_V_Y];
    d.P = f[FEM_FIELD_P]; d.RP = f[FEM_FIELD_RP]; d.Pnum = f[FEM_FIELD_PNUM];
    // solve!'s P_old keyword (a nodal field) in place of dr.P0 (DR.jl:220-226)
    d.Pold = (s->in.P_old && !s->in.P_old_at_points) ? const_cast<double*>(s->in.P_old) : f[FEM_FIELD_P0];
    d.gammaP = f[FEM_FIELD_GAMMA_P]; d.MP = f[FEM_FIELD_M_P]; d.pscale = s->d_pscale;
    d.tau_old[0] = f[FEM_FIELD_TAU_OLD_XX]; d.tau_old[1] = f[FEM_FIELD_TAU_OLD_YY];
    d.tau_old[2] = f[FEM_FIELD_TAU_OLD_XY];
    d.tau[0] = f[FEM_FIELD_TAU_XX]; d.tau[1] = f[FEM_FIELD_TAU_YY]; d.tau[2] = f[FEM_FIELD_TAU_XY];
    d.T = f[FEM_FIELD_T]; d.T0 = f[FEM_FIELD_T0]; d.Q = f[FEM_FIELD_Q]; d.Pf = f[FEM_FIELD_PF];
    d.bcval_x = s->d_bcval_x; d.bcval_y = s->d_bcval_y;
    d.ptab = s->d_ptab; d.geo = s->mesh->d_geo;
    d.Re = s->d_Re;
    return d;
}

static DRSet<float> set32(fem_stokes* s) {
    DRSet<float> d;
    fem_stokes::W32& w = *s->w32;
    d.vx = w.vx; d.vy = w.vy; d.rate_x = w.rate_x; d.rate_y = w.rate_y;
    d.Rv_x = w.Rv_x; d.Rv_y = w.Rv_y; d.Rv0_x = w.Rv0_x; d.Rv0_y = w.Rv0_y; d.PC_x = w.PC_x; d.PC_y = w.PC_y;
    d.P = w.P; d.RP = w.RP; d.Pnum = w.Pnum; d.gammaP = w.gammaP; d.MP = w.MP; d.pscale = w.pscale;
    d.bcval_x = w.zero_v; d.bcval_y = w.zero_v;   // the correction vanishes on Dirichlet nodes
    d.ptab = s->d_ptab; d.geo = w.geo;
    d.Re = w.Re;
    d.rhs_P = w.rhs_P; d.rhs_vx = w.rhs_vx; d.rhs_vy = w.rhs_vy;
    d.xvx = s->f[FEM_FIELD_V_X]; d.xvy = s->f[FEM_FIELD_V_Y]; d.xP = s->f[FEM_FIELD_P];
    d.xPf = s->pf_nonzero ? s->f[FEM_FIELD_PF] : nullptr;
    d.tau_old[0] = w.tau_old0; d.tau_old[1] = w.tau_old1; d.tau_old[2] = w.tau_old2;
    d.tile = w.bpart != nullptr && tile_ok(s->mesh);
    d.bpart = w.bpart;
    d.Rn = w.Rn;
    d.lin = true;
    return d;
}

template <class F, int FL>
static void launch_elem_t(fem_stokes* s, const ElemArgs<F>& a) {
    if constexpr ((FL & EF_TILE) != 0) {
        const dim3 grid(grid_xcd(a.ntiles, 1));
        const int rows = tile_rows(s->mesh);
        if (rows == 4) hipLaunchKernelGGL((k_t7_element<F, FL, 128>), grid, dim3(128), 0, s->ctx->stream, a);
        else if constexpr (std::is_same<F, float>::value) {
            if (rows == 16) hipLaunchKernelGGL((k_t7_element<F, FL, 512>), grid, dim3(512), 0, s->ctx->stream, a);
            else hipLaunchKernelGGL((k_t7_element<F, FL, 256>), grid, dim3(256), 0, s->ctx->stream, a);
        } else
            hipLaunchKernelGGL((k_t7_element<F, FL, 256>), grid, dim3(256), 0, s->ctx->stream, a);
    } else
        hipLaunchKernelGGL((k_t7_element<F, FL, EBLOCK>), dim3(grid_xcd(s->mesh->nels, EBLOCK)), dim3(EBLOCK), 0,
                           s->ctx->stream, a);
}

template <int FL>
static void dispatch_elem(int key, fem_stokes* s, const ElemArgs<double>& a) {
    if (key == FL) launch_elem_t<double, FL>(s, a);
    else if constexpr (FL > 0) dispatch_elem<FL - 1>(key, s, a);
}

// FP64 inner iterations in tiles: EF_ITER|EF_STRUCT|EF_TILE with any of
// EF_ELASTIC, EF_PLASTIC, EF_BULK
template <int I>
static void dispatch_elem_tile(int key, fem_stokes* s, const ElemArgs<double>& a) {
    constexpr int FL = EF_ITER | EF_STRUCT | EF_TILE | ((I & 1) ? EF_ELASTIC : 0) | ((I & 2) ? EF_PLASTIC : 0) |
                       ((I & 4) ? EF_BULK : 0);
    if (key == FL) launch_elem_t<double, FL>(s, a);
    else if constexpr (I > 0) dispatch_elem_tile<I - 1>(key, s, a);
}

// Whether an inner-iteration launch runs in tiles: the element kernel and the
// node update decide alike. FP64 tiles run the plain inner iteration only (no
// per-point outputs, no temperature, sources, fluid pressure, cap or sampled
// viscosity).
template <class F>
static bool tile_launch(const fem_stokes* s, const DRSet<F>& d, const LaunchOpts& o) {
    const fem_mesh* m = s->mesh;
    if (!d.tile || !o.use_pnum || m->nx <= 0) return false;
    const bool nodeph = s->layout_v == FEM_PHASE_ELEMENT || s->layout_P == FEM_PHASE_ELEMENT;
    if (nodeph) return false;
    if (d.lin) return true;
    const bool gen = o.thermal || s->pf_nonzero || s->q_nonzero || o.cap || s->in.viscosity ||
                     (s->in.P_old && s->in.P_old_at_points);
    return !gen && !o.store_tau && !o.phi_out && !o.tangent_out && !o.yield_count && !s->in.body_force[0];
}

// inner iterations on a correction: EF_LIN|EF_ITER with any of EF_BULK,
// EF_GEN, EF_STRUCT, and the return-map operator EF_DELTA|EF_PLASTIC with or
// without EF_ELASTIC; EF_TILE on structured meshes without EF_GEN
template <int I>
static void dispatch_lin(int key, fem_stokes* s, const ElemArgs<float>& a) {
    constexpr bool el = I & 8, delta = I & 16, tile = I & 32, tan = I & 64;
    constexpr int FL = EF_LIN | EF_ITER | ((I & 1) ? EF_BULK : 0) | ((I & 2) ? EF_GEN : 0) |
                       ((I & 4) ? EF_STRUCT : 0) | (el ? EF_ELASTIC : 0) | (delta ? (EF_DELTA | EF_PLASTIC) : 0) |
                       (tile ? EF_TILE : 0) | (tan ? EF_TANGENT : 0);
    if constexpr ((!el || delta) && (!tile || ((I & 4) && !(I & 2))) && (!tan || (delta && !el))) {
        if (key == FL) { launch_elem_t<float, FL>(s, a); return; }
    }
    if constexpr (I > 0) dispatch_lin<I - 1>(key, s, a);
}

// The element kernel's arguments for one launch and its template key.
template <class F>
static int build_elem_args(fem_stokes* s, const DRSet<F>& d, const LaunchOpts& o, ElemArgs<F>& a, int& fl) {
    fem_mesh* m = s->mesh;
    std::memset(&a, 0, sizeof a);
    a.nels = m->nels;
    a.el2n = m->d_el2n;
    a.geo = d.geo;
    a.vx = d.vx;
    a.vy = d.vy;
    a.P = d.P;
    a.Pold = d.Pold;
    a.pscale = d.pscale;
    a.tau_old0 = d.tau_old[0];
    a.tau_old1 = d.tau_old[1];
    a.tau_old2 = d.tau_old[2];
    a.phase_v = s->d_phase_v;
    a.phase_P = s->d_phase_P;
    a.ptab = d.ptab;
    a.nphases = s->mat.nphases;
    a.gx = (F)s->mat.g[0];
    a.gy = (F)s->mat.g[1];
    a.Tref = (F)s->mat.Tref;
    a.dt = (F)o.dt;
    a.Re = d.Re;
    a.rate_x = d.rate_x;
    a.rate_y = d.rate_y;
    a.PCx = d.PC_x;
    a.PCy = d.PC_y;
    a.bcflag = s->d_bcflag;
    a.bcval_x = d.bcval_x;
    a.bcval_y = d.bcval_y;
    a.alpha_x = (F)o.alpha[0]; a.alpha_y = (F)o.alpha[1];
    a.beta_x = (F)o.beta[0]; a.beta_y = (F)o.beta[1];
    a.cheb = (const F*)o.cheb;
    a.do_update = o.do_update;
    a.RP_out = o.store_RP ? d.RP : nullptr;
    a.Pnum_out = o.store_Pnum ? d.Pnum : nullptr;
    if (o.store_Rv) { a.Rv_out_x = d.Rv_x; a.Rv_out_y = d.Rv_y; }
    if (o.store_Rv0) { a.Rv0_out_x = d.Rv0_x; a.Rv0_out_y = d.Rv0_y; }
    if (o.store_tau) {
        a.tau_out0 = d.tau[0];
        a.tau_out1 = d.tau[1];
        a.tau_out2 = d.tau[2];
    }
    a.rhs_P = d.rhs_P;
    a.rhs_vx = d.rhs_vx;
    a.rhs_vy = d.rhs_vy;
    a.xvx = d.xvx;
    a.xvy = d.xvy;
    a.xP = d.xP;
    a.xPf = d.xPf;
    a.yield_count = o.yield_count;
    a.yield_flag = o.yield_flag;
    a.hot = d.hot;
    a.tangent = d.tangent;
    // phase arrays are always stored [nloc][nels]; row 0 is the cell phase
    const bool nodeph = s->layout_v == FEM_PHASE_ELEMENT || s->layout_P == FEM_PHASE_ELEMENT;
    // the correction operator has no temperature, source or fluid-pressure terms
    const bool gen = nodeph || (!d.lin && (o.thermal || s->pf_nonzero || s->q_nonzero || o.cap ||
                                          s->in.viscosity || (s->in.P_old && s->in.P_old_at_points)));
    a.nodeph = nodeph;
    a.bulk_all_inf = o.bulk_inf;
    a.thermal = o.thermal;
    a.T = d.T;
    a.T0 = d.T0;
    a.Q = s->q_nonzero ? d.Q : nullptr;
    a.Pf = s->pf_nonzero ? d.Pf : nullptr;
    const bool strct = m->nx > 0;
    if (strct) {
        const double dx = (m->box[1] - m->box[0]) / m->nx, dy = (m->box[3] - m->box[2]) / m->ny;
        a.nx = m->nx;
        a.sx = 2 * m->nx + 1;
        a.nT6 = (int64_t)(2 * m->nx + 1) * (2 * m->ny + 1);
        a.idx = (F)(1.0 / dx);
        a.idy = (F)(1.0 / dy);
        a.detc = (F)(dx * dy);
    }
    a.phase_same = s->phase_same && !nodeph;
    const int key = (!o.bulk_inf ? EF_BULK : 0) | (o.use_pnum ? EF_ITER : 0) | (gen ? EF_GEN : 0) |
                    (strct ? EF_STRUCT : 0);
    if constexpr (std::is_same<F, double>::value) {
        if (d.lin) return set_err(s->ctx, FEM_EARG, "no FP64 correction operator");
        a.phi_out = o.phi_out;
        a.tangent_out = o.tangent_out;
        a.tangent_bad = o.tangent_out ? o.tangent_bad : nullptr;
        if (o.cap) {
            a.cap = 1;
            a.gamma = s->f[FEM_FIELD_GAMMA_PL];
            if (o.commit) {
                a.hist_gamma = s->f[FEM_FIELD_GAMMA_PL];
                a.hist_theta = s->f[FEM_FIELD_THETA_PL];
            }
        }
        if (o.store_tau && s->dim == 2) a.P_out = s->f[FEM_FIELD_P_IP];
        a.bf_x = s->in.body_force[0];
        a.bf_y = s->in.body_force[1];
        a.eta_ip = s->in.viscosity;
        a.Pold_ip = s->in.P_old_at_points ? s->in.P_old : nullptr;
        const bool tile = tile_launch(s, d, o);
        if (tile) {
            a.ntx = m->nx / 16;
            a.ntiles = tile_count(m, tile_rows(m));
            a.bpart = d.bpart;
            a.Rn = d.Rn;
        }
        fl = key | (o.elastic ? EF_ELASTIC : 0) | (o.plastic ? EF_PLASTIC : 0) | (tile ? EF_TILE : 0);
    } else {
        if (!d.lin || !o.use_pnum)
            return set_err(s->ctx, FEM_EARG, "FP32 element residuals exist for inner iterations on a correction only");
        const bool tile = tile_launch(s, d, o) && strct && !gen;
        if (tile) {
            a.ntx = m->nx / 16;
            a.ntiles = tile_count(m, tile_rows(m));
            a.bpart = d.bpart;
            a.Rn = d.Rn;
        }
        // the linearised return needs no stress history (x's trial stress is in it)
        fl = key | EF_LIN | (tile ? EF_TILE : 0) |
             (d.delta ? (EF_DELTA | EF_PLASTIC | (o.elastic && !d.tangent ? EF_ELASTIC : 0) |
                         (d.tangent ? EF_TANGENT : 0))
                      : 0);
    }
    return FEM_OK;
}

template <class F>
static int launch_element(fem_stokes* s, const DRSet<F>& d, const LaunchOpts& o) {
    ElemArgs<F> a;
    int fl = 0;
    CHK(build_elem_args(s, d, o, a, fl));
    if constexpr (std::is_same<F, double>::value) {
        if (fl & EF_TILE) dispatch_elem_tile<7>(fl, s, a);
        else dispatch_elem<63>(fl, s, a);
    } else dispatch_lin<127>(fl, s, a);
    return check_launch(s->ctx, "k_t7_element");
}

template <class F>
static void build_gather_args(fem_stokes* s, const DRSet<F>& d, const LaunchOpts& o, GatherArgs<F>& g) {
    fem_mesh* m = s->mesh;
    std::memset(&g, 0, sizeof g);
    g.ns = m->ns;
    g.node_of = m->d_node_of;
    g.off = m->d_off;
    g.ent = m->d_ent;
    g.Re = d.Re;
    g.vx = d.vx;
    g.vy = d.vy;
    g.rate_x = d.rate_x;
    g.rate_y = d.rate_y;
    g.PCx = d.PC_x;
    g.PCy = d.PC_y;
    g.bcflag = s->d_bcflag;
    g.bcval_x = d.bcval_x;
    g.bcval_y = d.bcval_y;
    if (o.store_Rv) { g.Rv_out_x = d.Rv_x; g.Rv_out_y = d.Rv_y; }
    if (o.store_Rv0) { g.Rv0_out_x = d.Rv0_x; g.Rv0_out_y = d.Rv0_y; }
    g.alpha_x = (F)o.alpha[0]; g.alpha_y = (F)o.alpha[1];
    g.beta_x = (F)o.beta[0]; g.beta_y = (F)o.beta[1];
    g.cheb = (const F*)o.cheb;
    g.do_update = o.do_update;
    g.rhs_vx = d.rhs_vx;
    g.rhs_vy = d.rhs_vy;
    if (!d.lin) {
        g.bf_x = (const F*)s->in.body_force[0];
        g.bf_y = (const F*)s->in.body_force[1];
    }
}

// the grid-node update after an EF_TILE element launch (the tile kernel has
// summed every grid node's slots)
template <class F>
static void build_tile_update_args(fem_stokes* s, const DRSet<F>& d, const GatherArgs<F>& g, TileUpdArgs<F>& P) {
    fem_mesh* m = s->mesh;
    P.g = g;
    const int rows = tile_rows(m);
    P.ntx = m->nx / 16; P.nty = m->ny / rows;
    P.TH = 2 * rows;
    P.sx = 2 * m->nx + 1;
    P.nT6 = (int64_t)P.sx * (2 * m->ny + 1);
    P.bpart = d.bpart;
    P.Rn = d.Rn;
}

template <class F>
static int launch_gather(fem_stokes* s, const DRSet<F>& d, const LaunchOpts& o) {
    fem_mesh* m = s->mesh;
    GatherArgs<F> g;
    build_gather_args(s, d, o, g);
    if (tile_launch(s, d, o)) {
        TileUpdArgs<F> P;
        build_tile_update_args(s, d, g, P);
        hipLaunchKernelGGL((k_tile_update<F, GBLOCK>), dim3(grid_xcd(P.nT6, GBLOCK)), dim3(GBLOCK), 0,
                           s->ctx->stream, P);
        return check_launch(s->ctx, "k_tile_update");
    }
    // node -> element-slot lists are read from the CSR arrays for every mesh:
    // deriving them from indices on structured meshes lengthens the path to
    // the loads more than the list reads cost
    hipLaunchKernelGGL((k_gather_update<F, false, GBLOCK>), dim3(grid_xcd(m->ns, GBLOCK)), dim3(GBLOCK), 0,
                       s->ctx->stream, g);
    return check_launch(s->ctx, "k_gather_update");
}

template <bool EL, bool PL, bool NP_>
static void launch_jac_t(fem_stokes* s, const JacArgs<double>& a) {
    hipLaunchKernelGGL((k_t7_jacobian<double, EL, PL, NP_, 64>), dim3(grid_for(s->mesh->nels, 64)), dim3(64), 0,
                       s->ctx->stream, a);
}

static int launch_jacobian(fem_stokes* s, const LaunchOpts& o) {
    fem_mesh* m = s->mesh;
    JacArgs<double> a;
    std::memset(&a, 0, sizeof a);
    a.nels = m->nels;
    a.el2n = m->d_el2n;
    a.geo = m->d_geo;
    a.vx = s->f[FEM_FIELD_V_X];
    a.vy = s->f[FEM_FIELD_V_Y];
    a.P = s->f[FEM_FIELD_P];
    a.pscale = s->d_pscale;
    a.tau_old0 = s->f[FEM_FIELD_TAU_OLD_XX];
    a.tau_old1 = s->f[FEM_FIELD_TAU_OLD_YY];
    a.tau_old2 = s->f[FEM_FIELD_TAU_OLD_XY];
    a.Pf = s->f[FEM_FIELD_PF];
    a.phase_v = s->d_phase_v;
    a.ptab = s->d_ptab;
    a.nphases = s->mat.nphases;
    a.dt = o.dt;
    a.Jrow = s->d_Jrow;
    a.Jdiag = s->d_Jdiag;
    a.dRdv_x = s->f[FEM_FIELD_DRVDV_X];
    a.dRdv_y = s->f[FEM_FIELD_DRVDV_Y];
    a.PC_x = s->f[FEM_FIELD_PC_V_X];
    a.PC_y = s->f[FEM_FIELD_PC_V_Y];
    a.cap = o.cap;
    a.gamma = o.cap ? s->f[FEM_FIELD_GAMMA_PL] : nullptr;
    a.eta_ip = s->in.viscosity;
    a.Ablk = o.Ablk;
    const bool np_ = s->layout_v == FEM_PHASE_ELEMENT;
    const int key = (o.elastic ? 4 : 0) | (o.plastic ? 2 : 0) | (np_ ? 1 : 0);
    switch (key) {
        case 0: launch_jac_t<false, false, false>(s, a); break;
        case 1: launch_jac_t<false, false, true>(s, a); break;
        case 2: launch_jac_t<false, true, false>(s, a); break;
        case 3: launch_jac_t<false, true, true>(s, a); break;
        case 4: launch_jac_t<true, false, false>(s, a); break;
        case 5: launch_jac_t<true, false, true>(s, a); break;
        case 6: launch_jac_t<true, true, false>(s, a); break;
        case 7: launch_jac_t<true, true, true>(s, a); break;
    }
    CHK(check_launch(s->ctx, "k_t7_jacobian"));
    GatherJacArgs<double> g;
    g.ns = m->ns;
    g.node_of = m->d_node_of;
    g.off = m->d_off;
    g.ent = m->d_ent;
    g.Jrow = s->d_Jrow;
    g.Jdiag = s->d_Jdiag;
    g.dRdv_x = s->f[FEM_FIELD_DRVDV_X];
    g.dRdv_y = s->f[FEM_FIELD_DRVDV_Y];
    g.PC_x = s->f[FEM_FIELD_PC_V_X];
    g.PC_y = s->f[FEM_FIELD_PC_V_Y];
    hipLaunchKernelGGL((k_gather_jac<double, GBLOCK>), dim3(grid_for(m->ns, GBLOCK)), dim3(GBLOCK), 0,
                       s->ctx->stream, g);
    return check_launch(s->ctx, "k_gather_jac");
}

static int scaling_impl(fem_stokes* s, double gamma_fact, double dt) {
    fem_mesh* m = s->mesh;
    fem_ctx* c = s->ctx;
    // solve!'s scaling_viscosity defaults to the viscosity it solves with (DR.jl:315-325)
    const double* eta_ip = s->in.scaling_viscosity ? s->in.scaling_viscosity : s->in.viscosity;
    if (s->layout_v == FEM_PHASE_ELEMENT)
        hipLaunchKernelGGL((k_pressure_scaling<double, true, 256>), dim3(grid_for(m->nels, 256)), dim3(256), 0,
                           c->stream, m->nels, (const double*)m->d_geo, (const int32_t*)s->d_phase_v,
                           (const double*)s->d_ptab, s->mat.nphases, gamma_fact, dt, s->f[FEM_FIELD_M_P],
                           s->f[FEM_FIELD_GAMMA_P], s->d_pscale, eta_ip);
    else
        hipLaunchKernelGGL((k_pressure_scaling<double, false, 256>), dim3(grid_for(m->nels, 256)), dim3(256), 0,
                           c->stream, m->nels, (const double*)m->d_geo, (const int32_t*)s->d_phase_v,
                           (const double*)s->d_ptab, s->mat.nphases, gamma_fact, dt, s->f[FEM_FIELD_M_P],
                           s->f[FEM_FIELD_GAMMA_P], s->d_pscale, eta_ip);
    return check_launch(c, "k_pressure_scaling");
}

__global__ void k_pscale(int64_t n, const double* g, const double* M, double* ps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ps[i] = M[i] == 0.0 ? 0.0 : g[i] / M[i];
}

// reductions -> host scalars (synchronises)
template <class F>
static int reduce_check(fem_stokes* s, const DRSet<F>& d, const double al[2], double out[6]) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes;
    const unsigned g = std::min<unsigned>(RED_GRID, grid_for(nv, RED_BLOCK));
    hipLaunchKernelGGL(k_reduce_check<F>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, (const F*)d.Rv_x,
                       (const F*)d.Rv_y, (const F*)d.Rv0_x, (const F*)d.Rv0_y, (const F*)d.rate_x, (const F*)d.rate_y,
                       (const F*)d.PC_x, (const F*)d.PC_y, al[0], al[1], c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 6, 0,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_reduce_check"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 6 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_red, 6 * 8);
    return FEM_OK;
}

template <class F>
static int reduce_outer(fem_stokes* s, const DRSet<F>& d, double out[3]) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes, np = s->mesh->nnodesP;
    const unsigned g = std::min<unsigned>(RED_GRID, grid_for(std::max(nv, np), RED_BLOCK));
    hipLaunchKernelGGL(k_reduce_outer<F>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, np, (const F*)d.Rv_x,
                       (const F*)d.Rv_y, (const F*)d.RP, (const F*)d.MP, c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 3, 0,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_reduce_outer"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_red, 3 * 8);
    return FEM_OK;
}

static int reduce_lmax(fem_stokes* s, double out[2]) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes;
    const unsigned g = std::min<unsigned>(RED_GRID, grid_for(nv, RED_BLOCK));
    hipLaunchKernelGGL(k_reduce_lmax<double>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv,
                       (const double*)s->f[FEM_FIELD_DRVDV_X], (const double*)s->f[FEM_FIELD_DRVDV_Y],
                       (const double*)s->f[FEM_FIELD_PC_V_X], (const double*)s->f[FEM_FIELD_PC_V_Y], c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 3, 1,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_reduce_lmax"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = c->h_red[0];
    out[1] = c->h_red[1];
    if (c->h_red[2] > 0) out[0] = out[1] = std::numeric_limits<double>::quiet_NaN();
    return FEM_OK;
}

static int checked_lmax(fem_stokes* s, double lmax[2]) {
    CHK(reduce_lmax(s, lmax));
    const char* labels[2] = {"stokes vx", "stokes vy"};
    for (int k = 0; k < 2; ++k)
        if (!(std::isfinite(lmax[k]) && lmax[k] > 0))
            return set_err(s->ctx, FEM_EPRECOND,
                           "%s preconditioner produced invalid λmax = %g; check for zero or non-finite preconditioner entries",
                           labels[k], lmax[k]);
    return FEM_OK;
}

// compute_strain_rate_stress_postprocess (postprocess.jl:17-100), one thread per
// element: dΩ-weighted element averages of ε̇xx, ε̇yy, ε̇zz = 0, ε̇xy, ε̇II (of the
// deviatoric rate) and of the point stresses τxx, τyy, τzz = −(τxx + τyy), τxy,
// τII; out is [10][nels]
__global__ __launch_bounds__(256) void k_postprocess2(int64_t ne, const int32_t* el2n, const double* geo,
                                                      const double* vx, const double* vy, const double* txx,
                                                      const double* tyy, const double* txy, double* out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    const double i00 = geo[0 * ne + e], i01 = geo[1 * ne + e], i10 = geo[2 * ne + e], i11 = geo[3 * ne + e];
    const double det = geo[4 * ne + e];
    double ux[NV], uy[NV];
    for (int a = 0; a < NV; ++a) {
        const int32_t n = el2n[a * ne + e];
        ux[a] = vx[n];
        uy[a] = vy[n];
    }
    double tot[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, vol = 0;
    for (int q = 0; q < NQ; ++q) {
        double ax = 0, ay = 0, bx = 0, by = 0;
        for (int a = 0; a < NV; ++a) {
            const double d0 = kT.dN[q][a][0], d1 = kT.dN[q][a][1];
            ax += d0 * ux[a]; ay += d1 * ux[a]; bx += d0 * uy[a]; by += d1 * uy[a];
        }
        const double exx = ax * i00 + ay * i10, exy_x = ax * i01 + ay * i11;
        const double eyx = bx * i00 + by * i10, eyy = bx * i01 + by * i11;
        const double exy = (exy_x + eyx) / 2;
        const double tr = (exx + eyy) / 3;
        const double d0 = exx - tr, d1 = eyy - tr, d2 = -tr;
        const double tx = txx[e * NQ + q], ty = tyy[e * NQ + q], tz = -(tx + ty), txy_q = txy[e * NQ + q];
        const double val[10] = {exx, eyy, 0.0, exy, sqrt((d0 * d0 + d1 * d1 + d2 * d2) / 2 + exy * exy),
                                tx, ty, tz, txy_q, sqrt((tx * tx + ty * ty + tz * tz) / 2 + txy_q * txy_q)};
        const double dO = kT.w[q] * det;
        for (int k = 0; k < 10; ++k) tot[k] += val[k] * dO;
        vol += dO;
    }
    for (int k = 0; k < 10; ++k) out[k * ne + e] = tot[k] / vol;
}

// entries that are not finite (or, with positive, not > 0)
__global__ void k_count_bad(int64_t n, const double* x, int32_t positive, int32_t* flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double v = x[i];
        if (!std::isfinite(v) || (positive && !(v > 0))) atomicAdd(flag, 1);
    }
}

// per element: Σ_q (N^P_q · Qq) w_q |det J| and the count of negative or
// non-finite nodal and point values (_pressure_source_area_kernel!)
__global__ __launch_bounds__(RED_BLOCK) void k_source_area(int64_t ne, const double* Qq, const double* geo,
                                                           double* partial) {
    double acc[2] = {0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t e = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; e < ne; e += stride) {
        const double det = geo[4 * ne + e];
        double Ql[3];
        for (int i = 0; i < 3; ++i) {
            Ql[i] = Qq[e * 3 + i];
            if (!std::isfinite(Ql[i]) || Ql[i] < 0) acc[1] += 1;
        }
        for (int q = 0; q < NQ; ++q) {
            double qv = 0;
            for (int i = 0; i < 3; ++i) qv += kT.NP1[q][i] * Ql[i];
            if (!std::isfinite(qv) || qv < 0) acc[1] += 1;
            acc[0] += qv * (kT.w[q] * det);
        }
    }
    block_sum_store<2>(acc, partial);
}

__global__ void k_scale_copy(int64_t n, const double* x, double a, double* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] * a;
}

// number of bad entries of a device array (-1 when the check itself fails)
static int64_t count_bad(fem_stokes* s, const double* x, int64_t n, bool positive) {
    fem_ctx* c = s->ctx;
    if (n <= 0) return 0;
    int32_t h = 0;
    if (hipMemsetAsync(s->d_flag, 0, 4, c->stream) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_count_bad, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, x, (int32_t)positive,
                       s->d_flag);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&h, s->d_flag, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return -1;
    return h;
}

static int any_nonzero(fem_stokes* s, const double* x, int64_t n, bool* out) {
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipMemsetAsync(s->d_flag, 0, 4, c->stream));
    hipLaunchKernelGGL(k_any_nonzero<double>, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, x, s->d_flag);
    CHK(check_launch(c, "k_any_nonzero"));
    int32_t h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, s->d_flag, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = h != 0;
    return FEM_OK;
}

static int prepare(fem_stokes* s, const fem_drucker_prager* dp, double dt, int finite_K, LaunchOpts& o,
                   bool use_tau_old, const fem_drucker_prager_cap* cap = nullptr) {
    bool bulk_inf, thermal, eos_p, elastic;
    if (cap && s->dim != 2) return set_err(s->ctx, FEM_EARG, "the tensile cap is built for 2-D states");
    CHK(check_solve_inputs(s, dp != nullptr || cap != nullptr));
    CHK(any_nonzero(s, s->f[FEM_FIELD_PF], s->mesh->nnodesP, &s->pf_nonzero));
    CHK(any_nonzero(s, s->f[FEM_FIELD_Q], s->mesh->nnodesP, &s->q_nonzero));
    CHK(upload_ptab(s, dp, dt, finite_K, &bulk_inf, &thermal, &eos_p, &elastic, cap));
    o.plastic = dp != nullptr || cap != nullptr;
    o.cap = cap != nullptr;
    o.elastic = elastic && use_tau_old;
    o.bulk_inf = bulk_inf;
    o.thermal = thermal;
    o.eos_p = eos_p;
    o.dt = dt;
    if (!(dt > 0) || !std::isfinite(dt)) return set_err(s->ctx, FEM_EARG, "dt must be positive and finite");
    return FEM_OK;
}

// graph replay of check intervals (fem_tuning.graphs; off for tracing tools
// that do not follow graph launches)
static bool graphs_on(const fem_ctx* c) { return c->tune.graphs != 0; }


// One check interval (ncheck inner iterations: element + gather launches with
// the interval's fixed store pattern) captured once per solve and replayed
// from a graph. The Chebyshev scalars change at every check, so replayed
// launches read them from device memory, staged by a stream-ordered copy
// before each replay (the host writes the pinned stage only after the
// previous check's synchronisation, when that copy has completed).
struct IntervalGraph {
    hipGraphExec_t exec = nullptr;
    ~IntervalGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
    template <class F, class L>
    int run(fem_stokes* s, int64_t ncheck, const double al[2], const double be[2], L launch_iter) {
        fem_ctx* c = s->ctx;
        F* h = reinterpret_cast<F*>(c->h_cheb);
        h[0] = (F)al[0]; h[1] = (F)al[1]; h[2] = (F)be[0]; h[3] = (F)be[1];
        HIPCHK(c, hipMemcpyAsync(s->d_cheb, h, 4 * sizeof(F), hipMemcpyHostToDevice, c->stream));
        if (!exec) {
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = FEM_OK;
            for (int64_t k = 1; k <= ncheck && rc == FEM_OK; ++k) rc = launch_iter(k);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rc != FEM_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            HIPCHK(c, e);
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(c, ei);
        }
        HIPCHK(c, hipGraphLaunch(exec, c->stream));
        return FEM_OK;
    }
    // the same with the launches of the whole segment in one call
    template <class F, class L>
    int run_once(fem_stokes* s, const double al[2], const double be[2], L launch_all) {
        return run<F>(s, 1, al, be, [&](int64_t) { return launch_all(); });
    }
};

// Control and result of one PH/DYREL run (_solve_stokes_dyrel!, DR.jl:348-802).
// ---------------------------------------------------------------------------
// The frozen-operator adjoint of a 2-D StokesDR state (solve_adjoint! with
// operator = :blocks, DR_adjoint.jl:1-268, 396-427; FrozenAdjointOperator,
// adjoint_operator.jl:1-660) for models without a yield model. At the
// converged forward state the adjoint residual is affine in λ, so the element
// blocks are assembled once: A (14×14, the augmented velocity Jacobian) by
// the Jacobian-diagnostics kernel, B = ∂Rv/∂P, C = ∂RP/∂v, D = ∂RP/∂P (the ηb
// storage) and W = ∂Rv/∂Pnum·Γ here. All blocks are stored
// structure-of-arrays, [entry][nels], so a wave's loads of one entry are
// contiguous. An adjoint iteration is the rate/λ update, the element apply
// (gather λ, dense element products, shared-node slots) and the slot gather.
// ---------------------------------------------------------------------------
namespace {

constexpr int ANU = 2 * NV, ANP = NP;

struct AdjBlkArgs {
    int64_t nels;
    const int32_t* el2n;            // [7][nels]
    const double* geo;              // [5][nels]
    const double *vx, *vy, *P, *Pf;                  // the state (PLASTIC: the return's pressure derivative)
    const double *tau_old0, *tau_old1, *tau_old2;    // nullable: no stress history
    const int32_t* phase_v; int32_t nodeph_v;   // cell [nels] or element-node [7][nels], 0-based
    const int32_t* phase_P; int32_t nodeph_P;   // cell [nels] or [3][nels]
    const double* ptab; int32_t nphases;
    const double* pscale;           // Γ = γP/M_P, [nels][3]
    double gx, gy, dt;
    int32_t bulk_inf;               // every ηb infinite: D is zero
    double *B, *C, *D, *W;          // [14·3][nels], [3·14][nels], [3·3][nels], [14·3][nels]
};

// B, C, D, W of one element (element_adjoint_operator_blocks,
// adjoint_operator.jl:318-373): P enters the momentum residual through the
// total pressure, the density ρ0(1 − α(T − Tref) + βP) and, with the
// Drucker-Prager model, the return (∂τ/∂P at each point from the return's
// forward-mode derivative at the trial stress of the current state, with the
// per-phase viscosity as the residual uses it), and the continuity residual
// through the storage term.
template <bool PLASTIC>
__global__ __launch_bounds__(128) void k_adj_blocks(AdjBlkArgs A) {
    const int64_t e = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (e >= A.nels) return;
    const int64_t ne = A.nels;
    const double i00 = A.geo[0 * ne + e], i01 = A.geo[1 * ne + e], i10 = A.geo[2 * ne + e],
                 i11 = A.geo[3 * ne + e], det = A.geo[4 * ne + e];
    int pv[NV], pp[NP];
    for (int a = 0; a < NV; ++a) pv[a] = A.nodeph_v ? A.phase_v[a * ne + e] : A.phase_v[e];
    for (int i = 0; i < NP; ++i) pp[i] = A.nodeph_P ? A.phase_P[i * ne + e] : A.phase_P[e];
    double B[ANU][ANP], Bn[ANU][ANP], C[ANP][ANU], D[ANP][ANP];
    for (int r = 0; r < ANU; ++r)
        for (int i = 0; i < ANP; ++i) { B[r][i] = 0; Bn[r][i] = 0; C[i][r] = 0; }
    for (int i = 0; i < ANP; ++i)
        for (int j = 0; j < ANP; ++j) D[i][j] = 0;
    double vx[NV], vy[NV], Pl[NP], Pfl[NP];
    if (PLASTIC) {
        for (int a = 0; a < NV; ++a) {
            const int32_t n = A.el2n[a * ne + e];
            vx[a] = A.vx[n];
            vy[a] = A.vy[n];
        }
        for (int i = 0; i < NP; ++i) {
            Pl[i] = A.P[e * NP + i];
            Pfl[i] = A.Pf ? A.Pf[e * NP + i] : 0.0;
        }
    }
    for (int q = 0; q < NQ; ++q) {
        const double dOm = kT.w[q] * det;
        double r0 = 0, be = 0;
        for (int a = 0; a < NV; ++a) {
            r0 += kT.N[q][a] * A.ptab[PC_RHO0 * MAXPH + pv[a]];
            be += kT.N[q][a] * A.ptab[PC_BETA * MAXPH + pv[a]];
        }
        double dtxx = 0, dtyy = 0, dtxy = 0;   // ∂τ/∂P at the point
        if (PLASTIC) {
            double ax = 0, ay = 0, bx = 0, by = 0;
            for (int a = 0; a < NV; ++a) {
                const double d0 = kT.dN[q][a][0], d1 = kT.dN[q][a][1];
                ax += d0 * vx[a]; ay += d1 * vx[a]; bx += d0 * vy[a]; by += d1 * vy[a];
            }
            const double gxx = ax * i00 + ay * i10, gxy = ax * i01 + ay * i11;
            const double gyx = bx * i00 + by * i10, gyy = bx * i01 + by * i11;
            auto at = [&](int col) {
                double acc = 0;
                for (int a = 0; a < NV; ++a) acc += kT.N[q][a] * A.ptab[col * MAXPH + pv[a]];
                return acc;
            };
            IpMat<double> m;
            const double etaq = at(PC_ETA), invG = at(PC_INVG);
            m.etave = 1.0 / (1.0 / etaq + invG / A.dt);
            m.inv2Gdt = invG / (2 * A.dt);
            const double sphi = at(PC_SINPHI), spsi = at(PC_SINPSI);
            m.yield0 = at(PC_COSPHI) * at(PC_C);
            m.sphi = sphi;
            m.invden = 1.0 / ((m.etave + at(PC_EREG)) + ((at(PC_KB) * A.dt) * (-spsi)) * (-sphi));
            double txo = 0, tyo = 0, txyo = 0;
            if (A.tau_old0) {
                txo = A.tau_old0[e * NQ + q]; tyo = A.tau_old1[e * NQ + q]; txyo = A.tau_old2[e * NQ + q];
            }
            double Pq = 0;
            for (int i = 0; i < NP; ++i) Pq += kT.NP1[q][i] * Pl[i];
            for (int i = 0; i < NP; ++i) Pq -= kT.NP1[q][i] * Pfl[i];
            D3<double> txx, tyy, txy;
            stress2d<D3<double>, double, true, D3<double>>(D3<double>(gxx), D3<double>(gyy),
                                                           D3<double>((gxy + gyx) * 0.5), m, txo, tyo, txyo,
                                                           D3<double>(Pq, 1.0, 0.0, 0.0), txx, tyy, txy);
            dtxx = txx.d0;
            dtyy = tyy.d0;
            dtxy = txy.d0;
        }
        double eb = INFINITY;
        if (!A.bulk_inf) {
            eb = 0;
            for (int i = 0; i < NP; ++i) eb += kT.NP1[q][i] * A.ptab[PC_BULK * MAXPH + pp[i]];
        }
        for (int a = 0; a < NV; ++a) {
            const double d0 = kT.dN[q][a][0], d1 = kT.dN[q][a][1];
            const double Gx = d0 * i00 + d1 * i10, Gy = d0 * i01 + d1 * i11, Na = kT.N[q][a];
            for (int i = 0; i < NP; ++i) {
                const double Np = kT.NP1[q][i];
                const double drho = r0 * (be * Np);
                if (PLASTIC) {
                    B[a][i] += ((Gx * (dtxx * Np - Np) + Gy * (dtxy * Np)) - Na * (drho * A.gx)) * dOm;
                    B[NV + a][i] += ((Gx * (dtxy * Np) + Gy * (dtyy * Np - Np)) - Na * (drho * A.gy)) * dOm;
                } else {
                    B[a][i] += (Gx * (-Np) - Na * (drho * A.gx)) * dOm;
                    B[NV + a][i] += (Gy * (-Np) - Na * (drho * A.gy)) * dOm;
                }
                Bn[a][i] += (Gx * (-Np)) * dOm;
                Bn[NV + a][i] += (Gy * (-Np)) * dOm;
                C[i][a] += (Np * (-Gx)) * dOm;
                C[i][NV + a] += (Np * (-Gy)) * dOm;
            }
        }
        for (int i = 0; i < NP; ++i)
            for (int j = 0; j < NP; ++j) D[i][j] += (kT.NP1[q][i] * (-(kT.NP1[q][j] / (eb * A.dt)))) * dOm;
    }
    for (int r = 0; r < ANU; ++r)
        for (int i = 0; i < ANP; ++i) {
            A.B[(r * ANP + i) * ne + e] = B[r][i];
            A.W[(r * ANP + i) * ne + e] = Bn[r][i] * A.pscale[e * NP + i];
            A.C[(i * ANU + r) * ne + e] = C[i][r];
        }
    for (int i = 0; i < ANP; ++i)
        for (int j = 0; j < ANP; ++j) A.D[(i * ANP + j) * ne + e] = D[i][j];
}

struct AdjApplyArgs {
    int64_t nels;
    const int32_t* el2n;                        // [7][nels]
    const double *A, *B, *C, *D, *W;     // B null: velocity rows only (no pressure coupling)
    int32_t sym, stores;
    int32_t forward;                            // 1: A x (the forward velocity operator), not Aᵀ x
    const double *lx, *ly, *lP;
    double2* slot;                              // [nels][6]
    double *outx, *outy;                        // bubble rows written here
    double* dP;                                 // [nels][3]
    const double *objx, *objy;                  // nullable: added to the velocity rows
    const double *PCx, *PCy;                    // nullable: rows divided by √PC (power iteration)
    const uint8_t* bcflag;
};

__device__ inline double adj_row_out(double v, const double* obj, const double* PC, int64_t n, bool fixed) {
    if (fixed) return 0.0;
    if (PC) return v / sqrt(PC[n]);
    return obj ? obj[n] + v : v;
}

// apply_adjoint_operator! (adjoint_operator.jl:512-535, 631-660) for one
// element: resv = Aᵀλv + CᵀλP (the upper triangle of A and B λP for a
// symmetric tangent), resp = Bᵀλv + Dᵀ(Wᵀλv + λP)
__global__ __launch_bounds__(128) void k_adj_apply(AdjApplyArgs A) {
    const int64_t e = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (e >= A.nels) return;
    const int64_t ne = A.nels;
    int32_t n[NV];
    double lv[ANU], lp[ANP];
    for (int a = 0; a < NV; ++a) {
        n[a] = A.el2n[a * ne + e];
        lv[a] = A.lx[n[a]];
        lv[NV + a] = A.ly[n[a]];
    }
    for (int i = 0; i < ANP; ++i) lp[i] = A.B ? A.lP[e * NP + i] : 0.0;
    for (int i = 0; i < ANP && A.B; ++i) {
        double b = 0;
        for (int k = 0; k < ANU; ++k) b += A.B[(k * ANP + i) * ne + e] * lv[k];
        double st = 0;
        if (A.stores) {
            double w[ANP];
            for (int j = 0; j < ANP; ++j) {
                double acc = 0;
                for (int k = 0; k < ANU; ++k) acc += A.W[(k * ANP + j) * ne + e] * lv[k];
                w[j] = acc + lp[j];
            }
            for (int j = 0; j < ANP; ++j) st += A.D[(j * ANP + i) * ne + e] * w[j];
        }
        A.dP[e * NP + i] = b + st;   // discontinuous pressure: unshared
    }
    double rv[ANU];
    for (int r = 0; r < ANU; ++r) {
        double a = 0, b = 0;
        for (int k = 0; k < ANU; ++k) {
            const int64_t idx = A.forward ? (int64_t)r * ANU + k
                                          : (A.sym ? (int64_t)(r < k ? r : k) * ANU + (r < k ? k : r)
                                                   : (int64_t)k * ANU + r);
            a += A.A[idx * ne + e] * lv[k];
        }
        for (int i = 0; i < ANP && A.B; ++i)
            b += (A.sym ? A.B[(r * ANP + i) * ne + e] : A.C[(i * ANU + r) * ne + e]) * lp[i];
        rv[r] = a + b;
    }
    for (int a = 0; a < NSH; ++a) A.slot[e * NSH + a] = make_double2(rv[a], rv[NV + a]);
    const int32_t bn = n[NSH];
    const uint8_t f = A.bcflag[bn];
    A.outx[bn] = adj_row_out(rv[NSH], A.objx, A.PCx, bn, f & 1);
    A.outy[bn] = adj_row_out(rv[NV + NSH], A.objy, A.PCy, bn, f & 2);
}

struct AdjGatherArgs {
    int64_t ns;
    const int32_t *node_of, *off, *ent;
    const double2* slot;
    double *outx, *outy;
    const double *objx, *objy, *PCx, *PCy;
    const uint8_t* bcflag;
};

// shared-node rows in ascending element order
__global__ __launch_bounds__(256) void k_adj_gather(AdjGatherArgs A) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.ns) return;
    const int64_t node = A.node_of ? A.node_of[i] : i;
    double x = 0, y = 0;
    for (int32_t k = A.off[i]; k < A.off[i + 1]; ++k) {
        const double2 v = A.slot[A.ent[k]];
        x += v.x;
        y += v.y;
    }
    const uint8_t f = A.bcflag[node];
    A.outx[node] = adj_row_out(x, A.objx, A.PCx, node, f & 1);
    A.outy[node] = adj_row_out(y, A.objy, A.PCy, node, f & 2);
}

// R0 ← R, then the rate and λ update of every node (stokes_update_rate!,
// stokes_update_variable! with −α, homogeneous Dirichlet rows; DR_adjoint.jl:234-249)
__global__ void k_adj_update(int64_t n, const double* Rx, const double* Ry, double* R0x, double* R0y, double* rx,
                             double* ry, double* lx, double* ly, const double* PCx, const double* PCy,
                             const uint8_t* bcflag, double ax, double ay, double bx, double by) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t f = bcflag[i];
    const double Rxi = Rx[i], Ryi = Ry[i];
    R0x[i] = Rxi;
    R0y[i] = Ryi;
    if (f & 1) { rx[i] = 0; lx[i] = 0; }
    else { const double nr = bx * rx[i] + Rxi / PCx[i]; rx[i] = nr; lx[i] += -ax * nr; }
    if (f & 2) { ry[i] = 0; ly[i] = 0; }
    else { const double nr = by * ry[i] + Ryi / PCy[i]; ry[i] = nr; ly[i] += -ay * nr; }
}

// power-iteration step: x ← x/‖x‖, z = x/√PC (estimate_adjoint_λmax, adjoint_operator.jl:589-596)
__global__ void k_adj_power(int64_t n, double* xx, double* xy, double* zx, double* zy, const double* PCx,
                            const double* PCy, double nrm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = xx[i] / nrm, b = xy[i] / nrm;
    xx[i] = a;
    xy[i] = b;
    zx[i] = a / sqrt(PCx[i]);
    zy[i] = b / sqrt(PCy[i]);
}

// Σ a², Σ b², Σ c² (c over its own length)
__global__ __launch_bounds__(RED_BLOCK) void k_adj_norms(int64_t n, int64_t nc, const double* a, const double* b,
                                                         const double* c, double* partial) {
    double acc[3] = {0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * RED_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += stride) {
        acc[0] += a[i] * a[i];
        acc[1] += b[i] * b[i];
    }
    if (c)
        for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < nc; i += stride) acc[2] += c[i] * c[i];
    block_sum_store<3>(acc, partial);
}

// Arrow-Hurwicz update of the pressure adjoint, λP += γP ResλP / M_P
__global__ void k_adj_lamP(int64_t n, double* lP, const double* gP, const double* RP, const double* MP) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lP[i] += gP[i] * RP[i] / MP[i];
}

struct AdjWork {
    fem_ctx* c = nullptr;
    std::vector<void*> bufs;
    template <class T>
    int get(T** p, size_t n) {
        int r = dalloc_t(c, p, n);
        if (r == FEM_OK) bufs.push_back(*p);
        return r;
    }
    ~AdjWork() {
        if (c) hipStreamSynchronize(c->stream);
        for (void* b : bufs) hipFree(b);
    }
};

struct AdjointOp {
    double *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr, *W = nullptr;
    double2* slot = nullptr;
    bool sym = false, stores = false;
};

// the blocks at the state in s; also fills ∂Rv∂v and PC_v from A
static int adjoint_assemble(fem_stokes* s, const fem_drucker_prager* dp, double dt, AdjWork& w, AdjointOp& op) {
    fem_ctx* c = s->ctx;
    fem_mesh* m = s->mesh;
    const int64_t ne = m->nels;
    LaunchOpts o;
    CHK(prepare(s, dp, dt, 0, o, true));   // PC_BULK = ηb: the operator's pressure bulk
    const fem_material& mt = s->mat;
    op.sym = dp == nullptr;                // a yield model makes the tangent non-normal
    op.stores = false;
    for (int p = 0; p < mt.nphases; ++p) {
        if (!std::isinf(mt.K[p])) op.sym = false;
        if (std::isfinite(mt.etab[p])) op.stores = true;
    }
    CHK(w.get(&op.A, (size_t)ANU * ANU * ne));
    CHK(w.get(&op.B, (size_t)ANU * ANP * ne));
    CHK(w.get(&op.C, (size_t)ANU * ANP * ne));
    CHK(w.get(&op.D, (size_t)ANP * ANP * ne));
    CHK(w.get(&op.W, (size_t)ANU * ANP * ne));
    CHK(w.get(&op.slot, (size_t)NSH * ne));
    hipLaunchKernelGGL(k_pscale, dim3(grid_for(m->nnodesP, 256)), dim3(256), 0, c->stream, m->nnodesP,
                       (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    CHK(check_launch(c, "k_pscale"));
    o.Ablk = op.A;
    CHK(launch_jacobian(s, o));
    AdjBlkArgs b;
    b.nels = ne;
    b.el2n = m->d_el2n;
    b.geo = m->d_geo;
    b.vx = s->f[FEM_FIELD_V_X];
    b.vy = s->f[FEM_FIELD_V_Y];
    b.P = s->f[FEM_FIELD_P];
    b.Pf = s->pf_nonzero ? s->f[FEM_FIELD_PF] : nullptr;
    const bool old = o.elastic;
    b.tau_old0 = old ? s->f[FEM_FIELD_TAU_OLD_XX] : nullptr;
    b.tau_old1 = old ? s->f[FEM_FIELD_TAU_OLD_YY] : nullptr;
    b.tau_old2 = old ? s->f[FEM_FIELD_TAU_OLD_XY] : nullptr;
    b.phase_v = s->d_phase_v;
    b.nodeph_v = s->layout_v == FEM_PHASE_ELEMENT;
    b.phase_P = s->d_phase_P;
    b.nodeph_P = s->layout_P == FEM_PHASE_ELEMENT;
    b.ptab = s->d_ptab;
    b.nphases = mt.nphases;
    b.pscale = s->d_pscale;
    b.gx = mt.g[0];
    b.gy = mt.g[1];
    b.dt = dt;
    b.bulk_inf = o.bulk_inf;
    b.B = op.B; b.C = op.C; b.D = op.D; b.W = op.W;
    if (dp) hipLaunchKernelGGL(k_adj_blocks<true>, dim3(grid_for(ne, 128)), dim3(128), 0, c->stream, b);
    else hipLaunchKernelGGL(k_adj_blocks<false>, dim3(grid_for(ne, 128)), dim3(128), 0, c->stream, b);
    CHK(check_launch(c, "k_adj_blocks"));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}

// out = op applied to (lx, ly, lP): velocity rows into (ox, oy) with obj added
// (or divided by √PC when PCx is set), homogeneous Dirichlet rows; dP pressure rows
static int adjoint_apply(fem_stokes* s, const AdjointOp& op, const double* lx, const double* ly, const double* lP,
                         double* ox, double* oy, double* dP, const double* objx, const double* objy,
                         const double* PCx, const double* PCy, bool forward = false) {
    fem_ctx* c = s->ctx;
    fem_mesh* m = s->mesh;
    AdjApplyArgs a;
    a.nels = m->nels;
    a.el2n = m->d_el2n;
    a.A = op.A; a.B = op.B; a.C = op.C; a.D = op.D; a.W = op.W;
    a.sym = op.sym; a.stores = op.stores;
    a.forward = forward;
    a.lx = lx; a.ly = ly; a.lP = lP;
    a.slot = op.slot;
    a.outx = ox; a.outy = oy; a.dP = dP;
    a.objx = objx; a.objy = objy; a.PCx = PCx; a.PCy = PCy;
    a.bcflag = s->d_bcflag;
    hipLaunchKernelGGL(k_adj_apply, dim3(grid_for(m->nels, 128)), dim3(128), 0, c->stream, a);
    CHK(check_launch(c, "k_adj_apply"));
    AdjGatherArgs g;
    g.ns = m->ns;
    g.node_of = m->d_node_of; g.off = m->d_off; g.ent = m->d_ent;
    g.slot = op.slot;
    g.outx = ox; g.outy = oy;
    g.objx = objx; g.objy = objy; g.PCx = PCx; g.PCy = PCy;
    g.bcflag = s->d_bcflag;
    hipLaunchKernelGGL(k_adj_gather, dim3(grid_for(m->ns, 256)), dim3(256), 0, c->stream, g);
    return check_launch(c, "k_adj_gather");
}

// Σa², Σb², Σc² to the host
static int adjoint_norms(fem_stokes* s, const double* a, const double* b, const double* cc, int64_t nc,
                         double out[3]) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes;
    const unsigned g = std::min<unsigned>(RED_GRID, grid_for(std::max(nv, nc), RED_BLOCK));
    hipLaunchKernelGGL(k_adj_norms, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, nc, a, b, cc, c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 3, 0,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_adj_norms"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_red, 3 * 8);
    return FEM_OK;
}

// _assert_frozen_symmetry (adjoint_operator.jl:149-160) on the stored blocks
static int adjoint_check_symmetry(fem_stokes* s, const AdjointOp& op) {
    fem_ctx* c = s->ctx;
    const int64_t ne = s->mesh->nels;
    std::vector<double> A((size_t)ANU * ANU * ne), B((size_t)ANU * ANP * ne), C((size_t)ANU * ANP * ne);
    HIPCHK(c, hipMemcpy(A.data(), op.A, A.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(B.data(), op.B, B.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(C.data(), op.C, C.size() * 8, hipMemcpyDeviceToHost));
    double wA = 0, wC = 0;
    for (int64_t e = 0; e < ne; ++e) {
        double nA = 0, dA = 0, nC = 0, dC = 0;
        for (int i = 0; i < ANU; ++i)
            for (int j = 0; j < ANU; ++j) {
                const double x = A[(i * ANU + j) * ne + e], y = A[(j * ANU + i) * ne + e];
                nA += x * x;
                dA += (x - y) * (x - y);
            }
        for (int i = 0; i < ANP; ++i)
            for (int k = 0; k < ANU; ++k) {
                const double x = C[(i * ANU + k) * ne + e], y = B[(k * ANP + i) * ne + e];
                nC += x * x;
                dC += (x - y) * (x - y);
            }
        wA = std::max(wA, nA == 0 ? 0.0 : std::sqrt(dA) / std::sqrt(nA));
        wC = std::max(wC, nC == 0 ? 0.0 : std::sqrt(dC) / std::sqrt(nC));
    }
    const double tol = std::sqrt(std::numeric_limits<double>::epsilon());
    if (wA > tol || wC > tol)
        return set_err(c, FEM_EARG,
                       "the frozen adjoint operator assumes the element tangent is symmetric, but the assembled "
                       "blocks disagree: max ‖A - Aᵀ‖/‖A‖ = %g, max ‖C - Bᵀ‖/‖C‖ = %g over %lld elements",
                       wA, wC, (long long)ne);
    return FEM_OK;
}

// measure_λmax of the forward solve (DR.jl:375-378, 459-490, 728-760): the
// Jacobian launch also stores the velocity blocks A, and a power iteration on
// the symmetrically Jacobi-scaled A (estimate_velocity_λmax,
// adjoint_operator.jl:240-281; A x, not Aᵀ x) measures λmax, warm-started from
// the previous estimate's vector at every later assembly. λmax_v = min(the
// Gershgorin bound, safety · measured) for both components.
struct PowerState {
    AdjWork w;
    AdjointOp op;
    double *xx = nullptr, *xy = nullptr, *zx = nullptr, *zy = nullptr, *yx = nullptr, *yy = nullptr;
    bool seeded = false;
    int64_t its = 0;        // λmax_iterations: power steps over the whole solve
    int max_it = 12;
    double rtol = 1e-2, safety = 1.1;
};

static int power_init(fem_stokes* s, PowerState& ps, const fem_solve_params* p, bool plastic) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes, ne = s->mesh->nels;
    ps.w.c = c;
    ps.max_it = p->lmax_power_iterations > 0 ? p->lmax_power_iterations : 12;
    ps.rtol = p->lmax_power_rtol > 0 ? p->lmax_power_rtol : 1e-2;
    const double safety = p->lmax_safety > 0 ? p->lmax_safety : 1.1;
    ps.safety = plastic ? std::max(safety, 1.5) : safety;   // _stokes_spectral_safety (DR.jl:931-933)
    CHK(ps.w.get(&ps.op.A, (size_t)ANU * ANU * ne));
    CHK(ps.w.get(&ps.op.slot, (size_t)NSH * ne));
    for (double** v : {&ps.xx, &ps.xy, &ps.zx, &ps.zy, &ps.yx, &ps.yy}) CHK(ps.w.get(v, nv));
    return FEM_OK;
}

// after a Jacobian assembly with PowerState.op.A set and lmax holding the
// Gershgorin bounds of the two components
static int power_lmax_fwd(fem_stokes* s, PowerState& ps, double lmax[2]) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes;
    double* PCx = s->f[FEM_FIELD_PC_V_X];
    double* PCy = s->f[FEM_FIELD_PC_V_Y];
    std::vector<uint8_t> flag(nv);
    HIPCHK(c, hipMemcpy(flag.data(), s->d_bcflag, nv, hipMemcpyDeviceToHost));
    std::vector<double> hx(nv), hy(nv);
    if (!ps.seeded) {
        for (int64_t i = 0; i < nv; ++i) {   // x .= sin.(i .* 0.7), y .= cos.(i .* 1.3), 1-based i
            hx[i] = std::sin((double)(i + 1) * 0.7);
            hy[i] = std::cos((double)(i + 1) * 1.3);
        }
        ps.seeded = true;
    } else {
        HIPCHK(c, hipMemcpy(hx.data(), ps.xx, nv * 8, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(hy.data(), ps.xy, nv * 8, hipMemcpyDeviceToHost));
    }
    for (int64_t i = 0; i < nv; ++i) {
        if (flag[i] & 1) hx[i] = 0;
        if (flag[i] & 2) hy[i] = 0;
    }
    HIPCHK(c, hipMemcpy(ps.xx, hx.data(), nv * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(ps.xy, hy.data(), nv * 8, hipMemcpyHostToDevice));
    const unsigned gv = grid_for(nv, 256);
    double lam = 0, nrm[3], meas = 0;
    int64_t used = ps.max_it;
    bool done = false;
    CHK(adjoint_norms(s, ps.xx, ps.xy, nullptr, 0, nrm));
    for (int it = 1; it <= ps.max_it; ++it) {
        const double n = std::sqrt(nrm[0] + nrm[1]);
        if (!(n > 0)) return set_err(c, FEM_EARG, "power iteration collapsed to the zero vector");
        hipLaunchKernelGGL(k_adj_power, dim3(gv), dim3(256), 0, c->stream, nv, ps.xx, ps.xy, ps.zx, ps.zy,
                           (const double*)PCx, (const double*)PCy, n);
        CHK(check_launch(c, "k_adj_power"));
        CHK(adjoint_apply(s, ps.op, ps.zx, ps.zy, nullptr, ps.yx, ps.yy, nullptr, nullptr, nullptr, PCx, PCy, true));
        CHK(adjoint_norms(s, ps.yx, ps.yy, nullptr, 0, nrm));
        const double ln = std::sqrt(nrm[0] + nrm[1]);
        std::swap(ps.xx, ps.yx);
        std::swap(ps.xy, ps.yy);
        if (it > 1 && std::fabs(ln - lam) <= ps.rtol * ln) {
            meas = ln;
            used = it;
            done = true;
            break;
        }
        lam = ln;
    }
    if (!done) {
        if (!(std::isfinite(lam) && lam > 0))
            return set_err(c, FEM_EARG, "power iteration produced a non-positive λmax: %g", lam);
        meas = lam;
    }
    ps.its += used;
    const double g = std::max(lmax[0], lmax[1]);
    lmax[0] = lmax[1] = std::min(g, ps.safety * meas);
    return FEM_OK;
}

}  // namespace

struct DRControl {
    double tol = 0;            // converged when min(err_abs, err_rel) < tol
    int64_t max_iterations = 0, iter_max = 0, max_ph = 0, ncheck = 1;
    double rel_drop0 = 1e-2;
    bool freeze = true;
    int verbose = 0;
    const char* tag = "";
    // coupled solve (DR.jl:597-661): run before the Stokes launches of every
    // inner iteration; both loop tests also wait for *thermal_converged
    std::function<int()> thermal_step;
    const bool* thermal_converged = nullptr;
    PowerState* power = nullptr;   // measure_λmax: re-measured at every Jacobian assembly
    double lmax_g0 = 0;            // the initial Gershgorin bound when lmax holds a measured value
};
struct DRResult {
    int64_t itPH = 0, iter = 0, launches = 0, jac = 0;
    double err = INFINITY, err_abs = INFINITY, err_rel = INFINITY, err_v = INFINITY, err_P = INFINITY;
    double lmax_g = 0;
    bool converged = false;
};

// one row (iter, err_v, err_v_x, err_v_y, err_v_z, err_P) of collect_history
// (DR.jl:558-561, 722-727) into the caller's buffer; rows past its capacity are dropped
static void history_row(fem_stokes* s, int64_t iter, double ev, double ex, double ey, double eP) {
    const fem_solve_inputs& in = s->in;
    if (!in.history || !in.history_len || *in.history_len >= in.history_cap) return;
    double* h = in.history + 6 * *in.history_len;
    h[0] = (double)iter; h[1] = ev; h[2] = ex; h[3] = ey; h[4] = 0.0; h[5] = eP;
    ++*in.history_len;
}

// The outer Powell-Hestenes loop (outer residuals without Pnum, errors,
// rel_drop, P += γP RP / M_P) around the inner DYREL loop (Chebyshev-damped
// pseudo-transient steps, λmin from the residual change at each check, λmax
// from re-assembled Jacobian diagnostics unless frozen). lmax holds the
// current λmax on entry and on return.
template <class F>
static int ph_dyrel(fem_stokes* s, const DRSet<F>& d, const LaunchOpts& o, const DRControl& ctl, double lmax[2],
                    KernelTimer& timer, DRResult& r) {
    fem_ctx* c = s->ctx;
    const int64_t nv = s->mesh->nnodes, np = s->mesh->nnodesP, ncheck = ctl.ncheck;
    double al[2], be[2];
    for (int k = 0; k < 2; ++k) cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, 0.0, s->opts.c_fact, &al[k], &be[k]);
    r.lmax_g = ctl.lmax_g0 > 0 ? ctl.lmax_g0 : std::max(lmax[0], lmax[1]);
    const double eps_tol = ctl.tol;
    double err_min = INFINITY, err = 2 * eps_tol;
    double err_v0 = 1.0, err_P0 = 1.0, err_v00 = 1.0;
    double rel_drop = ctl.rel_drop0;
    int64_t iter = 0;
    int rc = FEM_OK;
    IntervalGraph graph;
    const bool coupled = (bool)ctl.thermal_step;
    auto thermal_ok = [&]() { return !coupled || *ctl.thermal_converged; };

    for (int64_t itPH = 1; itPH <= ctl.max_ph; ++itPH) {
        r.itPH = itPH;
        // outer residuals without Pnum (DR.jl:535-554)
        LaunchOpts oo = o;
        oo.use_pnum = false; oo.do_update = false; oo.store_RP = true; oo.store_Rv = true;
        CHK(launch_element(s, d, oo));
        CHK(launch_gather(s, d, oo));
        double r3[3];
        CHK(reduce_outer(s, d, r3));
        timer.collect();
        r.err_P = std::sqrt(r3[0]) / std::sqrt((double)np);
        r.err_v = std::max(std::sqrt(r3[1]), std::sqrt(r3[2])) / (2 * std::sqrt((double)nv));
        history_row(s, iter, r.err_v, std::sqrt(r3[1]) / (2 * std::sqrt((double)nv)),
                    std::sqrt(r3[2]) / (2 * std::sqrt((double)nv)), r.err_P);
        if (itPH == 1) { err_P0 = r.err_P + julia_eps(r.err_P); err_v0 = r.err_v + julia_eps(r.err_v); }
        else if (itPH == 2) { err_P0 = r.err_P + julia_eps(r.err_P); }
        r.err_abs = std::max(r.err_v, r.err_P);
        r.err_rel = std::max(r.err_v / err_v0, r.err_P / err_P0);
        err = std::min(r.err_abs, r.err_rel);
        if (std::isnan(r.err_v) || std::isnan(r.err_P) || std::isnan(err)) {
            rc = set_err(c, FEM_ENAN, "NaN detected in outer loop at PH=%lld", (long long)itPH);
            break;
        }
        if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in outer loop at PH=%lld", (long long)itPH); break; }
        if (ctl.verbose)
            printf("%sitPH = %02lld iter = %06lld err = %.3e abs = %.3e rel = %.3e - norm[Rv=%.3e %.3e, Rp=%.3e %.3e]\n",
                   ctl.tag, (long long)itPH, (long long)iter, err, r.err_abs, r.err_rel, r.err_v, r.err_v / err_v0,
                   r.err_P, r.err_P / err_P0);
        if (err < ctl.tol && thermal_ok()) { r.converged = true; break; }
        if (err > err_min * 1.05) rel_drop = std::max(rel_drop * 0.1, 1.0e-3);
        err_min = std::min(err_min, err);
        const double eps_vel = err * rel_drop;
        int64_t itPT = 0;
        bool stop_total = false;
        while ((err > eps_vel || !thermal_ok()) && itPT <= ctl.iter_max) {
            // a whole check interval that cannot end early is replayed from the
            // graph (every 16th interval stays on the host when kernels are timed)
            const bool whole = graphs_on(s->ctx) && !coupled && ncheck > 1 && iter % ncheck == 0 &&
                               itPT + ncheck - 1 <= ctl.iter_max && iter + ncheck - 1 <= ctl.max_iterations &&
                               !(timer.on && (iter / ncheck) % 16 == 0);
            if (whole) {
                auto launch_iter = [&](int64_t k) -> int {
                    LaunchOpts oi = o;
                    oi.use_pnum = true;
                    oi.do_update = true;
                    oi.store_RP = oi.store_Pnum = oi.store_Rv = k == ncheck;
                    oi.store_Rv0 = k == ncheck - 1;
                    oi.cheb = s->d_cheb;
                    CHK(launch_element(s, d, oi));
                    return launch_gather(s, d, oi);
                };
                CHK(graph.template run<F>(s, ncheck, al, be, launch_iter));
                itPT += ncheck;
                iter += ncheck;
                r.launches += ncheck;
            } else {
            ++itPT;
            ++iter;
            const bool do_check = iter % ncheck == 0;
            // an iteration after which the loop may end stores what the end needs
            // (a coupled loop can also end after the thermal check of iteration 1)
            const bool last_possible = do_check || itPT > ctl.iter_max || iter > ctl.max_iterations ||
                                       (coupled && iter == 1);
            if (do_check && ncheck == 1) {
                HIPCHK(c, hipMemcpyAsync(d.Rv0_x, d.Rv_x, nv * sizeof(F), hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(d.Rv0_y, d.Rv_y, nv * sizeof(F), hipMemcpyDeviceToDevice, c->stream));
            }
            if (coupled) CHK(ctl.thermal_step());
            LaunchOpts oi = o;
            oi.use_pnum = true;
            oi.do_update = true;
            oi.store_RP = last_possible;
            oi.store_Pnum = last_possible;
            oi.store_Rv = last_possible;
            oi.store_Rv0 = ncheck > 1 && ((iter + 1) % ncheck == 0);
            oi.alpha[0] = al[0]; oi.alpha[1] = al[1];
            oi.beta[0] = be[0]; oi.beta[1] = be[1];
            const bool tk = timer.take(iter);
            hipEvent_t e0 = tk ? timer.ev() : nullptr, e1 = tk ? timer.ev() : nullptr, e2 = tk ? timer.ev() : nullptr;
            if (tk) timer.weight.push_back(1);
            timer.mark(e0);
            CHK(launch_element(s, d, oi));
            timer.mark(e1);
            CHK(launch_gather(s, d, oi));
            timer.mark(e2);
            ++r.launches;
            }
            const bool do_check = iter % ncheck == 0;
            if (do_check) {
                double r6[6];
                CHK(reduce_check(s, d, al, r6));
                timer.collect();
                const double evi = std::max(std::sqrt(r6[0]), std::sqrt(r6[1])) / (2 * std::sqrt((double)nv));
                if (iter == ncheck) err_v00 = evi + julia_eps(evi);
                err = evi <= eps_tol ? evi : std::max(evi / err_v00, evi);
                if (std::isnan(err)) { rc = set_err(c, FEM_ENAN, "NaN detected in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (s->in.history) {
                    // the pressure residual of this check's iteration (stored by the check launch)
                    double h3[3];
                    CHK(reduce_outer(s, d, h3));
                    history_row(s, iter, evi, std::sqrt(r6[0]) / (2 * std::sqrt((double)nv)),
                                std::sqrt(r6[1]) / (2 * std::sqrt((double)nv)), std::sqrt(h3[0]) / std::sqrt((double)np));
                }
                if (ctl.verbose > 1) printf("%s  it = %lld, iter = %lld, err = %.3e\n", ctl.tag, (long long)itPT, (long long)iter, err);
                double lmin[2];
                for (int k = 0; k < 2; ++k) {
                    const double den = r6[2 + 2 * k], num = r6[3 + 2 * k];
                    lmin[k] = den == 0 ? 0.0 : std::fabs(num) / den;
                }
                if constexpr (std::is_same<F, double>::value) {
                    if (!ctl.freeze) {
                        ++r.jac;
                        CHK(launch_jacobian(s, o));
                        rc = checked_lmax(s, lmax);
                        if (rc != FEM_OK) break;
                        r.lmax_g = std::max(lmax[0], lmax[1]);
                        if (ctl.power) {
                            rc = power_lmax_fwd(s, *ctl.power, lmax);
                            if (rc != FEM_OK) break;
                        }
                    }
                }
                for (int k = 0; k < 2; ++k)
                    cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, lmin[k], s->opts.c_fact, &al[k], &be[k]);
            }
            if (ctl.verbose && itPT == ctl.iter_max) printf("%s  inner: max iters (%lld) reached at PH=%lld\n", ctl.tag, (long long)ctl.iter_max, (long long)itPH);
            if (iter > ctl.max_iterations) { stop_total = true; break; }
        }
        if (rc != FEM_OK) break;
        hipLaunchKernelGGL(k_pressure_update<F>, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np, d.P,
                           (const F*)d.gammaP, (const F*)d.RP, (const F*)d.MP);
        CHK(check_launch(c, "k_pressure_update"));
        if (stop_total || iter > ctl.max_iterations) break;
    }
    r.iter = iter;
    r.err = err;
    return rc;
}

__global__ void k_pin_bc(int64_t n, const uint8_t* flag, const double* bx, const double* by, double* vx, double* vy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        if (flag[i] & 1) vx[i] = bx[i];
        if (flag[i] & 2) vy[i] = by[i];
    }
}

static int ensure_w32(fem_stokes* s, bool delta) {
    fem_mesh* m = s->mesh;
    fem_ctx* c = s->ctx;
    const int64_t nv = m->nnodes, np = m->nnodesP, ne = m->nels;
    if (s->w32) {
        if (!delta || s->w32->tau_old0) return FEM_OK;
        fem_stokes::W32& w = *s->w32;
        CHK(dalloc_t(c, &s->d_yflag, ne));
        HIPCHK(c, hipMemsetAsync(s->d_yflag, 1, ne, c->stream));
        // +16: 16-byte global->LDS loads of the last block read past the end
        for (float** q : {&w.tau_old0, &w.tau_old1, &w.tau_old2}) {
            CHK(dalloc_t(c, q, NQ * ne + 16));
            HIPCHK(c, hipMemsetAsync(*q, 0, (NQ * ne + 16) * 4, c->stream));
        }
        return FEM_OK;
    }
    s->w32 = new fem_stokes::W32();
    std::memset(s->w32, 0, sizeof(fem_stokes::W32));
    fem_stokes::W32& w = *s->w32;
    float** vl[] = {&w.vx, &w.vy, &w.rate_x, &w.rate_y, &w.Rv_x, &w.Rv_y, &w.Rv0_x,
                    &w.Rv0_y, &w.PC_x, &w.PC_y, &w.rhs_vx, &w.rhs_vy, &w.zero_v};
    for (float** q : vl) CHK(dalloc_t(c, q, nv));
    float** pl[] = {&w.P, &w.RP, &w.Pnum, &w.gammaP, &w.MP, &w.pscale, &w.rhs_P};
    for (float** q : pl) CHK(dalloc_t(c, q, np));
    CHK(dalloc_t(c, &w.geo, 5 * ne));
    CHK(dalloc_t(c, &w.Re, NSH * ne));
    // sized for the smallest tiles (4 rows: the most perimeter nodes), so
    // that any tile_rows the mesh admits fits
    if (tile_shape_ok(m, 4)) {
        CHK(dalloc_t(c, &w.bpart, tile_count(m, 4) * tile_perimeter_nodes(4)));
        CHK(dalloc_t(c, &w.Rn, (int64_t)(2 * m->nx + 1) * (2 * m->ny + 1)));
    }
    HIPCHK(c, hipMemsetAsync(w.zero_v, 0, nv * 4, c->stream));
    hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(5 * ne, 256)), dim3(256), 0, c->stream, 5 * ne,
                       (const double*)m->d_geo, w.geo, 1.0);
    CHK(check_launch(c, "k_cast"));
    return delta ? ensure_w32(s, true) : FEM_OK;
}

static int cast_to32(fem_ctx* c, const double* x, float* y, int64_t n, double scale = 1.0) {
    hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, x, y, scale);
    return check_launch(c, "k_cast");
}
static int add_correction(fem_ctx* c, double* x, const float* dx, int64_t n, double scale) {
    hipLaunchKernelGGL(k_add_correction, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, x, dx, scale);
    return check_launch(c, "k_add_correction");
}

// Mixed precision (precision = 1): the reference's Powell-Hestenes/DYREL
// iteration with its state and outer residuals in FP64 and its inner
// pseudo-transient iterations in FP32 on a correction.
//
// Each outer iteration evaluates the true residual R(x) (no Pnum, DR.jl:535-554)
// in FP64 and applies the reference's errors, criterion and rel_drop rule to it
// unchanged. The inner loop then iterates on a correction δ (zero at the start
// of every outer iteration) of the linear problem R(x) + L δ = 0, L the
// visco-elastic operator linearised at x, in FP32 with the reference's inner
// criterion, Chebyshev parameters, λmin estimates and Pnum augmentation; the
// pseudo-velocities (rates) carry over between outer iterations as in the
// reference. At every convergence check δv is folded into the FP64 state and
// the correction restarts from the FP64 residual there (one FP64 residual
// evaluation per check interval); the pressure update acts on δP, folded in
// before the next outer residual.
//
// Without plasticity L is the visco-elastic operator, exact, and its Jacobian
// diagnostics (preconditioner, λmax) are assembled once. With a Drucker-
// Prager rheology the correction operator keeps the return map: its stress
// term is τ(x + δ) - τ(x), both returned in FP64 about the FP64 state (read in
// blocks with points near yield only) with τ_old an FP32 copy of the stress
// history, and the Jacobian
// diagnostics are re-assembled in FP64 at every check unless frozen, as the
// reference does (DR.jl:726-745). Every FP64 residual evaluation refreshes the
// per-element near-yield flags and counts the elements with a yielding point;
// the count is read back at once, and iterations after an evaluation with a
// count of zero run the cheaper visco-elastic operator, exact there. A point
// that starts to yield inside such a stretch is corrected by the next fold,
// at most fem_tuning.fold_interval iterations later. The solution x meets the
// Dirichlet values before the first iteration.
// With a yield model δ is also folded every fem_tuning.fold_interval
// iterations between checks: the PH/DYREL trajectory under plasticity is
// sensitive to perturbations, and smaller FP32 corrections keep the mixed
// iteration closer to the FP64 one (512² VEP, plastic step: 203k iterations
// folding at checks only, 151k every 25, 144k-165k every 10, 139k in FP64), for
// one FP64 residual per fold.
static int solve_mixed(fem_stokes* s, const fem_solve_params* p, const LaunchOpts& o, bool freeze, KernelTimer& timer,
                       fem_solve_stats* st) {
    fem_ctx* c = s->ctx;
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = m->nnodesP, nip = NQ * m->nels, ncheck = p->check_interval;
    // the return-map operator is used on the check intervals that start from a
    // state with yielding points; otherwise the visco-elastic operator is exact
    const bool delta = o.plastic;
    const int64_t fold_every = c->tune.fold_interval;
    const bool tangent = delta && p->precision == 2;
    CHK(ensure_w32(s, delta));
    fem_stokes::W32& w = *s->w32;
    if (tangent && !w.tangent) CHK(dalloc_t(c, &w.tangent, (int64_t)NTAN * NQ * m->nels));
    const DRSet<double> d64 = set64(s);
    DRSet<float> d32 = set32(s);
    d32.delta = delta;
    d32.hot = s->d_yflag;
    d32.tangent = tangent ? w.tangent : nullptr;
    // yielding elements of the last FP64 evaluation: counted on the device,
    // copied to pinned memory behind it, read at the next synchronisation
    int32_t* d_yield = s->d_flag + 1;
    int32_t* h_yield = reinterpret_cast<int32_t*>(c->h_red + (RED_MAXK - 1));
    *h_yield = 1;
    CHK(cast_to32(c, d64.gammaP, w.gammaP, np));
    CHK(cast_to32(c, d64.MP, w.MP, np));
    CHK(cast_to32(c, s->d_pscale, w.pscale, np));
    hipLaunchKernelGGL(k_pin_bc, dim3(grid_for(nv, 256)), dim3(256), 0, c->stream, nv, (const uint8_t*)s->d_bcflag,
                       (const double*)s->d_bcval_x, (const double*)s->d_bcval_y, d64.vx, d64.vy);
    CHK(check_launch(c, "k_pin_bc"));
    if (delta && o.elastic) {
        CHK(cast_to32(c, d64.tau_old[0], w.tau_old0, nip));
        CHK(cast_to32(c, d64.tau_old[1], w.tau_old1, nip));
        CHK(cast_to32(c, d64.tau_old[2], w.tau_old2, nip));
    }
    // Jacobian diagnostics of the correction operator
    LaunchOpts oj = o;
    oj.plastic = delta;
    CHK(launch_jacobian(s, oj));
    double lmax[2];
    CHK(checked_lmax(s, lmax));
    double lmax_g = std::max(lmax[0], lmax[1]);
    int64_t jac = 1;
    bool jac_elastic = false;   // the last assembly was at a state without yielding
    CHK(cast_to32(c, d64.PC_x, w.PC_x, nv));
    CHK(cast_to32(c, d64.PC_y, w.PC_y, nv));
    for (float* z : {w.vx, w.vy, w.rate_x, w.rate_y, w.Rv0_x, w.Rv0_y}) HIPCHK(c, hipMemsetAsync(z, 0, nv * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(w.P, 0, np * 4, c->stream));
    if (p->verbose) printf("Initial momentum preconditioner λmax_v = (%g, %g)\n", lmax[0], lmax[1]);

    double al[2], be[2];
    for (int k = 0; k < 2; ++k) cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, 0.0, s->opts.c_fact, &al[k], &be[k]);
    const double eps_tol = p->tolerance;
    double err_min = INFINITY, err = 2 * eps_tol, err_abs = INFINITY, err_rel = INFINITY;
    double err_v0 = 1.0, err_P0 = 1.0, err_v00 = 1.0, err_v = INFINITY, err_P = INFINITY;
    double rel_drop = p->rel_drop0;
    int64_t iter = 0, itPH_done = 0, launches = 0;
    bool converged = false;
    int rc = FEM_OK;
    LaunchOpts o32 = o;
    o32.plastic = delta;
    o32.elastic = delta && o.elastic;
    std::vector<std::pair<int64_t, std::unique_ptr<IntervalGraph>>> graphs;   // by segment shape
    hipEvent_t timer_e1 = nullptr;
    // FP64 residual at x; with plasticity it counts the yielding elements
    auto residual64 = [&](const LaunchOpts& oo) -> int {
        if (delta) HIPCHK(c, hipMemsetAsync(d_yield, 0, 8, c->stream));
        CHK(launch_element(s, d64, oo));
        CHK(launch_gather(s, d64, oo));
        if (delta) {
            HIPCHK(c, hipMemcpyAsync(h_yield, d_yield, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            d32.delta = h_yield[0] > 0;
            // the linearised return where every yielding point's linearisation
            // is positive, the two FP64 returns otherwise
            d32.tangent = tangent && h_yield[1] == 0 ? w.tangent : nullptr;
        }
        return FEM_OK;
    };
    // δv into the FP64 state, δv = 0
    auto fold_v = [&]() -> int {
        hipLaunchKernelGGL(k_fold2, dim3(grid_for(nv, 256)), dim3(256), 0, c->stream, nv, d64.vx, d64.vy, w.vx, w.vy);
        return check_launch(c, "k_fold2");
    };
    // the right-hand sides of the correction problem from the FP64 residual
    auto rhs32 = [&]() -> int {
        hipLaunchKernelGGL(k_cast_rhs, dim3(grid_for(std::max(nv, np), 256)), dim3(256), 0, c->stream, nv, np,
                           (const double*)d64.Rv_x, (const double*)d64.Rv_y, (const double*)d64.RP, w.rhs_vx,
                           w.rhs_vy, w.rhs_P);
        return check_launch(c, "k_cast_rhs");
    };
    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        LaunchOpts oo = o;
        oo.use_pnum = false; oo.do_update = false; oo.store_RP = true; oo.store_Rv = true;
        oo.yield_count = delta ? d_yield : nullptr;
        oo.yield_flag = delta ? s->d_yflag : nullptr;
        oo.tangent_out = tangent ? w.tangent : nullptr;
        oo.tangent_bad = tangent ? d_yield + 1 : nullptr;
        CHK(residual64(oo));
        double r3[3];
        CHK(reduce_outer(s, d64, r3));
        timer.collect();
        err_P = std::sqrt(r3[0]) / std::sqrt((double)np);
        err_v = std::max(std::sqrt(r3[1]), std::sqrt(r3[2])) / (2 * std::sqrt((double)nv));
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
            printf("itPH = %02lld iter = %06lld err = %.3e abs = %.3e rel = %.3e - norm[Rv=%.3e %.3e, Rp=%.3e %.3e]\n",
                   (long long)itPH, (long long)iter, err, err_abs, err_rel, err_v, err_v / err_v0, err_P, err_P / err_P0);
        if (err < eps_tol) { converged = true; break; }
        if (err > err_min * 1.05) rel_drop = std::max(rel_drop * 0.1, 1.0e-3);
        err_min = std::min(err_min, err);
        const double eps_vel = err * rel_drop;
        // right-hand side of the correction problem: the FP64 residual
        CHK(rhs32());
        int64_t itPT = 0;
        bool stop_total = false;
        while (err > eps_vel && itPT <= p->iter_max) {
            // the next segment: the inner iterations up to the next check, or
            // fold on intervals with the return map, or iteration cap. Segments
            // are replayed from graphs; every 16th interval runs one iteration
            // at a time from the host, timed.
            const bool timed = timer.on && (iter / ncheck) % 16 == 0;
            int64_t nseg = 1;
            if (ncheck > 1 && !timed) {
                nseg = ncheck - iter % ncheck;
                if (delta && fold_every > 0)
                    nseg = std::min<int64_t>(nseg, fold_every - iter % fold_every);
                nseg = std::max<int64_t>(1, std::min(nseg, std::min(p->iter_max - itPT + 1, p->max_iterations - iter + 1)));
            }
            const int64_t iter_end = iter + nseg;
            // an iteration after which the loop may end stores what the end needs
            const bool lastp = iter_end % ncheck == 0 || itPT + nseg > p->iter_max || iter_end > p->max_iterations;
            // position in the segment of the iteration before a check (its residual is Rv0)
            const int64_t mj = ncheck - 1 - iter % ncheck;
            const int64_t mpos = ncheck > 1 && mj >= 1 && mj <= nseg ? mj : 0;
            if (ncheck == 1 && iter_end % ncheck == 0) {
                HIPCHK(c, hipMemcpyAsync(w.Rv0_x, w.Rv_x, nv * 4, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(w.Rv0_y, w.Rv_y, nv * 4, hipMemcpyDeviceToDevice, c->stream));
            }
            auto seg = [&](const void* chb) -> int {
                LaunchOpts oi = o32;
                oi.use_pnum = true;
                oi.do_update = true;
                oi.alpha[0] = al[0]; oi.alpha[1] = al[1];
                oi.beta[0] = be[0]; oi.beta[1] = be[1];
                oi.cheb = chb;
                for (int64_t j = 1; j <= nseg; ++j) {
                    LaunchOpts oj = oi;
                    oj.store_RP = oj.store_Pnum = oj.store_Rv = lastp && j == nseg;
                    oj.store_Rv0 = j == mpos;
                    CHK(launch_element(s, d32, oj));
                    if (timed) timer.mark(timer_e1);
                    CHK(launch_gather(s, d32, oj));
                }
                return FEM_OK;
            };
            if (timed) {
                hipEvent_t e0 = timer.ev(), e2;
                timer_e1 = timer.ev();
                e2 = timer.ev();
                timer.weight.push_back(nseg);
                timer.kind.push_back(d32.delta ? 1 : 0);
                timer.mark(e0);
                CHK(seg(nullptr));
                timer.mark(e2);
            } else if (graphs_on(c) && nseg > 1) {
                const int64_t key = (((nseg * 2 + (lastp ? 1 : 0)) * (ncheck + 1) + mpos) * 2 + (d32.delta ? 1 : 0)) * 2 +
                                    (d32.tangent ? 1 : 0);
                IntervalGraph* gph = nullptr;
                for (auto& kv : graphs)
                    if (kv.first == key) gph = kv.second.get();
                if (!gph) {
                    graphs.emplace_back(key, std::unique_ptr<IntervalGraph>(new IntervalGraph()));
                    gph = graphs.back().second.get();
                }
                CHK(gph->run_once<float>(s, al, be, [&]() { return seg(s->d_cheb); }));
            } else {
                CHK(seg(nullptr));
            }
            itPT += nseg;
            iter = iter_end;
            launches += nseg;
            const bool do_check = iter % ncheck == 0;
            if (do_check) {
                double r6[6];
                CHK(reduce_check(s, d32, al, r6));
                timer.collect();
                const double evi = std::max(std::sqrt(r6[0]), std::sqrt(r6[1])) / (2 * std::sqrt((double)nv));
                if (iter == ncheck) err_v00 = evi + julia_eps(evi);
                err = evi <= eps_tol ? evi : std::max(evi / err_v00, evi);
                if (std::isnan(err)) { rc = set_err(c, FEM_ENAN, "NaN detected in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (err > 1.0e10) { rc = set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in inner loop PH=%lld PT=%lld", (long long)itPH, (long long)itPT); break; }
                if (p->verbose > 1) printf("  it = %lld, iter = %lld, err = %.3e\n", (long long)itPT, (long long)iter, err);
                double lmin[2];
                for (int k = 0; k < 2; ++k) {
                    const double den = r6[2 + 2 * k], num = r6[3 + 2 * k];
                    lmin[k] = den == 0 ? 0.0 : std::fabs(num) / den;
                }
                // fold δv into the FP64 state and restart the correction from the
                // FP64 residual there: FP32 rounding of L δ is relative to the
                // size of δ, so the FP32 iterations only ever reduce the residual
                // of the last ncheck steps, far above their rounding floor, and
                // the operator is renewed at the current state; the yield count
                // of this evaluation picks the next iterations' operator
                CHK(fold_v());
                CHK(residual64(oo));
                const int32_t ycount = delta ? *h_yield : 0;
                CHK(rhs32());
                // without yielding points the diagnostics are the visco-elastic
                // ones: re-assemble only when yielding was seen since the last
                // assembly
                const bool skip_jac = ycount == 0 && jac_elastic;
                if (delta && !freeze && !skip_jac) {
                    jac_elastic = ycount == 0;
                    ++jac;
                    CHK(launch_jacobian(s, o));
                    rc = checked_lmax(s, lmax);
                    if (rc != FEM_OK) break;
                    lmax_g = std::max(lmax[0], lmax[1]);
                    CHK(cast_to32(c, d64.PC_x, w.PC_x, nv));
                    CHK(cast_to32(c, d64.PC_y, w.PC_y, nv));
                }
                for (int k = 0; k < 2; ++k)
                    cheb(2 / std::sqrt(lmax[k]) * s->opts.CFL_v, lmin[k], s->opts.c_fact, &al[k], &be[k]);
            } else if (fold_every > 0 && delta && iter % fold_every == 0) {
                // extra folds between checks with a yield model
                CHK(fold_v());
                CHK(residual64(oo));
                CHK(rhs32());
            }
            if (iter > p->max_iterations) { stop_total = true; break; }
        }
        if (rc != FEM_OK) break;
        // P += γP RP / M_P on the correction, then fold δ into the FP64 state
        hipLaunchKernelGGL(k_pressure_update<float>, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np, w.P,
                           (const float*)w.gammaP, (const float*)w.RP, (const float*)w.MP);
        CHK(check_launch(c, "k_pressure_update"));
        CHK(fold_v());
        CHK(add_correction(c, d64.P, w.P, np, 1.0));
        HIPCHK(c, hipMemsetAsync(w.P, 0, np * 4, c->stream));
        if (stop_total || iter > p->max_iterations) break;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    timer.collect();
    st->kernel_ms_element = timer.ms_elem;
    st->kernel_ms_gather = timer.ms_gather;
    st->timed_launches = timer.n;
    st->kernel_ms_element_rm = timer.ms_elem_rm;
    st->kernel_ms_gather_rm = timer.ms_gather_rm;
    st->timed_launches_rm = timer.n_rm;
    st->itPH = itPH_done;
    st->iter = iter;
    st->err = err;
    st->err_abs = err_abs;
    st->err_rel = err_rel;
    st->err_v = err_v;
    st->err_P = err_P;
    st->converged = converged;
    st->reached_total_iter = iter > p->max_iterations;
    st->lambda_max = std::max(lmax[0], lmax[1]);
    st->lambda_max_gershgorin = lmax_g;
    st->lambda_max_iterations = 0;
    st->jacobian_assemblies = jac;
    st->element_launches = launches;
    return rc;
}

// rotate_stress! (src/stokes/stress_rotation.jl:28-61) in plane strain: at
// every point ω = ½(∂vx/∂y − ∂vy/∂x) and τ_old = R(θ) τ R(θ)ᵀ with θ = ωΔt,
// R = [cos θ  −sin θ; sin θ  cos θ] (GeoParams' 2-D convention). One thread
// per element and its seven points.
__global__ __launch_bounds__(128) void k_rotate_stress2(int64_t ne, const int32_t* el2n, const double* geo,
                                                        const double* vx, const double* vy, const double* t0,
                                                        const double* t1, const double* t2, double* o0, double* o1,
                                                        double* o2, double dt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double ux[NV], uy[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        const int32_t n = el2n[a * ne + e];
        ux[a] = vx[n];
        uy[a] = vy[n];
    }
    const double i00 = geo[0 * ne + e], i01 = geo[1 * ne + e], i10 = geo[2 * ne + e], i11 = geo[3 * ne + e];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double ax = 0, ay = 0, bx = 0, by = 0;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const double d0 = kT.dN[q][a][0], d1 = kT.dN[q][a][1];
            ax += d0 * ux[a]; ay += d1 * ux[a];
            bx += d0 * uy[a]; by += d1 * uy[a];
        }
        const double dvx_dy = ax * i01 + ay * i11, dvy_dx = bx * i00 + by * i10;
        const double th = 0.5 * (dvx_dy - dvy_dx) * dt;
        const double c = cos(th), sn = sin(th);
        const int64_t k = e * NQ + q;
        const double xx = t0[k], yy = t1[k], xy = t2[k];
        o0[k] = c * c * xx - 2 * c * sn * xy + sn * sn * yy;
        o1[k] = sn * sn * xx + 2 * c * sn * xy + c * c * yy;
        o2[k] = c * sn * (xx - yy) + (c * c - sn * sn) * xy;
    }
}



// _transfer_temperature_kernel! (DR.jl:44-47): T_P[DoFsP[i]] = T[el2nP[i]];
// the pressure DoFs 3e + a of element e sit at its vertex nodes a = 0..2
__global__ void k_transfer_T(int64_t nels, const int32_t* el2n, const double* T, double* TP) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nels) return;
    const int64_t e = i / 3;
    const int a = (int)(i - 3 * e);
    TP[i] = T[el2n[a * nels + e]];
}

static int transfer_T(fem_stokes* s, const double* T, double* TP) {
    const int64_t ne = s->mesh->nels;
    hipLaunchKernelGGL(k_transfer_T, dim3(grid_for(3 * ne, 256)), dim3(256), 0, s->ctx->stream, ne,
                       (const int32_t*)s->mesh->d_el2n, T, TP);
    return check_launch(s->ctx, "k_transfer_T");
}

// Shear dissipation at the velocity quadrature points: one residual launch
// of the element kernel (v, P without Pnum, τ_old) storing Φ only; the slot
// residuals it also writes are overwritten by the next assembly
static int shear_heating_launch(fem_stokes* s, const LaunchOpts& o, double* phi) {
    LaunchOpts oo = o;
    oo.use_pnum = false;
    oo.do_update = false;
    oo.store_RP = oo.store_Pnum = oo.store_Rv = oo.store_Rv0 = oo.store_tau = false;
    oo.yield_count = nullptr;
    oo.yield_flag = nullptr;
    oo.phi_out = phi;
    return launch_element(s, set64(s), oo);
}

// solve! on a 2-D state (DR.jl:302-337 and the loop of DR.jl:348-802); with
// cp, the coupled loop of solve_coupled! (DR.jl:514-661)
static int solve_2d(fem_stokes* s, const fem_bc bcs[2], const fem_drucker_prager* dp, const fem_solve_params* p,
                    fem_solve_stats* st, CoupledSide* cp, const fem_drucker_prager_cap* cap = nullptr) {
    fem_ctx* c = s->ctx;
    std::memset(st, 0, sizeof *st);
    HIPCHK(c, hipSetDevice(c->device));
    if (p->check_interval < 1) return set_err(c, FEM_EARG, "check_interval must be positive");
    if (p->precision < 0 || p->precision > 2)
        return set_err(c, FEM_EARG,
                       "precision must be 0 (FP64), 1 (FP32 corrections of FP64 residuals) or 2 (the same with "
                       "the return map linearised about the FP64 state), got %d",
                       p->precision);
    if (cap && s->dim != 2) return set_err(c, FEM_EARG, "the tensile cap is built for 2-D states");
    if (cap && p->precision != 0)
        return set_err(c, FEM_EARG, "solves with the tensile cap run in FP64 (precision 0), got precision %d",
                       p->precision);
    if (s->in.viscosity && cp) return set_err(c, FEM_EARG, "quadrature viscosity is not supported by coupled solves");
    if (s->in.viscosity && p->precision != 0)
        return set_err(c, FEM_EARG, "quadrature viscosity is solved in FP64 (precision 0)");
    if (s->in.history && (p->precision != 0 || s->dim != 2))
        return set_err(c, FEM_EARG, "collect_history is recorded by 2-D FP64 solves (precision 0)");
    if (s->in.history_len) *s->in.history_len = 0;
    if (p->measure_lmax) {   // DR.jl:394, 425-426
        if (s->dim != 2) return set_err(c, FEM_EARG, "measure_λmax is currently supported only in plane strain");
        if (s->in.viscosity) return set_err(c, FEM_EARG, "quadrature viscosity requires measure_λmax=false");
        if (p->precision != 0)
            return set_err(c, FEM_EARG, "measure_λmax runs with the FP64 solve (precision 0), got precision %d",
                           p->precision);
    }
    if (s->dim == 3) {
        if (cp && p->precision != 0)
            return set_err(c, FEM_EARG, "coupled solves run in FP64 (precision 0), got precision %d", p->precision);
        return hex27_solve(s, bcs, dp, p, st, cp);
    }
    const auto t0 = std::chrono::steady_clock::now();
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = m->nnodesP;
    const double dt = p->dt;
    if (cp) {
        // the thermal field on the pressure DoFs before the Stokes set-up (DR.jl:123-124)
        CHK(transfer_T(s, thermal_field_ptr(cp->tc.t, FEM_TFIELD_T), s->f[FEM_FIELD_T]));
        CHK(transfer_T(s, thermal_field_ptr(cp->tc.t, FEM_TFIELD_T0), s->f[FEM_FIELD_T0]));
    }
    LaunchOpts o;
    CHK(prepare(s, dp, dt, p->finite_K, o, p->use_tau_old != 0, cap));
    const bool freeze = p->freeze_jacobian < 0 ? (dp == nullptr && cap == nullptr) : (p->freeze_jacobian != 0);
    CHK(upload_bcs(s, bcs));
    if (!p->skip_scaling) CHK(scaling_impl(s, p->pressure_factor, dt));
    else
        hipLaunchKernelGGL(k_pscale, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np,
                           (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    KernelTimer timer(c, p->time_kernels != 0);
    if (p->precision >= 1) {
        const int rc = solve_mixed(s, p, o, freeze, timer, st);
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc != FEM_OK) return rc;
        if (!st->converged && p->throw_on_failure)
            return set_err(c, FEM_ENOCONV, "Stokes solve did not converge after %lld iterations (residual = %g, tolerance = %g)",
                           (long long)st->iter, st->err_abs, p->tolerance);
        return FEM_OK;
    }
    HIPCHK(c, hipMemsetAsync(s->f[FEM_FIELD_DVDTAU_X], 0, nv * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(s->f[FEM_FIELD_DVDTAU_Y], 0, nv * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(s->f[FEM_FIELD_RV0_X], 0, nv * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(s->f[FEM_FIELD_RV0_Y], 0, nv * 8, c->stream));

    std::unique_ptr<PowerState> power;
    if (p->measure_lmax) {
        power.reset(new PowerState());
        CHK(power_init(s, *power, p, dp != nullptr || cap != nullptr));
        o.Ablk = power->op.A;
    }
    CHK(launch_jacobian(s, o));
    double lmax[2];
    CHK(checked_lmax(s, lmax));
    const double lmax_g0 = std::max(lmax[0], lmax[1]);
    if (power) CHK(power_lmax_fwd(s, *power, lmax));
    if (p->verbose) printf("Initial momentum preconditioner λmax_v = (%g, %g)\n", lmax[0], lmax[1]);

    DRControl ctl;
    ctl.power = power.get();
    ctl.lmax_g0 = lmax_g0;
    ctl.tol = p->tolerance;
    ctl.max_iterations = p->max_iterations;
    ctl.iter_max = p->iter_max;
    ctl.max_ph = p->max_ph_iterations;
    ctl.ncheck = p->check_interval;
    ctl.rel_drop0 = p->rel_drop0;
    ctl.freeze = freeze;
    ctl.verbose = p->verbose;
    if (cp) {
        // ∂T∂τ = R0 = 0 and the dissipation of the initial velocity (DR.jl:524-530)
        CHK(thermal_coupling_begin(cp->tc, cp->bc_T));
        if (cp->shear) CHK(shear_heating_launch(s, o, cp->d_phi));
        double* Tth = thermal_field_ptr(cp->tc.t, FEM_TFIELD_T);
        ctl.thermal_step = [&, cp, Tth]() -> int {
            CHK(thermal_coupling_step(cp->tc));
            CHK(transfer_T(s, Tth, s->f[FEM_FIELD_T]));
            // refreshed after the thermal spectral estimate (DR.jl:659-661)
            if (cp->shear && cp->tc.checked) CHK(shear_heating_launch(s, o, cp->d_phi));
            if (p->verbose > 1 && cp->tc.checked)
                printf("  thermal iter = %lld, err_T = %.3e\n", (long long)cp->tc.iter, cp->tc.err_T);
            return FEM_OK;
        };
        ctl.thermal_converged = &cp->tc.converged;
    }
    DRSet<double> d64 = set64(s);
    if (tile64_ok(m)) {
        // sized for 4-row tiles (the most perimeter nodes): any admitted row count fits
        if (!s->d_bpart64) CHK(dalloc_t(c, &s->d_bpart64, tile_count(m, 4) * tile_perimeter_nodes(4)));
        if (!s->d_Rn64) CHK(dalloc_t(c, &s->d_Rn64, (int64_t)(2 * m->nx + 1) * (2 * m->ny + 1)));
        d64.tile = true;
        d64.bpart = s->d_bpart64;
        d64.Rn = s->d_Rn64;
    }
    DRResult r;
    const int rc = ph_dyrel<double>(s, d64, o, ctl, lmax, timer, r);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    timer.collect();
    st->kernel_ms_element = timer.ms_elem;
    st->kernel_ms_gather = timer.ms_gather;
    st->timed_launches = timer.n;
    st->itPH = r.itPH;
    st->iter = r.iter;
    st->err = r.err;
    st->err_abs = r.err_abs;
    st->err_rel = r.err_rel;
    st->err_v = r.err_v;
    st->err_P = r.err_P;
    st->converged = r.converged;
    st->reached_total_iter = r.iter > p->max_iterations;
    st->lambda_max = std::max(lmax[0], lmax[1]);
    st->lambda_max_gershgorin = r.lmax_g;
    st->lambda_max_iterations = power ? power->its : 0;
    st->jacobian_assemblies = 1 + r.jac;
    st->element_launches = r.launches;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != FEM_OK) return rc;
    if (!r.converged && p->throw_on_failure)
        return set_err(c, FEM_ENOCONV, "Stokes solve did not converge after %lld iterations (residual = %g, tolerance = %g)",
                       (long long)r.iter, r.err_abs, p->tolerance);
    return FEM_OK;
}

static int adjoint_checks(fem_stokes* s) {
    if (s->dim != 2) return set_err(s->ctx, FEM_EARG, "the frozen-operator adjoint runs on 2-D (T7/P1-disc) states");
    const int64_t np = s->mesh->nnodesP;
    std::vector<double> g(np);
    HIPCHK(s->ctx, hipMemcpy(g.data(), s->f[FEM_FIELD_GAMMA_P], np * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < np; ++i)
        if (g[i] != 0) return FEM_OK;
    return set_err(s->ctx, FEM_EARG, "dr.γP is unassembled; run the forward solve! on this state before solve_adjoint!");
}

extern "C" int fem_stokes_adjoint_blocks(fem_stokes* s, const fem_drucker_prager* plastic, double dt, double* host_A,
                                         double* host_B, double* host_C, double* host_D, double* host_W) {
    if (!s || !host_A || !host_B || !host_C || !host_D || !host_W) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim != 2) return set_err(c, FEM_EARG, "the frozen-operator adjoint runs on 2-D (T7/P1-disc) states");
    AdjWork w;
    w.c = c;
    AdjointOp op;
    CHK(adjoint_assemble(s, plastic, dt, w, op));
    const int64_t ne = s->mesh->nels;
    struct { const double* d; double* h; int rows; } out[5] = {
        {op.A, host_A, ANU * ANU}, {op.B, host_B, ANU * ANP}, {op.C, host_C, ANU * ANP}, {op.D, host_D, ANP * ANP},
        {op.W, host_W, ANU * ANP}};
    std::vector<double> t;
    for (auto& o : out) {   // [entry][nels] on the device -> [nels][entry]
        t.resize((size_t)o.rows * ne);
        HIPCHK(c, hipMemcpy(t.data(), o.d, t.size() * 8, hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < o.rows; ++k) o.h[e * o.rows + k] = t[(size_t)k * ne + e];
    }
    return FEM_OK;
}

// solve_adjoint!(dr, mesh, bc_v; dt, objective_v, λv, λP, ...) with operator =
// :blocks (DR_adjoint.jl:396-427 -> _solve_stokes_adjoint_dyrel!, :1-268)
extern "C" int fem_stokes_solve_adjoint(fem_stokes* s, const fem_bc* bc_v, const fem_drucker_prager* plastic,
                                        const double* objective_x, const double* objective_y, double* lam_x,
                                        double* lam_y, double* lam_P, const fem_adjoint_params* p,
                                        fem_adjoint_stats* st) {
    if (!s || !bc_v || !objective_x || !objective_y || !lam_x || !lam_y || !lam_P || !p || !st) return FEM_EARG;
    fem_ctx* c = s->ctx;
    std::memset(st, 0, sizeof *st);
    HIPCHK(c, hipSetDevice(c->device));
    if (p->check_interval < 1) return set_err(c, FEM_EARG, "check_interval must be positive");
    const auto t0 = std::chrono::steady_clock::now();
    CHK(adjoint_checks(s));
    CHK(upload_bcs(s, bc_v));
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = m->nnodesP;
    AdjWork w;
    w.c = c;
    AdjointOp op;
    CHK(adjoint_assemble(s, plastic, p->dt, w, op));
    if (op.sym) CHK(adjoint_check_symmetry(s, op));
    double* PCx = s->f[FEM_FIELD_PC_V_X];
    double* PCy = s->f[FEM_FIELD_PC_V_Y];
    double lmc[2];
    CHK(reduce_lmax(s, lmc));
    const char* labels[2] = {"adjoint vx", "adjoint vy"};
    for (int k = 0; k < 2; ++k)
        if (!(std::isfinite(lmc[k]) && lmc[k] > 0))
            return set_err(c, FEM_EPRECOND,
                           "%s preconditioner produced invalid λmax = %g; check for zero or non-finite preconditioner entries",
                           labels[k], lmc[k]);
    st->lmax_gershgorin = std::max(lmc[0], lmc[1]);
    double lmax_v[2] = {lmc[0], lmc[1]};
    double *Rx, *Ry, *R0x, *R0y, *rx, *ry, *RP;
    CHK(w.get(&Rx, nv)); CHK(w.get(&Ry, nv)); CHK(w.get(&R0x, nv)); CHK(w.get(&R0y, nv));
    CHK(w.get(&rx, nv)); CHK(w.get(&ry, nv)); CHK(w.get(&RP, np));
    const unsigned gv = grid_for(nv, 256);
    if (p->measure_lmax) {
        // estimate_adjoint_λmax (adjoint_operator.jl:564-607): symmetric Jacobi
        // scaling, λP = 0, Dirichlet rows projected out, 100 steps, rtol 1e-3
        double *xx, *xy, *zx, *zy, *yx, *yy, *zP, *sP;
        CHK(w.get(&xx, nv)); CHK(w.get(&xy, nv)); CHK(w.get(&zx, nv)); CHK(w.get(&zy, nv));
        CHK(w.get(&yx, nv)); CHK(w.get(&yy, nv)); CHK(w.get(&zP, np)); CHK(w.get(&sP, np));
        HIPCHK(c, hipMemsetAsync(zP, 0, np * 8, c->stream));
        std::vector<double> sx(nv), sy(nv);
        for (int64_t i = 0; i < nv; ++i) {   // _power_seed(i, d), 1-based i
            sx[i] = std::sin(0.7 * (double)(i + 1));
            sy[i] = std::cos(1.3 * (double)(i + 1));
        }
        std::vector<uint8_t> flag(nv);
        HIPCHK(c, hipMemcpy(flag.data(), s->d_bcflag, nv, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nv; ++i) {
            if (flag[i] & 1) sx[i] = 0;
            if (flag[i] & 2) sy[i] = 0;
        }
        HIPCHK(c, hipMemcpy(xx, sx.data(), nv * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(xy, sy.data(), nv * 8, hipMemcpyHostToDevice));
        double lam = 0, nrm[3];
        int64_t its = 0;
        CHK(adjoint_norms(s, xx, xy, nullptr, 0, nrm));
        for (int64_t it = 1; it <= 100; ++it) {
            its = it;
            const double nx = std::sqrt(nrm[0] + nrm[1]);
            if (!(nx > 0)) return set_err(c, FEM_EARG, "power iteration collapsed to the zero vector");
            hipLaunchKernelGGL(k_adj_power, dim3(gv), dim3(256), 0, c->stream, nv, xx, xy, zx, zy,
                               (const double*)PCx, (const double*)PCy, nx);
            CHK(check_launch(c, "k_adj_power"));
            CHK(adjoint_apply(s, op, zx, zy, zP, yx, yy, sP, nullptr, nullptr, PCx, PCy));
            CHK(adjoint_norms(s, yx, yy, nullptr, 0, nrm));
            const double ln = std::sqrt(nrm[0] + nrm[1]);
            std::swap(xx, yx);
            std::swap(xy, yy);
            const bool conv = it > 1 && std::fabs(ln - lam) <= 1e-3 * ln;
            lam = ln;
            if (conv) break;
        }
        if (!(std::isfinite(lam) && lam > 0))
            return set_err(c, FEM_EARG, "power iteration produced a non-positive λmax: %g", lam);
        lmax_v[0] = lmax_v[1] = lam;
        st->lmax_iterations = its;
    }
    st->lmax = lmax_v[0];
    double dtau[2], al[2], be[2];
    for (int k = 0; k < 2; ++k) {
        dtau[k] = 2 / std::sqrt(lmax_v[k]) * s->opts.CFL_v;
        cheb(dtau[k], 0.0, s->opts.c_fact, &al[k], &be[k]);
    }
    HIPCHK(c, hipMemsetAsync(rx, 0, nv * 8, c->stream));   // the rates start from rest
    HIPCHK(c, hipMemsetAsync(ry, 0, nv * 8, c->stream));
    DRSet<double> d;
    d.Rv_x = Rx; d.Rv_y = Ry; d.Rv0_x = R0x; d.Rv0_y = R0y; d.rate_x = rx; d.rate_y = ry;
    d.PC_x = PCx; d.PC_y = PCy;
    auto assemble = [&]() {
        return adjoint_apply(s, op, lam_x, lam_y, lam_P, Rx, Ry, RP, objective_x, objective_y, nullptr, nullptr);
    };
    const double snv = std::sqrt((double)nv), snp = std::sqrt((double)np);
    int64_t iter = 0, itPH_done = 0;
    double err = INFINITY, err_v0 = 1, err_P0 = 1, err_v = INFINITY, err_P = INFINITY;
    bool converged = false;
    double nrm[3], red[6];
    for (int64_t itPH = 1; itPH <= p->max_ph_iterations; ++itPH) {
        itPH_done = itPH;
        CHK(assemble());
        CHK(adjoint_norms(s, Rx, Ry, RP, np, nrm));
        err_v = std::max(std::sqrt(nrm[0]), std::sqrt(nrm[1])) / snv;
        err_P = std::sqrt(nrm[2]) / snp;
        if (itPH == 1) {
            err_v0 = err_v + julia_eps(err_v);
            err_P0 = err_P + julia_eps(err_P);
        }
        err = std::max(std::min(err_v, err_v / err_v0), std::min(err_P, err_P / err_P0));
        if (p->verbose)
            printf("adj PH=%03lld iter=%06lld err=%.3e Rv=%.3e RP=%.3e\n", (long long)itPH, (long long)iter, err, err_v,
                   err_P);
        if (err < p->tolerance) {
            converged = true;
            break;
        }
        const double target = std::max(err_v * p->rel_drop, p->tolerance);
        int64_t inner = 0;
        while (err_v > target && inner < p->iter_max && iter < p->max_iterations) {
            ++inner;
            ++iter;
            hipLaunchKernelGGL(k_adj_update, dim3(gv), dim3(256), 0, c->stream, nv, (const double*)Rx,
                               (const double*)Ry, R0x, R0y, rx, ry, lam_x, lam_y, (const double*)PCx,
                               (const double*)PCy, (const uint8_t*)s->d_bcflag, al[0], al[1], be[0], be[1]);
            CHK(check_launch(c, "k_adj_update"));
            CHK(assemble());
            if (iter % p->check_interval == 0) {
                CHK(reduce_check(s, d, al, red));
                err_v = std::max(std::sqrt(red[0]), std::sqrt(red[1])) / snv;
                for (int k = 0; k < 2; ++k) {
                    const double den = red[2 + 2 * k], num = red[3 + 2 * k];
                    const double lmin = den == 0 ? 0.0 : std::fabs(num) / den;
                    cheb(dtau[k], lmin, s->opts.c_fact, &al[k], &be[k]);
                }
            }
        }
        hipLaunchKernelGGL(k_adj_lamP, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np, lam_P,
                           (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)RP,
                           (const double*)s->f[FEM_FIELD_M_P]);
        CHK(check_launch(c, "k_adj_lamP"));
        if (iter >= p->max_iterations) break;
    }
    if (!converged) {
        CHK(assemble());
        CHK(adjoint_norms(s, Rx, Ry, RP, np, nrm));
        err_v = std::max(std::sqrt(nrm[0]), std::sqrt(nrm[1])) / snv;
        err_P = std::sqrt(nrm[2]) / snp;
        err = std::max(std::min(err_v, err_v / err_v0), std::min(err_P, err_P / err_P0));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    st->itPH = itPH_done;
    st->iter = iter;
    st->err = err;
    st->err_v = err_v;
    st->err_P = err_P;
    st->converged = converged || err < p->tolerance;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!st->converged && p->throw_on_failure)
        return set_err(c, FEM_ENOCONV,
                       "Stokes adjoint solve did not converge after %lld iterations (residual = %g, tolerance = %g)",
                       (long long)iter, err, p->tolerance);
    return FEM_OK;
}

extern "C" {

int fem_stokes_pressure_scaling(fem_stokes* s, double pressure_factor, double dt, int32_t finite_K) {
    if (!s) return FEM_EARG;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (s->dim == 3) {
        CHK(hex27_pressure_scaling(s, pressure_factor, dt, finite_K));
        return fem_sync(s->ctx);
    }
    LaunchOpts o;
    CHK(prepare(s, nullptr, dt, finite_K, o, true));
    CHK(scaling_impl(s, pressure_factor, dt));
    return fem_sync(s->ctx);
}

int fem_stokes_postprocess(fem_stokes* s, double* out) {
    if (!s || !out) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim == 3) {
        CHK(hex27_postprocess(s, out));
        return fem_sync(c);
    }
    const fem_mesh* m = s->mesh;
    double** f = s->f;
    hipLaunchKernelGGL(k_postprocess2, dim3(grid_for(m->nels, 256)), dim3(256), 0, c->stream, m->nels,
                       (const int32_t*)m->d_el2n, (const double*)m->d_geo, (const double*)f[FEM_FIELD_V_X],
                       (const double*)f[FEM_FIELD_V_Y], (const double*)f[FEM_FIELD_TAU_XX],
                       (const double*)f[FEM_FIELD_TAU_YY], (const double*)f[FEM_FIELD_TAU_XY], out);
    CHK(check_launch(c, "k_postprocess2"));
    return fem_sync(c);
}

int fem_stokes_set_solve_inputs(fem_stokes* s, const fem_solve_inputs* in) {
    if (!s) return FEM_EARG;
    fem_ctx* c = s->ctx;
    if (!in) { s->in = fem_solve_inputs{}; return FEM_OK; }
    HIPCHK(c, hipSetDevice(c->device));
    const fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, nip = m->nels * (s->dim == 2 ? NQ : 27);
    // 0: clean, 1: bad entries, -1: the device check failed
    auto bad = [&](const double* x, int64_t n, bool positive) -> int {
        const int64_t k = count_bad(s, x, n, positive);
        return k < 0 ? -1 : (k > 0 ? 1 : 0);
    };
    const int nbf = (in->body_force[0] != nullptr) + (in->body_force[1] != nullptr) + (in->body_force[2] != nullptr);
    if (nbf != 0 && (nbf != s->dim || in->body_force[s->dim - 1] == nullptr))
        return set_err(c, FEM_EARG, "body_force must contain one assembled load vector per velocity component");
    int b = 0;
    for (int k = 0; k < nbf && b == 0; ++k) b = bad(in->body_force[k], nv, false);
    if (b < 0) return set_err(c, FEM_EHIP, "device check of body_force failed");
    if (b > 0) return set_err(c, FEM_EARG, "body_force must contain finite values");
    for (const double* eta : {in->viscosity, in->scaling_viscosity}) {
        if (!eta) continue;
        b = bad(eta, nip, true);
        if (b < 0) return set_err(c, FEM_EHIP, "device check of the viscosity failed");
        if (b > 0) return set_err(c, FEM_EARG, "viscosity must contain finite positive values");
    }
    if (in->history && (in->history_cap < 0 || !in->history_len))
        return set_err(c, FEM_EARG, "a history buffer needs a nonnegative capacity and a length output");
    s->in = *in;
    return FEM_OK;
}

int fem_stokes_set_pressure_source(fem_stokes* s, const double* Qq, double Q2D, int32_t has_Q2D) {
    if (!s || !Qq) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim != 2) return set_err(c, FEM_EARG, "the pressure source is built for 2-D states");
    const int64_t np = s->mesh->nnodesP, ne = s->mesh->nels;
    double* Q = s->f[FEM_FIELD_Q];
    if (!has_Q2D) {
        const int64_t k = count_bad(s, Qq, np, false);
        if (k < 0) return set_err(c, FEM_EHIP, "device check of Qq failed");
        if (k > 0) return set_err(c, FEM_EARG, "Qq must be finite");
        HIPCHK(c, hipMemcpyAsync(Q, Qq, np * 8, hipMemcpyDeviceToDevice, c->stream));
        return fem_sync(c);
    }
    if (!std::isfinite(Q2D)) return set_err(c, FEM_EARG, "Q2D must be finite");
    // ∫ N^P·Qq dΩ and the count of negative or non-finite values (pressure_residual.jl:32-47)
    const unsigned g = std::min<unsigned>(RED_GRID, grid_for(ne, RED_BLOCK));
    hipLaunchKernelGGL(k_source_area, dim3(g), dim3(RED_BLOCK), 0, c->stream, ne, Qq,
                       (const double*)s->mesh->d_geo, c->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 2, 0,
                       (const double*)c->red_partial, c->red_out);
    CHK(check_launch(c, "k_source_area"));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 2 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double area = c->h_red[0], invalid = c->h_red[1];
    if (invalid != 0) return set_err(c, FEM_EARG, "Qq must be finite and nonnegative");
    if (!(std::isfinite(area) && area > 0)) {
        if (Q2D != 0) return set_err(c, FEM_EARG, "nonzero Q2D requires nonempty positive Qq support");
        HIPCHK(c, hipMemsetAsync(Q, 0, np * 8, c->stream));
        return fem_sync(c);
    }
    hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np, Qq, Q2D / area, Q);
    CHK(check_launch(c, "k_scale_copy"));
    return fem_sync(c);
}

int fem_stokes_residual(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t with_pnum, int32_t finite_K) {
    if (!s) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim == 3) {
        CHK(hex27_residual(s, dp, dt, with_pnum != 0, finite_K));
        return fem_sync(c);
    }
    LaunchOpts o;
    CHK(prepare(s, dp, dt, finite_K, o, true));
    hipLaunchKernelGGL(k_pscale, dim3(grid_for(s->mesh->nnodesP, 256)), dim3(256), 0, c->stream, s->mesh->nnodesP,
                       (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, c->stream));
    o.use_pnum = with_pnum != 0;
    o.store_RP = true;
    o.store_Pnum = with_pnum != 0;
    o.store_Rv = true;
    CHK(launch_element(s, set64(s), o));
    CHK(launch_gather(s, set64(s), o));
    return fem_sync(c);
}

int fem_stokes_jacobian_diagnostics(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t finite_K) {
    if (!s) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim == 3) {
        CHK(hex27_jacobian(s, dp, dt, finite_K));
        return fem_sync(c);
    }
    LaunchOpts o;
    CHK(prepare(s, dp, dt, finite_K, o, true));
    hipLaunchKernelGGL(k_pscale, dim3(grid_for(s->mesh->nnodesP, 256)), dim3(256), 0, c->stream, s->mesh->nnodesP,
                       (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    CHK(launch_jacobian(s, o));
    return fem_sync(c);
}

int fem_stokes_update_current_stress(fem_stokes* s, const fem_drucker_prager* dp, double dt) {
    if (!s) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim == 3) {
        CHK(hex27_update_current_stress(s, dp, dt));
        return fem_sync(c);
    }
    LaunchOpts o;
    CHK(prepare(s, dp, dt, 0, o, true));
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, c->stream));
    o.store_tau = true;
    o.store_Rv = true;
    CHK(launch_element(s, set64(s), o));
    CHK(launch_gather(s, set64(s), o));
    return fem_sync(c);
}

int fem_stokes_time_element_kernel(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t n, double* avg_ms) {
    if (!s || n <= 0 || !avg_ms) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim == 3) return hex27_time_element(s, dp, dt, n, avg_ms);
    LaunchOpts o;
    CHK(prepare(s, dp, dt, 0, o, true));
    o.use_pnum = true;
    o.do_update = true;
    o.alpha[0] = o.alpha[1] = 0.0;
    o.beta[0] = o.beta[1] = 1.0;
    CHK(launch_element(s, set64(s), o));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < n; ++i) CHK(launch_element(s, set64(s), o));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / n;
    return FEM_OK;
}

int fem_stokes_time_correction_kernel(fem_stokes* s, const fem_drucker_prager* dp, double dt, int32_t n,
                                      double* avg_ms) {
    if (!s || n <= 0 || !avg_ms) return FEM_EARG;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (s->dim != 3)
        return set_err(s->ctx, FEM_EARG, "the correction-step timer is built for 3-D states (2-D solves time theirs "
                                         "with time_kernels)");
    return hex27_time_lin(s, dp, dt, n, avg_ms);
}

int fem_stokes_solve(fem_stokes* s, const fem_bc bcs[2], const fem_drucker_prager* dp, const fem_solve_params* p,
                     fem_solve_stats* st);

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched independent solves on one stream. Every problem keeps the control
// flow of the single solve (the mixed-precision solve: FP64 state and outer
// residuals, FP32 inner iterations on a correction; its decisions, criteria
// and iteration counts are those of solve_mixed), but the inner iterations of
// all problems run as grouped launches: one element launch and one grid-node
// update launch per iteration cover every problem's tiles (k_t7_element_group,
// k_tile_update_group, per-problem arguments and Chebyshev scalars in device
// tables), and one check interval of those launches is captured once per
// active set and replayed as a graph. The per-problem work at checks and
// outer iterations (reductions, folds, FP64 residuals) is enqueued problem by
// problem on the same stream with one host synchronisation per round for all
// of them. Problems leave the group as they finish.
// ---------------------------------------------------------------------------
struct BatchItem {
    int32_t idx = 0;
    fem_stokes* s = nullptr;
    const fem_solve_params* p = nullptr;
    LaunchOpts o, oo;              // inner (correction) and outer-residual options
    DRSet<double> d64;
    DRSet<float> d32;
    double lmax[2] = {0, 0}, al[2] = {0, 0}, be[2] = {0, 0};
    double err_min = INFINITY, err = 0, err_abs = INFINITY, err_rel = INFINITY, err_v0 = 1, err_P0 = 1, err_v00 = 1;
    double err_v = INFINITY, err_P = INFINITY, rel_drop = 0, eps_vel = 0;
    int64_t iter = 0, itPH = 0, itPT = 0, launches = 0;
    bool need_outer = true, done = false, converged = false, stop_total = false;
    int rc = FEM_OK;
    double seconds = 0;
};

// the problems a grouped round can take: 2-D structured meshes with the FP32
// tile kernels, mixed precision, no yield model, no array keywords, one context
static bool batch_groupable(int32_t n, fem_stokes* const* states, const fem_drucker_prager* const* plastic,
                            const fem_solve_params* params) {
    for (int32_t i = 0; i < n; ++i) {
        fem_stokes* s = states[i];
        const fem_solve_inputs& in = s->in;
        if (s->ctx != states[0]->ctx || s->dim != 2 || !tile_ok(s->mesh) || params[i].precision < 1 ||
            (plastic && plastic[i]) || params[i].check_interval < 2 || in.viscosity || in.scaling_viscosity ||
            in.P_old || in.history || in.body_force[0] || in.body_force[1] ||
            s->layout_v == FEM_PHASE_ELEMENT || s->layout_P == FEM_PHASE_ELEMENT)
            return false;
    }
    return true;
}

// set-up of one problem: solve_2d's and solve_mixed's work before the loop
static int batch_setup(BatchItem& b, const fem_bc* bcs) {
    fem_stokes* s = b.s;
    const fem_solve_params* p = b.p;
    fem_ctx* c = s->ctx;
    fem_mesh* m = s->mesh;
    const int64_t nv = m->nnodes, np = m->nnodesP;
    CHK(prepare(s, nullptr, p->dt, p->finite_K, b.o, p->use_tau_old != 0));
    CHK(upload_bcs(s, bcs));
    if (!p->skip_scaling) CHK(scaling_impl(s, p->pressure_factor, p->dt));
    else
        hipLaunchKernelGGL(k_pscale, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np,
                           (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    CHK(ensure_w32(s, false));
    fem_stokes::W32& w = *s->w32;
    b.d64 = set64(s);
    b.d32 = set32(s);
    CHK(cast_to32(c, b.d64.gammaP, w.gammaP, np));
    CHK(cast_to32(c, b.d64.MP, w.MP, np));
    CHK(cast_to32(c, s->d_pscale, w.pscale, np));
    hipLaunchKernelGGL(k_pin_bc, dim3(grid_for(nv, 256)), dim3(256), 0, c->stream, nv, (const uint8_t*)s->d_bcflag,
                       (const double*)s->d_bcval_x, (const double*)s->d_bcval_y, b.d64.vx, b.d64.vy);
    CHK(check_launch(c, "k_pin_bc"));
    LaunchOpts oj = b.o;
    oj.plastic = false;
    CHK(launch_jacobian(s, oj));
    CHK(checked_lmax(s, b.lmax));
    CHK(cast_to32(c, b.d64.PC_x, w.PC_x, nv));
    CHK(cast_to32(c, b.d64.PC_y, w.PC_y, nv));
    for (float* z : {w.vx, w.vy, w.rate_x, w.rate_y, w.Rv0_x, w.Rv0_y}) HIPCHK(c, hipMemsetAsync(z, 0, nv * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(w.P, 0, np * 4, c->stream));
    for (int k = 0; k < 2; ++k) cheb(2 / std::sqrt(b.lmax[k]) * s->opts.CFL_v, 0.0, s->opts.c_fact, &b.al[k], &b.be[k]);
    b.err = 2 * p->tolerance;
    b.rel_drop = p->rel_drop0;
    b.oo = b.o;
    b.oo.use_pnum = false; b.oo.do_update = false; b.oo.store_RP = true; b.oo.store_Rv = true;
    b.o.plastic = false;
    b.o.elastic = false;   // the visco-elastic correction operator reads no stress history
    return FEM_OK;
}

// FP64 residual at x and its right-hand-side casts (solve_mixed's residual64 + casts)
static int batch_residual64(BatchItem& b) {
    fem_stokes* s = b.s;
    fem_ctx* c = s->ctx;
    fem_stokes::W32& w = *s->w32;
    const int64_t nv = s->mesh->nnodes, np = s->mesh->nnodesP;
    CHK(launch_element(s, b.d64, b.oo));
    CHK(launch_gather(s, b.d64, b.oo));
    CHK(cast_to32(c, b.d64.Rv_x, w.rhs_vx, nv));
    CHK(cast_to32(c, b.d64.Rv_y, w.rhs_vy, nv));
    return cast_to32(c, b.d64.RP, w.rhs_P, np);
}

// the inner-loop segment of one problem launched on its own (segments cut
// short by an iteration cap); solve_mixed's seg without graphs
static int batch_segment_single(BatchItem& b, int64_t nseg) {
    fem_stokes* s = b.s;
    const int64_t ncheck = b.p->check_interval;
    const int64_t iter_end = b.iter + nseg;
    const bool lastp = iter_end % ncheck == 0 || b.itPT + nseg > b.p->iter_max || iter_end > b.p->max_iterations;
    const int64_t mj = ncheck - 1 - b.iter % ncheck;
    const int64_t mpos = mj >= 1 && mj <= nseg ? mj : 0;
    LaunchOpts oi = b.o;
    oi.use_pnum = true;
    oi.do_update = true;
    oi.alpha[0] = b.al[0]; oi.alpha[1] = b.al[1];
    oi.beta[0] = b.be[0]; oi.beta[1] = b.be[1];
    for (int64_t j = 1; j <= nseg; ++j) {
        LaunchOpts oj = oi;
        oj.store_RP = oj.store_Pnum = oj.store_Rv = lastp && j == nseg;
        oj.store_Rv0 = j == mpos;
        CHK(launch_element(s, b.d32, oj));
        CHK(launch_gather(s, b.d32, oj));
    }
    return FEM_OK;
}

// device tables of one grouped round: element and update arguments of the
// three launch variants (plain, the iteration before a check that stores Rv0,
// the check iteration that stores RP, Pnum and Rv), block offsets, and the
// per-problem Chebyshev scalars the launches read
struct BatchGroup {
    int fl = 0;
    int rows = 8;                          // tile rows of the group's launches
    std::vector<int32_t> members;          // indices into the item list
    ElemArgs<float>* d_ea = nullptr;       // [3][n]
    TileUpdArgs<float>* d_ua = nullptr;    // [3][n]
    int32_t* d_eoff = nullptr;             // [n + 1] tile offsets
    int32_t* d_uoff = nullptr;             // [n + 1] update-block offsets
    int64_t eblocks = 0, ublocks = 0;
    hipGraphExec_t exec = nullptr;
    int64_t ncheck = 0;
};

// the grouped FP32 tile kernels: the visco-elastic correction operator with or without finite bulk terms
static bool group_key_ok(int key) {
    return key == (EF_LIN | EF_ITER | EF_STRUCT | EF_TILE) || key == (EF_LIN | EF_ITER | EF_STRUCT | EF_TILE | EF_BULK);
}
template <int I>
static void launch_group_lin(int key, int rows, hipStream_t st, unsigned grid, const ElemArgs<float>* ea,
                             const int32_t* off, int32_t n) {
    constexpr int FL = EF_LIN | EF_ITER | EF_STRUCT | EF_TILE | ((I & 1) ? EF_BULK : 0);
    if (key == FL) {
        if (rows == 4) hipLaunchKernelGGL((k_t7_element_group<float, FL, 128>), dim3(grid), dim3(128), 0, st, ea, off, n);
        else if (rows == 16) hipLaunchKernelGGL((k_t7_element_group<float, FL, 512>), dim3(grid), dim3(512), 0, st, ea, off, n);
        else hipLaunchKernelGGL((k_t7_element_group<float, FL, 256>), dim3(grid), dim3(256), 0, st, ea, off, n);
        return;
    }
    if constexpr (I > 0) launch_group_lin<I - 1>(key, rows, st, grid, ea, off, n);
}

static void free_group(BatchGroup& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (void* q : {(void*)g.d_ea, (void*)g.d_ua, (void*)g.d_eoff, (void*)g.d_uoff})
        if (q) (void)hipFree(q);
    g = BatchGroup();
}

// builds the tables and the interval graph of the active inner problems with key fl
static int build_group(fem_ctx* c, std::vector<BatchItem>& items, const std::vector<int32_t>& mem, int fl,
                       float* d_cheb, int64_t ncheck, BatchGroup& g) {
    free_group(g);
    if (!group_key_ok(fl)) return set_err(c, FEM_EARG, "no grouped element kernel for key %d", fl);
    g.fl = fl;
    g.rows = c->tune.tile_rows;
    g.members = mem;
    g.ncheck = ncheck;
    const int32_t n = (int32_t)mem.size();
    std::vector<ElemArgs<float>> ea(3 * n);
    std::vector<TileUpdArgs<float>> ua(3 * n);
    std::vector<int32_t> eoff(n + 1, 0), uoff(n + 1, 0);
    for (int32_t k = 0; k < n; ++k) {
        BatchItem& b = items[mem[k]];
        for (int v = 0; v < 3; ++v) {
            LaunchOpts oj = b.o;
            oj.use_pnum = true;
            oj.do_update = true;
            oj.store_Rv0 = v == 1;
            oj.store_RP = oj.store_Pnum = oj.store_Rv = v == 2;
            oj.cheb = d_cheb + 4 * (int64_t)b.idx;
            int key = 0;
            CHK(build_elem_args(b.s, b.d32, oj, ea[v * n + k], key));
            if (key != fl) return set_err(c, FEM_EARG, "batched problem %d has another element kernel", b.idx);
            GatherArgs<float> ga;
            build_gather_args(b.s, b.d32, oj, ga);
            build_tile_update_args(b.s, b.d32, ga, ua[v * n + k]);
        }
        eoff[k + 1] = eoff[k] + (int32_t)ea[k].ntiles;
        uoff[k + 1] = uoff[k] + (int32_t)grid_for(ua[k].nT6, GBLOCK);
    }
    g.eblocks = eoff[n];
    g.ublocks = uoff[n];
    CHK(dalloc_t(c, &g.d_ea, 3 * n));
    CHK(dalloc_t(c, &g.d_ua, 3 * n));
    CHK(dalloc_t(c, &g.d_eoff, n + 1));
    CHK(dalloc_t(c, &g.d_uoff, n + 1));
    HIPCHK(c, hipMemcpyAsync(g.d_ea, ea.data(), ea.size() * sizeof(ea[0]), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(g.d_ua, ua.data(), ua.size() * sizeof(ua[0]), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(g.d_eoff, eoff.data(), eoff.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(g.d_uoff, uoff.data(), uoff.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));   // the host vectors go out of scope
    // one check interval: ncheck (element, update) pairs
    const unsigned ge = grid_xcd(g.eblocks, 1), gu = grid_xcd(g.ublocks, 1);
    auto launches = [&]() -> int {
        for (int64_t j = 1; j <= ncheck; ++j) {
            const int v = j == ncheck ? 2 : (j == ncheck - 1 ? 1 : 0);
            launch_group_lin<1>(fl, g.rows, c->stream, ge, g.d_ea + v * n, g.d_eoff, n);
            hipLaunchKernelGGL((k_tile_update_group<float, GBLOCK>), dim3(gu), dim3(GBLOCK), 0, c->stream,
                               (const TileUpdArgs<float>*)(g.d_ua + v * n), (const int32_t*)g.d_uoff, n);
        }
        return check_launch(c, "k_t7_element_group");
    };
    if (!graphs_on(c)) return FEM_OK;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launches();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != FEM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    HIPCHK(c, e);
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(c, ei);
    return FEM_OK;
}

static int run_group(fem_ctx* c, BatchGroup& g) {
    if (g.exec) {
        HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
        return FEM_OK;
    }
    const int32_t n = (int32_t)g.members.size();
    const unsigned ge = grid_xcd(g.eblocks, 1), gu = grid_xcd(g.ublocks, 1);
    for (int64_t j = 1; j <= g.ncheck; ++j) {
        const int v = j == g.ncheck ? 2 : (j == g.ncheck - 1 ? 1 : 0);
        launch_group_lin<1>(g.fl, g.rows, c->stream, ge, g.d_ea + v * n, g.d_eoff, n);
        hipLaunchKernelGGL((k_tile_update_group<float, GBLOCK>), dim3(gu), dim3(GBLOCK), 0, c->stream,
                           (const TileUpdArgs<float>*)(g.d_ua + v * n), (const int32_t*)g.d_uoff, n);
    }
    return check_launch(c, "k_t7_element_group");
}

static void batch_finish(BatchItem& b, fem_solve_stats* st, double t) {
    std::memset(st, 0, sizeof *st);
    st->itPH = b.itPH;
    st->iter = b.iter;
    st->err = b.err;
    st->err_abs = b.err_abs;
    st->err_rel = b.err_rel;
    st->err_v = b.err_v;
    st->err_P = b.err_P;
    st->converged = b.converged;
    st->reached_total_iter = b.iter > b.p->max_iterations;
    st->lambda_max = std::max(b.lmax[0], b.lmax[1]);
    st->lambda_max_gershgorin = st->lambda_max;
    st->jacobian_assemblies = 1;
    st->element_launches = b.launches;
    st->seconds = t;
    b.done = true;
}

static int solve_batch_grouped(int32_t n, fem_stokes* const* states, const fem_bc* const* bcs,
                               const fem_solve_params* params, fem_solve_stats* stats, int32_t* status) {
    fem_ctx* c = states[0]->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    std::vector<BatchItem> items(n);
    for (int32_t i = 0; i < n; ++i) {
        BatchItem& b = items[i];
        b.idx = i;
        b.s = states[i];
        b.p = &params[i];
        std::memset(&stats[i], 0, sizeof stats[i]);
        status[i] = batch_setup(b, bcs[i]);
        if (status[i] != FEM_OK) { b.rc = status[i]; b.done = true; }
    }
    // per-problem Chebyshev scalars (α_x, α_y, β_x, β_y), staged in pinned memory
    float* d_cheb = nullptr;
    float* h_cheb = nullptr;
    double* d_red = nullptr;   // per-problem partial sums (RED_GRID x 6) and results (6)
    double* h_red = nullptr;
    CHK(dalloc_t(c, &d_cheb, 4 * (int64_t)n));
    CHK(dalloc_t(c, &d_red, (int64_t)n * (RED_GRID * 6 + 6)));
    if (hipHostMalloc((void**)&h_cheb, 16 * (size_t)n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h_red, 48 * (size_t)n, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(d_cheb);
        (void)hipFree(d_red);
        if (h_cheb) (void)hipHostFree(h_cheb);
        return set_err(c, FEM_EHIP, "hipHostMalloc failed");
    }
    std::vector<BatchGroup> groups;
    std::vector<int32_t> group_key_members;   // last active grouped set (for rebuilds)
    int rc = FEM_OK;
    auto fail = [&](BatchItem& b, int r) {
        b.rc = r;
        status[b.idx] = r;
        batch_finish(b, &stats[b.idx], now());
    };
    // the reductions of one phase: every listed problem's partial and final
    // kernels, one copy of all results, one synchronisation
    auto reduce_all = [&](const std::vector<int32_t>& which, bool outer) -> int {
        for (int32_t i : which) {
            BatchItem& b = items[i];
            fem_stokes* s = b.s;
            const int64_t nv = s->mesh->nnodes, np = s->mesh->nnodesP;
            double* part = d_red + (int64_t)i * (RED_GRID * 6 + 6);
            double* out = part + RED_GRID * 6;
            if (outer) {
                const unsigned g = std::min<unsigned>(RED_GRID, grid_for(std::max(nv, np), RED_BLOCK));
                hipLaunchKernelGGL(k_reduce_outer<double>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv, np,
                                   (const double*)b.d64.Rv_x, (const double*)b.d64.Rv_y, (const double*)b.d64.RP,
                                   (const double*)b.d64.MP, part);
                hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 3, 0,
                                   (const double*)part, out);
            } else {
                const DRSet<float>& d = b.d32;
                const unsigned g = std::min<unsigned>(RED_GRID, grid_for(nv, RED_BLOCK));
                hipLaunchKernelGGL(k_reduce_check<float>, dim3(g), dim3(RED_BLOCK), 0, c->stream, nv,
                                   (const float*)d.Rv_x, (const float*)d.Rv_y, (const float*)d.Rv0_x,
                                   (const float*)d.Rv0_y, (const float*)d.rate_x, (const float*)d.rate_y,
                                   (const float*)d.PC_x, (const float*)d.PC_y, b.al[0], b.al[1], part);
                hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)g, 6, 0,
                                   (const double*)part, out);
            }
            CHK(check_launch(c, "batched reductions"));
            HIPCHK(c, hipMemcpyAsync(h_red + 6 * i, out, 48, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return FEM_OK;
    };
    while (rc == FEM_OK) {
        // --- outer iterations (solve_mixed's itPH head) of the problems that need one
        std::vector<int32_t> outer;
        for (int32_t i = 0; i < n; ++i)
            if (!items[i].done && items[i].need_outer) outer.push_back(i);
        for (int32_t i : outer) {
            BatchItem& b = items[i];
            const int r = batch_residual64(b);
            if (r != FEM_OK) fail(b, r);
        }
        if ((rc = reduce_all(outer, true)) != FEM_OK) break;
        for (int32_t i : outer) {
            BatchItem& b = items[i];
            if (b.done) continue;
            const double* r3 = h_red + 6 * i;
            const int64_t nv = b.s->mesh->nnodes, np = b.s->mesh->nnodesP;
            ++b.itPH;
            b.err_P = std::sqrt(r3[0]) / std::sqrt((double)np);
            b.err_v = std::max(std::sqrt(r3[1]), std::sqrt(r3[2])) / (2 * std::sqrt((double)nv));
            if (b.itPH == 1) { b.err_P0 = b.err_P + julia_eps(b.err_P); b.err_v0 = b.err_v + julia_eps(b.err_v); }
            else if (b.itPH == 2) { b.err_P0 = b.err_P + julia_eps(b.err_P); }
            b.err_abs = std::max(b.err_v, b.err_P);
            b.err_rel = std::max(b.err_v / b.err_v0, b.err_P / b.err_P0);
            b.err = std::min(b.err_abs, b.err_rel);
            if (std::isnan(b.err_v) || std::isnan(b.err_P) || std::isnan(b.err)) {
                fail(b, set_err(c, FEM_ENAN, "NaN detected in outer loop at PH=%lld", (long long)b.itPH));
                continue;
            }
            if (b.err > 1.0e10) {
                fail(b, set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in outer loop at PH=%lld", (long long)b.itPH));
                continue;
            }
            if (b.err < b.p->tolerance) {
                b.converged = true;
                batch_finish(b, &stats[i], now());
                continue;
            }
            if (b.err > b.err_min * 1.05) b.rel_drop = std::max(b.rel_drop * 0.1, 1.0e-3);
            b.err_min = std::min(b.err_min, b.err);
            b.eps_vel = b.err * b.rel_drop;
            b.itPT = 0;
            b.need_outer = false;   // the right-hand sides were cast with the residual
        }
        // --- one check interval of inner iterations for every problem in its inner loop
        std::vector<int32_t> inner, grouped;
        for (int32_t i = 0; i < n; ++i)
            if (!items[i].done && !items[i].need_outer) inner.push_back(i);
        if (inner.empty()) {
            bool all = true;
            for (auto& b : items) all = all && b.done;
            if (all) break;
            continue;
        }
        std::vector<int64_t> nseg(n, 0);
        for (int32_t i : inner) {
            BatchItem& b = items[i];
            const int64_t ncheck = b.p->check_interval;
            int64_t k = ncheck - b.iter % ncheck;
            k = std::max<int64_t>(1, std::min(k, std::min(b.p->iter_max - b.itPT + 1, b.p->max_iterations - b.iter + 1)));
            nseg[i] = k;
            if (k == ncheck && ncheck == items[inner[0]].p->check_interval) grouped.push_back(i);
        }
        // Chebyshev scalars of this interval
        for (int32_t i : inner) {
            const BatchItem& b = items[i];
            h_cheb[4 * i] = (float)b.al[0]; h_cheb[4 * i + 1] = (float)b.al[1];
            h_cheb[4 * i + 2] = (float)b.be[0]; h_cheb[4 * i + 3] = (float)b.be[1];
        }
        if ((rc = [&]() -> int {
                HIPCHK(c, hipMemcpyAsync(d_cheb, h_cheb, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
                return FEM_OK;
            }()) != FEM_OK)
            break;
        if (!grouped.empty()) {
            // problems of one element-kernel key share a group; rebuild when the set changed
            std::map<int, std::vector<int32_t>> bykey;
            for (int32_t i : grouped) {
                ElemArgs<float> a;
                int key = 0;
                LaunchOpts oj = items[i].o;
                oj.use_pnum = true;
                if ((rc = build_elem_args(items[i].s, items[i].d32, oj, a, key)) != FEM_OK) break;
                bykey[key].push_back(i);
            }
            if (rc != FEM_OK) break;
            bool same = groups.size() == bykey.size();
            if (same) {
                size_t gi = 0;
                for (auto& kv : bykey) same = same && groups[gi].fl == kv.first && groups[gi++].members == kv.second;
            }
            if (!same) {
                for (auto& g : groups) free_group(g);
                groups.assign(bykey.size(), BatchGroup());
                size_t gi = 0;
                for (auto& kv : bykey) {
                    if ((rc = build_group(c, items, kv.second, kv.first, d_cheb, items[kv.second[0]].p->check_interval,
                                          groups[gi++])) != FEM_OK)
                        break;
                }
                if (rc != FEM_OK) break;
                if (graphs_on(c))
                    HIPCHK(c, hipMemcpyAsync(d_cheb, h_cheb, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
            }
            for (auto& g : groups)
                if ((rc = run_group(c, g)) != FEM_OK) break;
            if (rc != FEM_OK) break;
        }
        for (int32_t i : inner) {
            if (std::find(grouped.begin(), grouped.end(), i) != grouped.end()) continue;
            const int r = batch_segment_single(items[i], nseg[i]);
            if (r != FEM_OK) fail(items[i], r);
        }
        for (int32_t i : inner) {
            BatchItem& b = items[i];
            if (b.done) continue;
            b.itPT += nseg[i];
            b.iter += nseg[i];
            b.launches += nseg[i];
        }
        // --- checks
        std::vector<int32_t> checks;
        for (int32_t i : inner)
            if (!items[i].done && items[i].iter % items[i].p->check_interval == 0) checks.push_back(i);
        if ((rc = reduce_all(checks, false)) != FEM_OK) break;
        for (int32_t i : checks) {
            BatchItem& b = items[i];
            fem_stokes* s = b.s;
            fem_stokes::W32& w = *s->w32;
            const int64_t nv = s->mesh->nnodes;
            const double* r6 = h_red + 6 * i;
            const double evi = std::max(std::sqrt(r6[0]), std::sqrt(r6[1])) / (2 * std::sqrt((double)nv));
            if (b.iter == b.p->check_interval) b.err_v00 = evi + julia_eps(evi);
            b.err = evi <= b.p->tolerance ? evi : std::max(evi / b.err_v00, evi);
            if (std::isnan(b.err)) {
                fail(b, set_err(c, FEM_ENAN, "NaN detected in inner loop PH=%lld PT=%lld", (long long)b.itPH,
                                (long long)b.itPT));
                continue;
            }
            if (b.err > 1.0e10) {
                fail(b, set_err(c, FEM_EKABOOM, "Kaboom! Error > 1e10 in inner loop PH=%lld PT=%lld",
                                (long long)b.itPH, (long long)b.itPT));
                continue;
            }
            double lmin[2];
            for (int k = 0; k < 2; ++k) {
                const double den = r6[2 + 2 * k], num = r6[3 + 2 * k];
                lmin[k] = den == 0 ? 0.0 : std::fabs(num) / den;
            }
            const int r = [&]() -> int {
                CHK(add_correction(c, b.d64.vx, w.vx, nv, 1.0));
                CHK(add_correction(c, b.d64.vy, w.vy, nv, 1.0));
                HIPCHK(c, hipMemsetAsync(w.vx, 0, nv * 4, c->stream));
                HIPCHK(c, hipMemsetAsync(w.vy, 0, nv * 4, c->stream));
                return batch_residual64(b);
            }();
            if (r != FEM_OK) { fail(b, r); continue; }
            for (int k = 0; k < 2; ++k)
                cheb(2 / std::sqrt(b.lmax[k]) * s->opts.CFL_v, lmin[k], s->opts.c_fact, &b.al[k], &b.be[k]);
        }
        // --- ends of inner loops: P += γP RP / M_P on the correction, fold δ
        for (int32_t i : inner) {
            BatchItem& b = items[i];
            if (b.done) continue;
            if (b.iter > b.p->max_iterations) b.stop_total = true;
            if (!b.stop_total && b.err > b.eps_vel && b.itPT <= b.p->iter_max) continue;
            fem_stokes* s = b.s;
            fem_stokes::W32& w = *s->w32;
            const int64_t nv = s->mesh->nnodes, np = s->mesh->nnodesP;
            const int r = [&]() -> int {
                hipLaunchKernelGGL(k_pressure_update<float>, dim3(grid_for(np, 256)), dim3(256), 0, c->stream, np,
                                   w.P, (const float*)w.gammaP, (const float*)w.RP, (const float*)w.MP);
                CHK(check_launch(c, "k_pressure_update"));
                CHK(add_correction(c, b.d64.vx, w.vx, nv, 1.0));
                CHK(add_correction(c, b.d64.vy, w.vy, nv, 1.0));
                CHK(add_correction(c, b.d64.P, w.P, np, 1.0));
                HIPCHK(c, hipMemsetAsync(w.vx, 0, nv * 4, c->stream));
                HIPCHK(c, hipMemsetAsync(w.vy, 0, nv * 4, c->stream));
                HIPCHK(c, hipMemsetAsync(w.P, 0, np * 4, c->stream));
                return FEM_OK;
            }();
            if (r != FEM_OK) { fail(b, r); continue; }
            if (b.stop_total || b.itPH >= b.p->max_ph_iterations) batch_finish(b, &stats[i], now());
            else b.need_outer = true;
        }
    }
    const hipError_t es = hipStreamSynchronize(c->stream);
    for (auto& g : groups) free_group(g);
    (void)hipFree(d_cheb);
    (void)hipFree(d_red);
    (void)hipHostFree(h_cheb);
    (void)hipHostFree(h_red);
    if (rc != FEM_OK) return rc;
    HIPCHK(c, es);
    for (int32_t i = 0; i < n; ++i) {
        BatchItem& b = items[i];
        if (status[i] == FEM_OK && !b.converged && b.p->throw_on_failure)
            status[i] = set_err(c, FEM_ENOCONV,
                                "Stokes solve did not converge after %lld iterations (residual = %g, tolerance = %g)",
                                (long long)b.iter, b.err_abs, b.p->tolerance);
    }
    return FEM_OK;
}

extern "C" {

int fem_stokes_solve(fem_stokes* s, const fem_bc bcs[2], const fem_drucker_prager* dp, const fem_solve_params* p,
                     fem_solve_stats* st);

int fem_stokes_solve_batch(int32_t n, fem_stokes* const* states, const fem_bc* const* bcs,
                           const fem_drucker_prager* const* plastic, const fem_solve_params* params,
                           fem_solve_stats* stats, int32_t* status, int32_t nthreads) {
    (void)nthreads;
    if (n < 0) return FEM_EARG;
    if (n == 0) return FEM_OK;
    if (!states || !bcs || !params || !stats || !status) return FEM_EARG;
    for (int32_t i = 0; i < n; ++i) {
        if (!states[i] || !bcs[i]) return FEM_EARG;
        if (states[i]->ctx != states[0]->ctx)
            return set_err(states[i]->ctx, FEM_EARG,
                           "batched solves share one context (stream): state %d has another", i);
        status[i] = FEM_OK;
    }
    int rc = FEM_OK;
    if (batch_groupable(n, states, plastic, params)) {
        rc = solve_batch_grouped(n, states, bcs, params, stats, status);
    } else {
        // the other problem kinds run one after the other on the shared stream
        for (int32_t i = 0; i < n; ++i)
            status[i] = fem_stokes_solve(states[i], bcs[i], plastic ? plastic[i] : nullptr, &params[i], &stats[i]);
    }
    if (rc != FEM_OK) return rc;
    for (int32_t i = 0; i < n; ++i)
        if (status[i] != FEM_OK) return status[i];
    return FEM_OK;
}

// solve! + _solve_stokes_dyrel! (DR.jl:302-337, 348-802)
int fem_stokes_solve(fem_stokes* s, const fem_bc bcs[2], const fem_drucker_prager* dp, const fem_solve_params* p,
                     fem_solve_stats* st) {
    if (!s || !bcs || !p || !st) return FEM_EARG;
    return solve_2d(s, bcs, dp, p, st, nullptr);
}

// rotate_stress!(dr, mesh, Δt) (stress_rotation.jl:16-19)
int fem_stokes_rotate_stress(fem_stokes* s, double dt) {
    if (!s) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (!std::isfinite(dt)) return set_err(c, FEM_EARG, "dt must be finite");
    if (s->dim == 3) {
        CHK(hex27_rotate_stress(s, dt));
        return fem_sync(c);
    }
    const int64_t ne = s->mesh->nels;
    double** f = s->f;
    hipLaunchKernelGGL(k_rotate_stress2, dim3(grid_for(ne, 128)), dim3(128), 0, c->stream, ne,
                       (const int32_t*)s->mesh->d_el2n, (const double*)s->mesh->d_geo, (const double*)f[FEM_FIELD_V_X],
                       (const double*)f[FEM_FIELD_V_Y], (const double*)f[FEM_FIELD_TAU_XX],
                       (const double*)f[FEM_FIELD_TAU_YY], (const double*)f[FEM_FIELD_TAU_XY], f[FEM_FIELD_TAU_OLD_XX],
                       f[FEM_FIELD_TAU_OLD_YY], f[FEM_FIELD_TAU_OLD_XY], dt);
    CHK(check_launch(c, "k_rotate_stress2"));
    return fem_sync(c);
}

// ---- the tensile cap (rheology.jl:510-787) through the 2-D entry points
int fem_stokes_solve_cap(fem_stokes* s, const fem_bc bcs[2], const fem_drucker_prager_cap* cap,
                         const fem_solve_params* p, fem_solve_stats* st) {
    if (!s || !bcs || !cap || !p || !st) return FEM_EARG;
    return solve_2d(s, bcs, nullptr, p, st, nullptr, cap);
}

int fem_stokes_residual_cap(fem_stokes* s, const fem_drucker_prager_cap* cap, double dt, int32_t with_pnum,
                            int32_t finite_K) {
    if (!s || !cap) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    LaunchOpts o;
    CHK(prepare(s, nullptr, dt, finite_K, o, true, cap));
    hipLaunchKernelGGL(k_pscale, dim3(grid_for(s->mesh->nnodesP, 256)), dim3(256), 0, c->stream, s->mesh->nnodesP,
                       (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, c->stream));
    o.use_pnum = with_pnum != 0;
    o.store_RP = true;
    o.store_Pnum = with_pnum != 0;
    o.store_Rv = true;
    CHK(launch_element(s, set64(s), o));
    CHK(launch_gather(s, set64(s), o));
    return fem_sync(c);
}

int fem_stokes_jacobian_diagnostics_cap(fem_stokes* s, const fem_drucker_prager_cap* cap, double dt,
                                        int32_t finite_K) {
    if (!s || !cap) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    LaunchOpts o;
    CHK(prepare(s, nullptr, dt, finite_K, o, true, cap));
    hipLaunchKernelGGL(k_pscale, dim3(grid_for(s->mesh->nnodesP, 256)), dim3(256), 0, c->stream, s->mesh->nnodesP,
                       (const double*)s->f[FEM_FIELD_GAMMA_P], (const double*)s->f[FEM_FIELD_M_P], s->d_pscale);
    CHK(launch_jacobian(s, o));
    return fem_sync(c);
}

int fem_stokes_update_current_stress_cap(fem_stokes* s, const fem_drucker_prager_cap* cap, double dt) {
    if (!s || !cap) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    LaunchOpts o;
    CHK(prepare(s, nullptr, dt, 1, o, true, cap));
    HIPCHK(c, hipMemsetAsync(s->d_bcflag, 0, s->mesh->nnodes, c->stream));
    o.store_tau = true;
    o.store_Rv = true;
    CHK(launch_element(s, set64(s), o));
    CHK(launch_gather(s, set64(s), o));
    return fem_sync(c);
}

// commit_stokes_plastic_history! (DR.jl:899-958): one residual launch of the
// element kernel that writes γ + Δt γdot and θ + Δt θdot at every point
int fem_stokes_commit_plastic_history(fem_stokes* s, const fem_drucker_prager_cap* cap, double dt) {
    if (!s || !cap) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    LaunchOpts o;
    CHK(prepare(s, nullptr, dt, 1, o, true, cap));
    o.commit = true;
    CHK(launch_element(s, set64(s), o));
    const int64_t n = s->mesh->nels * NQ;
    std::vector<double> h(2 * n);
    HIPCHK(c, hipMemcpyAsync(h.data(), s->f[FEM_FIELD_GAMMA_PL], n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(h.data() + n, s->f[FEM_FIELD_THETA_PL], n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (double x : h)
        if (!std::isfinite(x))
            return set_err(c, FEM_ENAN, "non-finite plastic history: a tensile-cap local return map failed");
    return FEM_OK;
}

// assemble_shear_heating! (shear_heating.jl:18-46)
int fem_stokes_shear_heating(fem_stokes* s, const fem_drucker_prager* dp, double dt, double* phi) {
    if (!s || !phi) return FEM_EARG;
    fem_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (s->dim == 3) {
        CHK(hex27_shear_heating(s, dp, dt, phi));
        return fem_sync(c);
    }
    LaunchOpts o;
    CHK(prepare(s, dp, dt, 0, o, true));
    CHK(shear_heating_launch(s, o, phi));
    return fem_sync(c);
}

// solve_coupled! (DR.jl:81-131) over the coupled loop of _solve_stokes_dyrel!
int fem_coupled_solve(fem_thermal* th, fem_stokes* s, const fem_bc* bc_T, const fem_bc* bc_v,
                      const fem_drucker_prager* dp, const fem_solve_params* p, const fem_coupled_params* cp,
                      fem_solve_stats* st, fem_coupled_stats* cst) {
    if (!th || !s || !bc_v || !p || !st || !cst) return FEM_EARG;
    fem_ctx* c = s->ctx;
    std::memset(st, 0, sizeof *st);
    std::memset(cst, 0, sizeof *cst);
    cst->err_T = INFINITY;
    if (thermal_ctx(th) != c)
        return set_err(c, FEM_EARG, "the thermal and Stokes states must share one context (stream)");
    fem_mesh* tm = thermal_mesh(th);
    if (tm->dim != s->dim) return set_err(c, FEM_EDIM, "the thermal and Stokes meshes must have the same dimension");
    if (tm->nnodes != s->mesh->nnodes || tm->nels != s->mesh->nels || tm->nloc != s->mesh->nloc)
        return set_err(c, FEM_EDIM, "thermal mesh nodes must match the Stokes velocity-node layout");
    if (p->precision != 0)
        return set_err(c, FEM_EARG, "coupled solves run in FP64 (precision 0), got precision %d", p->precision);
    HIPCHK(c, hipSetDevice(c->device));
    CoupledSide side;
    side.tc.t = th;
    side.tc.dt = p->dt;
    side.tc.Tref = cp ? cp->Tref : 273.0;
    side.tc.nout = p->check_interval;
    side.bc_T = bc_T;
    side.shear = cp && cp->shear_heating;
    if (side.shear) {
        // Φ at the velocity quadrature points (7 per T7, 27 per Hex27), the
        // thermal element's own points
        CHK(dalloc_t(c, &side.d_phi, (int64_t)s->mesh->nloc * s->mesh->nels));
        side.tc.phi = side.d_phi;
    }
    const int rc = solve_2d(s, bc_v, dp, p, st, &side);
    hipStreamSynchronize(c->stream);
    hipFree(side.d_phi);
    cst->err_T = side.tc.err_T;
    cst->thermal_iterations = side.tc.iter;
    cst->thermal_converged = side.tc.converged;
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kernel-level residual norms of a DR check (DR.jl:556-557, 700): block b sums
// the squares of entries [b*NORM_CHUNK, (b+1)*NORM_CHUNK) of RP./M_P, Rv_x and
// Rv_y, one block then adds the block partials in block order, so the result
// is run-to-run reproducible and the same for every entry point that uses this
// partition.
// ---------------------------------------------------------------------------
constexpr int64_t NORM_CHUNK = (int64_t)RED_BLOCK * 16;

__global__ __launch_bounds__(RED_BLOCK) void k_norms_chunk(int64_t nv, int64_t np, const double* Rx,
                                                           const double* Ry, const double* RP, const double* MP,
                                                           double* partial) {
    double acc[3] = {0, 0, 0};
    const int64_t base = (int64_t)blockIdx.x * NORM_CHUNK + threadIdx.x;
#pragma unroll 4
    for (int j = 0; j < (int)(NORM_CHUNK / RED_BLOCK); ++j) {
        const int64_t i = base + (int64_t)j * RED_BLOCK;
        if (i < np) {
            const double r = RP[i] / MP[i];
            acc[0] += r * r;
        }
        if (i < nv) {
            const double a = Rx[i], b = Ry[i];
            acc[1] += a * a;
            acc[2] += b * b;
        }
    }
    block_sum_store<3>(acc, partial);
}

// the number of block partials of a norms launch over nv velocity and np pressure entries
static int64_t norms_blocks(int64_t nv, int64_t np) { return (std::max(nv, np) + NORM_CHUNK - 1) / NORM_CHUNK; }

// the two launches of the norms: partials into `partial` ([nb][3]), sums into out[3]
// (Σ(RP/M_P)², ΣRv_x², ΣRv_y²) on `stream`
static hipError_t norms_launch(hipStream_t stream, int64_t nv, int64_t np, const double* Rx, const double* Ry,
                               const double* RP, const double* MP, double* partial, double* out) {
    const int64_t nb = norms_blocks(nv, np);
    hipLaunchKernelGGL(k_norms_chunk, dim3((unsigned)nb), dim3(RED_BLOCK), 0, stream, nv, np, Rx, Ry, RP, MP,
                       partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, stream, (int)nb, 3, 0,
                       (const double*)partial, out);
    return hipGetLastError();
}

extern "C" {

int fem_norms(fem_ctx* c, const double* Rv_x, const double* Rv_y, int64_t nv, const double* RP, const double* MP,
              int64_t np, double* host_out) {
    if (!c || !host_out || nv < 0 || np < 0 || (nv > 0 && (!Rv_x || !Rv_y)) || (np > 0 && (!RP || !MP)))
        return set_err(c, FEM_EARG, "fem_norms: null argument or negative size");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nb = norms_blocks(nv, np);
    double sums[3] = {0, 0, 0};
    if (nb > 0) {
        // the context's reduction workspace holds RED_GRID * RED_MAXK partial
        // values; larger inputs get a workspace of their own for this call
        double* partial = c->red_partial;
        const bool own = nb * 3 > (int64_t)RED_GRID * RED_MAXK;
        if (own) CHK(dalloc_t(c, &partial, nb * 3));
        const hipError_t e = norms_launch(c->stream, nv, np, Rv_x, Rv_y, RP, MP, partial, c->red_out);
        hipError_t e2 = hipSuccess;
        if (e == hipSuccess)
            e2 = hipMemcpyAsync(c->h_red, c->red_out, 3 * 8, hipMemcpyDeviceToHost, c->stream);
        const hipError_t e3 = hipStreamSynchronize(c->stream);
        if (own) (void)hipFree(partial);
        HIPCHK(c, e);
        HIPCHK(c, e2);
        HIPCHK(c, e3);
        std::memcpy(sums, c->h_red, 3 * 8);
    }
    host_out[0] = std::sqrt(sums[1]);
    host_out[1] = std::sqrt(sums[2]);
    host_out[2] = std::sqrt(sums[0]);
    host_out[3] = nv > 0 ? std::max(host_out[0], host_out[1]) / (2 * std::sqrt((double)nv)) : 0.0;
    host_out[4] = np > 0 ? host_out[2] / std::sqrt((double)np) : 0.0;
    return FEM_OK;
}

}  // extern "C"

// out[5] from the three sums (Σ(RP/M_P)², ΣRv_x², ΣRv_y²), as fem_norms forms them on the host
__global__ void k_norms_finish(int64_t nv, int64_t np, const double* sums, double* out) {
    if (threadIdx.x != 0) return;
    const double nx = sqrt(sums[1]), ny = sqrt(sums[2]), nP = sqrt(sums[0]);
    out[0] = nx;
    out[1] = ny;
    out[2] = nP;
    out[3] = nv > 0 ? fmax(nx, ny) / (2 * sqrt((double)nv)) : 0.0;
    out[4] = np > 0 ? nP / sqrt((double)np) : 0.0;
}

extern "C" {

int fem_norms_noalloc(fem_ctx* c, const double* Rv_x, const double* Rv_y, int64_t nv, const double* RP,
                      const double* MP, int64_t np, double* dev_out) {
    if (!c || !dev_out || nv < 0 || np < 0 || (nv > 0 && (!Rv_x || !Rv_y)) || (np > 0 && (!RP || !MP)))
        return set_err(c, FEM_EARG, "fem_norms_noalloc: null argument or negative size");
    const int64_t nb = norms_blocks(nv, np);
    if (nb > NORMS_SCRATCH_BLOCKS)
        return set_err(c, FEM_EARG,
                       "fem_norms_noalloc: %lld entries exceed the static scratch limit of %lld per array "
                       "(use fem_norms)",
                       (long long)std::max(nv, np), (long long)(NORMS_SCRATCH_BLOCKS * NORM_CHUNK));
    HIPCHK(c, hipSetDevice(c->device));
    double* ptr = g_scratch_ptr[c->device];
    if (!ptr) return set_err(c, FEM_EHIP, "fem_norms_noalloc: device scratch not resolved");
    double* sums = ptr + NORMS_SCRATCH_BLOCKS * 3;
    if (nb > 0) {
        HIPCHK(c, norms_launch(c->stream, nv, np, Rv_x, Rv_y, RP, MP, ptr, sums));
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, c->stream, (int64_t)3, sums, 0.0);
        HIPCHK(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_norms_finish, dim3(1), dim3(64), 0, c->stream, nv, np, (const double*)sums, dev_out);
    HIPCHK(c, hipGetLastError());
    return FEM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kernel-level pieces of the inner DR step, for tests and for callers that
// assemble their own loop.
// ---------------------------------------------------------------------------
// update_stokes_velocity_kernel! (helpers.jl:31-41) with the negated step the
// solver passes (DR.jl:690-695): rate = β rate + R / PC, v -= α rate
__global__ void k_dr_velocity_update(int64_t n, double* rate, double* v, const double* R, const double* PC,
                                     double alpha, double beta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = beta * rate[i] + R[i] / PC[i];
    rate[i] = r;
    v[i] += (-alpha) * r;
}

// the visco-elasto-plastic stress with the plane-strain Drucker-Prager return
// at independent points (rheology.jl:35-42, 226-245, 272-334): each point
// brings its strain rate, old stress, pressure and material
__global__ void k_dp_points(int64_t n, double dt, fem_dp_points P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double invG = isinf(P.G[i]) ? 0.0 : 1.0 / P.G[i];
    IpMat<double> m;
    m.etave = 1.0 / (1.0 / P.eta[i] + invG / dt);
    m.inv2Gdt = invG / (2.0 * dt);
    const double sphi = sin(P.phi[i]), spsi = sin(P.psi[i]);
    m.yield0 = cos(P.phi[i]) * P.C[i];
    m.sphi = sphi;
    m.invden = 1.0 / ((m.etave + P.eta_reg[i]) + ((P.Kb[i] * dt) * (-spsi)) * (-sphi));
    double txx, tyy, txy;
    const bool y = stress2d<double, double, true>(P.exx[i], P.eyy[i], P.exy[i], m, P.txx_old[i], P.tyy_old[i],
                                                  P.txy_old[i], P.P[i], txx, tyy, txy);
    P.txx[i] = txx;
    P.tyy[i] = tyy;
    P.txy[i] = txy;
    if (P.yields) P.yields[i] = y ? 1 : 0;
}

extern "C" {

int fem_dr_velocity_update(fem_ctx* c, int64_t n, double* rate, double* v, const double* R, const double* PC,
                           double alpha, double beta) {
    if (!c || n < 0 || (n > 0 && (!rate || !v || !R || !PC))) return set_err(c, FEM_EARG, "fem_dr_velocity_update: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    if (n == 0) return FEM_OK;
    hipLaunchKernelGGL(k_dr_velocity_update, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, rate, v, R, PC,
                       alpha, beta);
    return check_launch(c, "k_dr_velocity_update");
}

int fem_dp_return_map_batch(fem_ctx* c, int64_t n, double dt, const fem_dp_points* pts) {
    if (!c || !pts || n < 0) return set_err(c, FEM_EARG, "fem_dp_return_map_batch: bad arguments");
    if (!(dt > 0) || !std::isfinite(dt)) return set_err(c, FEM_EARG, "dt must be positive and finite");
    const void* need[] = {pts->exx, pts->eyy, pts->exy, pts->txx_old, pts->tyy_old, pts->txy_old, pts->P, pts->eta,
                          pts->G, pts->C, pts->phi, pts->psi, pts->eta_reg, pts->Kb, pts->txx, pts->tyy, pts->txy};
    for (const void* q : need)
        if (n > 0 && !q) return set_err(c, FEM_EARG, "fem_dp_return_map_batch: a required array is NULL");
    HIPCHK(c, hipSetDevice(c->device));
    if (n == 0) return FEM_OK;
    hipLaunchKernelGGL(k_dp_points, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, dt, *pts);
    return check_launch(c, "k_dp_points");
}

}  // extern "C"
