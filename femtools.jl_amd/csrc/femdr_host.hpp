// femdr_host.hpp -- host-side internals shared by the translation units of
// libfemdr.so: the context and mesh objects, kernel timing, error and
// allocation helpers, launch-shape helpers.
#pragma once
#include "../../include/femdr.h"
#include "femdr_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <set>
#include <string>
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
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> tev;     // kernel-timing events (3 per timed iteration)
    size_t tev_used = 0;
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
                ms_elem += a;
                ms_gather += b;
                ++n;
            }
        }
        c->tev_used = 0;
    }
};

struct fem_mesh {
    fem_ctx* ctx = nullptr;
    int64_t nnodes = 0, nels = 0, nnodesP = 0, ns = 0;
    bool structured = false;
    double box[4] = {0, 0, 0, 0};
    std::vector<double> h_coords;      // [nnodes][2]
    std::vector<int32_t> h_el2n;       // [nels][7] 0-based
    std::vector<int32_t> h_boundary;   // 0-based
    int32_t* d_el2n = nullptr;         // [7][nels]
    double* d_geo = nullptr;           // [5][nels]
    int32_t* d_off = nullptr;          // [ns+1]
    int32_t* d_ent = nullptr;
    int32_t* d_node_of = nullptr;      // nullable
    int32_t nx = 0, ny = 0;            // structured box mesh cell counts
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

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }
static inline unsigned grid_xcd(int64_t n, int block) {
    unsigned g = grid_for(n, block);
    return (g + 7u) & ~7u;
}

