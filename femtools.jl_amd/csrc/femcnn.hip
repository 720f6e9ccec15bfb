// femcnn.hip -- the learned warm start's network on gfx950: a 4-level
// encoder-decoder of 3x3 convolutions that maps per-cell (log viscosity,
// phase) features of a 2-D box to an initial velocity perturbation and
// pressure. The reference has no learned component; what it fixes is the
// contract the prediction feeds: any initial guess is accepted and the DR
// iteration alone decides the answer (stokes_2D_pure_shear_triangle.jl:159-163).
//
// Storage: NHWC images and weights in bfloat16, master weights, gradients and
// Adam moments in FP32. Every convolution is a GEMM on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulate):
//   forward  y[P][Co]  = im2col(x)[P][9 Ci] . W[Co][9 Ci]^T      (+ bias, ReLU)
//   dgrad    gx[P][Ci] = im2col(gy)[P][9 Co] . Wf[Ci][9 Co]^T     (Wf: taps flipped, in/out swapped)
//   wgrad    gW[Co][9 Ci] = gy[P][Co]^T . im2col(x)[P][9 Ci]      (split over P, fixed-order sum)
// Layouts: a weight is [Co][ky][kx][Ci]; tap (ky, kx) reads pixel
// (y + ky - 1, x + kx - 1), zero outside the image.
#include "femdr_host.hpp"

#include <cstdint>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bfbits;   // bfloat16 storage

__device__ __forceinline__ bfbits f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;          // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                              // round to nearest, ties to even
    return (bfbits)(u >> 16);
}
__device__ __forceinline__ float bf2f(bfbits h) { return __uint_as_float((uint32_t)h << 16); }

union Frag {
    uint4 u;
    bf16x8 v;
    bfbits h[8];
};

// ---------------------------------------------------------------------------
// GEMM, both operands k-contiguous: C[m][n] = sum_k A[m lda + k] Bt[n ldb + k].
// A 256-thread block holds four waves stacked along m; a wave computes a
// (32 RM) x (32 RN) tile with RM x RN 32x32x16 MFMA accumulators. Lane l
// loads 8 consecutive k of row (l mod 32) at k offset 8 (l / 32) for A and
// the same for the columns of Bt, so A and B agree on which k each register
// element carries. K is a multiple of 16, lda and ldb multiples of 8.
// Accumulator register r of lane l holds C[8 (r / 4) + 4 (l / 32) + r mod 4][l mod 32].
// ---------------------------------------------------------------------------
struct Epilogue {
    const float* bias = nullptr;   // [N], nullable
    int relu = 0;
    float* cf = nullptr;           // fp32 out [M][ldc], nullable
    bfbits* cb = nullptr;          // bf16 out [M][ldc], nullable
    int64_t ldc = 0;
};

template <int RM, int RN>
__global__ __launch_bounds__(256) void k_gemm_nt(int64_t M, int64_t N, int64_t K, const bfbits* __restrict__ A,
                                                 int64_t lda, const bfbits* __restrict__ Bt, int64_t ldb,
                                                 Epilogue ep) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t m0 = ((int64_t)blockIdx.x * 4 + wave) * 32 * RM;
    const int64_t n0 = (int64_t)blockIdx.y * 32 * RN;
    const int r32 = lane & 31, kg = (lane >> 5) * 8;
    f32x16 acc[RM][RN];
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) acc[i][j] = f32x16{};
    const bfbits* ap[RM];
    const bfbits* bp[RN];
    bool aok[RM], bok[RN];
#pragma unroll
    for (int i = 0; i < RM; ++i) {
        const int64_t m = m0 + i * 32 + r32;
        aok[i] = m < M;
        ap[i] = A + (aok[i] ? m : 0) * lda + kg;
    }
#pragma unroll
    for (int j = 0; j < RN; ++j) {
        const int64_t n = n0 + j * 32 + r32;
        bok[j] = n < N;
        bp[j] = Bt + (bok[j] ? n : 0) * ldb + kg;
    }
    for (int64_t k = 0; k < K; k += 16) {
        Frag a[RM], b[RN];
#pragma unroll
        for (int i = 0; i < RM; ++i) a[i].u = aok[i] ? *reinterpret_cast<const uint4*>(ap[i] + k) : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < RN; ++j) b[j].u = bok[j] ? *reinterpret_cast<const uint4*>(bp[j] + k) : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RM; ++i)
#pragma unroll
            for (int j = 0; j < RN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i].v, b[j].v, acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) {
            const int64_t n = n0 + j * 32 + r32;
            if (n >= N) continue;
            const float bn = ep.bias ? ep.bias[n] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t m = m0 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m >= M) continue;
                float v = acc[i][j][r] + bn;
                if (ep.relu) v = fmaxf(v, 0.0f);
                if (ep.cf) ep.cf[m * ep.ldc + n] = v;
                if (ep.cb) ep.cb[m * ep.ldc + n] = f2bf(v);
            }
        }
}

// ---------------------------------------------------------------------------
// GEMM with both operands m/n-contiguous and the reduction over their rows:
// C[m][n] = sum_k A[k lda + m] B[k ldb + n], k in this block's split
// [z Ks, (z + 1) Ks). Register element e of lane l carries k = k0 + 8 (l / 32)
// + e; lanes 0..31 read 32 consecutive m (or n) of one k row, so each load is
// one 64-byte coalesced row segment. The split partials go to
// part[z][M][N] and are summed in split order by k_split_sum.
// ---------------------------------------------------------------------------
template <int RM, int RN>
__global__ __launch_bounds__(256) void k_gemm_tn(int64_t M, int64_t N, int64_t K, int64_t Ks,
                                                 const bfbits* __restrict__ A, int64_t lda,
                                                 const bfbits* __restrict__ B, int64_t ldb, float* __restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t m0 = ((int64_t)blockIdx.x * 4 + wave) * 32 * RM;
    const int64_t n0 = (int64_t)blockIdx.y * 32 * RN;
    const int64_t kb = (int64_t)blockIdx.z * Ks, ke = std::min<int64_t>(K, kb + Ks);
    const int r32 = lane & 31, kg = (lane >> 5) * 8;
    f32x16 acc[RM][RN];
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) acc[i][j] = f32x16{};
    for (int64_t k = kb; k < ke; k += 16) {
        Frag a[RM], b[RN];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t kk = k + kg + e;
            const bool kin = kk < ke;
#pragma unroll
            for (int i = 0; i < RM; ++i) {
                const int64_t m = m0 + i * 32 + r32;
                a[i].h[e] = (kin && m < M) ? A[kk * lda + m] : (bfbits)0;
            }
#pragma unroll
            for (int j = 0; j < RN; ++j) {
                const int64_t n = n0 + j * 32 + r32;
                b[j].h[e] = (kin && n < N) ? B[kk * ldb + n] : (bfbits)0;
            }
        }
#pragma unroll
        for (int i = 0; i < RM; ++i)
#pragma unroll
            for (int j = 0; j < RN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i].v, b[j].v, acc[i][j], 0, 0, 0);
    }
    float* out = part + (int64_t)blockIdx.z * M * N;
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) {
            const int64_t n = n0 + j * 32 + r32;
            if (n >= N) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t m = m0 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < M) out[m * N + n] = acc[i][j][r];
            }
        }
}

// out[i] = sum_z part[z][i] in split order (deterministic)
__global__ void k_split_sum(int64_t n, int splits, const float* part, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(int64_t)z * n + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------
// im2col of an NHWC bf16 image: col[p][t C + c] = x[b][y + ky - 1][x + kx - 1][c]
// (zero outside), p = (b H + y) W + x, t = 3 ky + kx; C a multiple of 8, one
// work-item per 8 channels of one tap of one pixel (16-byte copies).
// ---------------------------------------------------------------------------
__global__ void k_im2col(int64_t B, int H, int W, int C, const bfbits* __restrict__ x, bfbits* __restrict__ col) {
    const int c8 = C / 8;
    const int64_t total = B * H * W * 9 * c8;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int cc = (int)(i % c8);
    const int t = (int)((i / c8) % 9);
    const int64_t p = i / (9 * c8);
    const int xx = (int)(p % W), yy = (int)((p / W) % H);
    const int64_t b = p / ((int64_t)H * W);
    const int sy = yy + t / 3 - 1, sx = xx + t % 3 - 1;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (sy >= 0 && sy < H && sx >= 0 && sx < W)
        v = *reinterpret_cast<const uint4*>(x + (((b * H + sy) * W + sx) * C + cc * 8));
    *reinterpret_cast<uint4*>(col + p * (9 * C) + t * C + cc * 8) = v;
}

// features fp32 [n] -> bf16
__global__ void k_to_bf16(int64_t n, const float* x, bfbits* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// 2x2 average pool, NHWC bf16 (sum in fp32, one rounding)
__global__ void k_pool2(int64_t B, int H, int W, int C, const bfbits* x, bfbits* y) {
    const int h = H / 2, w = W / 2;
    const int64_t n = B * h * w * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t q = i / C;
    const int xx = (int)(q % w), yy = (int)((q / w) % h);
    const int64_t b = q / ((int64_t)h * w);
    const bfbits* s = x + ((b * H + 2 * yy) * W + 2 * xx) * C + c;
    const float v = 0.25f * ((bf2f(s[0]) + bf2f(s[C])) + (bf2f(s[(int64_t)W * C]) + bf2f(s[(int64_t)W * C + C])));
    y[i] = f2bf(v);
}

// u = up2(x_small) + skip, NHWC bf16 at the fine resolution H x W
__global__ void k_up2_add(int64_t B, int H, int W, int C, const bfbits* xs, const bfbits* skip, bfbits* u) {
    const int64_t n = B * H * W * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t q = i / C;
    const int xx = (int)(q % W), yy = (int)((q / W) % H);
    const int64_t b = q / ((int64_t)H * W);
    const float v = bf2f(xs[((b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c]) + bf2f(skip[i]);
    u[i] = f2bf(v);
}

// backward of the pool2 feeding a layer plus the skip gradient of the same
// tensor: g[b][y][x][c] = gs[b][y/2][x/2][c] / 4 + gskip[b][y][x][c] (gskip nullable)
__global__ void k_pool2_back(int64_t B, int H, int W, int C, const float* gs, const float* gskip, float* g) {
    const int64_t n = B * H * W * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t q = i / C;
    const int xx = (int)(q % W), yy = (int)((q / W) % H);
    const int64_t b = q / ((int64_t)H * W);
    float v = 0.25f * gs[((b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
    if (gskip) v += gskip[i];
    g[i] = v;
}

// backward of up2: gs[b][y][x][c] = sum of the 2x2 fine gradients it fed
__global__ void k_up2_back(int64_t B, int h, int w, int C, const float* g, float* gs) {
    const int64_t n = B * h * w * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t q = i / C;
    const int xx = (int)(q % w), yy = (int)((q / w) % h);
    const int64_t b = q / ((int64_t)h * w);
    const int W = 2 * w;
    const float* s = g + ((b * 2 * h + 2 * yy) * W + 2 * xx) * C + c;
    gs[i] = (s[0] + s[C]) + (s[(int64_t)W * C] + s[(int64_t)W * C + C]);
}

// gradient at a layer's pre-activation: gm = g (act > 0) (or g without ReLU), to bf16
__global__ void k_relu_back(int64_t n, const float* g, const bfbits* act, int relu, bfbits* gm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = g[i];
    if (relu && !(bf2f(act[i]) > 0.0f)) v = 0.0f;
    gm[i] = f2bf(v);
}

// bias gradient: db[k] = sum_p gm[p][k] (bf16 values, fp32 sums in a fixed
// partition: block j sums rows [j R, (j + 1) R), then k_split_sum)
constexpr int COLSUM_ROWS = 1024;
__global__ void k_colsum_part(int64_t P, int K, const bfbits* gm, float* part) {
    const int k = threadIdx.x;   // blockDim = K (<= 256)
    const int64_t p0 = (int64_t)blockIdx.x * COLSUM_ROWS, p1 = std::min<int64_t>(P, p0 + COLSUM_ROWS);
    float s = 0.0f;
    for (int64_t p = p0; p < p1; ++p) s += bf2f(gm[p * K + k]);
    part[(int64_t)blockIdx.x * K + k] = s;
}

// mean-squared error of the first 3 of C output channels: loss partials per
// block and the gradient g = 2 (y - t) / n (zero on the padded channels)
constexpr int MSE_BLOCK = 256;
__global__ __launch_bounds__(MSE_BLOCK) void k_mse(int64_t P, int C, const float* y, const float* t, float* g,
                                                  double* part) {
    const int64_t i = (int64_t)blockIdx.x * MSE_BLOCK + threadIdx.x;
    const double inv_n = 1.0 / (3.0 * (double)P);
    double acc[1] = {0.0};
    if (i < P) {
        for (int c = 0; c < C; ++c) {
            float gv = 0.0f;
            if (c < 3) {
                const float d = y[i * C + c] - t[i * 3 + c];
                acc[0] += (double)d * d;
                gv = (float)(2.0 * d * inv_n);
            }
            g[i * C + c] = gv;
        }
    }
    block_sum_store<1>(acc, part);
}

// Adam (Kingma & Ba 2015, Algorithm 1) on fp32 masters; c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 weights for the forward GEMM [Co][9][Ci] and the flipped, transposed
// copy for dgrad: Wf[ci][t][co] = W[co][8 - t][ci]
__global__ void k_weights_bf16(int Co, int Ci, const float* W, bfbits* Wb, bfbits* Wf) {
    const int64_t n = (int64_t)Co * 9 * Ci;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ci = (int)(i % Ci);
    const int t = (int)((i / Ci) % 9);
    const int co = (int)(i / (9 * Ci));
    const bfbits h = f2bf(W[i]);
    Wb[i] = h;
    Wf[((int64_t)ci * 9 + (8 - t)) * Co + co] = h;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

constexpr int CNN_LAYERS = 14;
constexpr int CNN_IN = 16, CNN_OUT = 32;
// (Ci, Co, level) in forward order: e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a head
constexpr int kCi[CNN_LAYERS] = {16, 32, 32, 64, 64, 128, 128, 128, 128, 128, 64, 64, 32, 32};
constexpr int kCo[CNN_LAYERS] = {32, 32, 64, 64, 128, 128, 128, 128, 128, 64, 64, 32, 32, 32};
constexpr int kLev[CNN_LAYERS] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};

int64_t weight_count(int l) { return (int64_t)kCo[l] * 9 * kCi[l]; }

}  // namespace

struct fem_cnn {
    fem_ctx* ctx = nullptr;
    int B = 0, H = 0, W = 0;
    // per layer: fp32 master weights/biases, their gradients and Adam moments,
    // bf16 forward and flipped weights
    float *Wm[CNN_LAYERS] = {}, *bm[CNN_LAYERS] = {}, *gW[CNN_LAYERS] = {}, *gb[CNN_LAYERS] = {};
    float *mW[CNN_LAYERS] = {}, *vW[CNN_LAYERS] = {}, *mb[CNN_LAYERS] = {}, *vb[CNN_LAYERS] = {};
    bfbits *Wb[CNN_LAYERS] = {}, *Wf[CNN_LAYERS] = {};
    // activations (bf16): input, layer outputs, pooled and upsampled tensors
    bfbits* xin = nullptr;
    bfbits* out[CNN_LAYERS] = {};
    bfbits *pool[3] = {}, *up[3] = {};
    const bfbits* lin[CNN_LAYERS] = {};   // each layer's input tensor
    float* yhead = nullptr;               // fp32 head output [P0][CNN_OUT]
    // backward work: fp32 gradients at two ping-pong buffers and three skips,
    // the bf16 pre-activation gradient, im2col and split partials
    float *g0 = nullptr, *g1 = nullptr, *gskip[3] = {};
    bfbits* gm = nullptr;
    bfbits* col = nullptr;
    float* part = nullptr;
    int64_t part_n = 0;
    double* lpart = nullptr;
    double* h_loss = nullptr;
    std::vector<void*> allocs;
    bool weights_ready = false;
};

namespace {

int fill_t(fem_ctx* c, float* p, float v, int64_t n) {
    if (n <= 0) return FEM_OK;
    hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, p, v);
    return check_launch(c, "k_fill");
}

int64_t pixels(const fem_cnn* n, int lev, int B) { return (int64_t)B * (n->H >> lev) * (n->W >> lev); }

template <class T>
int cnn_alloc(fem_cnn* n, T** p, int64_t count) {
    CHK(dalloc_t(n->ctx, p, (size_t)std::max<int64_t>(count, 1)));
    n->allocs.push_back(*p);
    return FEM_OK;
}

// C = A . Bt^T with the epilogue, both operands k-contiguous
int gemm_nt(fem_ctx* c, int64_t M, int64_t N, int64_t K, const bfbits* A, int64_t lda, const bfbits* Bt, int64_t ldb,
            const Epilogue& ep) {
    if (K % 16 || lda % 8 || ldb % 8) return set_err(c, FEM_EARG, "gemm: K must be a multiple of 16 and lda, ldb of 8");
    if (N > 32) {
        dim3 g((unsigned)((M + 255) / 256), (unsigned)((N + 63) / 64));
        hipLaunchKernelGGL((k_gemm_nt<2, 2>), g, dim3(256), 0, c->stream, M, N, K, A, lda, Bt, ldb, ep);
    } else {
        dim3 g((unsigned)((M + 255) / 256), (unsigned)((N + 31) / 32));
        hipLaunchKernelGGL((k_gemm_nt<2, 1>), g, dim3(256), 0, c->stream, M, N, K, A, lda, Bt, ldb, ep);
    }
    return check_launch(c, "k_gemm_nt");
}

// C[M][N] = A^T B with A [K][M], B [K][N], split over K, summed in fixed order into out
int gemm_tn(fem_cnn* n, int64_t M, int64_t N, int64_t K, const bfbits* A, const bfbits* B, float* out) {
    fem_ctx* c = n->ctx;
    const int64_t mt = (M + 127) / 128, nt = (N + 63) / 64;
    // about 2048 blocks in total, each split a multiple of 16 rows of K
    int64_t splits = std::max<int64_t>(1, std::min<int64_t>(2048 / std::max<int64_t>(1, mt * nt), (K + 255) / 256));
    while (splits * M * N > n->part_n) --splits;
    if (splits < 1) return set_err(c, FEM_EARG, "gemm_tn: split workspace too small");
    int64_t Ks = (K + splits - 1) / splits;
    Ks = (Ks + 15) / 16 * 16;
    splits = (K + Ks - 1) / Ks;
    dim3 g((unsigned)mt, (unsigned)nt, (unsigned)splits);
    hipLaunchKernelGGL((k_gemm_tn<1, 2>), g, dim3(256), 0, c->stream, M, N, K, Ks, A, M, B, N, n->part);
    CHK(check_launch(c, "k_gemm_tn"));
    hipLaunchKernelGGL(k_split_sum, dim3(grid_for(M * N, 256)), dim3(256), 0, c->stream, M * N, (int)splits,
                       (const float*)n->part, out);
    return check_launch(c, "k_split_sum");
}

int im2col(fem_ctx* c, int64_t B, int H, int W, int C, const bfbits* x, bfbits* col) {
    const int64_t total = B * H * W * 9 * (C / 8);
    hipLaunchKernelGGL(k_im2col, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, B, H, W, C, x, col);
    return check_launch(c, "k_im2col");
}

int conv_forward(fem_cnn* n, int l, int B, const bfbits* x, bfbits* y, float* yf) {
    fem_ctx* c = n->ctx;
    const int lev = kLev[l], H = n->H >> lev, W = n->W >> lev;
    const int64_t P = (int64_t)B * H * W;
    CHK(im2col(c, B, H, W, kCi[l], x, n->col));
    Epilogue ep;
    ep.bias = n->bm[l];
    ep.relu = l != CNN_LAYERS - 1;
    ep.cb = y;
    ep.cf = yf;
    ep.ldc = kCo[l];
    return gemm_nt(c, P, kCo[l], 9 * kCi[l], n->col, 9 * kCi[l], n->Wb[l], 9 * kCi[l], ep);
}

int prepare_weights(fem_cnn* n) {
    fem_ctx* c = n->ctx;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        hipLaunchKernelGGL(k_weights_bf16, dim3(grid_for(weight_count(l), 256)), dim3(256), 0, c->stream, kCo[l],
                           kCi[l], (const float*)n->Wm[l], n->Wb[l], n->Wf[l]);
        CHK(check_launch(c, "k_weights_bf16"));
    }
    n->weights_ready = true;
    return FEM_OK;
}

int check_batch(fem_cnn* n, int B) {
    if (B < 1 || B > n->B) return set_err(n->ctx, FEM_EARG, "batch %d outside 1..%d", B, n->B);
    return FEM_OK;
}

// forward pass of B images (features fp32 [B][H][W][CNN_IN], device)
int forward(fem_cnn* n, const float* x, int B) {
    fem_ctx* c = n->ctx;
    if (!n->weights_ready) CHK(prepare_weights(n));
    const int64_t nin = pixels(n, 0, B) * CNN_IN;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(nin, 256)), dim3(256), 0, c->stream, nin, x, n->xin);
    CHK(check_launch(c, "k_to_bf16"));
    const bfbits* h = n->xin;
    int l = 0;
    for (int lev = 0; lev < 4; ++lev) {
        for (int k = 0; k < 2; ++k, ++l) {
            n->lin[l] = h;
            CHK(conv_forward(n, l, B, h, n->out[l], nullptr));
            h = n->out[l];
        }
        if (lev < 3) {
            const int H = n->H >> lev, W = n->W >> lev, C = kCo[l - 1];
            const int64_t cnt = (int64_t)B * (H / 2) * (W / 2) * C;
            hipLaunchKernelGGL(k_pool2, dim3(grid_for(cnt, 256)), dim3(256), 0, c->stream, (int64_t)B, H, W, C, h,
                               n->pool[lev]);
            CHK(check_launch(c, "k_pool2"));
            h = n->pool[lev];
        }
    }
    for (int lev = 2; lev >= 0; --lev) {
        const int H = n->H >> lev, W = n->W >> lev, C = kCo[l - 1];
        const bfbits* skip = n->out[2 * lev + 1];
        const int64_t cnt = (int64_t)B * H * W * C;
        hipLaunchKernelGGL(k_up2_add, dim3(grid_for(cnt, 256)), dim3(256), 0, c->stream, (int64_t)B, H, W, C, h, skip,
                           n->up[lev]);
        CHK(check_launch(c, "k_up2_add"));
        h = n->up[lev];
        for (int k = 0; k < 2; ++k, ++l) {
            n->lin[l] = h;
            CHK(conv_forward(n, l, B, h, n->out[l], l == CNN_LAYERS - 1 ? n->yhead : nullptr));
            h = n->out[l];
        }
    }
    return FEM_OK;
}

// backward pass from the head gradient in n->g0 ([P0][CNN_OUT] fp32): every
// layer's gW, gb
int backward(fem_cnn* n, int B) {
    fem_ctx* c = n->ctx;
    float* g = n->g0;      // gradient at the current layer's output
    float* gx = n->g1;     // gradient at its input
    for (int l = CNN_LAYERS - 1; l >= 0; --l) {
        const int lev = kLev[l], H = n->H >> lev, W = n->W >> lev, Ci = kCi[l], Co = kCo[l];
        const int64_t P = (int64_t)B * H * W;
        hipLaunchKernelGGL(k_relu_back, dim3(grid_for(P * Co, 256)), dim3(256), 0, c->stream, P * Co, (const float*)g,
                           (const bfbits*)n->out[l], (int)(l != CNN_LAYERS - 1), n->gm);
        CHK(check_launch(c, "k_relu_back"));
        // bias gradient
        const int64_t nb = (P + COLSUM_ROWS - 1) / COLSUM_ROWS;
        hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)nb), dim3(Co), 0, c->stream, P, Co, (const bfbits*)n->gm,
                           n->part);
        hipLaunchKernelGGL(k_split_sum, dim3(grid_for(Co, 256)), dim3(256), 0, c->stream, (int64_t)Co, (int)nb,
                           (const float*)n->part, n->gb[l]);
        CHK(check_launch(c, "k_colsum"));
        // weight gradient: gy^T . im2col(x)
        CHK(im2col(c, B, H, W, Ci, n->lin[l], n->col));
        CHK(gemm_tn(n, Co, 9 * Ci, P, n->gm, n->col, n->gW[l]));
        if (l == 0) break;
        // input gradient: im2col(gy) . Wf^T
        CHK(im2col(c, B, H, W, Co, n->gm, n->col));
        Epilogue ep;
        ep.cf = gx;
        ep.ldc = Ci;
        CHK(gemm_nt(c, P, Ci, 9 * Co, n->col, 9 * Co, n->Wf[l], 9 * Co, ep));
        // route it to the tensor that fed layer l
        const int64_t cnt = P * Ci;
        if (l == 12 || l == 10 || l == 8) {
            // input = up2(prev) + skip: the skip gradient is gx, prev gets up2_back(gx)
            const int slev = lev;   // 0, 1, 2
            HIPCHK(c, hipMemcpyAsync(n->gskip[slev], gx, cnt * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            const int64_t cs = (int64_t)B * (H / 2) * (W / 2) * Ci;
            hipLaunchKernelGGL(k_up2_back, dim3(grid_for(cs, 256)), dim3(256), 0, c->stream, (int64_t)B, H / 2, W / 2,
                               Ci, (const float*)gx, g);
            CHK(check_launch(c, "k_up2_back"));
            // g now holds the gradient at layer l-1's output
        } else if (l == 6 || l == 4 || l == 2) {
            // input = pool2(out[l-1]); out[l-1] is also the skip of its level
            const int flev = lev - 1, Hf = n->H >> flev, Wf = n->W >> flev;
            const int64_t cf = (int64_t)B * Hf * Wf * Ci;
            hipLaunchKernelGGL(k_pool2_back, dim3(grid_for(cf, 256)), dim3(256), 0, c->stream, (int64_t)B, Hf, Wf, Ci,
                               (const float*)gx, (const float*)n->gskip[flev], g);
            CHK(check_launch(c, "k_pool2_back"));
        } else {
            std::swap(g, gx);
        }
    }
    return FEM_OK;
}

}  // namespace

extern "C" {

int fem_cnn_create(fem_ctx* c, int32_t max_batch, int32_t H, int32_t W, fem_cnn** out) {
    if (!c || !out) return FEM_EARG;
    *out = nullptr;
    if (max_batch < 1 || H < 8 || W < 8 || H % 8 || W % 8)
        return set_err(c, FEM_EARG, "fem_cnn_create: batch >= 1 and image sides multiples of 8 (got %d, %dx%d)",
                       max_batch, H, W);
    HIPCHK(c, hipSetDevice(c->device));
    fem_cnn* n = new fem_cnn();
    n->ctx = c;
    n->B = max_batch;
    n->H = H;
    n->W = W;
    auto fail = [&](int rc) {
        for (void* p : n->allocs) (void)hipFree(p);
        delete n;
        return rc;
    };
    int64_t colmax = 0, actmax = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int64_t P = pixels(n, kLev[l], max_batch);
        colmax = std::max(colmax, P * 9 * std::max(kCi[l], kCo[l]));
        actmax = std::max(actmax, P * std::max(kCi[l], kCo[l]));
        const int64_t nw = weight_count(l);
        int rc = FEM_OK;
        for (float** q : {&n->Wm[l], &n->gW[l], &n->mW[l], &n->vW[l]}) rc = rc ? rc : cnn_alloc(n, q, nw);
        for (float** q : {&n->bm[l], &n->gb[l], &n->mb[l], &n->vb[l]}) rc = rc ? rc : cnn_alloc(n, q, kCo[l]);
        rc = rc ? rc : cnn_alloc(n, &n->Wb[l], nw);
        rc = rc ? rc : cnn_alloc(n, &n->Wf[l], nw);
        rc = rc ? rc : cnn_alloc(n, &n->out[l], P * kCo[l]);
        if (rc) return fail(rc);
    }
    int rc = cnn_alloc(n, &n->xin, pixels(n, 0, max_batch) * CNN_IN);
    for (int lev = 0; lev < 3 && !rc; ++lev) {
        const int64_t C = kCo[2 * lev + 1];
        rc = cnn_alloc(n, &n->pool[lev], pixels(n, lev + 1, max_batch) * C);
        rc = rc ? rc : cnn_alloc(n, &n->up[lev], pixels(n, lev, max_batch) * kCi[12 - 2 * lev]);
        rc = rc ? rc : cnn_alloc(n, &n->gskip[lev], pixels(n, lev, max_batch) * kCi[12 - 2 * lev]);
    }
    rc = rc ? rc : cnn_alloc(n, &n->yhead, pixels(n, 0, max_batch) * CNN_OUT);
    rc = rc ? rc : cnn_alloc(n, &n->g0, actmax);
    rc = rc ? rc : cnn_alloc(n, &n->g1, actmax);
    rc = rc ? rc : cnn_alloc(n, &n->gm, actmax);
    rc = rc ? rc : cnn_alloc(n, &n->col, colmax);
    // split partials: 2048 blocks' worth of the largest weight gradient, and
    // the bias column sums
    int64_t pn = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) pn = std::max(pn, 16 * weight_count(l));
    pn = std::max(pn, (pixels(n, 0, max_batch) / COLSUM_ROWS + 1) * 256);
    n->part_n = pn;
    rc = rc ? rc : cnn_alloc(n, &n->part, pn);
    rc = rc ? rc : cnn_alloc(n, &n->lpart, grid_for(pixels(n, 0, max_batch), MSE_BLOCK) + 1);
    if (!rc && hipHostMalloc((void**)&n->h_loss, 8, hipHostMallocDefault) != hipSuccess)
        rc = set_err(c, FEM_EHIP, "hipHostMalloc failed");
    if (rc) return fail(rc);
    for (int l = 0; l < CNN_LAYERS && !rc; ++l) {
        for (float* q : {n->mW[l], n->vW[l], n->gW[l]}) rc = rc ? rc : fill_t(c, q, 0.0f, weight_count(l));
        for (float* q : {n->mb[l], n->vb[l], n->bm[l], n->gb[l]}) rc = rc ? rc : fill_t(c, q, 0.0f, (int64_t)kCo[l]);
        rc = rc ? rc : fill_t(c, n->Wm[l], 0.0f, weight_count(l));
    }
    if (rc) return fail(rc);
    *out = n;
    return FEM_OK;
}

int fem_cnn_destroy(fem_cnn* n) {
    if (!n) return FEM_OK;
    (void)hipSetDevice(n->ctx->device);
    (void)hipStreamSynchronize(n->ctx->stream);
    for (void* p : n->allocs) (void)hipFree(p);
    if (n->h_loss) (void)hipHostFree(n->h_loss);
    delete n;
    return FEM_OK;
}

int fem_cnn_parameter_count(int64_t* weights, int64_t* biases) {
    if (!weights || !biases) return FEM_EARG;
    *weights = *biases = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        *weights += weight_count(l);
        *biases += kCo[l];
    }
    return FEM_OK;
}

static int cnn_params_io(fem_cnn* n, float* hW, float* hb, int which, bool to_device) {
    fem_ctx* c = n->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    int64_t ow = 0, ob = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        float* dw = which == 0 ? n->Wm[l] : n->gW[l];
        float* db = which == 0 ? n->bm[l] : n->gb[l];
        const size_t bw = weight_count(l) * sizeof(float), bb = kCo[l] * sizeof(float);
        if (to_device) {
            HIPCHK(c, hipMemcpyAsync(dw, hW + ow, bw, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(db, hb + ob, bb, hipMemcpyHostToDevice, c->stream));
        } else {
            HIPCHK(c, hipMemcpyAsync(hW + ow, dw, bw, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipMemcpyAsync(hb + ob, db, bb, hipMemcpyDeviceToHost, c->stream));
        }
        ow += weight_count(l);
        ob += kCo[l];
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (to_device && which == 0) {
        for (int l = 0; l < CNN_LAYERS; ++l) {
            CHK(fill_t(c, n->mW[l], 0.0f, weight_count(l)));
            CHK(fill_t(c, n->vW[l], 0.0f, weight_count(l)));
            CHK(fill_t(c, n->mb[l], 0.0f, (int64_t)kCo[l]));
            CHK(fill_t(c, n->vb[l], 0.0f, (int64_t)kCo[l]));
        }
        n->weights_ready = false;
    }
    return FEM_OK;
}

int fem_cnn_set_weights(fem_cnn* n, const float* hW, const float* hb) {
    if (!n || !hW || !hb) return FEM_EARG;
    return cnn_params_io(n, const_cast<float*>(hW), const_cast<float*>(hb), 0, true);
}
int fem_cnn_get_weights(fem_cnn* n, float* hW, float* hb) {
    if (!n || !hW || !hb) return FEM_EARG;
    return cnn_params_io(n, hW, hb, 0, false);
}
int fem_cnn_get_grads(fem_cnn* n, float* hW, float* hb) {
    if (!n || !hW || !hb) return FEM_EARG;
    return cnn_params_io(n, hW, hb, 1, false);
}

int fem_cnn_forward(fem_cnn* n, const float* x, int32_t B, float* y) {
    if (!n || !x || !y) return FEM_EARG;
    fem_ctx* c = n->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    CHK(check_batch(n, B));
    CHK(forward(n, x, B));
    // the three predicted channels of the fp32 head output
    const int64_t P = pixels(n, 0, B);
    HIPCHK(c, hipMemcpy2DAsync(y, 3 * sizeof(float), n->yhead, CNN_OUT * sizeof(float), 3 * sizeof(float), P,
                               hipMemcpyDeviceToDevice, c->stream));
    return FEM_OK;
}

int fem_cnn_train_step(fem_cnn* n, const float* x, const float* t, int32_t B, double lr, int32_t step,
                       double* loss) {
    if (!n || !x || !t || !loss) return FEM_EARG;
    fem_ctx* c = n->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    CHK(check_batch(n, B));
    if (lr > 0 && step < 1) return set_err(c, FEM_EARG, "fem_cnn_train_step: the Adam step count starts at 1");
    CHK(forward(n, x, B));
    const int64_t P = pixels(n, 0, B);
    const unsigned nb = grid_for(P, MSE_BLOCK);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(MSE_BLOCK), 0, c->stream, P, CNN_OUT, (const float*)n->yhead, t, n->g0,
                       n->lpart);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, (int)nb, 1, 0, (const double*)n->lpart,
                       c->red_out);
    CHK(check_launch(c, "k_mse"));
    HIPCHK(c, hipMemcpyAsync(n->h_loss, c->red_out, 8, hipMemcpyDeviceToHost, c->stream));
    CHK(backward(n, B));
    if (lr > 0) {
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 - std::pow(0.9, step)), c2 = (float)(1.0 - std::pow(0.999, step));
        for (int l = 0; l < CNN_LAYERS; ++l) {
            hipLaunchKernelGGL(k_adam, dim3(grid_for(weight_count(l), 256)), dim3(256), 0, c->stream, weight_count(l),
                               n->Wm[l], (const float*)n->gW[l], n->mW[l], n->vW[l], (float)lr, b1, b2, c1, c2, 1e-8f);
            hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, c->stream, (int64_t)kCo[l], n->bm[l],
                               (const float*)n->gb[l], n->mb[l], n->vb[l], (float)lr, b1, b2, c1, c2, 1e-8f);
        }
        CHK(check_launch(c, "k_adam"));
        n->weights_ready = false;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *loss = *n->h_loss / (3.0 * (double)P) ;
    return FEM_OK;
}

int fem_cnn_gemm_bf16(fem_ctx* c, int64_t M, int64_t N, int64_t K, const uint16_t* A, const uint16_t* Bt,
                      float* C) {
    if (!c || !A || !Bt || !C || M < 1 || N < 1 || K < 16) return FEM_EARG;
    HIPCHK(c, hipSetDevice(c->device));
    Epilogue ep;
    ep.cf = C;
    ep.ldc = N;
    return gemm_nt(c, M, N, K, A, K, Bt, K, ep);
}

}  // extern "C"
