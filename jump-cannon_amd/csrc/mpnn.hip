// Learned warm start for FA2: a 3-layer message-passing network (hidden 128)
// trained and run with this project's own HIP kernels on gfx950.
//
// No reference counterpart: the reference seeds layouts at random or from a
// coarsened solve (engines/multilevel.rs:358-470).  Here a network maps local
// structure to an initial layout that the plain FA2 engine then refines; the
// plain method stays the fallback and its stress judges the prediction.
//
//   input   x_i = [deg_i/mean_deg, log(1+deg_i), rw_1..rw_8, u_i(3), 0,0,0]  (16)
//           rw_k = Monte-Carlo probability that a k-step random walk from i is
//           back at i; u_i = noise in [-1,1]^3 that breaks symmetry
//   layer l h' = ReLU([h | mean_{j in N(i)} h_j] W_l + b_l), 3 layers, width 128
//   head    y_i = h_i W_h + b_h  (3 coordinates)
//
// Dense products run on MFMA (v_mfma_f32_16x16x16bf16_1k): activations and
// weights in BF16, accumulation in FP32.  The loss is invariant to rotation,
// translation and scale: over pairs (graph edges + pairs to a few anchor
// nodes) with target distances D from a layout the plain engines produced,
//   alpha = sum w D d / sum w d^2,  L = sum w (alpha d - D)^2 / sum w D^2,
//   w = 1/D^2, d = |y_i - y_j|  (alpha is the optimal scale, so dL/dalpha = 0).
// Every reduction is in a fixed order, so training runs repeat bit for bit.
// The CSR must be symmetric (as csr_from_pairs builds): the transpose of the
// mean aggregation is then a pull over the same rows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

struct jc_mpnn;

namespace jc {
namespace {

constexpr int kHid = 128;
constexpr int kIn = 16;
constexpr int kRw = 8;
constexpr int kLayers = 3;
constexpr int kOut = 3;
constexpr uint32_t kXtgRows = 64;  // nodes per weight-gradient block

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return 0x7FC0;  // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __host__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------- features
// W walks of kRw steps from every node; rw[i*8+k] = fraction back at i after k+1 steps.
__global__ __launch_bounds__(256) void k_rwpe(uint32_t n, const uint32_t *__restrict__ off,
                                              const uint32_t *__restrict__ nbr, uint64_t seed,
                                              uint32_t walks, float *__restrict__ rw) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t cnt[kRw] = {};
    for (uint32_t w = 0; w < walks; w++) {
        uint32_t v = i;
        uint64_t st = mix64(seed ^ ((uint64_t)i << 20) ^ w);
        for (int k = 0; k < kRw; k++) {
            const uint32_t s = off[v], d = off[v + 1] - s;
            st = mix64(st + 0x9E3779B97F4A7C15ull);
            if (d) v = nbr[s + (uint32_t)(st % d)];
            cnt[k] += (v == i);
        }
    }
    for (int k = 0; k < kRw; k++) rw[(size_t)i * kRw + k] = (float)cnt[k] / (float)walks;
}

__global__ __launch_bounds__(256) void k_features(uint32_t n, const uint32_t *__restrict__ off,
                                                  const float *__restrict__ rw, float mean_deg,
                                                  const float *__restrict__ noise,
                                                  uint16_t *__restrict__ x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = (float)(off[i + 1] - off[i]);
    float f[kIn];
    f[0] = d / mean_deg;
    f[1] = logf(1.0f + d);
    for (int k = 0; k < kRw; k++) f[2 + k] = rw[(size_t)i * kRw + k];
    for (int k = 0; k < 3; k++) f[10 + k] = noise[3 * (size_t)i + k];
    f[13] = f[14] = f[15] = 0.0f;
    for (int k = 0; k < kIn; k++) x[(size_t)i * kIn + k] = f2bf(f[k]);
}

// ---------------------------------------------------------------- aggregation
// A_i = mean_{j in N(i)} H_j over F bf16 features (F/2 threads per node, each
// owning two features); f32 sums in CSR order.
template <int F>
__global__ __launch_bounds__(256) void k_agg_mean(uint32_t n, const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ nbr,
                                                  const uint16_t *__restrict__ H,
                                                  uint16_t *__restrict__ A) {
    constexpr int TPN = F / 2;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = g / TPN, c = (g % TPN) * 2;
    if (i >= n) return;
    const uint32_t s = off[i], e = off[i + 1];
    float a0 = 0.0f, a1 = 0.0f;
    for (uint32_t k = s; k < e; k++) {
        const uint32_t pr = *reinterpret_cast<const uint32_t *>(H + (size_t)nbr[k] * F + c);
        a0 += bf2f((uint16_t)(pr & 0xFFFFu));
        a1 += bf2f((uint16_t)(pr >> 16));
    }
    const float inv = e > s ? 1.0f / (float)(e - s) : 0.0f;
    const uint32_t o = (uint32_t)f2bf(a0 * inv) | ((uint32_t)f2bf(a1 * inv) << 16);
    *reinterpret_cast<uint32_t *>(A + (size_t)i * F + c) = o;
}

// Transpose of the mean (symmetric CSR): dH_j += sum_{i in N(j)} dA_i / deg_i.
template <int F>
__global__ __launch_bounds__(256) void k_agg_mean_t(uint32_t n, const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ nbr,
                                                    const float *__restrict__ dX, int ldx,
                                                    float *__restrict__ dH) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = g / F, c = g % F;
    if (j >= n) return;
    float acc = 0.0f;
    for (uint32_t k = off[j]; k < off[j + 1]; k++) {
        const uint32_t i = nbr[k];
        const float inv = 1.0f / (float)(off[i + 1] - off[i]);
        acc += dX[(size_t)i * ldx + F + c] * inv;
    }
    dH[(size_t)j * F + c] = dX[(size_t)j * ldx + c] + acc;
}

// ---------------------------------------------------------------- MFMA GEMM
// C[M x N] = [A1 | A2][M x K] * Bt^T with Bt = [N x K] (both K-contiguous bf16),
// K = K1 + K2, multiples of 16; N = 16*NT.  One wave owns 16 rows and NT
// 16x16 accumulator tiles.  Fragment layout of v_mfma_f32_16x16x16bf16_1k:
// lane l holds A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16], and D rows
// 4(l/16)..+3 of column l%16.
// MODE 0: out_bf16 = ReLU(C + bias) (relu flag) in bf16; MODE 1: out_f32 = C.
template <int NT, int MODE>
__global__ __launch_bounds__(256) void k_gemm(uint32_t M, const uint16_t *__restrict__ A1, int K1,
                                              const uint16_t *__restrict__ A2, int K2,
                                              const uint16_t *__restrict__ Bt,
                                              const float *__restrict__ bias, int relu,
                                              uint16_t *__restrict__ out_bf, float *__restrict__ out_f) {
    constexpr int N = 16 * NT;
    const int K = K1 + K2;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t row0 = wave * 16;
    if (row0 >= M) return;
    const uint32_t r = lane & 15, kq = (lane >> 4) * 4;
    const uint32_t row = row0 + r;
    v4f acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a = v4s{0, 0, 0, 0};
        if (row < M) {
            const uint16_t *src = k0 < K1 ? A1 + (size_t)row * K1 + k0 + kq
                                          : A2 + (size_t)row * K2 + (k0 - K1) + kq;
            a = *reinterpret_cast<const v4s *>(src);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const v4s b = *reinterpret_cast<const v4s *>(Bt + (size_t)(t * 16 + r) * K + k0 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const uint32_t col = t * 16 + r;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t orow = row0 + (lane >> 4) * 4 + q;
            if (orow >= M) continue;
            if (MODE == 0) {
                float v = acc[t][q] + bias[col];
                if (relu) v = fmaxf(v, 0.0f);
                out_bf[(size_t)orow * N + col] = f2bf(v);
            } else {
                out_f[(size_t)orow * N + col] = acc[t][q];
            }
        }
    }
}

// ---------------------------------------------------------------- head
__global__ __launch_bounds__(256) void k_head(uint32_t n, const uint16_t *__restrict__ H,
                                              const float *__restrict__ Wh,
                                              const float *__restrict__ bh, float *__restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a0 = bh[0], a1 = bh[1], a2 = bh[2];
    const uint16_t *h = H + (size_t)i * kHid;
    for (int k = 0; k < kHid; k++) {
        const float v = bf2f(h[k]);
        a0 = fmaf(v, Wh[k * 3 + 0], a0);
        a1 = fmaf(v, Wh[k * 3 + 1], a1);
        a2 = fmaf(v, Wh[k * 3 + 2], a2);
    }
    y[3 * (size_t)i] = a0;
    y[3 * (size_t)i + 1] = a1;
    y[3 * (size_t)i + 2] = a2;
}

// dH = dY W_h^T (n x 128) then masked by the ReLU of the last layer (H > 0):
// dZ (bf16) for the MFMA backward and dZf (f32) for the weight gradient.
__global__ __launch_bounds__(256) void k_head_back(uint32_t n, const float *__restrict__ dY,
                                                   const float *__restrict__ Wh,
                                                   const uint16_t *__restrict__ H,
                                                   float *__restrict__ dZf,
                                                   uint16_t *__restrict__ dZ) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = g / kHid, k = g % kHid;
    if (i >= n) return;
    float v = dY[3 * (size_t)i] * Wh[k * 3] + dY[3 * (size_t)i + 1] * Wh[k * 3 + 1] +
              dY[3 * (size_t)i + 2] * Wh[k * 3 + 2];
    if (!(bf2f(H[(size_t)i * kHid + k]) > 0.0f)) v = 0.0f;
    dZf[(size_t)i * kHid + k] = v;
    dZ[(size_t)i * kHid + k] = f2bf(v);
}

// dZ_prev = dH_prev masked by ReLU(H_prev > 0) -> f32 and bf16 copies.
__global__ __launch_bounds__(256) void k_relu_back(uint32_t total, const float *__restrict__ dH,
                                                   const uint16_t *__restrict__ H,
                                                   float *__restrict__ dZf,
                                                   uint16_t *__restrict__ dZ) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const float v = bf2f(H[g]) > 0.0f ? dH[g] : 0.0f;
    dZf[g] = v;
    dZ[g] = f2bf(v);
}

// Weight gradient partials: P[blk][k][c] = sum over the block's node range of
// X[node][k] * G[node][c], X = [X1 | X2] (bf16), G f32 (n x NO); thread
// (k, c-group) holds CPT columns.  The bias gradient is row k = K (X = 1).
template <int NO, int CPT>
__global__ __launch_bounds__(1024) void k_xtg(uint32_t n, uint32_t per_block,
                                              const uint16_t *__restrict__ X1, int K1,
                                              const uint16_t *__restrict__ X2, int K2,
                                              const float *__restrict__ G, float *__restrict__ P) {
    constexpr int GROUPS = NO / CPT;
    const int K = K1 + K2;
    const int rows_per_pass = (int)(blockDim.x / GROUPS);
    const int c0 = (int)(threadIdx.x % GROUPS) * CPT;
    const uint32_t b0 = blockIdx.x * per_block, b1 = min(b0 + per_block, n);
    for (int k = (int)(threadIdx.x / GROUPS); k <= K; k += rows_per_pass) {
        float acc[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++) acc[c] = 0.0f;
        for (uint32_t i = b0; i < b1; i++) {
            float x = 1.0f;
            if (k < K1) x = bf2f(X1[(size_t)i * K1 + k]);
            else if (k < K) x = bf2f(X2[(size_t)i * K2 + (k - K1)]);
            const float *g = G + (size_t)i * NO + c0;
#pragma unroll
            for (int c = 0; c < CPT; c++) acc[c] = fmaf(x, g[c], acc[c]);
        }
        float *p = P + ((size_t)blockIdx.x * (K + 1) + k) * NO + c0;
#pragma unroll
        for (int c = 0; c < CPT; c++) p[c] = acc[c];
    }
}

// Sum of the block partials in block order -> grad[(K+1) x NO] (last row = bias).
__global__ __launch_bounds__(256) void k_reduce_parts(uint32_t nparts, uint32_t len,
                                                      const float *__restrict__ P,
                                                      float *__restrict__ grad) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.0f;
    for (uint32_t b = 0; b < nparts; b++) s += P[(size_t)b * len + j];
    grad[j] = s;
}

// ---------------------------------------------------------------- loss
struct LossAcc {
    double num, den, wd2, loss;
};

__device__ __forceinline__ void pair_terms(const float *y, const float *tgt, uint32_t i, uint32_t j,
                                           double &w, double &D, double &d, float dv[3]) {
    const double tx = (double)tgt[3 * (size_t)i] - tgt[3 * (size_t)j];
    const double ty = (double)tgt[3 * (size_t)i + 1] - tgt[3 * (size_t)j + 1];
    const double tz = (double)tgt[3 * (size_t)i + 2] - tgt[3 * (size_t)j + 2];
    D = sqrt(tx * tx + ty * ty + tz * tz);
    dv[0] = y[3 * (size_t)i] - y[3 * (size_t)j];
    dv[1] = y[3 * (size_t)i + 1] - y[3 * (size_t)j + 1];
    dv[2] = y[3 * (size_t)i + 2] - y[3 * (size_t)j + 2];
    d = sqrt((double)dv[0] * dv[0] + (double)dv[1] * dv[1] + (double)dv[2] * dv[2]);
    w = D > 0.0 ? 1.0 / (D * D) : 0.0;
}

// Pass 1 (alpha == 0) accumulates num/den/wd2; pass 2 accumulates the loss.
// Pairs: edges with i < j and (i, anchor) for every anchor != i.
__global__ __launch_bounds__(256) void k_loss(uint32_t n, const uint32_t *__restrict__ off,
                                              const uint32_t *__restrict__ nbr,
                                              const uint32_t *__restrict__ anchors, uint32_t na,
                                              const float *__restrict__ y,
                                              const float *__restrict__ tgt, double alpha,
                                              LossAcc *__restrict__ part) {
    __shared__ double red[4][256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double num = 0, den = 0, wd2 = 0, loss = 0;
    if (i < n) {
        auto add = [&](uint32_t j) {
            double w, D, d;
            float dv[3];
            pair_terms(y, tgt, i, j, w, D, d, dv);
            if (w == 0.0) return;
            num += w * D * d;
            den += w * d * d;
            wd2 += w * D * D;
            const double r = alpha * d - D;
            loss += w * r * r;
        };
        for (uint32_t k = off[i]; k < off[i + 1]; k++)
            if (nbr[k] > i) add(nbr[k]);
        for (uint32_t a = 0; a < na; a++)
            if (anchors[a] != i) add(anchors[a]);
    }
    red[0][threadIdx.x] = num;
    red[1][threadIdx.x] = den;
    red[2][threadIdx.x] = wd2;
    red[3][threadIdx.x] = loss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = LossAcc{red[0][0], red[1][0], red[2][0], red[3][0]};
}

__global__ void k_loss_final(uint32_t nb, const LossAcc *__restrict__ part, LossAcc *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    LossAcc s{0, 0, 0, 0};
    for (uint32_t b = 0; b < nb; b++) {
        s.num += part[b].num;
        s.den += part[b].den;
        s.wd2 += part[b].wd2;
        s.loss += part[b].loss;
    }
    *out = s;
}

// dL/dy_i: edge pairs pulled from i's own row (both directions) and i's own
// anchor pairs; anchors add the opposite contributions of every node (k_grad_anchor).
// coefficient of (y_i - y_j): 2 w (alpha d - D) alpha / d / wd2.
__global__ __launch_bounds__(256) void k_grad(uint32_t n, const uint32_t *__restrict__ off,
                                              const uint32_t *__restrict__ nbr,
                                              const uint32_t *__restrict__ anchors, uint32_t na,
                                              const float *__restrict__ y,
                                              const float *__restrict__ tgt,
                                              const LossAcc *__restrict__ tot,
                                              float *__restrict__ dy) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = tot->den > 0 ? tot->num / tot->den : 0.0;
    const double inv = tot->wd2 > 0 ? 1.0 / tot->wd2 : 0.0;
    double g0 = 0, g1 = 0, g2 = 0;
    auto add = [&](uint32_t j) {
        double w, D, d;
        float dv[3];
        pair_terms(y, tgt, i, j, w, D, d, dv);
        if (w == 0.0 || d <= 0.0) return;
        const double c = 2.0 * w * (alpha * d - D) * alpha / d * inv;
        g0 += c * dv[0];
        g1 += c * dv[1];
        g2 += c * dv[2];
    };
    for (uint32_t k = off[i]; k < off[i + 1]; k++) add(nbr[k]);
    for (uint32_t a = 0; a < na; a++)
        if (anchors[a] != i) add(anchors[a]);
    dy[3 * (size_t)i] = (float)g0;
    dy[3 * (size_t)i + 1] = (float)g1;
    dy[3 * (size_t)i + 2] = (float)g2;
}

// One block per anchor a: dy_a += sum_i c_ia (y_a - y_i) over all i != a.
__global__ __launch_bounds__(256) void k_grad_anchor(uint32_t n, const uint32_t *__restrict__ anchors,
                                                     const float *__restrict__ y,
                                                     const float *__restrict__ tgt,
                                                     const LossAcc *__restrict__ tot,
                                                     float *__restrict__ dy) {
    __shared__ double red[3][256];
    const uint32_t a = anchors[blockIdx.x];
    const double alpha = tot->den > 0 ? tot->num / tot->den : 0.0;
    const double inv = tot->wd2 > 0 ? 1.0 / tot->wd2 : 0.0;
    double g[3] = {0, 0, 0};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (i == a) continue;
        double w, D, d;
        float dv[3];
        pair_terms(y, tgt, a, i, w, D, d, dv);
        if (w == 0.0 || d <= 0.0) continue;
        const double c = 2.0 * w * (alpha * d - D) * alpha / d * inv;
        for (int q = 0; q < 3; q++) g[q] += c * dv[q];
    }
    for (int q = 0; q < 3; q++) red[q][threadIdx.x] = g[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) dy[3 * (size_t)a + q] += (float)red[q][0];
}

// Adam (fp32 master weights), then the bf16 copies both MFMA passes read:
// Wt[c][k] (forward Bt) and Wb[k][c] (backward Bt) for the dense layers.
__global__ __launch_bounds__(256) void k_adam(uint32_t count, float *__restrict__ p,
                                              const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const float gj = g[j];
    const float mj = b1 * m[j] + (1.0f - b1) * gj;
    const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    p[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

__global__ __launch_bounds__(256) void k_pack_w(uint32_t K, const float *__restrict__ W,
                                                uint16_t *__restrict__ Wt, uint16_t *__restrict__ Wb) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= K * kHid) return;
    const uint32_t k = j / kHid, c = j % kHid;
    const uint16_t b = f2bf(W[j]);
    Wt[(size_t)c * K + k] = b;
    Wb[j] = b;
}

inline uint32_t cd(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Centring and scaling of the predicted layout (n x 3), in f64 with a fixed
// summation order so the result is deterministic: per-block partial sums over
// a grid-stride of rows (kCenterBlocks blocks), then one thread adding the
// partials in block order.  Pass 0 sums the coordinates; pass 1 sums the
// squared distances from the mean.
constexpr uint32_t kCenterBlocks = 256;
__global__ __launch_bounds__(256) void k_center_part(uint32_t n, const float *__restrict__ y,
                                                     const double *__restrict__ mean, int pass,
                                                     double *__restrict__ part) {
    __shared__ double red[3][256];
    double a = 0.0, b = 0.0, c = 0.0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double x0 = y[3 * (size_t)i], x1 = y[3 * (size_t)i + 1], x2 = y[3 * (size_t)i + 2];
        if (pass == 0) {
            a += x0;
            b += x1;
            c += x2;
        } else {
            const double d0 = x0 - mean[0], d1 = x1 - mean[1], d2 = x2 - mean[2];
            a += d0 * d0 + d1 * d1 + d2 * d2;
        }
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// stat[0..2] = mean after pass 0; stat[3] = target / max(rms, 1e-30) after pass 1.
__global__ void k_center_final(uint32_t n, const double *__restrict__ part, int pass,
                               double target, double *__restrict__ stat) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t[3] = {0.0, 0.0, 0.0};
    for (uint32_t b = 0; b < kCenterBlocks; b++)
        for (int k = 0; k < 3; k++) t[k] += part[b * 3 + k];
    if (pass == 0) {
        for (int k = 0; k < 3; k++) stat[k] = t[k] / (double)n;
    } else {
        const double r = sqrt(t[0] / (double)n);
        stat[3] = target / fmax(r, 1e-30);
    }
}

__global__ __launch_bounds__(256) void k_center_apply(uint32_t n, const float *__restrict__ y,
                                                      const double *__restrict__ stat,
                                                      float *__restrict__ out) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= 3 * n) return;
    out[q] = (float)(((double)y[q] - stat[q % 3]) * stat[3]);
}

}  // namespace

// ====================================================================== network
class Mpnn {
  public:
    static constexpr uint32_t kK[kLayers] = {2 * kIn, 2 * kHid, 2 * kHid};
    // flat parameter layout: [W_l (K_l x 128), b_l (128)] x 3, W_h (128 x 3), b_h (3)
    static uint64_t layer_offset(int l) {
        uint64_t o = 0;
        for (int q = 0; q < l; q++) o += (uint64_t)(kK[q] + 1) * kHid;
        return o;
    }
    static uint64_t head_offset() { return layer_offset(kLayers); }
    static uint64_t param_count() { return head_offset() + (uint64_t)(kHid + 1) * kOut; }

    ~Mpnn() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    Status create(uint64_t seed) {
        JC_TRY(require_device(nullptr));
        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        const uint64_t P = param_count();
        std::vector<float> h(P, 0.0f);
        uint64_t s = seed;
        auto uni = [&]() {
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int l = 0; l < kLayers; l++) {  // Glorot-uniform weights, zero bias
            const double lim = std::sqrt(6.0 / (kK[l] + kHid));
            for (uint64_t j = 0; j < (uint64_t)kK[l] * kHid; j++)
                h[layer_offset(l) + j] = (float)((2.0 * uni() - 1.0) * lim);
        }
        const double limh = std::sqrt(6.0 / (kHid + kOut));
        for (uint64_t j = 0; j < (uint64_t)kHid * kOut; j++)
            h[head_offset() + j] = (float)((2.0 * uni() - 1.0) * limh);
        JC_TRY(params_.alloc(P));
        JC_TRY(grads_.alloc(P));
        JC_TRY(m_.alloc(P));
        JC_TRY(v_.alloc(P));
        JC_HIP(hipMemcpyAsync(params_.p, h.data(), P * 4, hipMemcpyHostToDevice, stream_));
        JC_HIP(hipMemsetAsync(grads_.p, 0, P * 4, stream_));
        JC_HIP(hipMemsetAsync(m_.p, 0, P * 4, stream_));
        JC_HIP(hipMemsetAsync(v_.p, 0, P * 4, stream_));
        for (int l = 0; l < kLayers; l++) {
            JC_TRY(wt_[l].alloc((size_t)kK[l] * kHid));
            JC_TRY(wb_[l].alloc((size_t)kK[l] * kHid));
        }
        JC_TRY(pack());
        JC_HIP(hipStreamSynchronize(stream_));
        adam_t_ = 0;
        return Status::ok();
    }

    Status get_params(float *out) {
        JC_TRY(copy_sync(out, params_.p, param_count() * 4, hipMemcpyDeviceToHost, stream_));
        return Status::ok();
    }
    Status set_params(const float *in) {
        JC_TRY(copy_sync(params_.p, in, param_count() * 4, hipMemcpyHostToDevice, stream_));
        JC_TRY(pack());
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }
    Status get_grads(float *out) {
        JC_TRY(copy_sync(out, grads_.p, param_count() * 4, hipMemcpyDeviceToHost, stream_));
        return Status::ok();
    }

    // Graph + features.  noise (3n, nullable -> drawn from feat_seed in [-1,1]).
    Status set_graph(uint32_t n, const uint32_t *off, const uint32_t *nbr, uint64_t nnz,
                     uint64_t feat_seed, uint32_t walks, const float *noise) {
        CsrView g{n, off, nbr, nnz};
        JC_TRY(validate_csr(g));
        n_ = n;
        has_target_ = false;
        const size_t nn = n ? n : 1;
        JC_TRY(off_.fit((size_t)n + 1));
        JC_TRY(nbr_.fit(nnz ? nnz : 1));
        JC_TRY(x_.fit(nn * kIn));
        JC_TRY(a0_.fit(nn * kIn));
        for (int l = 0; l < kLayers; l++) JC_TRY(h_[l].fit(nn * kHid));
        for (int l = 1; l < kLayers; l++) JC_TRY(a_[l].fit(nn * kHid));
        JC_TRY(y_.fit(nn * 3));
        JC_TRY(dy_.fit(nn * 3));
        JC_TRY(dzf_.fit(nn * kHid));
        JC_TRY(dz_.fit(nn * kHid));
        JC_TRY(dx_.fit(nn * 2 * kHid));
        JC_TRY(dh_.fit(nn * kHid));
        JC_TRY(rw_.fit(nn * kRw));
        JC_TRY(parts_.fit((size_t)cd(nn, kXtgRows) * (2 * kHid + 1) * kHid));
        DevBuf<float> noise_d;
        JC_TRY(noise_d.alloc(nn * 3));
        // the caller's noise goes to the device as it is; without one, the
        // seeded stream below fills a host copy
        std::vector<float> nz(noise ? 0 : nn * 3, 0.0f);
        if (!noise) {
            uint64_t s = feat_seed ^ 0x6E6F697365ull;
            for (size_t q = 0; q < (size_t)n * 3; q++) {
                s += 0x9E3779B97F4A7C15ull;
                uint64_t z = s;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                nz[q] = (float)((double)(z >> 40) / 16777216.0) * 2.0f - 1.0f;
            }
        }
        if (n) {
            JC_HIP(hipMemcpyAsync(off_.p, off, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
            if (nnz) JC_HIP(hipMemcpyAsync(nbr_.p, nbr, nnz * 4, hipMemcpyHostToDevice, stream_));
            JC_HIP(hipMemcpyAsync(noise_d.p, noise ? noise : nz.data(), (size_t)n * 3 * 4,
                                  hipMemcpyHostToDevice, stream_));
            const float mean_deg = nnz ? (float)((double)nnz / n) : 1.0f;
            k_rwpe<<<cd(n, 256), 256, 0, stream_>>>(n, off_.p, nbr_.p, feat_seed, walks, rw_.p);
            k_features<<<cd(n, 256), 256, 0, stream_>>>(n, off_.p, rw_.p, mean_deg, noise_d.p, x_.p);
            JC_HIP(hipGetLastError());
        }
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }

    Status features(float *out) {
        std::vector<uint16_t> h((size_t)n_ * kIn);
        JC_TRY(copy_sync(h.data(), x_.p, h.size() * 2, hipMemcpyDeviceToHost, stream_));
        for (size_t q = 0; q < h.size(); q++) out[q] = bf2f(h[q]);
        return Status::ok();
    }

    Status forward(float *out_xyz, double *ms) {
        if (n_ == 0) return Status::ok();
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        JC_HIP(hipEventRecord(a, stream_));
        JC_TRY(enqueue_forward());
        JC_HIP(hipEventRecord(b, stream_));
        JC_HIP(hipEventSynchronize(b));
        float t = 0;
        JC_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms) *ms = t;
        if (out_xyz) JC_TRY(copy_sync(out_xyz, y_.p, (size_t)n_ * 3 * 4, hipMemcpyDeviceToHost, stream_));
        return Status::ok();
    }

    // The forward pass, then the layout centred at the origin and scaled to
    // RMS radius `target` on the device (f64 sums, fixed order); `ms` is the
    // device time of both.
    Status forward_scaled(double target, float *out_xyz, double *ms) {
        if (n_ == 0) return Status::ok();
        JC_TRY(cpart_.fit(kCenterBlocks * 3));
        JC_TRY(cstat_.fit(4));
        JC_TRY(yout_.fit((size_t)n_ * 3));
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        JC_HIP(hipEventRecord(a, stream_));
        JC_TRY(enqueue_forward());
        for (int pass = 0; pass < 2; pass++) {
            k_center_part<<<kCenterBlocks, 256, 0, stream_>>>(n_, y_.p, cstat_.p, pass, cpart_.p);
            k_center_final<<<1, 64, 0, stream_>>>(n_, cpart_.p, pass, target, cstat_.p);
        }
        k_center_apply<<<cd((uint64_t)n_ * 3, 256), 256, 0, stream_>>>(n_, y_.p, cstat_.p, yout_.p);
        JC_HIP(hipGetLastError());
        JC_HIP(hipEventRecord(b, stream_));
        JC_HIP(hipEventSynchronize(b));
        float t = 0;
        JC_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms) *ms = t;
        if (out_xyz) JC_TRY(copy_sync(out_xyz, yout_.p, (size_t)n_ * 3 * 4, hipMemcpyDeviceToHost, stream_));
        return Status::ok();
    }

    Status activations(int layer, float *out) {  // debug: h_layer (n x 128) as f32
        if (layer < 0 || layer >= kLayers) return Status::err(JC_E_INVALID, "bad layer");
        std::vector<uint16_t> h((size_t)n_ * kHid);
        JC_TRY(copy_sync(h.data(), h_[layer].p, h.size() * 2, hipMemcpyDeviceToHost, stream_));
        for (size_t q = 0; q < h.size(); q++) out[q] = bf2f(h[q]);
        return Status::ok();
    }

    Status set_target(const float *tgt, uint32_t n_anchor, uint64_t seed) {
        if (n_ == 0) return Status::ok();
        JC_TRY(tgt_.alloc((size_t)n_ * 3));
        JC_TRY(copy_sync(tgt_.p, tgt, (size_t)n_ * 3 * 4, hipMemcpyHostToDevice, stream_));
        std::vector<uint32_t> anc;
        uint64_t s = seed;
        const uint32_t na = std::min(n_anchor, n_);
        while (anc.size() < na) {
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const uint32_t c = (uint32_t)(z % n_);
            if (std::find(anc.begin(), anc.end(), c) == anc.end()) anc.push_back(c);
        }
        na_ = na;
        JC_TRY(anchors_.alloc(na ? na : 1));
        if (na) JC_TRY(copy_sync(anchors_.p, anc.data(), na * 4, hipMemcpyHostToDevice, stream_));
        JC_TRY(lparts_.alloc(cd(n_, 256)));
        JC_TRY(ltot_.alloc(1));
        has_target_ = true;
        return Status::ok();
    }

    // Forward + loss + backward: grads_ holds dL/dparams afterwards.
    Status loss_grad(double *loss) {
        if (!has_target_) return Status::err(JC_E_STATE, "set_target first");
        JC_TRY(enqueue_forward());
        JC_TRY(enqueue_loss());
        JC_TRY(enqueue_backward());
        if (loss) {
            LossAcc h;
            JC_HIP(hipMemcpyAsync(&h, ltot_.p, sizeof h, hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
            const double alpha = h.den > 0 ? h.num / h.den : 0.0;
            // sum w (alpha d - D)^2 = alpha^2 den - 2 alpha num + wd2
            *loss = h.wd2 > 0 ? (alpha * alpha * h.den - 2.0 * alpha * h.num + h.wd2) / h.wd2 : 0.0;
        }
        return Status::ok();
    }

    Status adam(float lr) {
        adam_t_++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)adam_t_), c2 = 1.0f - std::pow(b2, (float)adam_t_);
        const uint32_t P = (uint32_t)param_count();
        k_adam<<<cd(P, 256), 256, 0, stream_>>>(P, params_.p, grads_.p, m_.p, v_.p, lr, b1, b2, 1e-8f, c1, c2);
        JC_HIP(hipGetLastError());
        JC_TRY(pack());
        return Status::ok();
    }

    Status train(uint32_t steps, float lr, double *last_loss) {
        double L = 0;
        for (uint32_t s = 0; s < steps; s++) {
            JC_TRY(loss_grad(s + 1 == steps ? &L : nullptr));
            JC_TRY(adam(lr));
        }
        JC_HIP(hipStreamSynchronize(stream_));
        if (last_loss) *last_loss = L;
        return Status::ok();
    }

  private:
    Status pack() {
        for (int l = 0; l < kLayers; l++)
            k_pack_w<<<cd((uint64_t)kK[l] * kHid, 256), 256, 0, stream_>>>(
                kK[l], params_.p + layer_offset(l), wt_[l].p, wb_[l].p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_forward() {
        const uint32_t n = n_;
        const uint32_t gw = cd((uint64_t)cd(n, 16) * 64, 256);  // one wave per 16 rows
        // layer 0: [x | mean x] (32) -> 128
        k_agg_mean<kIn><<<cd((uint64_t)n * (kIn / 2), 256), 256, 0, stream_>>>(n, off_.p, nbr_.p, x_.p, a0_.p);
        k_gemm<8, 0><<<gw, 256, 0, stream_>>>(n, x_.p, kIn, a0_.p, kIn, wt_[0].p,
                                               params_.p + layer_offset(0) + (uint64_t)kK[0] * kHid, 1,
                                               h_[0].p, nullptr);
        for (int l = 1; l < kLayers; l++) {
            k_agg_mean<kHid><<<cd((uint64_t)n * (kHid / 2), 256), 256, 0, stream_>>>(n, off_.p, nbr_.p,
                                                                                       h_[l - 1].p, a_[l].p);
            k_gemm<8, 0><<<gw, 256, 0, stream_>>>(n, h_[l - 1].p, kHid, a_[l].p, kHid, wt_[l].p,
                                                   params_.p + layer_offset(l) + (uint64_t)kK[l] * kHid,
                                                   1, h_[l].p, nullptr);
        }
        k_head<<<cd(n, 256), 256, 0, stream_>>>(n, h_[kLayers - 1].p, params_.p + head_offset(),
                                                params_.p + head_offset() + kHid * kOut, y_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_loss() {
        const uint32_t nb = cd(n_, 256);
        k_loss<<<nb, 256, 0, stream_>>>(n_, off_.p, nbr_.p, anchors_.p, na_, y_.p, tgt_.p, 0.0, lparts_.p);
        k_loss_final<<<1, 64, 0, stream_>>>(nb, lparts_.p, ltot_.p);
        k_grad<<<nb, 256, 0, stream_>>>(n_, off_.p, nbr_.p, anchors_.p, na_, y_.p, tgt_.p, ltot_.p, dy_.p);
        if (na_) k_grad_anchor<<<na_, 256, 0, stream_>>>(n_, anchors_.p, y_.p, tgt_.p, ltot_.p, dy_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_backward() {
        const uint32_t n = n_;
        const uint32_t per = kXtgRows;
        const uint32_t nparts = cd(n, per);
        // head: dWh = H2^T dY, dbh = sum dY ; dZ2 = (dY Wh^T) * relu'(H2)
        k_xtg<kOut, kOut><<<nparts, 1024, 0, stream_>>>(n, per, h_[kLayers - 1].p, kHid, nullptr, 0,
                                                         dy_.p, parts_.p);
        k_reduce_parts<<<cd((kHid + 1) * kOut, 256), 256, 0, stream_>>>(
            nparts, (kHid + 1) * kOut, parts_.p, grads_.p + head_offset());
        k_head_back<<<cd((uint64_t)n * kHid, 256), 256, 0, stream_>>>(n, dy_.p, params_.p + head_offset(),
                                                                      h_[kLayers - 1].p, dzf_.p, dz_.p);
        for (int l = kLayers - 1; l >= 0; l--) {
            const uint16_t *X1 = l == 0 ? x_.p : h_[l - 1].p;
            const uint16_t *X2 = l == 0 ? a0_.p : a_[l].p;
            const int F = l == 0 ? kIn : kHid;
            // dW_l = [X1 | X2]^T dZ (f32), db_l = sum dZ
            k_xtg<kHid, 32><<<nparts, 1024, 0, stream_>>>(n, per, X1, F, X2, F, dzf_.p, parts_.p);
            k_reduce_parts<<<cd((uint64_t)(kK[l] + 1) * kHid, 256), 256, 0, stream_>>>(
                nparts, (kK[l] + 1) * kHid, parts_.p, grads_.p + layer_offset(l));
            if (l == 0) break;
            // dX = dZ W_l^T on MFMA: A = dZ (n x 128), Bt = W_l (K_l x 128)
            k_gemm<16, 1><<<cd((uint64_t)cd(n, 16) * 64, 256), 256, 0, stream_>>>(
                n, dz_.p, kHid, nullptr, 0, wb_[l].p, nullptr, 0, nullptr, dx_.p);
            // dH_{l-1} = dX[:, :128] + mean^T(dX[:, 128:])
            k_agg_mean_t<kHid><<<cd((uint64_t)n * kHid, 256), 256, 0, stream_>>>(n, off_.p, nbr_.p, dx_.p,
                                                                                  2 * kHid, dh_.p);
            k_relu_back<<<cd((uint64_t)n * kHid, 256), 256, 0, stream_>>>(n * kHid, dh_.p, h_[l - 1].p,
                                                                          dzf_.p, dz_.p);
        }
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    hipStream_t stream_ = nullptr;
    uint32_t n_ = 0, na_ = 0;
    uint64_t adam_t_ = 0;
    bool has_target_ = false;
    DevBuf<float> params_, grads_, m_, v_, parts_;
    DevBuf<uint16_t> wt_[kLayers], wb_[kLayers];
    DevBuf<uint32_t> off_, nbr_, anchors_;
    DevBuf<uint16_t> x_, a0_, h_[kLayers], a_[kLayers], dz_;
    DevBuf<float> y_, dy_, dzf_, dx_, dh_, rw_, tgt_, yout_;
    DevBuf<double> cpart_, cstat_;  // centring partial sums and {mean xyz, scale}
    DevBuf<LossAcc> lparts_, ltot_;
};

constexpr uint32_t Mpnn::kK[kLayers];

}  // namespace jc

struct jc_mpnn {
    jc::Mpnn impl;
};

namespace {
int done(jc::Status s) {
    if (!s.failed()) jc::set_error("");
    return s.code;
}
int nullh() {
    jc::set_error("NULL network handle");
    return JC_E_INVALID;
}
}  // namespace

extern "C" {

uint64_t jc_mpnn_param_count(void) { return jc::Mpnn::param_count(); }

int jc_mpnn_create(uint64_t seed, jc_mpnn **out) {
    if (!out) return nullh();
    auto *m = new jc_mpnn();
    jc::Status s = m->impl.create(seed);
    if (s.failed()) {
        delete m;
        return s.code;
    }
    *out = m;
    return done(s);
}
void jc_mpnn_destroy(jc_mpnn *m) { delete m; }
int jc_mpnn_get_params(jc_mpnn *m, float *out) { return m ? done(m->impl.get_params(out)) : nullh(); }
int jc_mpnn_set_params(jc_mpnn *m, const float *in) { return m ? done(m->impl.set_params(in)) : nullh(); }
int jc_mpnn_get_grads(jc_mpnn *m, float *out) { return m ? done(m->impl.get_grads(out)) : nullh(); }
int jc_mpnn_set_graph(jc_mpnn *m, uint32_t n, const uint32_t *off, const uint32_t *nbr, uint64_t nnz,
                      uint64_t feat_seed, uint32_t walks, const float *noise) {
    return m ? done(m->impl.set_graph(n, off, nbr, nnz, feat_seed, walks, noise)) : nullh();
}
int jc_mpnn_features(jc_mpnn *m, float *out) { return m ? done(m->impl.features(out)) : nullh(); }
int jc_mpnn_activations(jc_mpnn *m, int layer, float *out) {
    return m ? done(m->impl.activations(layer, out)) : nullh();
}
int jc_mpnn_forward(jc_mpnn *m, float *out_xyz, double *gpu_ms) {
    return m ? done(m->impl.forward(out_xyz, gpu_ms)) : nullh();
}
int jc_mpnn_forward_scaled(jc_mpnn *m, double target_rms, float *out_xyz, double *gpu_ms) {
    return m ? done(m->impl.forward_scaled(target_rms, out_xyz, gpu_ms)) : nullh();
}
int jc_mpnn_set_target(jc_mpnn *m, const float *xyz, uint32_t n_anchor, uint64_t seed) {
    return m ? done(m->impl.set_target(xyz, n_anchor, seed)) : nullh();
}
int jc_mpnn_loss_grad(jc_mpnn *m, double *loss) { return m ? done(m->impl.loss_grad(loss)) : nullh(); }
int jc_mpnn_adam(jc_mpnn *m, float lr) { return m ? done(m->impl.adam(lr)) : nullh(); }
int jc_mpnn_train(jc_mpnn *m, uint32_t steps, float lr, double *last_loss) {
    return m ? done(m->impl.train(steps, lr, last_loss)) : nullh();
}

}  // extern "C"
