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
// Dense products run on MFMA (v_mfma_f32_16x16x32_bf16): activations and
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
constexpr uint32_t kXtgRows = 64;    // fewest nodes per weight-gradient block
constexpr uint32_t kXtgParts = 512;  // most weight-gradient blocks (partials summed in block order)
// Nodes per weight-gradient block: a fixed function of n, so the partial sums
// and their order are the same on every run.
inline uint32_t xtg_rows(uint32_t n) { return std::max(kXtgRows, (n + kXtgParts - 1) / kXtgParts); }

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
// A_i = mean_{j in N(i)} H_j over F bf16 features.  A node's TPN = F/2 lanes
// each own two features.  Rows of at most kAggLight entries are summed by
// their node's lanes in CSR order, agg_round() neighbours per round: the group's
// first lanes load the round's neighbour ids, shuffle them to the group, and
// every lane then has the round's row loads (agg_round) in flight together.  Longer rows
// are left to k_agg_mean_hub.  f32 sums, one bf16 rounding at the end.
constexpr uint32_t kAggLight = 64;   // longer rows: one block each (k_agg_mean_hub)
// neighbour rows in flight per lane: 16, or the group size when smaller
template <int TPN>
constexpr int agg_round() { return TPN < 16 ? TPN : 16; }
constexpr int kHubThreads = 1024;

template <int TPN>
__device__ __forceinline__ void agg_rows(const uint32_t *__restrict__ nbr, const uint16_t *__restrict__ H,
                                         int F, uint32_t c, uint32_t j, uint32_t s, uint32_t e, float &a0,
                                         float &a1) {
    constexpr int R = agg_round<TPN>();
    for (uint32_t base = s; base < e; base += R) {
        const uint32_t id = (j < (uint32_t)R && base + j < e) ? nbr[base + j] : 0u;
        uint32_t v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t row = (uint32_t)__shfl((int)id, r, TPN);
            v[r] = base + r < e ? *reinterpret_cast<const uint32_t *>(H + (size_t)row * F + c) : 0u;
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (base + r < e) {
                a0 += bf2f((uint16_t)(v[r] & 0xFFFFu));
                a1 += bf2f((uint16_t)(v[r] >> 16));
            }
    }
}

template <int F>
__global__ __launch_bounds__(256) void k_agg_mean(uint32_t n, const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ nbr,
                                                  const uint16_t *__restrict__ H,
                                                  uint16_t *__restrict__ A) {
    constexpr int TPN = F / 2;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = g / TPN, j = g % TPN, c = j * 2;
    uint32_t s = 0, e = 0;
    if (i < n) {
        s = off[i];
        e = off[i + 1];
    }
    if (i >= n || e - s > kAggLight) return;  // whole groups leave together
    float a0 = 0.0f, a1 = 0.0f;
    agg_rows<TPN>(nbr, H, F, c, j, s, e, a0, a1);
    const float inv = e > s ? 1.0f / (float)(e - s) : 0.0f;
    const uint32_t o = (uint32_t)f2bf(a0 * inv) | ((uint32_t)f2bf(a1 * inv) << 16);
    *reinterpret_cast<uint32_t *>(A + (size_t)i * F + c) = o;
}

// One row longer than kAggLight per block: the block's G = kHubThreads/TPN
// lane groups take consecutive equal chunks of the row (multiples of
// agg_round() entries), and group 0 adds the G partial sums in group order, so
// the result is the same on every run.
template <int F>
__global__ __launch_bounds__(kHubThreads) void k_agg_mean_hub(const uint32_t *__restrict__ hubs,
                                                              const uint32_t *__restrict__ off,
                                                              const uint32_t *__restrict__ nbr,
                                                              const uint16_t *__restrict__ H,
                                                              uint16_t *__restrict__ A) {
    constexpr int TPN = F / 2, G = kHubThreads / TPN;
    __shared__ float2 part[G][TPN];
    const uint32_t v = hubs[blockIdx.x];
    const uint32_t s = off[v], e = off[v + 1];
    const uint32_t q = threadIdx.x / TPN, j = threadIdx.x % TPN, c = j * 2;
    constexpr uint32_t R = agg_round<TPN>();
    const uint32_t per = (((e - s + G - 1) / G) + R - 1) / R * R;
    const uint32_t cs = min(s + q * per, e), ce = min(cs + per, e);
    float a0 = 0.0f, a1 = 0.0f;
    agg_rows<TPN>(nbr, H, F, c, j, cs, ce, a0, a1);
    part[q][j] = make_float2(a0, a1);
    __syncthreads();
    if (q == 0) {
        float b0 = 0.0f, b1 = 0.0f;
        for (int k = 0; k < G; k++) {
            b0 += part[k][j].x;
            b1 += part[k][j].y;
        }
        const float inv = 1.0f / (float)(e - s);
        const uint32_t o = (uint32_t)f2bf(b0 * inv) | ((uint32_t)f2bf(b1 * inv) << 16);
        *reinterpret_cast<uint32_t *>(A + (size_t)v * F + c) = o;
    }
}

// Transpose of the mean (symmetric CSR): dH_j = dX_j[:F] + sum_{i in N(j)} dX_i[F:] / deg_i,
// with inv[i] = 1/deg_i; the same lane groups, rounds and hub split as the
// forward aggregation, two f32 features per lane.
template <int TPN>
__device__ __forceinline__ void agg_t_rows(const uint32_t *__restrict__ nbr, const float *__restrict__ inv,
                                           const float *__restrict__ dX, int ldx, int F, uint32_t c,
                                           uint32_t j, uint32_t s, uint32_t e, float &a0, float &a1) {
    constexpr int R = agg_round<TPN>();
    for (uint32_t base = s; base < e; base += R) {
        const uint32_t id = (j < (uint32_t)R && base + j < e) ? nbr[base + j] : 0u;
        float2 v[R];
        float w[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t row = (uint32_t)__shfl((int)id, r, TPN);
            const bool ok = base + r < e;
            v[r] = ok ? *reinterpret_cast<const float2 *>(dX + (size_t)row * ldx + F + c) : make_float2(0.0f, 0.0f);
            w[r] = ok ? inv[row] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (base + r < e) {
                a0 += v[r].x * w[r];
                a1 += v[r].y * w[r];
            }
    }
}

template <int F>
__global__ __launch_bounds__(256) void k_agg_mean_t(uint32_t n, const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ nbr,
                                                    const float *__restrict__ inv,
                                                    const float *__restrict__ dX, int ldx,
                                                    float *__restrict__ dH) {
    constexpr int TPN = F / 2;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = g / TPN, j = g % TPN, c = j * 2;
    uint32_t s = 0, e = 0;
    if (i < n) {
        s = off[i];
        e = off[i + 1];
    }
    if (i >= n || e - s > kAggLight) return;
    float a0 = 0.0f, a1 = 0.0f;
    agg_t_rows<TPN>(nbr, inv, dX, ldx, F, c, j, s, e, a0, a1);
    const float2 self = *reinterpret_cast<const float2 *>(dX + (size_t)i * ldx + c);
    *reinterpret_cast<float2 *>(dH + (size_t)i * F + c) = make_float2(self.x + a0, self.y + a1);
}

template <int F>
__global__ __launch_bounds__(kHubThreads) void k_agg_mean_t_hub(const uint32_t *__restrict__ hubs,
                                                                const uint32_t *__restrict__ off,
                                                                const uint32_t *__restrict__ nbr,
                                                                const float *__restrict__ inv,
                                                                const float *__restrict__ dX, int ldx,
                                                                float *__restrict__ dH) {
    constexpr int TPN = F / 2, G = kHubThreads / TPN;
    __shared__ float2 part[G][TPN];
    const uint32_t v = hubs[blockIdx.x];
    const uint32_t s = off[v], e = off[v + 1];
    const uint32_t q = threadIdx.x / TPN, j = threadIdx.x % TPN, c = j * 2;
    constexpr uint32_t R = agg_round<TPN>();
    const uint32_t per = (((e - s + G - 1) / G) + R - 1) / R * R;
    const uint32_t cs = min(s + q * per, e), ce = min(cs + per, e);
    float a0 = 0.0f, a1 = 0.0f;
    agg_t_rows<TPN>(nbr, inv, dX, ldx, F, c, j, cs, ce, a0, a1);
    part[q][j] = make_float2(a0, a1);
    __syncthreads();
    if (q == 0) {
        float b0 = 0.0f, b1 = 0.0f;
        for (int k = 0; k < G; k++) {
            b0 += part[k][j].x;
            b1 += part[k][j].y;
        }
        const float2 self = *reinterpret_cast<const float2 *>(dX + (size_t)v * ldx + c);
        *reinterpret_cast<float2 *>(dH + (size_t)v * F + c) = make_float2(self.x + b0, self.y + b1);
    }
}

__global__ __launch_bounds__(256) void k_inv_deg(uint32_t n, const uint32_t *__restrict__ off,
                                                 float *__restrict__ inv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = off[i + 1] - off[i];
    inv[i] = d ? 1.0f / (float)d : 0.0f;
}

// ---------------------------------------------------------------- MFMA GEMM
// C[M x N] = [A1 | A2][M x K] * Bt^T with Bt = [N x K] (both K-contiguous bf16),
// K = K1 + K2 <= KMAX, K1 and K2 multiples of 32, N = 16*NT, on
// v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16] and D rows 4(l/16)..+3 of column l%16.
// A persistent block keeps Bt in LDS and walks row tiles of kGemmRows rows
// (16 per wave); each tile's rows are contiguous in A1 and in A2, so they come
// in as coalesced 16-byte loads, and the next tile's loads are issued into
// registers before the current tile's products.
// MODE 0: out_bf16 = ReLU(C + bias) (relu flag), staged through the wave's
// own LDS rows and written as whole rows; MODE 1: out_f32 = C.
constexpr int kGemmThreads = 512;
constexpr uint32_t kGemmRows = 16 * (kGemmThreads / 64);
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int NT, int MODE, int KMAX>
__global__ __launch_bounds__(kGemmThreads) void k_gemm(uint32_t M, const uint16_t *__restrict__ A1, int K1,
                                                       const uint16_t *__restrict__ A2, int K2,
                                                       const uint16_t *__restrict__ Bt,
                                                       const float *__restrict__ bias, int relu,
                                                       uint16_t *__restrict__ out_bf,
                                                       float *__restrict__ out_f) {
    constexpr int N = 16 * NT;
    constexpr int LDK = KMAX + 8;  // padded row: 16 lanes' 16-byte reads hit distinct banks
    constexpr int PER = (int)kGemmRows * (KMAX / 8) / kGemmThreads;  // 16-byte chunks per thread
    static_assert(MODE == 1 || LDK >= N, "bf16 output rows are staged in the wave's A rows");
    __shared__ __align__(16) uint16_t Bs[N * LDK];
    __shared__ __align__(16) uint16_t As[kGemmRows * LDK];
    const int K = K1 + K2, c1 = K1 / 8, cr = K / 8;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int q = (int)tid; q < N * cr; q += kGemmThreads) {
        const int row = q / cr, kc = (q % cr) * 8;
        *reinterpret_cast<uint4 *>(&Bs[row * LDK + kc]) =
            *reinterpret_cast<const uint4 *>(Bt + (size_t)row * K + kc);
    }
    const uint32_t tiles = (M + kGemmRows - 1) / kGemmRows;
    uint4 pre[PER];
    auto fetch = [&](uint32_t t) {
        const uint32_t r0 = t * kGemmRows;
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int q = (int)tid + p * kGemmThreads;
            const int row = q / cr, ch = q % cr;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (t < tiles && q < (int)kGemmRows * cr && r0 + row < M)
                v = ch < c1 ? *reinterpret_cast<const uint4 *>(A1 + (size_t)(r0 + row) * K1 + ch * 8)
                            : *reinterpret_cast<const uint4 *>(A2 + (size_t)(r0 + row) * K2 + (ch - c1) * 8);
            pre[p] = v;
        }
    };
    fetch(blockIdx.x);
    const uint32_t m = lane & 15, kq = (lane >> 4) * 8;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        __syncthreads();  // Bs is written; every wave is done with the previous tile's rows
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int q = (int)tid + p * kGemmThreads;
            if (q < (int)kGemmRows * cr)
                *reinterpret_cast<uint4 *>(&As[(q / cr) * LDK + (q % cr) * 8]) = pre[p];
        }
        __syncthreads();
        fetch(t + gridDim.x);  // in flight while this tile's products run
        v4f acc[NT];
#pragma unroll
        for (int tt = 0; tt < NT; tt++) acc[tt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = 0; k0 < K; k0 += 32) {
            const v8bf a = *reinterpret_cast<const v8bf *>(&As[(w * 16 + m) * LDK + k0 + kq]);
#pragma unroll
            for (int tt = 0; tt < NT; tt++) {
                const v8bf b = *reinterpret_cast<const v8bf *>(&Bs[(tt * 16 + m) * LDK + k0 + kq]);
                acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[tt], 0, 0, 0);
            }
        }
        const uint32_t r0 = t * kGemmRows + w * 16;
        if constexpr (MODE == 0) {
            // the wave's own 16 rows of As take its outputs, then go out as rows
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int tt = 0; tt < NT; tt++) {
                const uint32_t col = tt * 16 + m;
                const float b = bias[col];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    float v = acc[tt][q] + b;
                    if (relu) v = fmaxf(v, 0.0f);
                    As[(w * 16 + (lane >> 4) * 4 + q) * LDK + col] = f2bf(v);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            constexpr int CH = N / 8;  // 16-byte chunks per output row
            for (int q = (int)lane; q < 16 * CH; q += 64) {
                const int row = q / CH, ch = q % CH;
                if (r0 + row < M)
                    *reinterpret_cast<uint4 *>(out_bf + (size_t)(r0 + row) * N + ch * 8) =
                        *reinterpret_cast<const uint4 *>(&As[(w * 16 + row) * LDK + ch * 8]);
            }
        } else {
#pragma unroll
            for (int tt = 0; tt < NT; tt++) {
                const uint32_t col = tt * 16 + m;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t orow = r0 + (lane >> 4) * 4 + q;
                    if (orow < M) out_f[(size_t)orow * N + col] = acc[tt][q];
                }
            }
        }
    }
}

// ---------------------------------------------------------------- head
// y_i = h_i W_h + b_h, k = 0..127 in order (fmaf), each row read as 16 16-byte
// loads; the weights and bias come from LDS.
__global__ __launch_bounds__(256) void k_head(uint32_t n, const uint16_t *__restrict__ H,
                                              const float *__restrict__ Wh,
                                              const float *__restrict__ bh, float *__restrict__ y) {
    __shared__ float w[kHid * kOut + kOut];
    for (uint32_t q = threadIdx.x; q < kHid * kOut + kOut; q += blockDim.x)
        w[q] = q < kHid * kOut ? Wh[q] : bh[q - kHid * kOut];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a0 = w[kHid * kOut], a1 = w[kHid * kOut + 1], a2 = w[kHid * kOut + 2];
    const uint4 *h = reinterpret_cast<const uint4 *>(H + (size_t)i * kHid);
#pragma unroll 4
    for (int kk = 0; kk < kHid / 8; kk++) {
        const uint4 v = h[kk];
        const uint32_t word[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float x = bf2f((uint16_t)(word[j >> 1] >> (16 * (j & 1))));
            const int k = kk * 8 + j;
            a0 = fmaf(x, w[k * 3 + 0], a0);
            a1 = fmaf(x, w[k * 3 + 1], a1);
            a2 = fmaf(x, w[k * 3 + 2], a2);
        }
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
    // Adam's first and second moments and step count, for a training run that
    // stops and resumes where it left off.
    Status get_adam(float *m, float *v, uint64_t *t) {
        JC_TRY(copy_sync(m, m_.p, param_count() * 4, hipMemcpyDeviceToHost, stream_));
        JC_TRY(copy_sync(v, v_.p, param_count() * 4, hipMemcpyDeviceToHost, stream_));
        *t = adam_t_;
        return Status::ok();
    }
    Status set_adam(const float *m, const float *v, uint64_t t) {
        JC_TRY(copy_sync(m_.p, m, param_count() * 4, hipMemcpyHostToDevice, stream_));
        JC_TRY(copy_sync(v_.p, v, param_count() * 4, hipMemcpyHostToDevice, stream_));
        adam_t_ = t;
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
        JC_TRY(rw_.fit(nn * kRw));
        JC_TRY(inv_.fit(nn));
        std::vector<uint32_t> hubs;
        for (uint32_t v = 0; v < n; v++)
            if (off[v + 1] - off[v] > kAggLight) hubs.push_back(v);
        n_hubs_ = (uint32_t)hubs.size();
        JC_TRY(hubs_.fit(hubs.empty() ? 1 : hubs.size()));
        if (!hubs.empty())
            JC_HIP(hipMemcpyAsync(hubs_.p, hubs.data(), hubs.size() * 4, hipMemcpyHostToDevice, stream_));
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
            k_inv_deg<<<cd(n, 256), 256, 0, stream_>>>(n, off_.p, inv_.p);
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
        JC_TRY(fit_train());
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
    // The buffers only the loss and backward pass use, sized for the current
    // graph at the first loss_grad after set_graph: a graph that is only
    // predicted on never allocates them (about 1.8 GB at 1M nodes).
    Status fit_train() {
        const size_t nn = n_ ? n_ : 1;
        JC_TRY(dy_.fit(nn * 3));
        JC_TRY(dzf_.fit(nn * kHid));
        JC_TRY(dz_.fit(nn * kHid));
        JC_TRY(dx_.fit(nn * 2 * kHid));
        JC_TRY(dh_.fit(nn * kHid));
        JC_TRY(parts_.fit((size_t)cd(nn, xtg_rows((uint32_t)nn)) * (2 * kHid + 1) * kHid));
        return Status::ok();
    }

    Status pack() {
        for (int l = 0; l < kLayers; l++)
            k_pack_w<<<cd((uint64_t)kK[l] * kHid, 256), 256, 0, stream_>>>(
                kK[l], params_.p + layer_offset(l), wt_[l].p, wb_[l].p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_forward() {
        const uint32_t n = n_;
        const uint32_t gw = gemm_grid(n);
        // layer 0: [x | mean x] (32) -> 128
        k_agg_mean<kIn><<<cd((uint64_t)n * (kIn / 2), 256), 256, 0, stream_>>>(n, off_.p, nbr_.p, x_.p, a0_.p);
        if (n_hubs_)
            k_agg_mean_hub<kIn><<<n_hubs_, kHubThreads, 0, stream_>>>(hubs_.p, off_.p, nbr_.p, x_.p, a0_.p);
        k_gemm<8, 0, 2 * kHid><<<gw, kGemmThreads, 0, stream_>>>(
            n, x_.p, kIn, a0_.p, kIn, wt_[0].p, params_.p + layer_offset(0) + (uint64_t)kK[0] * kHid, 1, h_[0].p,
            nullptr);
        for (int l = 1; l < kLayers; l++) {
            k_agg_mean<kHid><<<cd((uint64_t)n * (kHid / 2), 256), 256, 0, stream_>>>(n, off_.p, nbr_.p,
                                                                                       h_[l - 1].p, a_[l].p);
            if (n_hubs_)
                k_agg_mean_hub<kHid><<<n_hubs_, kHubThreads, 0, stream_>>>(hubs_.p, off_.p, nbr_.p, h_[l - 1].p,
                                                                           a_[l].p);
            k_gemm<8, 0, 2 * kHid><<<gw, kGemmThreads, 0, stream_>>>(
                n, h_[l - 1].p, kHid, a_[l].p, kHid, wt_[l].p, params_.p + layer_offset(l) + (uint64_t)kK[l] * kHid,
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
        const uint32_t per = xtg_rows(n);
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
            k_gemm<16, 1, kHid><<<gemm_grid(n), kGemmThreads, 0, stream_>>>(
                n, dz_.p, kHid, nullptr, 0, wb_[l].p, nullptr, 0, nullptr, dx_.p);
            // dH_{l-1} = dX[:, :128] + mean^T(dX[:, 128:])
            k_agg_mean_t<kHid><<<cd((uint64_t)n * (kHid / 2), 256), 256, 0, stream_>>>(
                n, off_.p, nbr_.p, inv_.p, dx_.p, 2 * kHid, dh_.p);
            if (n_hubs_)
                k_agg_mean_t_hub<kHid><<<n_hubs_, kHubThreads, 0, stream_>>>(hubs_.p, off_.p, nbr_.p, inv_.p, dx_.p,
                                                                             2 * kHid, dh_.p);
            k_relu_back<<<cd((uint64_t)n * kHid, 256), 256, 0, stream_>>>(n * kHid, dh_.p, h_[l - 1].p,
                                                                          dzf_.p, dz_.p);
        }
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    // Persistent GEMM blocks: one per CU, at most one per row tile.
    static uint32_t gemm_grid(uint32_t n) {
        return std::max<uint32_t>(1, std::min<uint32_t>(cd(n, kGemmRows), 256));
    }

    hipStream_t stream_ = nullptr;
    uint32_t n_ = 0, na_ = 0, n_hubs_ = 0;
    DevBuf<uint32_t> hubs_;  // rows longer than kAggLight, summed one block each
    DevBuf<float> inv_;      // 1 / degree per node (the backward aggregation)
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
int jc_mpnn_get_adam(jc_mpnn *m, float *m1, float *m2, uint64_t *step) {
    if (!m) return nullh();
    if (!m1 || !m2 || !step) { jc::set_error("jc_mpnn_get_adam: NULL output"); return JC_E_INVALID; }
    return done(m->impl.get_adam(m1, m2, step));
}
int jc_mpnn_set_adam(jc_mpnn *m, const float *m1, const float *m2, uint64_t step) {
    if (!m) return nullh();
    if (!m1 || !m2) { jc::set_error("jc_mpnn_set_adam: NULL input"); return JC_E_INVALID; }
    return done(m->impl.set_adam(m1, m2, step));
}
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
