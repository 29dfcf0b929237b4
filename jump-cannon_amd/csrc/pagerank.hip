// PageRank on gfx950: pull power iteration over the symmetrised CSR.
//
// Reference: crates/graph-compute/src/analytics/pagerank.rs:67-287 (host set-up,
// ping-pong ranks, one submit) and shaders/pagerank.wgsl:30-46 (one thread per
// row: rank_out[v] = teleport + d * sum_{u in N(v)} rank_in[u] * inv_deg[u]).
//
// Layout per iteration: every row writes rank_out[v] and x_out[v] =
// rank_out[v] * inv_deg[v], so the next iteration gathers one f32 per CSR entry
// (the product is the one the reference forms per edge, rounded identically).
// Rows are summed in CSR order exactly as the reference thread does, in one
// launch per iteration (k_pr_iter):
//   light rows (<= kHeavy entries): one lane per row;
//   hub rows: one wave per row gathers the x values into LDS in parallel, then
//   one lane adds them in order -- the hub's sum stays bit-identical while its
//   gathers run 64-wide.
// The L1 residual sum_v |rank_out - rank_in| is accumulated exactly (max into
// one fixed-point accumulator, min into another) and rounded once to f64, so
// the stopping decision is deterministic and identical to the oracle's.  The
// convergence flag lives on the device: iterations after convergence exit at
// once, so a whole chunk of iterations runs without a host round trip.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "exact_sum.hpp"

struct jc_pagerank_plan;

namespace jc {
namespace {

constexpr uint32_t kPrTB = 256;
constexpr uint32_t kHeavy = 32;     // rows with more CSR entries go to the hub kernel
constexpr uint32_t kHubChunk = 1024;  // x values per hub-row round: 16 gathers per lane
constexpr uint32_t kAccCols = 256;    // residual limb accumulator columns (few atomics per address)
constexpr uint32_t kChunkIters = 10;

struct PrState {
    uint32_t converged;
    uint32_t iters;
    double resid;
    double tol;
};

__device__ __forceinline__ void resid_add(unsigned long long *acc, float r_new, float r_old) {
    // |r_new - r_old| = max - min, both non-negative ranks: exact as two sums
    exact_acc_add(acc, fmaxf(r_new, r_old));
    exact_acc_add(acc + kLimbs, fminf(r_new, r_old));
}

// The same for one row per lane of a wave (valid: the lane has a row): each
// lane places its two values in register limbs, the wave sums every limb that
// some lane touched, and one lane adds the sums to acc.  The 64 lanes' values
// share an exponent range, so adding them one by one into acc would queue 64
// atomics on the same few LDS words.
__device__ __forceinline__ void resid_add_wave(unsigned long long *acc, bool valid, float r_new,
                                               float r_old, uint32_t lane) {
    unsigned long long p[kLimbs] = {}, q[kLimbs] = {};
    if (valid) {
        exact_reg_add(p, fmaxf(r_new, r_old));
        exact_reg_add(q, fminf(r_new, r_old));
    }
#pragma unroll
    for (uint32_t j = 0; j < kLimbs; j++) {
        if (__any(p[j] != 0ull)) {
            unsigned long long t = p[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            if (lane == 0 && t) atomicAdd(&acc[j], t);
        }
        if (__any(q[j] != 0ull)) {
            unsigned long long t = q[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            if (lane == 0 && t) atomicAdd(&acc[kLimbs + j], t);
        }
    }
}

// Also clears the residual accumulator (k_pr_reduce clears it again after
// every read, so each iteration starts from zero).
__global__ __launch_bounds__(kPrTB) void k_pr_init(uint32_t n, float inv_n,
                                                   const float *__restrict__ inv,
                                                   float *__restrict__ rank0,
                                                   float *__restrict__ x0, PrState *st, double tol,
                                                   unsigned long long *__restrict__ acc_cols) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < 2 * kLimbs * kAccCols) acc_cols[v] = 0ull;
    if (v == 0) {
        st->converged = 0;
        st->iters = 0;
        st->resid = 0.0;
        st->tol = tol;
    }
    if (v >= n) return;
    rank0[v] = inv_n;
    x0[v] = inv_n * inv[v];
}

// One iteration over all rows in one launch.  The first hub_blocks blocks take
// the hub rows (more than kHeavy entries), one row per wave, longest first; the
// rest take the light rows, 64 consecutive rows per wave.  Both halves run side
// by side: the hub rows' serial sums are latency chains that leave most of the
// machine to the light rows.
//   light: the wave's rows have contiguous entries in the light CSR (offl/nbrl,
//     where hub rows are empty); it reads the index run coalesced, gathers all
//     x values at once into LDS, and each lane then adds its own row in CSR
//     order.
//   hub: each round gathers up to kHubChunk x values into the wave's LDS
//     segment (all sixteen index loads of a lane, then all sixteen gathers, in
//     flight together), then lane 0 adds them in CSR order.
__global__ __launch_bounds__(kPrTB) void k_pr_iter(
    uint32_t n, uint32_t hub_blocks, uint32_t n_hubs, const uint32_t *__restrict__ hubs,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
    const uint32_t *__restrict__ offl, const uint32_t *__restrict__ nbrl,
    const float *__restrict__ inv, const float *__restrict__ x_in,
    const float *__restrict__ rank_in, float *__restrict__ rank_out, float *__restrict__ x_out,
    float teleport, float damping, const PrState *st, unsigned long long *__restrict__ partial) {
    if (st->converged) return;
    constexpr uint32_t W = kPrTB / 64, SEG = 64 * kHeavy;
    static_assert(kHubChunk <= SEG, "a hub round fits the wave's segment");
    __shared__ __align__(16) float xs[W][SEG];
    __shared__ unsigned long long acc[2 * kLimbs];
    if (threadIdx.x < 2 * kLimbs) acc[threadIdx.x] = 0ull;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (blockIdx.x < hub_blocks) {
        const uint32_t hi = blockIdx.x * W + w;
        if (hi < n_hubs) {
            const uint32_t v = hubs[hi];
            const uint32_t s = off[v], e = off[v + 1];
            float sum = 0.0f;
            for (uint32_t c = s; c < e; c += kHubChunk) {
                const uint32_t len = min(kHubChunk, e - c);
                constexpr uint32_t R = kHubChunk / 64;
                uint32_t j[R];
#pragma unroll
                for (uint32_t r = 0; r < R; r++) {
                    const uint32_t k = lane + 64u * r;
                    // column ids are read once per iteration: non-temporal, so they
                    // do not push the reused x lines out of the L2 (same-box A/B
                    // at 1M nodes: 0.0878 -> 0.0826 ms per iteration, same ranks)
                    j[r] = k < len ? __builtin_nontemporal_load(nbr + c + k) : 0u;
                }
                float a[R];
#pragma unroll
                for (uint32_t r = 0; r < R; r++) a[r] = x_in[j[r]];
#pragma unroll
                for (uint32_t r = 0; r < R; r++) {
                    const uint32_t k = lane + 64u * r;
                    if (k < len) xs[w][k] = a[r];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (lane == 0) {
                    uint32_t k = 0;
                    for (; k + 4 <= len; k += 4) {
                        const float4 q = *reinterpret_cast<const float4 *>(&xs[w][k]);
                        sum = sum + q.x;
                        sum = sum + q.y;
                        sum = sum + q.z;
                        sum = sum + q.w;
                    }
                    for (; k < len; k++) sum = sum + xs[w][k];
                }
                // lane 0 has read the round before the next one overwrites it
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (lane == 0) {
                const float r = teleport + damping * sum;
                rank_out[v] = r;
                x_out[v] = r * inv[v];
                resid_add(acc, r, rank_in[v]);
            }
        }
    } else {
        const uint32_t r0 = ((blockIdx.x - hub_blocks) * W + w) * 64;
        const uint32_t v = r0 + lane;
        uint32_t seg0 = 0, seg1 = 0;
        if (r0 < n) {
            seg0 = offl[r0];
            seg1 = offl[min(r0 + 64u, n)];
        }
        for (uint32_t base = seg0; base < seg1; base += 8 * 64) {
            uint32_t u[8];
            float a[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) {
                const uint32_t k = base + lane + 64u * q;
                u[q] = k < seg1 ? __builtin_nontemporal_load(nbrl + k) : 0u;  // read once per iteration
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) a[q] = x_in[u[q]];
#pragma unroll
            for (uint32_t q = 0; q < 8; q++) {
                const uint32_t k = base + lane + 64u * q;
                if (k < seg1) xs[w][k - seg0] = a[q];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool mine = v < n && off[v + 1] - off[v] <= kHeavy;
        float r = 0.0f, r_old = 0.0f;
        if (mine) {
            const uint32_t s = offl[v] - seg0, e = offl[v + 1] - seg0;
            float sum = 0.0f;
            for (uint32_t k = s; k < e; k++) sum = sum + xs[w][k];
            r = teleport + damping * sum;
            rank_out[v] = r;
            x_out[v] = r * inv[v];
            r_old = rank_in[v];
        }
        resid_add_wave(acc, mine, r, r_old, lane);
    }
    __syncthreads();
    if (threadIdx.x < 2 * kLimbs && acc[threadIdx.x])
        atomicAdd(&partial[threadIdx.x * kAccCols + blockIdx.x % kAccCols], acc[threadIdx.x]);
}

// Sums the accumulator columns per limb (32 threads per limb, then wave
// shuffles), clears them for the next iteration, and updates the state.
__global__ __launch_bounds__(640) void k_pr_reduce(unsigned long long *__restrict__ partial,
                                                   PrState *st) {
    if (st->converged) return;
    constexpr uint32_t G = 32;
    static_assert(2 * kLimbs * G == 640 && kAccCols % G == 0 && 64 % G == 0, "layout");
    __shared__ unsigned long long red[2 * kLimbs];
    const uint32_t t = threadIdx.x;
    {
        const uint32_t k = t / G, r = t % G;
        unsigned long long *row = partial + (size_t)k * kAccCols;
        unsigned long long sum = 0;
#pragma unroll
        for (uint32_t c = r; c < kAccCols; c += G) {
            sum += row[c];
            row[c] = 0ull;
        }
#pragma unroll
        for (uint32_t o = G / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, (int)o, 64);
        if (r == 0) red[k] = sum;
    }
    __syncthreads();
    if (t != 0) return;
    unsigned long long p[kLimbs], q[kLimbs];
#pragma unroll
    for (uint32_t k = 0; k < kLimbs; k++) {
        p[k] = red[k];
        q[k] = red[kLimbs + k];
    }
    exact_normalise(p);
    exact_normalise(q);
    exact_sub(p, q);
    const double r = exact_to_double(p);
    st->resid = r;
    st->iters += 1;
    if (r < st->tol) st->converged = 1;
}

inline uint32_t cdiv32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- the allocation-free variant (jc_pagerank_noalloc) ----
// The CSR is the caller's, in device memory, so nothing is prepared on the
// host: the CSR is checked and 1/deg formed on the device, and each iteration
// runs over the plain CSR (no hub list, no light CSR).  Every row is still
// summed in CSR order, so the ranks are bit-identical to k_pr_iter's.

// Set by k_pr_check_csr: bad = 1 offsets[0] != 0, 2 offsets[n] != nnz,
// 4 decreasing offsets, 8 a neighbour id >= n; dangling = smallest node of
// degree 0 (0xFFFFFFFF: none).
struct CsrCheck {
    uint32_t bad;
    uint32_t dangling;
};

__global__ void k_pr_check_clear(CsrCheck *c) {
    if (threadIdx.x == 0) {
        c->bad = 0u;
        c->dangling = 0xFFFFFFFFu;
    }
}

__global__ __launch_bounds__(kPrTB) void k_pr_check_csr(uint32_t n, const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ nbr, uint64_t nnz,
                                                        float *__restrict__ inv, CsrCheck *c) {
    const uint32_t stride = gridDim.x * kPrTB;
    for (uint32_t v = blockIdx.x * kPrTB + threadIdx.x; v <= n; v += stride) {
        const uint32_t a = off[v];
        if (v == 0 && a != 0u) atomicOr(&c->bad, 1u);
        if (v == n) {
            if ((uint64_t)a != nnz) atomicOr(&c->bad, 2u);
            continue;
        }
        const uint32_t b = off[v + 1];
        if (b < a) {
            atomicOr(&c->bad, 4u);
            continue;
        }
        if (b == a) atomicMin(&c->dangling, v);
        else inv[v] = __fdiv_rn(1.0f, (float)(b - a));  // pagerank.rs:105, correctly rounded
    }
    for (uint64_t j = (uint64_t)blockIdx.x * kPrTB + threadIdx.x; j < nnz; j += stride)
        if (nbr[j] >= n) atomicOr(&c->bad, 8u);
}

// One iteration over the plain CSR, 64 consecutive rows per wave.  A row of at
// most kHeavy entries is summed by its own lane straight from global memory;
// longer rows are taken one after another by the whole wave, which gathers
// kHubChunk x values per round into LDS for lane 0 to add in order (the hub
// path of k_pr_iter).
__global__ __launch_bounds__(kPrTB) void k_pr_iter_csr(
    uint32_t n, const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
    const float *__restrict__ inv, const float *__restrict__ x_in, const float *__restrict__ rank_in,
    float *__restrict__ rank_out, float *__restrict__ x_out, float teleport, float damping,
    const PrState *st, unsigned long long *__restrict__ partial) {
    if (st->converged) return;
    constexpr uint32_t W = kPrTB / 64;
    __shared__ __align__(16) float xs[W][kHubChunk];
    __shared__ unsigned long long acc[2 * kLimbs];
    if (threadIdx.x < 2 * kLimbs) acc[threadIdx.x] = 0ull;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    const uint32_t v = (blockIdx.x * W + w) * 64 + lane;
    uint32_t s = 0, e = 0;
    if (v < n) {
        s = off[v];
        e = off[v + 1];
    }
    const bool light = v < n && e - s <= kHeavy;
    float r = 0.0f, r_old = 0.0f;
    if (light) {
        float sum = 0.0f;
        for (uint32_t j = s; j < e; j++) sum = sum + x_in[nbr[j]];
        r = teleport + damping * sum;
        rank_out[v] = r;
        x_out[v] = r * inv[v];
        r_old = rank_in[v];
    }
    resid_add_wave(acc, light, r, r_old, lane);
    unsigned long long heavy = __ballot(v < n && !light);
    while (heavy) {
        const uint32_t L = (uint32_t)__builtin_ctzll(heavy);
        heavy &= heavy - 1ull;
        const uint32_t hv = __shfl(v, (int)L, 64), hs = __shfl(s, (int)L, 64), he = __shfl(e, (int)L, 64);
        float sum = 0.0f;
        for (uint32_t c = hs; c < he; c += kHubChunk) {
            const uint32_t len = min(kHubChunk, he - c);
            constexpr uint32_t R = kHubChunk / 64;
            uint32_t j[R];
#pragma unroll
            for (uint32_t q = 0; q < R; q++) {
                const uint32_t k = lane + 64u * q;
                j[q] = k < len ? nbr[c + k] : 0u;
            }
            float a[R];
#pragma unroll
            for (uint32_t q = 0; q < R; q++) a[q] = x_in[j[q]];
#pragma unroll
            for (uint32_t q = 0; q < R; q++) {
                const uint32_t k = lane + 64u * q;
                if (k < len) xs[w][k] = a[q];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) {
                uint32_t k = 0;
                for (; k + 4 <= len; k += 4) {
                    const float4 q = *reinterpret_cast<const float4 *>(&xs[w][k]);
                    sum = sum + q.x;
                    sum = sum + q.y;
                    sum = sum + q.z;
                    sum = sum + q.w;
                }
                for (; k < len; k++) sum = sum + xs[w][k];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane == 0) {
            const float rh = teleport + damping * sum;
            rank_out[hv] = rh;
            x_out[hv] = rh * inv[hv];
            resid_add(acc, rh, rank_in[hv]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kLimbs && acc[threadIdx.x])
        atomicAdd(&partial[threadIdx.x * kAccCols + blockIdx.x % kAccCols], acc[threadIdx.x]);
}

// Workspace of jc_pagerank_noalloc, carved from one static device array:
// residual accumulator, iteration state, CSR check, then five node vectors
// (1/deg, two rank and two x buffers) of kNoAllocMaxN floats each.
constexpr uint32_t kNoAllocMaxN = JC_PAGERANK_NOALLOC_MAX_N;
constexpr size_t kNaPartial = 0;
constexpr size_t kNaState = kNaPartial + (size_t)kAccCols * 2 * kLimbs * sizeof(unsigned long long);
constexpr size_t kNaCheck = kNaState + 256;
constexpr size_t kNaVec = kNaCheck + 256;
constexpr size_t kNaVecBytes = (size_t)kNoAllocMaxN * sizeof(float);
constexpr size_t kNaBytes = kNaVec + 5 * kNaVecBytes;
static_assert(sizeof(PrState) <= 256 && sizeof(CsrCheck) <= 256 && kNaVec % 256 == 0, "layout");

}  // namespace

class PageRankPlan {
  public:
    Status create(uint32_t n, const uint32_t *off, const uint32_t *nbr, uint64_t nnz, float damping) {
        JC_TRY(require_device(nullptr));
        CsrView g{n, off, nbr, nnz};
        JC_TRY(validate_csr(g));
        n_ = n;
        damping_ = damping;
        std::vector<float> inv(n ? n : 1, 0.0f);
        std::vector<uint32_t> hubs, offl((size_t)n + 1, 0), nbrl;
        for (uint32_t v = 0; v < n; v++) {
            const uint32_t d = off[v + 1] - off[v];
            if (d == 0)
                return Status::err(JC_E_DANGLING,
                                   "gpu_pagerank requires no dangling nodes (node " +
                                       std::to_string(v) + " has degree 0)");
            inv[v] = 1.0f / (float)d;  // pagerank.rs:105
            if (d > kHeavy) hubs.push_back(v);
            else nbrl.insert(nbrl.end(), nbr + off[v], nbr + off[v + 1]);
            offl[v + 1] = (uint32_t)nbrl.size();
        }
        // longest hub rows first: they are the critical path of an iteration
        std::stable_sort(hubs.begin(), hubs.end(), [&](uint32_t a, uint32_t b) {
            return off[a + 1] - off[a] > off[b + 1] - off[b];
        });
        n_hubs_ = (uint32_t)hubs.size();
        if (n == 0) return Status::ok();
        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        JC_TRY(off_.alloc((size_t)n + 1));
        JC_TRY(nbr_.alloc(nnz));
        JC_TRY(inv_.alloc(n));
        JC_TRY(hubs_.alloc(hubs.size()));
        JC_TRY(offl_.alloc((size_t)n + 1));
        JC_TRY(nbrl_.alloc(nbrl.size() ? nbrl.size() : 1));
        for (int k = 0; k < 2; k++) {
            JC_TRY(rank_[k].alloc(n));
            JC_TRY(x_[k].alloc(n));
        }
        nb_light_ = cdiv32(n, kPrTB);
        JC_TRY(partial_.alloc((size_t)kAccCols * 2 * kLimbs));
        JC_TRY(st_.alloc(1));
        JC_HIP(hipMemcpyAsync(off_.p, off, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
        if (nnz) JC_HIP(hipMemcpyAsync(nbr_.p, nbr, nnz * 4, hipMemcpyHostToDevice, stream_));
        JC_HIP(hipMemcpyAsync(inv_.p, inv.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream_));
        JC_HIP(hipMemcpyAsync(offl_.p, offl.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
        if (!nbrl.empty())
            JC_HIP(hipMemcpyAsync(nbrl_.p, nbrl.data(), nbrl.size() * 4, hipMemcpyHostToDevice, stream_));
        if (!hubs.empty())
            JC_HIP(hipMemcpyAsync(hubs_.p, hubs.data(), hubs.size() * 4, hipMemcpyHostToDevice, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }

    ~PageRankPlan() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (chunk_exec_) (void)hipGraphExecDestroy(chunk_exec_);
        if (chunk_graph_) (void)hipGraphDestroy(chunk_graph_);
        for (auto &e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    Status run(uint32_t max_iters, double tol, float *out, uint32_t *iters, double *resid, double *ms) {
        if (n_ == 0) {
            if (iters) *iters = 0;
            if (resid) *resid = 0.0;
            if (ms) *ms = 0.0;
            return Status::ok();
        }
        const float inv_n = 1.0f / (float)n_;           // pagerank.rs:108
        const float teleport = (1.0f - damping_) * inv_n;  // pagerank.rs:109
        teleport_ = teleport;
        for (auto &e : ev_)
            if (!e) JC_HIP(hipEventCreate(&e));
        JC_HIP(hipEventRecord(ev_[0], stream_));
        k_pr_init<<<cdiv32(std::max<uint32_t>(n_, 2 * kLimbs * kAccCols), kPrTB), kPrTB, 0, stream_>>>(
            n_, inv_n, inv_.p, rank_[0].p, x_[0].p, st_.p, tol, partial_.p);
        JC_HIP(hipGetLastError());
        PrState hs{};
        uint32_t launched = 0;
        while (launched < max_iters) {
            const uint32_t chunk = std::min(kChunkIters, max_iters - launched);
            if (chunk == kChunkIters) {
                JC_TRY(ensure_chunk_graph());
                JC_HIP(hipGraphLaunch(chunk_exec_, stream_));
            } else {
                for (uint32_t k = 0; k < chunk; k++) JC_TRY(enqueue_iteration(k & 1u));
            }
            launched += chunk;
            JC_HIP(hipMemcpyAsync(&hs, st_.p, sizeof hs, hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
            if (hs.converged) break;
        }
        JC_HIP(hipEventRecord(ev_[1], stream_));
        JC_HIP(hipEventSynchronize(ev_[1]));
        float t = 0;
        JC_HIP(hipEventElapsedTime(&t, ev_[0], ev_[1]));
        if (ms) *ms = t;
        if (iters) *iters = hs.iters;
        if (resid) *resid = hs.resid;
        if (out)
            JC_TRY(copy_sync(out, rank_[hs.iters & 1u].p, (size_t)n_ * 4, hipMemcpyDeviceToHost, stream_));
        return Status::ok();
    }

    // Kernel-only timing of n_iters iterations with convergence disabled
    // (tol = -1), events around each iteration's kernels.
    Status time_iterations(uint32_t n_iters, double *ms_per_iter) {
        if (n_ == 0) return Status::ok();
        for (auto &e : ev_)
            if (!e) JC_HIP(hipEventCreate(&e));
        const float inv_n = 1.0f / (float)n_;
        teleport_ = (1.0f - damping_) * inv_n;
        k_pr_init<<<cdiv32(std::max<uint32_t>(n_, 2 * kLimbs * kAccCols), kPrTB), kPrTB, 0, stream_>>>(
            n_, inv_n, inv_.p, rank_[0].p, x_[0].p, st_.p, -1.0, partial_.p);
        JC_HIP(hipGetLastError());
        JC_HIP(hipEventRecord(ev_[0], stream_));
        for (uint32_t k = 0; k < n_iters; k++) JC_TRY(enqueue_iteration(k & 1u));
        JC_HIP(hipEventRecord(ev_[1], stream_));
        JC_HIP(hipEventSynchronize(ev_[1]));
        float t = 0;
        JC_HIP(hipEventElapsedTime(&t, ev_[0], ev_[1]));
        if (ms_per_iter) *ms_per_iter = n_iters ? t / n_iters : 0.0;
        return Status::ok();
    }

  private:
    Status enqueue_iteration(uint32_t c) {
        const uint32_t hub_blocks = cdiv32(n_hubs_, kPrTB / 64);
        k_pr_iter<<<hub_blocks + nb_light_, kPrTB, 0, stream_>>>(
            n_, hub_blocks, n_hubs_, hubs_.p, off_.p, nbr_.p, offl_.p, nbrl_.p, inv_.p, x_[c].p,
            rank_[c].p, rank_[c ^ 1].p, x_[c ^ 1].p, teleport_, damping_, st_.p, partial_.p);
        k_pr_reduce<<<1, 640, 0, stream_>>>(partial_.p, st_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status ensure_chunk_graph() {
        if (chunk_exec_) return Status::ok();
        JC_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        Status s = Status::ok();
        for (uint32_t k = 0; k < kChunkIters && !s.failed(); k++) s = enqueue_iteration(k & 1u);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream_, &g);
        if (s.failed()) return s;
        if (e != hipSuccess) return Status::err(JC_E_HIP, hipGetErrorString(e));
        chunk_graph_ = g;
        JC_HIP(hipGraphInstantiate(&chunk_exec_, g, nullptr, nullptr, 0));
        return Status::ok();
    }

    uint32_t n_ = 0, n_hubs_ = 0, nb_light_ = 0;
    float damping_ = 0.85f, teleport_ = 0.0f;
    hipStream_t stream_ = nullptr;
    DevBuf<uint32_t> off_, nbr_, hubs_, offl_, nbrl_;
    DevBuf<float> inv_, rank_[2], x_[2];
    DevBuf<unsigned long long> partial_;
    DevBuf<PrState> st_;
    hipGraph_t chunk_graph_ = nullptr;
    hipGraphExec_t chunk_exec_ = nullptr;
    hipEvent_t ev_[2] = {};
};

}  // namespace jc

struct jc_pagerank_plan {
    jc::PageRankPlan impl;
};

// The workspace of jc_pagerank_noalloc.  Each device holds its own copy at its
// own address: host code uses the address hipGetSymbolAddress gives for the
// current device (noalloc_scratch), never this name.
static __device__ __align__(256) unsigned char scratch[jc::kNaBytes];

namespace jc {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
Status noalloc_scratch(unsigned char **out) {
    int dev = 0;
    JC_TRY(require_device(&dev));
    if (dev < 0 || dev >= kMaxDevices) return Status::err(JC_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        JC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return Status::ok();
}

Status pagerank_noalloc(uint32_t n, const uint32_t *off, const uint32_t *nbr, uint64_t nnz, float damping,
                        uint32_t max_iters, double tol, float *out, uint32_t *iters, double *resid,
                        hipStream_t st) {
    if (n > kNoAllocMaxN)
        return Status::err(JC_E_INVALID, "jc_pagerank_noalloc holds at most " + std::to_string(kNoAllocMaxN) +
                                             " nodes in its static workspace (n = " + std::to_string(n) + ")");
    if (nnz > 0xFFFFFFFFull) return Status::err(JC_E_INVALID, "CSR nnz exceeds the u32 offset range");
    if (n && (!off || !out || (nnz && !nbr))) return Status::err(JC_E_INVALID, "NULL CSR or output pointer");
    unsigned char *base = nullptr;
    JC_TRY(noalloc_scratch(&base));
    if (n == 0) {
        if (iters) *iters = 0;
        if (resid) *resid = 0.0;
        return Status::ok();
    }
    auto *partial = reinterpret_cast<unsigned long long *>(base + kNaPartial);
    auto *state = reinterpret_cast<PrState *>(base + kNaState);
    auto *check = reinterpret_cast<CsrCheck *>(base + kNaCheck);
    float *vec[5];
    for (int k = 0; k < 5; k++) vec[k] = reinterpret_cast<float *>(base + kNaVec + k * kNaVecBytes);
    float *inv = vec[0], *rank[2] = {vec[1], vec[2]}, *x[2] = {vec[3], vec[4]};

    k_pr_check_clear<<<1, 64, 0, st>>>(check);
    k_pr_check_csr<<<std::min<uint32_t>(cdiv32((uint64_t)n + 1, kPrTB), 2048), kPrTB, 0, st>>>(n, off, nbr, nnz,
                                                                                            inv, check);
    JC_HIP(hipGetLastError());
    CsrCheck hc{};
    JC_HIP(hipMemcpyAsync(&hc, check, sizeof hc, hipMemcpyDeviceToHost, st));
    JC_HIP(hipStreamSynchronize(st));
    if (hc.bad & 1u) return Status::err(JC_E_INVALID, "CSR offsets[0] != 0");
    if (hc.bad & 2u) return Status::err(JC_E_INVALID, "CSR offsets[n] != nnz");
    if (hc.bad & 4u) return Status::err(JC_E_INVALID, "CSR offsets decrease");
    if (hc.bad & 8u) return Status::err(JC_E_INVALID, "CSR neighbour id out of range");
    if (hc.dangling != 0xFFFFFFFFu)
        return Status::err(JC_E_DANGLING, "gpu_pagerank requires no dangling nodes (node " +
                                              std::to_string(hc.dangling) + " has degree 0)");

    const float inv_n = 1.0f / (float)n;             // pagerank.rs:108
    const float teleport = (1.0f - damping) * inv_n;  // pagerank.rs:109
    k_pr_init<<<cdiv32(std::max<uint32_t>(n, 2 * kLimbs * kAccCols), kPrTB), kPrTB, 0, st>>>(
        n, inv_n, inv, rank[0], x[0], state, tol, partial);
    JC_HIP(hipGetLastError());
    PrState hs{};
    uint32_t launched = 0;
    while (launched < max_iters) {
        const uint32_t chunk = std::min(kChunkIters, max_iters - launched);
        for (uint32_t k = 0; k < chunk; k++) {
            const uint32_t c = (launched + k) & 1u;
            k_pr_iter_csr<<<cdiv32(n, kPrTB), kPrTB, 0, st>>>(n, off, nbr, inv, x[c], rank[c], rank[c ^ 1],
                                                             x[c ^ 1], teleport, damping, state, partial);
            k_pr_reduce<<<1, 640, 0, st>>>(partial, state);
        }
        JC_HIP(hipGetLastError());
        launched += chunk;
        JC_HIP(hipMemcpyAsync(&hs, state, sizeof hs, hipMemcpyDeviceToHost, st));
        JC_HIP(hipStreamSynchronize(st));
        if (hs.converged) break;
    }
    JC_HIP(hipMemcpyAsync(out, rank[hs.iters & 1u], (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    JC_HIP(hipStreamSynchronize(st));
    if (iters) *iters = hs.iters;
    if (resid) *resid = hs.resid;
    return Status::ok();
}

}  // namespace
}  // namespace jc

extern "C" {

int jc_pagerank_plan_create(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                            uint64_t nnz, float damping, jc_pagerank_plan **out) {
    if (!out) {
        jc::set_error("NULL output handle");
        return JC_E_INVALID;
    }
    auto *p = new jc_pagerank_plan();
    jc::Status s = p->impl.create(n, offsets, neighbors, nnz, damping);
    if (s.failed()) {
        delete p;
        return s.code;
    }
    *out = p;
    jc::set_error("");
    return JC_OK;
}

int jc_pagerank_plan_run(jc_pagerank_plan *p, uint32_t max_iters, double tol_l1, float *out_ranks,
                         uint32_t *out_iters, double *out_residual_l1, double *out_gpu_ms) {
    if (!p) {
        jc::set_error("NULL plan");
        return JC_E_INVALID;
    }
    jc::Status s = p->impl.run(max_iters, tol_l1, out_ranks, out_iters, out_residual_l1, out_gpu_ms);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_pagerank_plan_time(jc_pagerank_plan *p, uint32_t n_iters, double *ms_per_iter) {
    if (!p) {
        jc::set_error("NULL plan");
        return JC_E_INVALID;
    }
    jc::Status s = p->impl.time_iterations(n_iters, ms_per_iter);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

void jc_pagerank_plan_destroy(jc_pagerank_plan *p) { delete p; }

int jc_pagerank(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                float damping, uint32_t max_iters, double tol_l1, float *out_ranks,
                uint32_t *out_iters, double *out_residual_l1) {
    jc_pagerank_plan *p = nullptr;
    int rc = jc_pagerank_plan_create(n, offsets, neighbors, nnz, damping, &p);
    if (rc) return rc;
    rc = jc_pagerank_plan_run(p, max_iters, tol_l1, out_ranks, out_iters, out_residual_l1, nullptr);
    jc_pagerank_plan_destroy(p);
    return rc;
}

int jc_pagerank_noalloc(uint32_t n, const uint32_t *d_offsets, const uint32_t *d_neighbors, uint64_t nnz,
                        float damping, uint32_t max_iters, double tol_l1, float *d_out_ranks,
                        uint32_t *out_iters, double *out_residual_l1, void *stream) {
    jc::Status s = jc::pagerank_noalloc(n, d_offsets, d_neighbors, nnz, damping, max_iters, tol_l1, d_out_ranks,
                                        out_iters, out_residual_l1, static_cast<hipStream_t>(stream));
    if (!s.failed()) jc::set_error("");
    return s.code;
}

}  // extern "C"
