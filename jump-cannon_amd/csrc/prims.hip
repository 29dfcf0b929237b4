// Device-wide primitives for gfx950: stable LSD radix sort of 64-bit keys with
// 32-bit payloads.  Each 8-bit pass is three launches: tile histograms ->
// per-digit scans of the tile counts (256 blocks in parallel) -> stable scatter,
// which also scans the 256 digit totals into digit bases in LDS.  Ranking inside a tile is wave-level (64 lanes): the lanes
// holding the same digit are found with eight ballots, the lowest of them
// reserves a run with one LDS atomic and broadcasts it, so equal digits keep
// their input order (stability) without any block barrier between rounds.
#include "jc_internal.hpp"

namespace jc {

namespace {

__global__ __launch_bounds__(kSortThreads) void k_rs_hist(const uint64_t *__restrict__ keys,
                                                          uint32_t n, uint32_t shift,
                                                          uint32_t *__restrict__ hist,
                                                          uint32_t nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; j++) {
        uint32_t i = base + j * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Per digit d (one block each): exclusive scan over the tile counts hist[d][*]
// in place, and the digit's total into dtot[d].
__global__ __launch_bounds__(256) void k_rs_digit_scan(uint32_t *__restrict__ hist, uint32_t nb,
                                                       uint32_t *__restrict__ dtot) {
    __shared__ uint32_t ws[4];
    __shared__ uint32_t carry;
    uint32_t *row = hist + (size_t)blockIdx.x * nb;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t base = 0; base < nb; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? row[i] : 0u;
        const uint32_t incl = wave_incl_scan(v);
        if (lane == 63) ws[w] = incl;
        __syncthreads();
        uint32_t wp = 0, tot = 0;
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            if (q < w) wp += ws[q];
            tot += ws[q];
        }
        const uint32_t c0 = carry;
        if (i < nb) row[i] = c0 + wp + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c0 + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// Exclusive scan of one value per thread over a 256-thread block (the caller
// synchronises before reusing ws).
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *ws) {
    const uint32_t incl = wave_incl_scan(v);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t wp = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; q++)
        if (q < w) wp += ws[q];
    return wp + incl - v;
}

// Stable scatter of one tile.  Keys are ranked per wave and digit (ballots),
// placed in LDS in (digit, input order) order, and then written out by
// consecutive threads, so each digit's run of the tile goes to global memory
// as one contiguous, coalesced span.
__global__ __launch_bounds__(kSortThreads) void k_rs_scatter(
    const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n, uint32_t shift,
    const uint32_t *__restrict__ hist, const uint32_t *__restrict__ dtot, uint32_t nb) {
    static_assert(kSortThreads == 256, "one thread per digit");
    constexpr uint32_t W = kSortThreads / 64;
    __shared__ uint32_t wcnt[W][256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t ws[4];
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (uint32_t q = 0; q < W; q++) wcnt[q][tid] = 0;
    const uint32_t dbase = block_excl_scan(dtot[tid], ws);   // digit base over all tiles
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t tile0 = blockIdx.x * kSortTile;
    const uint32_t base = tile0 + w * (64 * kSortItems);
    uint64_t k[kSortItems];
    uint32_t v[kSortItems], rank[kSortItems];
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        const uint32_t i = base + r * 64 + lane;
        const bool valid = i < n;
        k[r] = valid ? kin[i] : 0ull;
        v[r] = valid ? vin[i] : 0u;
        const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
        uint64_t mask = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            mask &= bit ? bb : ~bb;
        }
        const uint32_t before = (uint32_t)__popcll(mask & lt);
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t old = 0;
        if (valid && before == 0) old = atomicAdd(&wcnt[w][d], (uint32_t)__popcll(mask));
        old = __shfl(old, (int)leader, 64);
        rank[r] = old + before;
    }
    __syncthreads();
    {
        // thread tid owns digit tid: the tile's count, its local start (scan over
        // digits) and the per-wave local bases
        uint32_t c[W], tot = 0;
#pragma unroll
        for (uint32_t q = 0; q < W; q++) {
            c[q] = wcnt[q][tid];
            tot += c[q];
        }
        const uint32_t lstart = block_excl_scan(tot, ws);
        uint32_t run = lstart;
#pragma unroll
        for (uint32_t q = 0; q < W; q++) {
            wcnt[q][tid] = run;
            run += c[q];
        }
        gbase[tid] = dbase + hist[tid * nb + blockIdx.x] - lstart;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        if (base + r * 64 + lane < n) {
            const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
            const uint32_t loc = wcnt[w][d] + rank[r];
            skey[loc] = k[r];
            sval[loc] = v[r];
        }
    }
    __syncthreads();
    const uint32_t cnt = min(kSortTile, n - tile0);
    for (uint32_t j = tid; j < cnt; j += kSortThreads) {
        const uint64_t kk = skey[j];
        const uint32_t dst = gbase[(uint32_t)(kk >> shift) & 255u] + j;
        kout[dst] = kk;
        vout[dst] = sval[j];
    }
}

}  // namespace

Status RadixScratch::alloc(uint32_t count) {
    n = count;
    n_blocks = (count + kSortTile - 1) / kSortTile;
    if (n_blocks == 0) n_blocks = 1;
    JC_TRY(keys_alt.alloc(count));
    JC_TRY(vals_alt.alloc(count));
    JC_TRY(hist.alloc(256ull * n_blocks));
    JC_TRY(dbase.alloc(256));
    return Status::ok();
}

__global__ void k_copy_pairs(const uint64_t *__restrict__ ki, const uint32_t *__restrict__ vi,
                             uint64_t *__restrict__ ko, uint32_t *__restrict__ vo, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ko[i] = ki[i];
        vo[i] = vi[i];
    }
}

Status radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, RadixScratch &s, hipStream_t st,
                            uint32_t begin_bit, uint32_t end_bit) {
    if (s.n == 0 || end_bit <= begin_bit) return Status::ok();
    uint64_t *ka = keys, *kb = s.keys_alt.p;
    uint32_t *va = vals, *vb = s.vals_alt.p;
    for (uint32_t shift = begin_bit; shift < end_bit; shift += 8) {
        k_rs_hist<<<s.n_blocks, kSortThreads, 0, st>>>(ka, s.n, shift, s.hist.p, s.n_blocks);
        k_rs_digit_scan<<<256, 256, 0, st>>>(s.hist.p, s.n_blocks, s.dbase.p);
        k_rs_scatter<<<s.n_blocks, kSortThreads, 0, st>>>(ka, va, kb, vb, s.n, shift, s.hist.p,
                                                          s.dbase.p, s.n_blocks);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys)  // odd number of passes: result is in the scratch pair
        k_copy_pairs<<<(s.n + 255) / 256, 256, 0, st>>>(ka, va, keys, vals, s.n);
    JC_HIP(hipGetLastError());
    return Status::ok();
}

}  // namespace jc
