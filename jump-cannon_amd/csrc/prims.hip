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

// Also clears *zero_me (nullable): the long-run counter of sort_morton_pairs,
// reset by a kernel so the sequence can be captured into a graph.
__global__ __launch_bounds__(kSortThreads) void k_rs_hist(const uint64_t *__restrict__ keys,
                                                          uint32_t n, uint32_t shift,
                                                          uint32_t *__restrict__ hist,
                                                          uint32_t nb, uint32_t *zero_me) {
    __shared__ uint32_t h[256];
    if (zero_me && blockIdx.x == 0 && threadIdx.x == 0) *zero_me = 0u;
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
    JC_TRY(long_runs.alloc(count / (kShortRun + 1) + 1));
    JC_TRY(long_count.alloc(1));
    return Status::ok();
}

// ---- Morton keys: the upper 32 bits by radix passes, runs by the lower ----
//
// After stable passes over bits [32, 64) the keys are ordered by their upper
// halves, and keys with equal upper halves keep their input order.  Ordering
// each such run by the lower half (ties kept in order) then gives exactly the
// stable sort by the whole key.  Octree keys of real layouts leave short runs
// (at most a few keys share their upper 32 bits, i.e. their top 10 octree
// levels), so four passes and one run-sorting sweep replace eight passes.

// One thread per run start: runs of up to kShortRun keys are insertion-sorted in
// place (strict compare: stable); longer runs are listed for k_rs_long_runs.
__global__ __launch_bounds__(256) void k_rs_runs(uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                 uint32_t n, uint32_t *__restrict__ long_runs,
                                                 uint32_t *long_count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t hi = (uint32_t)(keys[i] >> 32);
    if (i > 0 && (uint32_t)(keys[i - 1] >> 32) == hi) return;      // not a run start
    if (i + 1 >= n || (uint32_t)(keys[i + 1] >> 32) != hi) return;  // run of one
    uint32_t len = 2;
    while (len <= kShortRun && i + len < n && (uint32_t)(keys[i + len] >> 32) == hi) len++;
    if (len > kShortRun) {
        long_runs[atomicAdd(long_count, 1u)] = i;
        return;
    }
    for (uint32_t a = 1; a < len; a++) {
        const uint64_t ka = keys[i + a];
        const uint32_t va = vals[i + a];
        uint32_t b = a;
        while (b > 0 && keys[i + b - 1] > ka) {
            keys[i + b] = keys[i + b - 1];
            vals[i + b] = vals[i + b - 1];
            b--;
        }
        keys[i + b] = ka;
        vals[i + b] = va;
    }
}

// After three passes over bits [40, 64) (the scratch pair holds keys ordered by
// their upper 24 bits, ties in input order), each key's final place: the start
// of its run of equal key >> 40 plus its rank in the run, the number of the
// run's keys smaller than it or equal and earlier.  The block's keys and
// kRankRun on each side are staged in LDS; every thread ranks its own key over
// its run, so no lane idles behind another's run.  A run longer than kRankRun
// is copied as it is and listed (by its first key) for k_rs_long_runs.
__global__ __launch_bounds__(256) void k_rs_rank40(const uint64_t *__restrict__ ki, const uint32_t *__restrict__ vi,
                                                   uint64_t *__restrict__ ko, uint32_t *__restrict__ vo,
                                                   uint32_t n, uint32_t *__restrict__ long_runs,
                                                   uint32_t *long_count) {
    constexpr uint32_t H = kRankRun;
    __shared__ uint64_t sk[256 + 2 * H];
    const uint32_t b0 = blockIdx.x * 256;
    for (uint32_t q = threadIdx.x; q < 256 + 2 * H; q += 256) {
        const uint32_t g = b0 + q - H;  // wraps below 0: not loaded
        sk[q] = (b0 + q >= H && g < n) ? ki[g] : 0ull;
    }
    __syncthreads();
    const uint32_t i = b0 + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = threadIdx.x + H;
    const uint64_t k = sk[w], hi = k >> 40;
    uint32_t back = 0, fwd = 0;
    while (back < H && i > back && (sk[w - back - 1] >> 40) == hi) back++;
    while (fwd < H && i + fwd + 1 < n && (sk[w + fwd + 1] >> 40) == hi) fwd++;
    const uint32_t v = vi[i];
    if (back + fwd + 1 > H) {  // a long run: copied, and listed by its first key
        ko[i] = k;
        vo[i] = v;
        if (back == 0) long_runs[atomicAdd(long_count, 1u)] = i;
        return;
    }
    uint32_t rank = 0;
    for (uint32_t j = w - back; j <= w + fwd; j++) {
        const uint64_t kj = sk[j];
        rank += (kj < k || (kj == k && j < w)) ? 1u : 0u;
    }
    ko[i - back + rank] = k;
    vo[i - back + rank] = v;
}

// Segmented form for many graphs at once: after radix passes over a composite
// key ck (graph id above the top bits of the raw key), each element's final
// place is the start of its run of equal ck plus its rank in the run by (raw
// key, body index): raw keys are gathered through the values, and a body index
// orders equal raw keys as a stable sort does.  Runs longer than kRankRun are
// copied and listed for k_rs_long_runs_seg.
__global__ __launch_bounds__(256) void k_rs_rank_seg(const uint64_t *__restrict__ ck, const uint32_t *__restrict__ vi,
                                                     const uint64_t *__restrict__ raw, uint64_t *__restrict__ ko,
                                                     uint32_t *__restrict__ vo, uint32_t n,
                                                     uint32_t *__restrict__ long_runs, uint32_t *long_count) {
    constexpr uint32_t H = kRankRun;
    __shared__ uint64_t sc[256 + 2 * H];
    __shared__ uint64_t sr[256 + 2 * H];
    __shared__ uint32_t sv[256 + 2 * H];
    const uint32_t b0 = blockIdx.x * 256;
    for (uint32_t q = threadIdx.x; q < 256 + 2 * H; q += 256) {
        const uint32_t g = b0 + q - H;  // wraps below 0: not loaded
        const bool in = b0 + q >= H && g < n;
        const uint32_t v = in ? vi[g] : 0u;
        sc[q] = in ? ck[g] : 0ull;
        sv[q] = v;
        sr[q] = in ? raw[v] : 0ull;
    }
    __syncthreads();
    const uint32_t i = b0 + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = threadIdx.x + H;
    const uint64_t c = sc[w], k = sr[w];
    const uint32_t v = sv[w];
    uint32_t back = 0, fwd = 0;
    while (back < H && i > back && sc[w - back - 1] == c) back++;
    while (fwd < H && i + fwd + 1 < n && sc[w + fwd + 1] == c) fwd++;
    if (back + fwd + 1 > H) {
        ko[i] = k;
        vo[i] = v;
        if (back == 0) long_runs[atomicAdd(long_count, 1u)] = i;
        return;
    }
    uint32_t rank = 0;
    for (uint32_t j = w - back; j <= w + fwd; j++) {
        const uint64_t kj = sr[j];
        rank += (kj < k || (kj == k && sv[j] < v)) ? 1u : 0u;
    }
    ko[i - back + rank] = k;
    vo[i - back + rank] = v;
}

// Ascending bitonic sort of m (key, value) pairs by key, then value (every
// pair distinct), in a barrier-synchronised array (LDS or, per block, global).
template <typename KArr, typename VArr>
__device__ void bitonic_pairs(KArr K, VArr V, uint32_t m) {
    uint32_t p = 1;
    while (p < m) p <<= 1;
    for (uint32_t k = 2; k <= p; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
                const uint32_t partner = j == (k >> 1) ? (i ^ (k - 1)) : (i ^ j);
                if (partner > i && partner < m) {
                    const uint64_t a = K[i], b = K[partner];
                    const uint32_t va = V[i], vb = V[partner];
                    if (a > b || (a == b && va > vb)) {
                        K[i] = b;
                        K[partner] = a;
                        V[i] = vb;
                        V[partner] = va;
                    }
                }
            }
            __syncthreads();
        }
}

// One block per listed run of equal composite key: its end by binary search
// over ck, then the run's (raw key, body index) pairs sorted (in LDS up to
// kLdsRun, otherwise in place in global memory).
__global__ __launch_bounds__(kLongRunThreads) void k_rs_long_runs_seg(uint64_t *__restrict__ keys,
                                                                      uint32_t *__restrict__ vals,
                                                                      const uint64_t *__restrict__ ck, uint32_t n,
                                                                      const uint32_t *__restrict__ long_runs,
                                                                      const uint32_t *long_count) {
    __shared__ uint64_t sk[kLdsRun];
    __shared__ uint32_t sv[kLdsRun];
    __shared__ uint32_t s_end;
    const uint32_t cnt = *long_count;
    for (uint32_t r = blockIdx.x; r < cnt; r += gridDim.x) {
        const uint32_t s = long_runs[r];
        const uint64_t c = ck[s];
        if (threadIdx.x == 0) {
            uint32_t lo = s + 1, up = n;  // first index past the run
            while (lo < up) {
                const uint32_t mid = lo + (up - lo) / 2;
                if (ck[mid] == c) lo = mid + 1;
                else up = mid;
            }
            s_end = lo;
        }
        __syncthreads();
        const uint32_t m = s_end - s;
        if (m <= kLdsRun) {
            for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
                sk[t] = keys[s + t];
                sv[t] = vals[s + t];
            }
            __syncthreads();
            bitonic_pairs(sk, sv, m);
            for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
                keys[s + t] = sk[t];
                vals[s + t] = sv[t];
            }
        } else {
            bitonic_pairs(keys + s, vals + s, m);
        }
        __syncthreads();
    }
}

// Ascending bitonic sort of m unique 64-bit values (with payloads) for any m:
// each merge stage first compares element i with its mirror i ^ (2k-1), then
// with i ^ j; partners past m act as +infinity and never move.  The caller
// puts a barrier-synchronised array behind `K`/`V` (LDS or, per block, global).
template <typename KArr, typename VArr>
__device__ void bitonic_any(KArr K, VArr V, uint32_t m) {
    uint32_t p = 1;
    while (p < m) p <<= 1;
    for (uint32_t k = 2; k <= p; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
                const uint32_t partner = j == (k >> 1) ? (i ^ (k - 1)) : (i ^ j);
                if (partner > i && partner < m) {
                    const uint64_t a = K[i], b = K[partner];
                    if (a > b) {
                        K[i] = b;
                        K[partner] = a;
                        const uint32_t t = V[i];
                        V[i] = V[partner];
                        V[partner] = t;
                    }
                }
            }
            __syncthreads();
        }
}

// One block per listed run of equal key >> LO: its end by binary search (keys
// are ordered by that prefix), then a bitonic sort of (low LO bits << (64 - LO)
// | rank in run), which makes every value unique and keeps equal low parts in
// input order (ranks need 64 - LO bits: runs shorter than 2^(64 - LO)).  Runs
// of up to kLdsRun keys sort in LDS; longer ones (only layouts with extreme
// outliers) in the scratch pair, in global memory.
template <uint32_t LO>
__global__ __launch_bounds__(kLongRunThreads) void k_rs_long_runs(uint64_t *__restrict__ keys,
                                                                  uint32_t *__restrict__ vals, uint32_t n,
                                                                  const uint32_t *__restrict__ long_runs,
                                                                  const uint32_t *long_count,
                                                                  uint64_t *__restrict__ gk) {
    __shared__ uint64_t sk[kLdsRun];
    __shared__ uint32_t sv[kLdsRun];
    __shared__ uint32_t s_end;
    constexpr uint32_t R = 64 - LO;
    const uint32_t cnt = *long_count;
    for (uint32_t r = blockIdx.x; r < cnt; r += gridDim.x) {
        const uint32_t s = long_runs[r];
        const uint64_t hi = keys[s] >> LO;
        if (threadIdx.x == 0) {
            uint32_t lo = s + 1, up = n;  // first index past the run
            while (lo < up) {
                const uint32_t mid = lo + (up - lo) / 2;
                if ((keys[mid] >> LO) == hi) lo = mid + 1;
                else up = mid;
            }
            s_end = lo;
        }
        __syncthreads();
        const uint32_t m = s_end - s;
        if (m <= kLdsRun) {
            for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
                sk[t] = (keys[s + t] << R) | t;
                sv[t] = vals[s + t];
            }
            __syncthreads();
            bitonic_any(sk, sv, m);
            for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
                keys[s + t] = (hi << LO) | (sk[t] >> R);
                vals[s + t] = sv[t];
            }
        } else {
            uint64_t *K = gk + s;
            for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) K[t] = (keys[s + t] << R) | t;
            __syncthreads();
            bitonic_any(K, vals + s, m);
            for (uint32_t t = threadIdx.x; t < m; t += blockDim.x)
                keys[s + t] = (hi << LO) | (K[t] >> R);
        }
        __syncthreads();
    }
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
        k_rs_hist<<<s.n_blocks, kSortThreads, 0, st>>>(ka, s.n, shift, s.hist.p, s.n_blocks, nullptr);
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

Status sort_morton_pairs(uint64_t *keys, uint32_t *vals, RadixScratch &s, hipStream_t st) {
    if (s.n == 0) return Status::ok();
    uint64_t *ka = keys, *kb = s.keys_alt.p;
    uint32_t *va = vals, *vb = s.vals_alt.p;
    for (uint32_t shift = 32; shift < 64; shift += 8) {
        k_rs_hist<<<s.n_blocks, kSortThreads, 0, st>>>(ka, s.n, shift, s.hist.p, s.n_blocks,
                                                       shift == 32 ? s.long_count.p : nullptr);
        k_rs_digit_scan<<<256, 256, 0, st>>>(s.hist.p, s.n_blocks, s.dbase.p);
        k_rs_scatter<<<s.n_blocks, kSortThreads, 0, st>>>(ka, va, kb, vb, s.n, shift, s.hist.p,
                                                          s.dbase.p, s.n_blocks);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // four passes: the result is back in keys/vals, the scratch pair is free
    k_rs_runs<<<(s.n + 255) / 256, 256, 0, st>>>(keys, vals, s.n, s.long_runs.p, s.long_count.p);
    k_rs_long_runs<32><<<kLongRunBlocks, kLongRunThreads, 0, st>>>(keys, vals, s.n, s.long_runs.p,
                                                                   s.long_count.p, s.keys_alt.p);
    JC_HIP(hipGetLastError());
    return Status::ok();
}

Status sort_segmented_pairs(uint64_t *keys, uint32_t *vals, const uint64_t *raw, RadixScratch &s,
                            hipStream_t st, uint32_t ck_bits) {
    if (s.n == 0) return Status::ok();
    if (ck_bits == 0 || ck_bits > 64 || ck_bits % 16)
        return Status::err(JC_E_INVALID, "sort_segmented_pairs: ck_bits must be 16, 32, 48 or 64");
    // an even number of passes from the scratch pair ends in it again
    uint64_t *ka = s.keys_alt.p, *kb = keys;
    uint32_t *va = s.vals_alt.p, *vb = vals;
    for (uint32_t shift = 0; shift < ck_bits; shift += 8) {
        k_rs_hist<<<s.n_blocks, kSortThreads, 0, st>>>(ka, s.n, shift, s.hist.p, s.n_blocks,
                                                       shift == 0 ? s.long_count.p : nullptr);
        k_rs_digit_scan<<<256, 256, 0, st>>>(s.hist.p, s.n_blocks, s.dbase.p);
        k_rs_scatter<<<s.n_blocks, kSortThreads, 0, st>>>(ka, va, kb, vb, s.n, shift, s.hist.p,
                                                          s.dbase.p, s.n_blocks);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    k_rs_rank_seg<<<(s.n + 255) / 256, 256, 0, st>>>(s.keys_alt.p, s.vals_alt.p, raw, keys, vals, s.n,
                                                     s.long_runs.p, s.long_count.p);
    k_rs_long_runs_seg<<<kLongRunBlocks, kLongRunThreads, 0, st>>>(keys, vals, s.keys_alt.p, s.n, s.long_runs.p,
                                                                   s.long_count.p);
    JC_HIP(hipGetLastError());
    return Status::ok();
}

// Up to kSmallSort pairs by ranking: the position of pair i in the stable order
// is the number of pairs j whose composite key (key_j, j) is below (key_i, i).
// Each block stages all keys in LDS and ranks kRankPerBlock of them, 32 lanes
// per key counting every 32nd key and summing the counts across the lanes;
// lane 0 writes the pair to its rank in the other buffer.  A launch and a copy
// instead of the radix passes' eleven launches, for trees small enough that
// launches, not bytes, set the sort's time.
constexpr uint32_t kRankPerBlock = 32;
__global__ __launch_bounds__(1024) void k_sort_small_rank(const uint64_t *__restrict__ ki,
                                                          const uint32_t *__restrict__ vi,
                                                          uint64_t *__restrict__ ko,
                                                          uint32_t *__restrict__ vo, uint32_t n) {
    __shared__ uint64_t sk[kSmallSort];
    for (uint32_t t = threadIdx.x; t < n; t += 1024) sk[t] = ki[t];
    __syncthreads();
    const uint32_t part = threadIdx.x & 31u;
    const uint32_t i = blockIdx.x * kRankPerBlock + (threadIdx.x >> 5);
    const bool live = i < n;
    const uint64_t k = live ? sk[i] : 0ull;
    uint32_t below = 0;
    for (uint32_t j = part; j < n; j += 32) {
        const uint64_t kj = sk[j];
        below += (kj < k || (kj == k && j < i)) ? 1u : 0u;
    }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) below += __shfl_xor(below, o, 64);  // within the 32 lanes
    if (live && part == 0) {
        ko[below] = k;
        vo[below] = vi[i];
    }
}

Status sort_morton_pairs24(uint64_t *keys, uint32_t *vals, RadixScratch &s, hipStream_t st) {
    if (s.n == 0) return Status::ok();
    if (s.n <= kSmallSort) {
        k_sort_small_rank<<<(s.n + kRankPerBlock - 1) / kRankPerBlock, 1024, 0, st>>>(
            keys, vals, s.keys_alt.p, s.vals_alt.p, s.n);
        k_copy_pairs<<<(s.n + 255) / 256, 256, 0, st>>>(s.keys_alt.p, s.vals_alt.p, keys, vals, s.n);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }
    if (s.n >= kSort24Max) return Status::err(JC_E_INVALID, "sort_morton_pairs24: 2^24 keys or more");
    uint64_t *ka = keys, *kb = s.keys_alt.p;
    uint32_t *va = vals, *vb = s.vals_alt.p;
    for (uint32_t shift = 40; shift < 64; shift += 8) {
        k_rs_hist<<<s.n_blocks, kSortThreads, 0, st>>>(ka, s.n, shift, s.hist.p, s.n_blocks,
                                                       shift == 40 ? s.long_count.p : nullptr);
        k_rs_digit_scan<<<256, 256, 0, st>>>(s.hist.p, s.n_blocks, s.dbase.p);
        k_rs_scatter<<<s.n_blocks, kSortThreads, 0, st>>>(ka, va, kb, vb, s.n, shift, s.hist.p,
                                                          s.dbase.p, s.n_blocks);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // three passes: the keys are in the scratch pair, ordered by their upper 24
    // bits; ranking within the runs writes the full order back to keys/vals
    k_rs_rank40<<<(s.n + 255) / 256, 256, 0, st>>>(ka, va, keys, vals, s.n, s.long_runs.p, s.long_count.p);
    k_rs_long_runs<40><<<kLongRunBlocks, kLongRunThreads, 0, st>>>(keys, vals, s.n, s.long_runs.p,
                                                                   s.long_count.p, s.keys_alt.p);
    JC_HIP(hipGetLastError());
    return Status::ok();
}

}  // namespace jc
