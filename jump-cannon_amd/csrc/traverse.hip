// Graph traversal on gfx950: single-source BFS hop distances, connected
// components, the largest component, and the max-min pivot selection of the
// sparse-stress engine built on the BFS.
//
// References:
//   crates/graph-compute/src/analytics/bfs.rs:44-66 cpu_bfs (queue BFS, u32 hops,
//     u32::MAX unreachable) and :71-300 gpu_bfs (the GPU entry point: empty graph
//     -> empty result, source >= n -> error).
//   crates/graph-compute/src/analytics/connected_components.rs:42-66
//     cpu_connected_components (union-find, union toward the smaller root, so
//     every node is labelled with the smallest node index of its component; CSR
//     entries are taken as undirected edges) and :74 gpu_connected_components.
//   crates/graph-compute/src/engines/sgd_stress.rs:143-175 select_pivots and
//     :372-394 bfs_distances (the pivot rows of the sparse-stress engines).
//
// BFS.  Hop distances are exact whatever the visiting order, so the GPU result
// is bit-identical to the queue BFS.  Frontiers are level-synchronous:
//   * k_bfs_block_ell: one 1024-thread workgroup runs level after level with
//     the frontier in LDS, the visited set as an LDS bitmap (graphs up to
//     kBfsBmNodes nodes) and every row padded to 32 entries, so a level costs
//     one global row load and one barrier: what long thin frontiers need (a
//     random geometric graph has hundreds of levels of a few hundred nodes).
//     k_bfs_block is the same loop for larger graphs, over the CSR, claiming
//     nodes by atomicCAS on their distance.
//   * k_bfs_level: one launch per level over the whole chip for frontiers that
//     outgrow the LDS queue (scale-free graphs reach half the graph within a
//     few hops); visited = atomicCAS on the distance array, appends aggregated
//     per wave.  Frontiers above n/32 of a symmetric CSR run bottom-up instead
//     (k_bfs_level_up: each unreached node stops at its first neighbour in the
//     frontier); a directed CSR stays top-down (k_csr_symmetric decides).
// The host switches between the two by the frontier size the block kernel
// leaves in BfsState.
//
// Components.  Lock-free union-find over the CSR entries (hook the larger root
// under the smaller with atomicCAS, intermediate pointer jumping), after
// Jaiganesh & Burtscher, "A High-Performance Connected Components
// Implementation for GPUs" (HPDC 2018).  A node's parent never exceeds the node,
// so the final root of every node is the smallest index of its component: the
// labels equal the reference union-find's exactly.
#include <algorithm>
#include <vector>

#include "engine.hpp"

struct jc_bfs_plan;

namespace jc {
namespace {

constexpr uint32_t kUnreached = 0xFFFFFFFFu;
constexpr uint32_t kBfsThreads = 1024;
constexpr uint32_t kBfsQ = 4096;            // frontier entries held in LDS
constexpr uint32_t kBfsBmWords = 16384;     // LDS visited bitmap (64 KB)
constexpr uint32_t kBfsBmNodes = kBfsBmWords * 32;
constexpr uint32_t kBfsLight = 32;          // longer rows are expanded by the whole block
constexpr uint32_t kLevelTB = 256;
constexpr uint32_t kWaveHeavy = 64;         // k_bfs_level / k_cc_hook: longer rows go wave-wide
constexpr uint32_t kBottomUpShare = 32;     // bottom-up levels for frontiers above n / 32
constexpr uint32_t kTB = 256;

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

struct BfsState {
    uint32_t qn;     // frontier length
    uint32_t level;  // hop distance of the frontier's nodes
    uint32_t cur;    // global queue (0/1) that holds the frontier
    uint32_t done;   // 1 once a level found no new node
};

// dist = unreached except dist[src] = 0; the frontier is {src}.  The source
// comes from *src_dev when given (the pivot chain), else from src_imm.  Also
// clears the pivot selection's best key for the row this BFS feeds.
__global__ __launch_bounds__(kTB) void k_bfs_init(uint32_t n, uint32_t *__restrict__ dist,
                                                  uint32_t *__restrict__ q0, BfsState *st,
                                                  const uint32_t *src_dev, uint32_t src_imm,
                                                  unsigned long long *best) {
    const uint32_t src = src_dev ? *src_dev : src_imm;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        dist[v] = v == src ? 0u : kUnreached;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        q0[0] = src;
        st->qn = 1;
        st->level = 0;
        st->cur = 0;
        st->done = 0;
        if (best) *best = 0ull;
    }
}

// Whole levels in one workgroup while the frontier fits kBfsQ, for graphs too
// large for the LDS bitmap: a node is claimed by atomicCAS on its distance.  On
// exit BfsState says whether the search ended (done) or which global queue
// holds the next, larger frontier.
__global__ __launch_bounds__(kBfsThreads) void k_bfs_block(const uint32_t *__restrict__ off,
                                                           const uint32_t *__restrict__ nbr,
                                                           uint32_t *__restrict__ dist,
                                                           uint32_t *__restrict__ q0,
                                                           uint32_t *__restrict__ q1, BfsState *st) {
    __shared__ uint32_t lq[2][kBfsQ];
    __shared__ uint32_t heavy[kBfsQ];
    __shared__ uint32_t s_cnt[2], s_nh[2];
    const uint32_t tid = threadIdx.x;
    uint32_t qn = st->qn, lvl = st->level;
    const uint32_t cur = st->cur;
    if (st->done || qn == 0 || qn > kBfsQ) return;
    const uint32_t *gin = cur ? q1 : q0;
    uint32_t *gout = cur ? q0 : q1;
    for (uint32_t i = tid; i < qn; i += kBfsThreads) lq[0][i] = gin[i];
    if (tid < 2) {
        s_cnt[tid] = 0;
        s_nh[tid] = 0;
    }
    __syncthreads();
    uint32_t c = 0;
    while (true) {
        const uint32_t lvl1 = lvl + 1;
        uint32_t *out = lq[c ^ 1];
        uint32_t *cnt = &s_cnt[c ^ 1];
        auto visit = [&](uint32_t u) {
            if (atomicCAS(&dist[u], kUnreached, lvl1) == kUnreached) {
                const uint32_t idx = atomicAdd(cnt, 1u);
                if (idx < kBfsQ) out[idx] = u;
                else gout[idx] = u;  // spills past the LDS queue; idx < n
            }
        };
        for (uint32_t i = tid; i < qn; i += kBfsThreads) {
            const uint32_t v = lq[c][i];
            const uint32_t b = off[v], e = off[v + 1];
            if (e - b > kBfsLight) {
                heavy[atomicAdd(&s_nh[c], 1u)] = v;
                continue;
            }
            for (uint32_t j = b; j < e; j += 4) {
                uint32_t u[4];
#pragma unroll
                for (int k = 0; k < 4; k++) u[k] = j + k < e ? nbr[j + k] : kUnreached;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (u[k] != kUnreached) visit(u[k]);
            }
        }
        __syncthreads();
        const uint32_t nh = s_nh[c];
        for (uint32_t h = 0; h < nh; h++) {
            const uint32_t v = heavy[h];
            const uint32_t e = off[v + 1];
            for (uint32_t j = off[v] + tid; j < e; j += kBfsThreads) visit(nbr[j]);
        }
        // every thread has read this level's counts (s_cnt[c] into qn before
        // the previous barrier, s_nh[c^1] in the previous level)
        if (tid == 0) {
            s_cnt[c] = 0;
            s_nh[c ^ 1] = 0;
        }
        __syncthreads();
        const uint32_t nq = *cnt;
        lvl = lvl1;
        if (nq == 0) {
            if (tid == 0) {
                st->qn = 0;
                st->level = lvl;
                st->done = 1;
            }
            return;
        }
        if (nq > kBfsQ) {
            for (uint32_t i = tid; i < kBfsQ; i += kBfsThreads) gout[i] = out[i];
            if (tid == 0) {
                st->qn = nq;
                st->level = lvl;
                st->cur = cur ^ 1;
            }
            return;
        }
        qn = nq;
        c ^= 1;
    }
}

// Padded rows for the workgroup kernel: slot 0 of row v holds v's degree, the
// next kEllW-1 slots its neighbours; a longer row holds kHeavyRow in slot 0 and
// is expanded from the CSR by the whole block.  One 128-byte row load per
// frontier node replaces the dependent offsets-then-neighbours loads.
constexpr uint32_t kEllW = 32;
constexpr uint32_t kHeavyRow = 0xFFFFFFFEu;

__global__ __launch_bounds__(kTB) void k_build_ell(uint32_t n, const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ nbr,
                                                   uint4 *__restrict__ ell) {
    for (uint32_t v = blockIdx.x * kTB + threadIdx.x; v < n; v += gridDim.x * kTB) {
        const uint32_t b = off[v], deg = off[v + 1] - b;
        uint32_t r[kEllW];
        r[0] = deg < kEllW ? deg : kHeavyRow;
#pragma unroll
        for (uint32_t k = 1; k < kEllW; k++) r[k] = k <= deg && deg < kEllW ? nbr[b + k - 1] : kUnreached;
#pragma unroll
        for (uint32_t k = 0; k < kEllW / 4; k++)
            ell[(size_t)v * (kEllW / 4) + k] = make_uint4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
    }
}

// Whole levels in one workgroup for graphs up to kBfsBmNodes nodes: visited
// set as an LDS bitmap, frontier in LDS, rows from the padded copy.  Per level:
// one row load per frontier node, the bitmap claims of all its neighbours
// issued together (as many as the wave's longest row), one queue reservation
// per wave (wave prefix sum of the claim counts), and one barrier (two when a
// row longer than kEllW needs the whole block).  A node's distance is written
// when its own row is processed, so claims touch LDS only.  Counters rotate
// over three slots so a single barrier orders every reset before the next use.
// On exit BfsState says whether the search ended or which global queue holds
// the next, larger frontier.
__global__ __launch_bounds__(kBfsThreads) void k_bfs_block_ell(const uint4 *__restrict__ ell,
                                                               const uint32_t *__restrict__ off,
                                                               const uint32_t *__restrict__ nbr, uint32_t n,
                                                               uint32_t *__restrict__ dist,
                                                               uint32_t *__restrict__ q0,
                                                               uint32_t *__restrict__ q1, BfsState *st) {
    __shared__ uint32_t bm[kBfsBmWords];
    __shared__ uint32_t lq[2][kBfsQ];
    __shared__ uint32_t heavy[kBfsQ];
    __shared__ uint32_t s_cnt[3], s_nh[3];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    uint32_t qn = st->qn, lvl = st->level;
    const uint32_t cur = st->cur;
    if (st->done || qn == 0 || qn > kBfsQ) return;
    const uint32_t *gin = cur ? q1 : q0;
    uint32_t *gout = cur ? q0 : q1;
    for (uint32_t i = tid; i < qn; i += kBfsThreads) lq[0][i] = gin[i];
    const uint32_t words = (n + 31) / 32;
    if (lvl == 0) {
        for (uint32_t w = tid; w < words; w += kBfsThreads) bm[w] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < qn; i += kBfsThreads) {
            const uint32_t u = lq[0][i];
            atomicOr(&bm[u >> 5], 1u << (u & 31));
        }
    } else {
        // visited = every node that already has a distance (earlier levels ran
        // in k_bfs_level): one ballot per 64 consecutive nodes
        for (uint32_t base = (tid >> 6) * 64; base < words * 32; base += kBfsThreads) {
            const uint32_t v = base + lane;
            const unsigned long long m = __ballot(v < n && dist[v] != kUnreached);
            if (lane == 0) {
                bm[base >> 5] = (uint32_t)m;
                if ((base >> 5) + 1 < words) bm[(base >> 5) + 1] = (uint32_t)(m >> 32);
            }
        }
    }
    if (tid < 3) {
        s_cnt[tid] = 0;
        s_nh[tid] = 0;
    }
    __syncthreads();
    uint32_t c = 0, r3 = 0;
    while (true) {
        const uint32_t lvl1 = lvl + 1;
        uint32_t *out = lq[c ^ 1];
        uint32_t *cnt = &s_cnt[r3 == 2 ? 0 : r3 + 1];
        uint32_t *nhp = &s_nh[r3];
        if (tid == 0) {
            // read for the last time two levels ago; used again next level
            s_cnt[r3 == 0 ? 2 : r3 - 1] = 0;
            s_nh[r3 == 2 ? 0 : r3 + 1] = 0;
        }
        for (uint32_t base = 0; base < qn; base += kBfsThreads) {
            // whole wave idle: a scalar branch, so it issues nothing
            if (__builtin_amdgcn_readfirstlane(base + (tid & ~63u)) >= qn) continue;
            const uint32_t i = base + tid;
            uint32_t u[kEllW];
            uint32_t deg = 0;
            if (i < qn) {
                const uint32_t v = lq[c][i];
                // the frontier's own distance, written by the node's processor
                // (claims only set the bitmap)
                dist[v] = lvl;
                const uint4 *row = ell + (size_t)v * (kEllW / 4);
#pragma unroll
                for (uint32_t k = 0; k < kEllW / 4; k++) {
                    const uint4 x = row[k];
                    u[4 * k] = x.x;
                    u[4 * k + 1] = x.y;
                    u[4 * k + 2] = x.z;
                    u[4 * k + 3] = x.w;
                }
                if (u[0] == kHeavyRow) heavy[atomicAdd(nhp, 1u)] = v;
                else deg = u[0];
            }
            // the wave's longest row (uniform, from ballots) ends the loops
            // below with scalar branches
            uint32_t dmax = 0;
#pragma unroll
            for (int bit = 4; bit >= 0; bit--) {
                const uint32_t cand = dmax | (1u << bit);
                if (__ballot(deg >= cand)) dmax = cand;
            }
            uint32_t fresh = 0;
#pragma unroll
            for (uint32_t k = 1; k < kEllW; k++) {
                if (k > dmax) break;
                if (k <= deg) {
                    const uint32_t bit = 1u << (u[k] & 31);
                    if (!(atomicOr(&bm[u[k] >> 5], bit) & bit)) fresh |= 1u << k;
                }
            }
            // one reservation per wave: exclusive prefix of the claim counts,
            // bit by bit from ballots (no cross-lane shuffles)
            const uint32_t mine = (uint32_t)__popc(fresh);
            uint32_t pre = 0, wtot = 0;
#pragma unroll
            for (int bit = 0; bit < 6; bit++) {
                const unsigned long long m = __ballot((mine >> bit) & 1u);
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                pre += below << bit;
                wtot += (uint32_t)__popcll(m) << bit;
            }
            uint32_t wbase = 0;
            if (wtot) {
                if (lane == 0) wbase = atomicAdd(cnt, wtot);
                wbase = __builtin_amdgcn_readfirstlane(wbase);
            }
            uint32_t o = wbase + pre;
#pragma unroll
            for (uint32_t k = 1; k < kEllW; k++) {
                if (k > dmax) break;
                if (fresh & (1u << k)) {
                    if (o < kBfsQ) out[o] = u[k];
                    else gout[o] = u[k];  // spills past the LDS queue; o < n
                    o++;
                }
            }
        }
        __syncthreads();
        const uint32_t nh = *nhp;
        if (nh) {
            for (uint32_t h = 0; h < nh; h++) {
                const uint32_t v = heavy[h];
                const uint32_t e = off[v + 1];
                for (uint32_t j = off[v] + tid; j < e; j += kBfsThreads) {
                    const uint32_t u = nbr[j];
                    const uint32_t bit = 1u << (u & 31);
                    if (!(atomicOr(&bm[u >> 5], bit) & bit)) {
                        const uint32_t idx = atomicAdd(cnt, 1u);
                        if (idx < kBfsQ) out[idx] = u;
                        else gout[idx] = u;
                    }
                }
            }
            __syncthreads();
        }
        const uint32_t nq = *cnt;
        lvl = lvl1;
        if (nq == 0) {
            if (tid == 0) {
                st->qn = 0;
                st->level = lvl;
                st->done = 1;
            }
            return;
        }
        if (nq > kBfsQ) {
            // the chip-wide levels that take over claim nodes by their distance:
            // the new frontier gets its distance here
            for (uint32_t i = tid; i < nq; i += kBfsThreads) {
                const uint32_t u = i < kBfsQ ? out[i] : gout[i];
                dist[u] = lvl;
                if (i < kBfsQ) gout[i] = u;
            }
            if (tid == 0) {
                st->qn = nq;
                st->level = lvl;
                st->cur = cur ^ 1;
            }
            return;
        }
        qn = nq;
        c ^= 1;
        r3 = r3 == 2 ? 0 : r3 + 1;
    }
}

// One append per wave: the leader reserves popcount(m) slots.
__device__ __forceinline__ void wave_append(bool fresh, uint32_t u, uint32_t *__restrict__ q,
                                            uint32_t *cnt) {
    const unsigned long long m = __ballot(fresh);
    if (!m) return;
    const uint32_t lane = __lane_id();
    const int leader = __ffsll(m) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (fresh) q[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = u;
}

// One level over the chip: qin[0..qn) -> qout, *qn_out counts the new nodes.
__global__ __launch_bounds__(kLevelTB) void k_bfs_level(const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ nbr,
                                                        uint32_t *__restrict__ dist,
                                                        const uint32_t *__restrict__ qin, uint32_t qn,
                                                        uint32_t *__restrict__ qout,
                                                        uint32_t *qn_out, uint32_t lvl1) {
    const uint32_t lane = __lane_id();
    const uint32_t stride = gridDim.x * kLevelTB;
    for (uint32_t base = blockIdx.x * kLevelTB + (threadIdx.x & ~63u); base < qn; base += stride) {
        const uint32_t i = base + lane;
        uint32_t b = 0, e = 0;
        if (i < qn) {
            const uint32_t v = qin[i];
            b = off[v];
            e = off[v + 1];
        }
        const bool hv = e - b > kWaveHeavy;
        uint32_t j = hv ? e : b;
        while (__any(j < e)) {
            const bool act = j < e;
            const uint32_t u = act ? nbr[j] : 0u;
            j += act ? 1u : 0u;
            const bool fresh = act && atomicCAS(&dist[u], kUnreached, lvl1) == kUnreached;
            wave_append(fresh, u, qout, qn_out);
        }
        unsigned long long hm = __ballot(hv);
        while (hm) {
            const int src = __ffsll(hm) - 1;
            hm &= hm - 1ull;
            const uint32_t hb = __shfl(b, src), he = __shfl(e, src);
            for (uint32_t jb = hb; jb < he; jb += 64) {
                const uint32_t jj = jb + lane;
                const bool act = jj < he;
                const uint32_t u = act ? nbr[jj] : 0u;
                const bool fresh = act && atomicCAS(&dist[u], kUnreached, lvl1) == kUnreached;
                wave_append(fresh, u, qout, qn_out);
            }
        }
    }
}

// One level bottom-up (direction-optimizing BFS, Beamer, Asanovic & Patterson,
// SC 2012), for frontiers holding a large share of the graph: every node still
// unreached scans its own row in CSR order and takes distance lvl + 1 at the
// first neighbour whose distance is lvl.  A frontier node's distance does not
// change during the level and a node written in it gets lvl + 1, so no claim
// races and the distances are the queue BFS's.  The new frontier is appended to
// qout (one reservation per wave).
__global__ __launch_bounds__(kLevelTB) void k_bfs_level_up(const uint32_t *__restrict__ off,
                                                           const uint32_t *__restrict__ nbr,
                                                           uint32_t *__restrict__ dist, uint32_t n,
                                                           uint32_t lvl, uint32_t *__restrict__ qout,
                                                           uint32_t *qn_out) {
    const uint32_t lane = __lane_id();
    const uint32_t stride = gridDim.x * kLevelTB;
    for (uint32_t base = blockIdx.x * kLevelTB + (threadIdx.x & ~63u); base < n; base += stride) {
        const uint32_t v = base + lane;
        bool fresh = false;
        if (v < n && dist[v] == kUnreached) {
            const uint32_t e = off[v + 1];
            for (uint32_t j = off[v]; j < e; j++)
                if (dist[nbr[j]] == lvl) {
                    fresh = true;
                    break;
                }
            if (fresh) dist[v] = lvl + 1;
        }
        wave_append(fresh, v, qout, qn_out);
    }
}

// Bottom-up levels scan a node's own row for a parent, which finds every
// parent only when each entry u -> v has its reverse v -> u.  One thread per
// CSR entry: the row u holding entry j (the last u with off[u] <= j), then a
// binary search for u in row v.  *not_sym is set by an entry without a
// reverse and by a row out of ascending order (its entries cannot be searched,
// and the CSR is then taken as directed).
__global__ __launch_bounds__(kTB) void k_csr_symmetric(const uint32_t *__restrict__ off,
                                                       const uint32_t *__restrict__ nbr, uint32_t n,
                                                       uint64_t nnz, uint32_t *not_sym) {
    for (uint64_t j = (uint64_t)blockIdx.x * kTB + threadIdx.x; j < nnz; j += (uint64_t)gridDim.x * kTB) {
        uint32_t lo = 0, hi = n;  // last u in [0, n) with off[u] <= j
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if ((uint64_t)off[mid] <= j) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t u = lo, v = nbr[j];
        bool bad = j > off[u] && nbr[j - 1] > v;
        if (!bad) {
            uint32_t a = off[v], b = off[v + 1];
            while (a < b) {
                const uint32_t mid = a + ((b - a) >> 1);
                if (nbr[mid] < u) a = mid + 1;
                else b = mid;
            }
            bad = a == off[v + 1] || nbr[a] != u;
        }
        if (bad) atomicOr(not_sym, 1u);
    }
}

// ---- max-min pivot selection (sgd_stress.rs:143-175) ----
// After the BFS of pivot t: mind[v] = min(mind[v], d[v]) over reachable v; the
// next pivot is the node of largest finite cover mind[v] > 0, smallest index on
// ties (the reference's strict '>' in node order), kept in *best as
// (cover << 32) | ~v.  Column t of the u16 distance matrix (chunks of eight
// pivots per node: dist8[((t/8) * n + v) * 8 + t%8]) gets d[v], 0 for the pivot
// itself and for unreachable nodes; *bad flags a finite d above 65535.
__global__ __launch_bounds__(kTB) void k_pivot_select(uint32_t n, const uint32_t *__restrict__ dist,
                                                      uint32_t *__restrict__ mind,
                                                      uint16_t *__restrict__ dist8, uint32_t t,
                                                      unsigned long long *best, uint32_t *dmax,
                                                      uint32_t *bad) {
    __shared__ unsigned long long s_key[kTB / 64];
    __shared__ uint32_t s_dm[kTB / 64], s_bad[kTB / 64];
    unsigned long long key = 0ull;
    uint32_t dm = 0, b = 0;
    for (uint32_t v = blockIdx.x * kTB + threadIdx.x; v < n; v += gridDim.x * kTB) {
        const uint32_t dv = dist[v];
        uint32_t m = mind[v];
        if (dv != kUnreached && dv < m) {
            m = dv;
            mind[v] = m;
        }
        if (m != kUnreached && m > 0u) key = max(key, ((unsigned long long)m << 32) | (unsigned long long)(~v));
        const bool fin = dv != kUnreached && dv != 0u;
        if (fin && dv > 65535u) b = 1u;
        if (fin) dm = max(dm, dv);
        dist8[((size_t)(t >> 3) * n + v) * 8 + (t & 7u)] = fin ? (uint16_t)dv : (uint16_t)0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        key = max(key, (unsigned long long)__shfl_xor((unsigned long long)key, o));
        dm = max(dm, (uint32_t)__shfl_xor((int)dm, o));
        b |= (uint32_t)__shfl_xor((int)b, o);
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_key[w] = key;
        s_dm[w] = dm;
        s_bad[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t q = 1; q < kTB / 64; q++) {
            key = max(key, s_key[q]);
            dm = max(dm, s_dm[q]);
            b |= s_bad[q];
        }
        if (key) atomicMax(best, key);
        if (dm) atomicMax(dmax, dm);
        if (b) atomicOr(bad, 1u);
    }
}

// pivots[t] = the source of row t; the next source is the cover argmax, or the
// same node when no node has a positive finite cover.
__global__ void k_pivot_next(const unsigned long long *best, uint32_t *src, uint32_t *pivot_t) {
    const uint32_t s = *src;
    *pivot_t = s;
    const unsigned long long k = *best;
    *src = k ? ~(uint32_t)(k & 0xFFFFFFFFull) : s;
}

__global__ __launch_bounds__(kTB) void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * kTB + threadIdx.x; i < n; i += gridDim.x * kTB) p[i] = v;
}

// ---- connected components ----
__device__ __forceinline__ uint32_t ld_par(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_par(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of v; every node on the way is pointed at its grandparent.  parent[x] <=
// x always, so x is a root iff parent[x] == x and the walk ends.
__device__ __forceinline__ uint32_t cc_root(uint32_t *parent, uint32_t v) {
    uint32_t cur = ld_par(&parent[v]);
    if (cur == v) return v;
    uint32_t prev = v, next;
    while (cur > (next = ld_par(&parent[cur]))) {
        st_par(&parent[prev], next);
        prev = cur;
        cur = next;
    }
    return cur;
}

// Union of the components of v and u: the larger root is hooked under the
// smaller one, retried from the current roots when another thread got there
// first (each retry strictly lowers a root, so it ends).
__device__ __forceinline__ void cc_union(uint32_t *parent, uint32_t v, uint32_t u) {
    uint32_t rv = cc_root(parent, v), ru = cc_root(parent, u);
    while (rv != ru) {
        if (rv < ru) {
            const uint32_t t = rv;
            rv = ru;
            ru = t;
        }
        const uint32_t old = atomicCAS(&parent[rv], rv, ru);
        if (old == rv) return;
        rv = cc_root(parent, old);
        ru = cc_root(parent, ru);
    }
}

// parent[v] = the row's first neighbour when it is smaller than v (an edge, so
// a valid first union), else v.
__global__ __launch_bounds__(kTB) void k_cc_init(uint32_t n, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ nbr,
                                                 uint32_t *__restrict__ parent) {
    for (uint32_t v = blockIdx.x * kTB + threadIdx.x; v < n; v += gridDim.x * kTB) {
        const uint32_t b = off[v];
        const uint32_t u = b < off[v + 1] ? nbr[b] : v;
        parent[v] = u < v ? u : v;
    }
}

__global__ __launch_bounds__(kTB) void k_cc_hook(uint32_t n, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ nbr,
                                                 uint32_t *parent) {
    const uint32_t lane = __lane_id();
    const uint32_t stride = gridDim.x * kTB;
    for (uint32_t base = blockIdx.x * kTB + (threadIdx.x & ~63u); base < n; base += stride) {
        const uint32_t v = base + lane;
        uint32_t b = 0, e = 0;
        if (v < n) {
            b = off[v];
            e = off[v + 1];
        }
        const bool hv = e - b > kWaveHeavy;
        if (!hv)
            for (uint32_t j = b; j < e; j++) {
                const uint32_t u = nbr[j];
                if (u != v) cc_union(parent, v, u);
            }
        unsigned long long hm = __ballot(hv);
        while (hm) {
            const int src = __ffsll(hm) - 1;
            hm &= hm - 1ull;
            const uint32_t hb = __shfl(b, src), he = __shfl(e, src);
            const uint32_t hvtx = base + (uint32_t)src;
            for (uint32_t j = hb + lane; j < he; j += 64) {
                const uint32_t u = nbr[j];
                if (u != hvtx) cc_union(parent, hvtx, u);
            }
        }
    }
}

__global__ __launch_bounds__(kTB) void k_cc_label(uint32_t n, uint32_t *parent,
                                                  uint32_t *__restrict__ label) {
    for (uint32_t v = blockIdx.x * kTB + threadIdx.x; v < n; v += gridDim.x * kTB)
        label[v] = cc_root(parent, v);
}

// size[label[v]] += 1, one atomic per distinct label of a wave.
__global__ __launch_bounds__(kTB) void k_cc_sizes(uint32_t n, const uint32_t *__restrict__ label,
                                                  uint32_t *size) {
    const uint32_t lane = __lane_id();
    for (uint32_t base = blockIdx.x * kTB + (threadIdx.x & ~63u); base < n; base += gridDim.x * kTB) {
        const uint32_t v = base + lane;
        const uint32_t l = v < n ? label[v] : kUnreached;
        bool pending = v < n;
        while (__any(pending)) {
            const int leader = __ffsll(__ballot(pending)) - 1;
            const uint32_t ll = __shfl(l, leader);
            const bool mine = pending && l == ll;
            const unsigned long long m = __ballot(mine);
            if ((int)lane == leader) atomicAdd(&size[ll], (uint32_t)__popcll(m));
            pending = pending && !mine;
        }
    }
}

// Largest component: max over roots of (size << 32) | ~root, so equal sizes go
// to the smallest root, i.e. the component holding the smallest node id.
__global__ __launch_bounds__(kTB) void k_cc_best(uint32_t n, const uint32_t *__restrict__ label,
                                                 const uint32_t *__restrict__ size,
                                                 unsigned long long *best) {
    __shared__ unsigned long long s_key[kTB / 64];
    unsigned long long key = 0ull;
    for (uint32_t v = blockIdx.x * kTB + threadIdx.x; v < n; v += gridDim.x * kTB)
        if (label[v] == v) key = max(key, ((unsigned long long)size[v] << 32) | (unsigned long long)(~v));
    for (int o = 32; o > 0; o >>= 1) key = max(key, (unsigned long long)__shfl_xor((unsigned long long)key, o));
    if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t q = 1; q < kTB / 64; q++) key = max(key, s_key[q]);
        if (key) atomicMax(best, key);
    }
}

// ---- exclusive scan of u32 values f(i), i < n (three passes over tiles) ----
constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanTile = kTB * kScanItems;

// Exclusive scan of one value per thread over the block; returns the prefix and
// leaves the block total in *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t *s_tmp, uint32_t *total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = x;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, o);
        if ((int)lane >= o) inc += y;
    }
    if (lane == 63) s_tmp[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (uint32_t q = 0; q < kTB / 64; q++) {
        if (q < w) wbase += s_tmp[q];
        tot += s_tmp[q];
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - x;
}

template <class F>
__global__ __launch_bounds__(kTB) void k_scan_tiles(uint32_t n, F f, uint32_t *__restrict__ tile_sum) {
    __shared__ uint32_t s_tmp[kTB / 64];
    const uint32_t t0 = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanItems; k++)
        if (t0 + k < n) s += f(t0 + k);
    uint32_t total;
    (void)block_exclusive_scan(s, s_tmp, &total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// In place exclusive scan of the tile sums (one block); grand total in *total.
__global__ __launch_bounds__(kTB) void k_scan_sums(uint32_t nt, uint32_t *__restrict__ tile_sum,
                                                   uint32_t *total) {
    __shared__ uint32_t s_tmp[kTB / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nt; base += kTB) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < nt ? tile_sum[i] : 0u;
        uint32_t chunk;
        const uint32_t pre = block_exclusive_scan(x, s_tmp, &chunk);
        if (i < nt) tile_sum[i] = carry + pre;
        carry += chunk;
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class F, class G>
__global__ __launch_bounds__(kTB) void k_scan_apply(uint32_t n, F f, const uint32_t *__restrict__ tile_off,
                                                    G g) {
    __shared__ uint32_t s_tmp[kTB / 64];
    const uint32_t t0 = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t vals[kScanItems], s = 0;
    for (uint32_t k = 0; k < kScanItems; k++) {
        vals[k] = t0 + k < n ? f(t0 + k) : 0u;
        s += vals[k];
    }
    uint32_t total;
    uint32_t pre = tile_off[blockIdx.x] + block_exclusive_scan(s, s_tmp, &total);
    for (uint32_t k = 0; k < kScanItems; k++)
        if (t0 + k < n) {
            g(t0 + k, vals[k], pre);
            pre += vals[k];
        }
}

template <class F, class G>
Status exclusive_scan(uint32_t n, F f, G g, DevBuf<uint32_t> &tiles, uint32_t *total_dev, hipStream_t st) {
    const uint32_t nt = std::max<uint32_t>(cdiv(n, kScanTile), 1);
    JC_TRY(tiles.fit(nt));
    k_scan_tiles<<<nt, kTB, 0, st>>>(n, f, tiles.p);
    k_scan_sums<<<1, kTB, 0, st>>>(nt, tiles.p, total_dev);
    k_scan_apply<<<nt, kTB, 0, st>>>(n, f, tiles.p, g);
    JC_HIP(hipGetLastError());
    return Status::ok();
}

struct KeepFlag {
    const uint32_t *label;
    const unsigned long long *best;
    __device__ uint32_t operator()(uint32_t v) const {
        return label[v] == ~(uint32_t)(*best & 0xFFFFFFFFull) ? 1u : 0u;
    }
};
struct KeepWrite {
    uint32_t *keep;
    __device__ void operator()(uint32_t v, uint32_t flag, uint32_t pre) const {
        keep[v] = flag ? pre : kUnreached;
    }
};
struct KeptDegree {
    const uint32_t *off;
    const uint32_t *keep;
    __device__ uint32_t operator()(uint32_t v) const {
        return keep[v] != kUnreached ? off[v + 1] - off[v] : 0u;
    }
};
struct SubOffset {
    const uint32_t *keep;
    uint32_t *sub_off;
    __device__ void operator()(uint32_t v, uint32_t deg, uint32_t pre) const {
        (void)deg;
        if (keep[v] != kUnreached) sub_off[keep[v]] = pre;
    }
};

// Rows of the kept nodes with every neighbour renumbered; a wave per 64 nodes,
// lanes striding over one row at a time for long rows.
__global__ __launch_bounds__(kTB) void k_sub_rows(uint32_t n, const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ nbr,
                                                  const uint32_t *__restrict__ keep,
                                                  const uint32_t *__restrict__ sub_off,
                                                  uint32_t *__restrict__ sub_nbr) {
    for (uint32_t v = blockIdx.x * kTB + threadIdx.x; v < n; v += gridDim.x * kTB) {
        const uint32_t k = keep[v];
        if (k == kUnreached) continue;
        uint32_t o = sub_off[k];
        for (uint32_t j = off[v]; j < off[v + 1]; j++) sub_nbr[o++] = keep[nbr[j]];
    }
}

uint32_t grid_for(uint32_t n) { return std::min<uint32_t>(std::max<uint32_t>(cdiv(n, kTB), 1), 2048); }

}  // namespace

// ---- resident CSR + BFS runner ----
class BfsPlan {
  public:
    ~BfsPlan() {
        if (stream_ && own_stream_) (void)hipStreamDestroy(stream_);
    }

    Status create(const CsrView &g, hipStream_t stream) {
        JC_TRY(require_device(nullptr));
        JC_TRY(validate_csr(g));
        n_ = g.n;
        nnz_ = g.nnz;
        if (stream) stream_ = stream;
        else {
            JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
            own_stream_ = true;
        }
        JC_TRY(off_.alloc((size_t)n_ + 1));
        JC_TRY(nbr_.alloc(nnz_ ? nnz_ : 1));
        JC_TRY(dist_.alloc(n_ ? n_ : 1));
        JC_TRY(q_[0].alloc(n_ ? n_ : 1));
        JC_TRY(q_[1].alloc(n_ ? n_ : 1));
        JC_TRY(st_.alloc(1));
        JC_TRY(cnt_.alloc(1));
        if (n_) JC_HIP(hipMemcpyAsync(off_.p, g.offsets, ((size_t)n_ + 1) * 4, hipMemcpyHostToDevice, stream_));
        if (nnz_) JC_HIP(hipMemcpyAsync(nbr_.p, g.neighbors, nnz_ * 4, hipMemcpyHostToDevice, stream_));
        if (n_ && n_ <= kBfsBmNodes) {
            // padded rows for the one-workgroup kernel (128 B per node)
            JC_TRY(ell_.alloc((size_t)n_ * (kEllW / 4)));
            k_build_ell<<<grid_for(n_), kTB, 0, stream_>>>(n_, off_.p, nbr_.p, ell_.p);
            JC_HIP(hipGetLastError());
        }
        symmetric_ = true;
        if (nnz_) {
            k_fill_u32<<<1, kTB, 0, stream_>>>(cnt_.p, 1, 0u);
            k_csr_symmetric<<<std::min<uint32_t>(cdiv(nnz_, kTB), 4096), kTB, 0, stream_>>>(off_.p, nbr_.p, n_,
                                                                                          nnz_, cnt_.p);
            JC_HIP(hipGetLastError());
            JC_HIP(hipMemcpyAsync(&h_qn_, cnt_.p, 4, hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
            symmetric_ = h_qn_ == 0;
        }
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }

    // BFS from *src_dev (device) or src; distances stay in dist() on return.
    // `best` (nullable) is cleared for the pivot selection that follows.
    Status run(const uint32_t *src_dev, uint32_t src, unsigned long long *best) {
        if (n_ == 0) return Status::ok();
        k_bfs_init<<<grid_for(n_), kTB, 0, stream_>>>(n_, dist_.p, q_[0].p, st_.p, src_dev, src, best);
        JC_HIP(hipGetLastError());
        while (true) {
            if (ell_.p)
                k_bfs_block_ell<<<1, kBfsThreads, 0, stream_>>>(ell_.p, off_.p, nbr_.p, n_, dist_.p, q_[0].p,
                                                                q_[1].p, st_.p);
            else
                k_bfs_block<<<1, kBfsThreads, 0, stream_>>>(off_.p, nbr_.p, dist_.p, q_[0].p, q_[1].p, st_.p);
            JC_HIP(hipGetLastError());
            JC_HIP(hipMemcpyAsync(&h_, st_.p, sizeof h_, hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
            if (h_.done) return Status::ok();
            // frontier larger than the LDS queue: whole-chip levels until it
            // fits again (or the search ends)
            while (h_.qn > kBfsQ) {
                k_fill_u32<<<1, kTB, 0, stream_>>>(cnt_.p, 1, 0u);
                if (symmetric_ && (uint64_t)h_.qn * kBottomUpShare > n_) {
                    // large frontier: every unreached node looks for a parent
                    // in its own row (every in-neighbour is in it: symmetric CSR)
                    const uint32_t blocks = std::min<uint32_t>(cdiv(n_, kLevelTB), 4096);
                    k_bfs_level_up<<<blocks, kLevelTB, 0, stream_>>>(off_.p, nbr_.p, dist_.p, n_, h_.level,
                                                                      q_[h_.cur ^ 1].p, cnt_.p);
                } else {
                    const uint32_t blocks = std::min<uint32_t>(cdiv(h_.qn, kLevelTB), 4096);
                    k_bfs_level<<<blocks, kLevelTB, 0, stream_>>>(off_.p, nbr_.p, dist_.p, q_[h_.cur].p,
                                                                  h_.qn, q_[h_.cur ^ 1].p, cnt_.p, h_.level + 1);
                }
                JC_HIP(hipGetLastError());
                JC_HIP(hipMemcpyAsync(&h_qn_, cnt_.p, 4, hipMemcpyDeviceToHost, stream_));
                JC_HIP(hipStreamSynchronize(stream_));
                h_.cur ^= 1;
                h_.level += 1;
                h_.qn = h_qn_;
            }
            if (h_.qn == 0) return Status::ok();
            h_.done = 0;
            // h_ is not touched again before the block kernel's state is read back
            JC_HIP(hipMemcpyAsync(st_.p, &h_, sizeof h_, hipMemcpyHostToDevice, stream_));
        }
    }

    Status read(uint32_t *out) {
        if (n_) {
            JC_HIP(hipMemcpyAsync(out, dist_.p, (size_t)n_ * 4, hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
        }
        return Status::ok();
    }

    Status timed(uint32_t src, uint32_t *out, double *gpu_ms) {
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        JC_HIP(hipEventRecord(a, stream_));
        Status s = run(nullptr, src, nullptr);
        JC_HIP(hipEventRecord(b, stream_));
        JC_HIP(hipEventSynchronize(b));
        float ms = 0;
        JC_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (s.failed()) return s;
        if (gpu_ms) *gpu_ms = ms;
        if (out) JC_TRY(read(out));
        return Status::ok();
    }

    uint32_t n() const { return n_; }
    uint64_t nnz() const { return nnz_; }
    uint32_t *dist() { return dist_.p; }
    const uint32_t *off() const { return off_.p; }
    const uint32_t *nbr() const { return nbr_.p; }
    hipStream_t stream() const { return stream_; }
    bool symmetric() const { return symmetric_; }

  private:
    uint32_t n_ = 0;
    bool symmetric_ = true;  // every entry u -> v has v -> u: bottom-up levels allowed
    uint64_t nnz_ = 0;
    BfsState h_{};
    uint32_t h_qn_ = 0;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    DevBuf<uint32_t> off_, nbr_, dist_, q_[2], cnt_;
    DevBuf<uint4> ell_;
    DevBuf<BfsState> st_;
};

// Max-min pivots on the device (sgd_stress.rs:143-175): the first pivot is
// `first`, each later one the node farthest (in hops) from those chosen so far.
// dist8[((t/8) * n + v) * 8 + t%8] = hop distance of v from pivot t as u16 (0:
// the pivot itself or unreachable); the caller zeroes the buffer.
Status select_pivots_device(const CsrView &g, uint32_t k, uint32_t first, hipStream_t stream,
                            uint32_t *pivots_dev, uint16_t *dist8_dev, std::vector<uint32_t> &h_pivots,
                            uint32_t *dmax_out) {
    h_pivots.clear();
    *dmax_out = 0;
    if (g.n == 0 || k == 0) return Status::ok();
    BfsPlan bfs;
    JC_TRY(bfs.create(g, stream));
    const uint32_t n = g.n;
    DevBuf<uint32_t> mind, small;
    DevBuf<unsigned long long> best;
    JC_TRY(mind.alloc(n));
    JC_TRY(small.alloc(3));  // src, dmax, bad
    JC_TRY(best.alloc(1));
    k_fill_u32<<<grid_for(n), kTB, 0, stream>>>(mind.p, n, kUnreached);
    k_fill_u32<<<1, kTB, 0, stream>>>(small.p, 3, 0u);
    k_fill_u32<<<1, kTB, 0, stream>>>(small.p, 1, first);
    for (uint32_t t = 0; t < k; t++) {
        JC_TRY(bfs.run(small.p, 0, best.p));
        k_pivot_select<<<grid_for(n), kTB, 0, stream>>>(n, bfs.dist(), mind.p, dist8_dev, t, best.p,
                                                        small.p + 1, small.p + 2);
        k_pivot_next<<<1, 1, 0, stream>>>(best.p, small.p, pivots_dev + t);
        JC_HIP(hipGetLastError());
    }
    uint32_t out3[3];
    h_pivots.resize(k);
    JC_HIP(hipMemcpyAsync(out3, small.p, sizeof out3, hipMemcpyDeviceToHost, stream));
    JC_HIP(hipMemcpyAsync(h_pivots.data(), pivots_dev, (size_t)k * 4, hipMemcpyDeviceToHost, stream));
    JC_HIP(hipStreamSynchronize(stream));
    if (out3[2]) return Status::err(JC_E_INVALID, "hop distance exceeds the u16 range");
    *dmax_out = out3[1];
    return Status::ok();
}

namespace {

Status components_on(BfsPlan &g, DevBuf<uint32_t> &label) {
    const uint32_t n = g.n();
    hipStream_t st = g.stream();
    JC_TRY(label.alloc(n ? n : 1));
    if (!n) return Status::ok();
    DevBuf<uint32_t> parent;
    JC_TRY(parent.alloc(n));
    k_cc_init<<<grid_for(n), kTB, 0, st>>>(n, g.off(), g.nbr(), parent.p);
    k_cc_hook<<<grid_for(n), kTB, 0, st>>>(n, g.off(), g.nbr(), parent.p);
    k_cc_label<<<grid_for(n), kTB, 0, st>>>(n, parent.p, label.p);
    JC_HIP(hipGetLastError());
    JC_HIP(hipStreamSynchronize(st));
    return Status::ok();
}

}  // namespace

Status connected_components(const CsrView &g, uint32_t *out_labels) {
    BfsPlan plan;
    JC_TRY(plan.create(g, nullptr));
    DevBuf<uint32_t> label;
    JC_TRY(components_on(plan, label));
    if (g.n) JC_TRY(copy_sync(out_labels, label.p, (size_t)g.n * 4, hipMemcpyDeviceToHost, plan.stream()));
    return Status::ok();
}

Status largest_component(const CsrView &g, uint32_t *keep_map, uint32_t *out_size, uint32_t *sub_offsets,
                         uint32_t *sub_neighbors, uint64_t *out_sub_nnz) {
    BfsPlan plan;
    JC_TRY(plan.create(g, nullptr));
    const uint32_t n = g.n;
    if (n == 0) {
        if (out_size) *out_size = 0;
        if (sub_offsets) sub_offsets[0] = 0;
        if (out_sub_nnz) *out_sub_nnz = 0;
        return Status::ok();
    }
    hipStream_t st = plan.stream();
    DevBuf<uint32_t> label, size, keep, tiles, totals, sub_off, sub_nbr;
    DevBuf<unsigned long long> best;
    JC_TRY(components_on(plan, label));
    JC_TRY(size.alloc(n));
    JC_TRY(keep.alloc(n));
    JC_TRY(totals.alloc(2));
    JC_TRY(best.alloc(1));
    k_fill_u32<<<grid_for(n), kTB, 0, st>>>(size.p, n, 0u);
    JC_HIP(hipMemsetAsync(best.p, 0, sizeof(unsigned long long), st));
    k_cc_sizes<<<grid_for(n), kTB, 0, st>>>(n, label.p, size.p);
    k_cc_best<<<grid_for(n), kTB, 0, st>>>(n, label.p, size.p, best.p);
    JC_HIP(hipGetLastError());
    JC_TRY(exclusive_scan(n, KeepFlag{label.p, best.p}, KeepWrite{keep.p}, tiles, totals.p, st));
    uint32_t n_keep = 0;
    JC_HIP(hipMemcpyAsync(&n_keep, totals.p, 4, hipMemcpyDeviceToHost, st));
    if (keep_map) JC_HIP(hipMemcpyAsync(keep_map, keep.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    JC_HIP(hipStreamSynchronize(st));
    if (out_size) *out_size = n_keep;
    if (!sub_offsets && !sub_neighbors && !out_sub_nnz) return Status::ok();
    JC_TRY(sub_off.alloc((size_t)n_keep + 1));
    JC_TRY(exclusive_scan(n, KeptDegree{plan.off(), keep.p}, SubOffset{keep.p, sub_off.p}, tiles,
                          totals.p + 1, st));
    uint32_t sub_nnz = 0;
    JC_HIP(hipMemcpyAsync(&sub_nnz, totals.p + 1, 4, hipMemcpyDeviceToHost, st));
    JC_HIP(hipStreamSynchronize(st));
    JC_HIP(hipMemcpyAsync(sub_off.p + n_keep, &sub_nnz, 4, hipMemcpyHostToDevice, st));
    if (out_sub_nnz) *out_sub_nnz = sub_nnz;
    if (sub_neighbors && sub_nnz) {
        JC_TRY(sub_nbr.alloc(sub_nnz));
        k_sub_rows<<<grid_for(n), kTB, 0, st>>>(n, plan.off(), plan.nbr(), keep.p, sub_off.p, sub_nbr.p);
        JC_HIP(hipGetLastError());
        JC_HIP(hipMemcpyAsync(sub_neighbors, sub_nbr.p, (size_t)sub_nnz * 4, hipMemcpyDeviceToHost, st));
    }
    if (sub_offsets)
        JC_HIP(hipMemcpyAsync(sub_offsets, sub_off.p, ((size_t)n_keep + 1) * 4, hipMemcpyDeviceToHost, st));
    JC_HIP(hipStreamSynchronize(st));
    return Status::ok();
}

}  // namespace jc

struct jc_bfs_plan {
    jc::BfsPlan impl;
};

namespace {
int finish(const jc::Status &s) {
    if (!s.failed()) jc::set_error("");
    return s.code;
}
}  // namespace

extern "C" {

int jc_bfs_plan_create(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                       jc_bfs_plan **out) {
    if (!out) {
        jc::set_error("NULL output handle");
        return JC_E_INVALID;
    }
    auto *p = new jc_bfs_plan();
    jc::CsrView g;
    g.n = n;
    g.offsets = offsets;
    g.neighbors = neighbors;
    g.nnz = nnz;
    jc::Status s = p->impl.create(g, nullptr);
    if (s.failed()) {
        delete p;
        return s.code;
    }
    *out = p;
    return finish(s);
}

int jc_bfs_plan_run(jc_bfs_plan *p, uint32_t source, uint32_t *out_dist, double *out_gpu_ms) {
    if (!p) {
        jc::set_error("NULL plan");
        return JC_E_INVALID;
    }
    if (p->impl.n() == 0) return finish(jc::Status::ok());
    if (source >= p->impl.n()) {
        return jc::Status::err(JC_E_INVALID, "source " + std::to_string(source) + " out of range (n = " +
                                                 std::to_string(p->impl.n()) + ")")
            .code;
    }
    return finish(p->impl.timed(source, out_dist, out_gpu_ms));
}

void jc_bfs_plan_destroy(jc_bfs_plan *p) { delete p; }

int jc_bfs(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz, uint32_t source,
           uint32_t *out_dist) {
    if (n == 0) return finish(jc::require_device(nullptr));  // empty result, as bfs.rs:82-84
    if (!out_dist) {
        jc::set_error("NULL output array");
        return JC_E_INVALID;
    }
    jc_bfs_plan *p = nullptr;
    int rc = jc_bfs_plan_create(n, offsets, neighbors, nnz, &p);
    if (rc) return rc;
    rc = jc_bfs_plan_run(p, source, out_dist, nullptr);
    jc_bfs_plan_destroy(p);
    return rc;
}

int jc_connected_components(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                            uint32_t *out_labels) {
    if (n && !out_labels) {
        jc::set_error("NULL output array");
        return JC_E_INVALID;
    }
    jc::CsrView g;
    g.n = n;
    g.offsets = offsets;
    g.neighbors = neighbors;
    g.nnz = nnz;
    return finish(jc::connected_components(g, out_labels));
}

int jc_largest_component(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                         uint32_t *keep_map, uint32_t *out_size, uint32_t *sub_offsets,
                         uint32_t *sub_neighbors, uint64_t *out_sub_nnz) {
    jc::CsrView g;
    g.n = n;
    g.offsets = offsets;
    g.neighbors = neighbors;
    g.nnz = nnz;
    return finish(jc::largest_component(g, keep_map, out_size, sub_offsets, sub_neighbors, out_sub_nnz));
}

}  // extern "C"
