// ForceAtlas2 + Barnes-Hut on gfx950: the whole iteration stays on the device.
//
// Reference: crates/graph-compute/src/engines/fa2_bh.rs (host octree, two GPU
// passes and two read-backs per step) and shaders/fa2_barnes_hut.wgsl.  Here one
// iteration is a fixed sequence of kernels on the engine's stream, captured once
// into a hipGraph:
//
//   bbox        min/max over finite coordinates (order-preserving u32 atomics)
//   keys        per body, the octant path through the reference's own f32 cell
//               chain (fa2_bh.rs:768-799, 841-861) for 21 levels -> 63-bit key
//   sort        stable LSD radix sort of (key, body)
//   count/scan  cells opened by each body from the common-prefix lengths of
//               neighbouring keys; exclusive scan -> pre-order positions
//   emit        leaves and internal cells written straight into pre-order, with
//               the skip rope (first node after a subtree) found by a galloping
//               search on the sorted keys
//   com         centre of mass level by level from the deepest cells, children
//               in octant order (fa2_bh.rs:925-967): no float atomics; the top
//               levels (depth <= 3) in one block, which also sets their ropes
//               and gathers them into the top tree the walk reads from LDS
//   force       rope walk (fa2_barnes_hut.wgsl:108-156), attraction over the
//               reference's edge-scan order (:158-181), gravity, stats; bodies
//               processed in Morton order so a wave walks one region of the tree
//   speed       exact swing/traction sums -> AdaptiveSpeed (fa2_speed.rs:48-100)
//               on one device thread
//   apply       fa2_barnes_hut.wgsl:203-226, ping-pong positions
//
// A point-region octree whose cells split only while they hold two bodies is
// unique for a point set, so this build equals the reference insertion tree
// (up to the 21-level key depth); with FMA contraction off and the reference's
// operation order every f32 value matches the restated host path bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "exact_sum.hpp"
#include "fa2_device.hpp"

namespace jc {
namespace {

constexpr uint32_t kTB = 256;             // threads per block for per-node kernels
constexpr uint32_t kScanItems = 16;       // count/scan tile = 4096 bodies
constexpr uint32_t kScanTile = kTB * kScanItems;
// lane walk, wave-union, 2 and 4 chains per lane, lane walk in 512-, 1024- and
// 64-thread blocks (the last spreads a small graph's few waves over as many CUs)
constexpr int kWalkModes = 7;
// fa2-brute-hip: one lane per body, 4 lanes per body, 16 lanes per body
constexpr int kBruteModes = 3;
constexpr uint32_t kTimedIters = 10;      // timed iterations per host read-back of events
constexpr uint32_t kHeavyAtt = 32;        // attraction rows longer than this go to k_force_heavy
constexpr uint32_t kHeavyStage = 1024;    // heavy-row terms staged in LDS per round
constexpr uint32_t kEmitTasks = 1024;     // cell tasks per k_emit block held in LDS
constexpr uint32_t kComGrid = 1024;       // k_com_fused workgroups (each strides over chunks)
constexpr uint32_t kApplyBlocks = 512;    // k_apply workgroups (each strides over bodies)
// Columns of the swing/traction limb accumulator.  Blocks add into column
// blockIdx.x % kAccCols; many columns keep the device-scope atomics on any one
// address few (they serialise at the memory side across XCDs), few columns keep
// k_speed's read of them short: 256 columns give a 1M-node force launch (977
// blocks) four blocks per address.
constexpr uint32_t kAccCols = 256;
constexpr uint32_t kSpeedThreads = 640;   // k_speed: 32 threads per limb

struct Frame {
    uint32_t bbmin[3];
    uint32_t bbmax[3];
    uint32_t n_invalid;
    uint32_t n_tree;
    float root[4];
    float speed;
    float speed_eff;
    double sum_swing;
    double sum_traction;
    uint32_t max_depth;
    uint32_t n_top;                   // nodes in the top tree (depth <= kTopDepth)
    uint32_t n_valid;                 // finite bodies of the last tree build
    unsigned long long interactions;  // accepted cells + leaf bodies summed by the walk
    uint32_t n_big;                   // cells below the top tree too large for k_com_fused
};

struct Params {
    uint32_t n;
    float gravity;
    uint32_t strong_gravity;
    float scaling_ratio;
    float jitter_tolerance;
    uint32_t lin_log_mode;
    float time_step;
    float theta2;
    float max_displacement;
    float ewi_weight;  // pow(max(1, 1e-6), edge_weight_influence): every fa2-bh edge weighs 1
    double est_jt;     // 0.05 * sqrt(n)           (fa2_speed.rs:72)
    double min_jt;     // sqrt(est_jt)             (fa2_speed.rs:73)
    uint32_t cap;      // node capacity of the octree arrays
    // scaling_ratio in [2^-10, 2^10]: with masses in [1, 2^34], degrees below
    // 2^31 and squared distances in [1e-4, 2^84] (root half-size below 2^40,
    // checked per launch), every repulsion quotient kr*deg*mass / r2 is in the
    // range of div_in_range
    uint32_t fastdiv;
};

// ---------------------------------------------------------------- kernels
// Also clears the limb accumulator the force kernels add into (every force
// launch follows a tree build, which starts with this kernel) and the top-tree
// slot table that k_emit fills.
__global__ void k_frame_reset(Frame *fr, unsigned long long *__restrict__ acc_cols,
                              uint32_t *__restrict__ top_slot) {
    for (uint32_t q = threadIdx.x; q < 2 * kLimbs * kAccCols; q += blockDim.x) acc_cols[q] = 0ull;
    for (uint32_t q = threadIdx.x; q < kTopSlots; q += blockDim.x) top_slot[q] = kTopAbsent;
    if (threadIdx.x == 0) {
        const uint32_t pinf = 0xFF800000u;  // f2ord(+inf)
        for (int k = 0; k < 3; k++) {
            fr->bbmin[k] = pinf;
            fr->bbmax[k] = ~pinf;  // f2ord(-inf)
        }
        fr->n_invalid = 0;
        fr->max_depth = 0;
        fr->interactions = 0ull;
        fr->n_big = 0;
    }
}

// Min/max over finite coordinates: per-thread, wave shuffles, then across the
// block's waves in LDS, so only one atomic per block and axis reaches L2.
__global__ __launch_bounds__(kTB) void k_bbox(const float4 *__restrict__ pos, uint32_t n,
                                              Frame *fr) {
    __shared__ float smn[3][kTB / 64], smx[3][kTB / 64];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = pos[i];
        float v[3] = {p.x + 0.0f, p.y + 0.0f, p.z + 0.0f};  // +0: canonical zero sign
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (finitef(v[k])) {
                mn[k] = fminf(mn[k], v[k]);
                mx[k] = fmaxf(mx[k], v[k]);
            }
        }
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
        }
        if (lane == 0) {
            smn[k][w] = mn[k];
            smx[k][w] = mx[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        float a = INFINITY, b = -INFINITY;
        for (uint32_t q = 0; q < kTB / 64; q++) {
            a = fminf(a, smn[k][q]);
            b = fmaxf(b, smx[k][q]);
        }
        if (finitef(a)) atomicMin(&fr->bbmin[k], f2ord(a));
        if (finitef(b)) atomicMax(&fr->bbmax[k], f2ord(b));
    }
}

// Octant path of every body through the reference cell chain: digit
// (x>=cx) | (y>=cy)<<1 | (z>=cz)<<2, child half = half*0.5, child centre =
// centre +- child half (fa2_bh.rs:841-861).  Non-finite bodies get an all-ones
// key and sort last (the reference skips them at insertion, fa2_bh.rs:810-812).
__global__ __launch_bounds__(kTB) void k_keys(const float4 *__restrict__ pos, uint32_t n,
                                              Frame *fr, uint64_t *__restrict__ keys,
                                              uint32_t *__restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float r4[4];
    compute_root(fr->bbmin, fr->bbmax, r4);
    if (i == 0) {
        for (int k = 0; k < 4; k++) fr->root[k] = r4[k];
    }
    if (i >= n) return;
    float4 p = pos[i];
    uint64_t key;
    if (finitef(p.x) && finitef(p.y) && finitef(p.z)) {
        key = octant_key(r4, p.x, p.y, p.z);
    } else {
        key = ~0ull;
        atomicAdd(&fr->n_invalid, 1u);
    }
    keys[i] = key;
    vals[i] = i;
}

// Cells opened by sorted body i: c_i = max(0, l_i - l_{i-1}) with l = common
// leading octal digits of neighbouring keys (-1 past either end); a lone body
// opens the root.  Block-local exclusive scan of c, plus the block total.
__global__ __launch_bounds__(kTB) void k_count_scan(const uint64_t *__restrict__ keys, uint32_t n,
                                                    const Frame *fr, uint32_t *__restrict__ loc,
                                                    uint32_t *__restrict__ blk_tot,
                                                    uint32_t *__restrict__ blk_depth) {
    __shared__ uint32_t ws[kTB / 64], wd[kTB / 64];
    const uint32_t nv = n - fr->n_invalid;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    // the thread's keys and one neighbour on each side, each loaded once
    uint64_t kk[kScanItems + 2];
#pragma unroll
    for (uint32_t j = 0; j < kScanItems + 2; j++) {
        const uint32_t q = base + j - 1u;  // base - 1 wraps for base 0: not loaded
        kk[j] = (base + j >= 1u && q < nv) ? keys[q] : 0ull;
    }
    uint32_t c[kScanItems];
    uint32_t s = 0;
    int md = 0;  // deepest cell: the largest common prefix of two neighbouring keys
#pragma unroll
    for (uint32_t j = 0; j < kScanItems; j++) {
        const uint32_t i = base + j;
        uint32_t cj = 0;
        if (i < nv) {
            if (nv == 1) {
                cj = 1;
            } else {
                const int lp = i > 0 ? lcp_digits(kk[j], kk[j + 1]) : -1;
                const int li = i + 1 < nv ? lcp_digits(kk[j + 1], kk[j + 2]) : -1;
                cj = li > lp ? (uint32_t)(li - lp) : 0u;
                md = li > md ? li : md;
            }
        }
        c[j] = cj;
        s += cj;
    }
    // block exclusive scan of s
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += t;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) md = max(md, __shfl_xor(md, o, 64));
    if (lane == 63) {
        ws[w] = incl;
        wd[w] = (uint32_t)md;
    }
    __syncthreads();
    uint32_t wp = 0, tot = 0, bd = 0;
#pragma unroll
    for (uint32_t q = 0; q < kTB / 64; q++) {
        if (q < w) wp += ws[q];
        tot += ws[q];
        bd = max(bd, wd[q]);
    }
    // the tile's deepest cell; k_scan_top takes the maximum (a per-block
    // device-scope atomic on one word in k_emit cost about 10 us at 1M bodies)
    if (threadIdx.x == 0) blk_depth[blockIdx.x] = bd;
    uint32_t run = incl - s + wp;
#pragma unroll
    for (uint32_t j = 0; j < kScanItems; j++) {
        if (base + j < n) loc[base + j] = run;
        run += c[j];
    }
    if (threadIdx.x == 0) blk_tot[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t *__restrict__ blk, uint32_t nb,
                                                   uint32_t n, Frame *fr,
                                                   const uint32_t *__restrict__ blk_depth) {
    __shared__ uint32_t ws[16], wd[16];
    const uint32_t per = (nb + 1023u) / 1024u;
    const uint32_t b0 = min(threadIdx.x * per, nb), b1 = min(b0 + per, nb);
    uint32_t s = 0, md = 0;
    for (uint32_t i = b0; i < b1; i++) {
        s += blk[i];
        md = max(md, blk_depth[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) md = max(md, __shfl_xor(md, o, 64));
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += t;
    }
    if (lane == 63) {
        ws[w] = incl;
        wd[w] = md;
    }
    __syncthreads();
    uint32_t wp = 0, tot = 0;
    for (uint32_t q = 0; q < 16; q++) {
        if (q < w) wp += ws[q];
        tot += ws[q];
        md = max(md, wd[q]);
    }
    uint32_t run = incl - s + wp;
    for (uint32_t i = b0; i < b1; i++) {
        uint32_t v = blk[i];
        blk[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) {
        fr->max_depth = md;
        const uint32_t nv = n - fr->n_invalid;
        fr->n_tree = nv ? tot + nv : 0u;
        fr->n_valid = nv;
    }
}

__device__ __forceinline__ uint32_t first_node_of(const uint32_t *loc, const uint32_t *blk,
                                                  uint32_t i) {
    return loc[i] + blk[i / kScanTile] + i;
}

// One internal cell: depth d, opened by sorted body i (one of the block's) at
// node index `node`.  Its half-size is the root's halved d times
// (fa2_bh.rs:855-861); its skip rope is the first node of the body after the
// last one sharing its d-digit prefix (for a cell of depth <= kTopDepth
// k_top_build sets it from the slot table).  Returns the cell's Node::b.
__device__ __forceinline__ uint4 cell_record_win(const uint64_t *__restrict__ keys, uint32_t nv,
                                                 uint32_t i, int d, uint32_t node, const Frame *fr,
                                                 const uint32_t *__restrict__ loc,
                                                 const uint32_t *__restrict__ blk,
                                                 uint32_t *__restrict__ top_slot,
                                                 const uint64_t *wk, const uint32_t *wf,
                                                 uint32_t w0, uint32_t wn) {
    float h = fr->root[3];
    for (int q = 0; q < d; q++) h = h * 0.5f;
    uint32_t skip = kEnd;
    const uint64_t ki = wk[i - w0];  // i is one of the block's bodies
    if (d <= kTopDepth) {
        top_slot[top_slot_of(ki, d)] = node;
    } else {
        const uint32_t j = cell_end_win(keys, wk, w0, wn, i, nv, ki, d);
        if (j + 1 < nv) skip = j + 1 - w0 < wn ? wf[j + 1 - w0] : first_node_of(loc, blk, j + 1);
    }
    const float s = h * 2.0f;
    const float s2 = s * s;
    return cell_b(s2, skip);
}

// Pre-order emission.  Body i (sorted) opens the cells at depths lp+1 .. li
// (lp, li: common octal digits with its neighbours), then its leaf at depth
// max(lp, li)+1.  The leaves are written by their body's thread; the cells go
// through a block-wide task list, so the prefix searches of a body that opens
// many cells run on many threads instead of one after another.  (Assembling a
// block's node range in LDS and writing it out as one span was measured: it is
// slower, the partial-line writes are not what bounds this kernel.)
__global__ __launch_bounds__(kTB) void k_emit(const uint64_t *__restrict__ keys,
                                              const uint32_t *__restrict__ perm,
                                              const float4 *__restrict__ pos,
                                              const uint32_t *__restrict__ deg, uint32_t n,
                                              Frame *fr, const uint32_t *__restrict__ loc,
                                              const uint32_t *__restrict__ blk,
                                              Node *__restrict__ nodes,
                                              uint8_t *__restrict__ node_depth, uint64_t cap,
                                              uint32_t *__restrict__ top_slot,
                                              uint32_t *__restrict__ cskip) {
    __shared__ uint32_t ntask;
    __shared__ uint32_t t_node[kEmitTasks], t_body[kEmitTasks];
    __shared__ uint8_t t_depth[kEmitTasks];
    // cell searches read the block's keys and the next kEmitWin from LDS
    // (cell_end_win): the sorted keys [w0, w0 + wn) and each one's first node
    __shared__ uint64_t wk[kTB + kEmitWin];
    __shared__ uint32_t wf[kTB + kEmitWin];
    const uint32_t nv = n - fr->n_invalid;
    const uint32_t w0 = blockIdx.x * kTB;
    const uint32_t wn = w0 < nv ? min(kTB + kEmitWin, nv - w0) : 0u;
    if (threadIdx.x == 0) {
        ntask = 0;
    }
    for (uint32_t q = threadIdx.x; q < wn; q += kTB) {
        wk[q] = keys[w0 + q];
        wf[q] = first_node_of(loc, blk, w0 + q);
    }
    __syncthreads();
    const uint32_t i = w0 + threadIdx.x;
    const uint64_t k = i < nv ? wk[threadIdx.x] : 0ull;
    const uint64_t kprev = (i > 0 && i < nv) ? (threadIdx.x > 0 ? wk[threadIdx.x - 1] : keys[i - 1]) : 0ull;
    const uint64_t knext = i + 1 < nv ? wk[threadIdx.x + 1] : 0ull;
    if (i < nv) {
        int lp, li;
        if (nv == 1) {
            lp = -1;
            li = 0;  // the lone body hangs below an internal root
        } else {
            lp = i > 0 ? lcp_digits(kprev, k) : -1;
            li = i + 1 < nv ? lcp_digits(k, knext) : -1;
        }
        const uint32_t c = li > lp ? (uint32_t)(li - lp) : 0u;
        const int a0 = lp + 1;
        const uint32_t pos0 = wf[threadIdx.x];
        if ((uint64_t)pos0 + c < cap) {  // always, for a consistent scan
            if (c) {
                const uint32_t slot = atomicAdd(&ntask, c);
                for (uint32_t t = 0; t < c; t++) {
                    const int d = a0 + (int)t;
                    if (slot + t < kEmitTasks) {
                        t_node[slot + t] = pos0 + t;
                        t_body[slot + t] = i;
                        t_depth[slot + t] = (uint8_t)d;
                    } else {
                        const uint4 rec = cell_record_win(keys, nv, i, d, pos0 + t, fr, loc, blk, top_slot,
                                                          wk, wf, w0, wn);
                        nodes[pos0 + t].b = rec;
                        cskip[pos0 + t] = rec.w;
                        node_depth[pos0 + t] = (uint8_t)d;
                    }
                }
            }
            const uint32_t leaf = pos0 + c;
            const uint32_t b = perm[i];
            if (b < n) {
                const float4 p = pos[b];  // .w: the body's mass max(deg + 1, 1e-3)
                nodes[leaf].a = p;
                nodes[leaf].b = leaf_b(b, (i + 1 < nv) ? leaf + 1 : kEnd);
                const int ld = (lp > li ? lp : li) + 1;
                node_depth[leaf] = (uint8_t)(0x80 | ld);
                if (ld <= kTopDepth) top_slot[top_slot_of(k, ld)] = leaf;
            }
        }
    }
    __syncthreads();
    const uint32_t nt = min(ntask, kEmitTasks);
    for (uint32_t q = threadIdx.x; q < nt; q += blockDim.x) {
        const uint4 rec = cell_record_win(keys, nv, t_body[q], t_depth[q], t_node[q], fr, loc, blk,
                                          top_slot, wk, wf, w0, wn);
        nodes[t_node[q]].b = rec;
        cskip[t_node[q]] = rec.w;
        node_depth[t_node[q]] = t_depth[q];
    }
}

// Centre of mass of the cells of depth >= kFuseDepth in one launch
// (com_fused_pass: each fitting subtree summed in LDS by one workgroup), the
// cells too large for it in one more (com_big_pass); the few wide cells above
// kFuseDepth take one k_com_level launch per depth, the top tree k_top_build.
constexpr uint32_t kFuseDepth = 6;
constexpr uint32_t kFuseChunk = 1024;
constexpr uint32_t kFuseCap = 512;
constexpr uint32_t kFuseThreads = 512;

// Graphs of at most kComSmallMax bodies fuse from just below the top tree
// (Depth = kTopDepth + 1): their cells there hold few nodes, and the two
// k_com_level launches of the larger graphs go away.  Larger graphs put too
// many depth-4 and depth-5 cells past the fused pass's capacity, onto the
// one-workgroup pass for such cells.
constexpr uint32_t kComSmallMax = 262144;
template <uint32_t Depth>
__global__ __launch_bounds__(kFuseThreads) void k_com_fused(Frame *fr, uint32_t cap,
                                                            Node *__restrict__ nodes,
                                                            const uint8_t *__restrict__ node_depth,
                                                            const uint32_t *__restrict__ cskip,
                                                            uint32_t *__restrict__ big, uint32_t big_max) {
    com_fused_pass<Depth, kFuseChunk, kFuseCap, kFuseThreads>(min(fr->n_tree, cap), nodes, node_depth,
                                                              cskip, big, big_max, &fr->n_big);
}

template <uint32_t Depth>
__global__ __launch_bounds__(1024) void k_com_big(const Frame *fr, uint32_t cap, Node *__restrict__ nodes,
                                                  const uint8_t *__restrict__ node_depth,
                                                  const uint32_t *__restrict__ big, uint32_t big_max) {
    const uint32_t nbig = min(fr->n_big, big_max);
    if (!nbig) return;
    com_big_pass<Depth>(min(fr->n_tree, cap), fr->max_depth, nodes, node_depth, big, nbig);
}

// The cells of depth d < kFuseDepth, one launch per depth (com_level_pass,
// shared with the batched engine): few cells with wide subtrees.
__global__ __launch_bounds__(kTB) void k_com_level(const Frame *fr, uint32_t d, uint32_t cap,
                                                   Node *__restrict__ nodes,
                                                   const uint8_t *__restrict__ node_depth) {
    if (d > fr->max_depth) return;
    __shared__ uint32_t list[kTB / 64][kComSpan];
    com_level_pass<kTB>(nodes, node_depth, min(fr->n_tree, cap), d, list);
}

// The top of the tree (the nodes of depth <= kTopDepth), one block, once the
// centres of mass of the deeper cells are final:
//  1. an exclusive scan over slot presence gives each top node its pre-order
//     rank; a node's subtree ends at the first present slot after its complete
//     subtree (rank ts);
//  2. the skip ropes of the top cells, which k_emit leaves to this kernel: the
//     node of rank ts, kEnd past the last (the cells' long prefix searches are
//     the longest dependent chains of k_emit);
//  3. the centres of mass of the top cells, deepest first, children in octant
//     order, m = sum m_c, com = sum com_c * m_c / m (fa2_bh.rs:944-965): a
//     frontier cell's children are read from the global array along their
//     ropes, a shallower cell's children are the top nodes u+1, ts(u+1), ...;
//  4. the top tree for the force kernels: centre of mass and mass; size, id,
//     subtree end rank | frontier flag, byte offset.
__global__ __launch_bounds__(1024) void k_top_build(Frame *fr, const uint32_t *__restrict__ top_slot,
                                                    Node *__restrict__ nodes, uint32_t cap,
                                                    float4 *__restrict__ top_a,
                                                    uint4 *__restrict__ top_b) {
    static_assert(kTopSlots < 1024, "one thread per slot");
    __shared__ uint32_t rk[kTopSlots + 1];
    __shared__ uint32_t ws[16];
    __shared__ uint32_t gof[kTopSlots];   // global node of each rank
    __shared__ uint32_t tsk[kTopSlots];   // subtree end rank of each rank
    __shared__ uint8_t dep[kTopSlots];
    __shared__ float4 A[kTopSlots];       // centre of mass and mass of each rank
    const uint32_t s = threadIdx.x, lane = s & 63, w = s >> 6;
    const uint32_t g = s < kTopSlots ? top_slot[s] : kTopAbsent;
    const uint32_t e = g != kTopAbsent ? 1u : 0u;
    uint32_t incl = e;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += t;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t wp = 0;
    for (uint32_t q = 0; q < w; q++) wp += ws[q];
    if (s <= kTopSlots) rk[s] = wp + incl - e;
    __syncthreads();
    if (e) {
        const uint32_t t = rk[s];
        const int d = top_slot_depth(s);
        gof[t] = g;
        tsk[t] = rk[s + top_sub(d)];
        dep[t] = (uint8_t)d;
    }
    __syncthreads();
    const uint32_t K = rk[kTopSlots];
    const uint32_t N = min(fr->n_tree, cap);
    const uint32_t u = threadIdx.x;
    uint4 nb = make_uint4(0u, 0u, 0u, 0u);
    uint32_t skip = kEnd;
    if (u < K) {
        const Node nd = nodes[gof[u]];
        nb = nd.b;
        A[u] = nd.a;  // final for leaves; cells are recomputed below
        if (nb.z == kCellId) {
            skip = tsk[u] < K ? gof[tsk[u]] : kEnd;
            nodes[gof[u]].b = make_uint4(skip_bytes(skip), nb.y, nb.z, skip);
        }
    }
    __syncthreads();
    for (int lvl = kTopDepth; lvl >= 0; lvl--) {
        if (u < K && dep[u] == lvl && nb.z == kCellId) {
            float tm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
            if (lvl == kTopDepth) {
                for (uint32_t ch = gof[u] + 1; ch < N && ch != skip; ch = nodes[ch].b.w) {
                    const float4 a = nodes[ch].a;
                    tm = tm + a.w;
                    cx = cx + a.x * a.w;
                    cy = cy + a.y * a.w;
                    cz = cz + a.z * a.w;
                }
            } else {
                for (uint32_t c = u + 1; c < K && c != tsk[u]; c = tsk[c]) {
                    const float4 a = A[c];
                    tm = tm + a.w;
                    cx = cx + a.x * a.w;
                    cy = cy + a.y * a.w;
                    cz = cz + a.z * a.w;
                }
            }
            if (tm > 0.0f) {
                cx = cx / tm;
                cy = cy / tm;
                cz = cz / tm;
            }
            const float4 r = make_float4(cx, cy, cz, tm);
            A[u] = r;
            nodes[gof[u]].a = r;
        }
        __syncthreads();
    }
    if (u < K) {
        const bool frontier = dep[u] == kTopDepth && nb.z == kCellId;
        top_a[u] = A[u];
        top_b[u] = make_uint4(nb.y, nb.z, tsk[u] | (frontier ? kTopFrontier : 0u), gof[u] << 5);
    }
    if (u == 0) fr->n_top = K;
}

// Attraction, gravity and the swing/traction stats of body b, after its walk
// (fa2_barnes_hut.wgsl:158-197), accumulated exactly into the block's limbs.
// One attraction term d * att (fa2_barnes_hut.wgsl:158-181), d = q - p.
__device__ __forceinline__ void attraction_term(const Params &P, float4 p, float4 q, float &tx,
                                                float &ty, float &tz) {
    const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    float att = P.ewi_weight;
    if (P.lin_log_mode) {
        const float r = fmaxf(len3(dx, dy, dz), 1.0e-3f);
        att = P.ewi_weight * logf(1.0f + r) / r;
    }
    tx = dx * att;
    ty = dy * att;
    tz = dz * att;
}

// Gravity (fa2_barnes_hut.wgsl:183-189), the force write and the swing /
// traction stats (:191-197) into the exact accumulator.
__device__ __forceinline__ void gravity_and_stats(const Params &P, uint32_t b, float4 p, float deg_i,
                                                  float fx, float fy, float fz,
                                                  const float4 *__restrict__ old_force,
                                                  float4 *__restrict__ force,
                                                  unsigned long long *acc) {
    const float gd = P.gravity * deg_i;
    if (P.strong_gravity) {
        fx = fx - p.x * gd;
        fy = fy - p.y * gd;
        fz = fz - p.z * gd;
    } else {
        const float r0 = fmaxf(len3(p.x, p.y, p.z), 1.0e-3f);
        fx = fx - (p.x / r0) * gd;
        fy = fy - (p.y / r0) * gd;
        fz = fz - (p.z / r0) * gd;
    }
    const float4 o = old_force[b];
    force[b] = make_float4(fx, fy, fz, 0.0f);
    const float swing = deg_i * len3(o.x - fx, o.y - fy, o.z - fz);
    const float tra = deg_i * 0.5f * len3(o.x + fx, o.y + fy, o.z + fz);
    if (finitef(swing) && finitef(tra)) {
        exact_acc_add(acc, swing);
        exact_acc_add(acc + kLimbs, tra);
    }
}

__device__ __forceinline__ void finish_body(const Params &P, uint32_t b, float4 p, float deg_i,
                                            float fx, float fy, float fz,
                                            const float4 *__restrict__ pos,
                                            const float4 *__restrict__ old_force,
                                            float4 *__restrict__ force,
                                            const uint32_t *__restrict__ att_off,
                                            const uint32_t *__restrict__ att_adj,
                                            unsigned long long *acc) {
        // attraction: the other endpoint of every edge touching b, in the
        // reference's edge-list order (fa2_barnes_hut.wgsl:158-181).  Rows
        // longer than kHeavyAtt leave the repulsion sum in force[b] and are
        // finished by k_force_heavy, one wave per row.
        const uint32_t e0 = att_off[b], e1 = att_off[b + 1];
        if (e1 - e0 > kHeavyAtt) {
            force[b] = make_float4(fx, fy, fz, 0.0f);
            return;
        }
        uint32_t e = e0;
        // four neighbours per round: the index loads, then the position
        // gathers, are issued back to back; the sum stays in list order
        for (; e + 4 <= e1; e += 4) {
            const uint32_t j0 = att_adj[e], j1 = att_adj[e + 1], j2 = att_adj[e + 2], j3 = att_adj[e + 3];
            const float4 q[4] = {pos[j0], pos[j1], pos[j2], pos[j3]};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                float tx, ty, tz;
                attraction_term(P, p, q[u], tx, ty, tz);
                fx = fx + tx;
                fy = fy + ty;
                fz = fz + tz;
            }
        }
        for (; e < e1; e++) {
            float tx, ty, tz;
            attraction_term(P, p, pos[att_adj[e]], tx, ty, tz);
            fx = fx + tx;
            fy = fy + ty;
            fz = fz + tz;
        }
        gravity_and_stats(P, b, p, deg_i, fx, fy, fz, old_force, force, acc);
    }

// Whether this launch may use the reduced walk and the in-range division
// (uniform over the launch: the tree size and root come from the frame, the
// scaling ratio from the host's range check in Params::fastdiv).
struct WalkForm {
    bool reduced;
    bool fastdiv;
};
__device__ __forceinline__ WalkForm walk_form(const Params &P, const Frame *fr, uint32_t N) {
    WalkForm w;
    w.reduced = tree_reducible(N, fr->root[3], P.theta2);
    w.fastdiv = w.reduced && P.fastdiv != 0u;
    return w;
}

// Interactions of the walks over the current tree (accepted cells plus leaf
// bodies, summed over all bodies) into fr->interactions.  A measurement hook:
// the force kernels do not count.
__global__ __launch_bounds__(kTB) void k_count_interactions(const Params P, Frame *fr,
                                                            const float4 *__restrict__ pos,
                                                            const uint32_t *__restrict__ deg,
                                                            const uint32_t *__restrict__ perm,
                                                            const Node *__restrict__ nodes) {
    __shared__ unsigned long long inter;
    if (threadIdx.x == 0) inter = 0ull;
    __syncthreads();
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const uint32_t b = t < P.n ? perm[t] : P.n;
    const uint32_t N = min(fr->n_tree, P.cap);
    uint32_t taken = 0;
    if (b < P.n && N > 0) {
        const float4 p = pos[b];
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        taken = walk_exact(nodes, 0u, N, p.x, p.y, p.z, b, P.scaling_ratio * ((float)deg[b] + 1.0f),
                           P.theta2, fx, fy, fz);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) taken += __shfl_xor(taken, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&inter, (unsigned long long)taken);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&fr->interactions, inter);
}

// MODE 0: every lane follows its own rope, the top tree from LDS.
// MODE 1: wave-union walk.  The wave steps through the union of its lanes'
// walks in pre-order: the current node v is wave-uniform (one scalar load),
// lanes whose own walk is at v process it, the others wait.  Each lane still
// sees exactly its own node sequence in its own order, so the sums are
// bit-identical to MODE 0; the next v is v+1 when some lane descended, else
// the minimum over the lanes' rope pointers.
template <int MODE, int BS = kTB>
__global__ __launch_bounds__(BS) void k_force(const Params P, Frame *fr,
                                               const float4 *__restrict__ pos,
                                               const float4 *__restrict__ old_force,
                                               float4 *__restrict__ force,
                                               const uint32_t *__restrict__ deg,
                                               const uint32_t *__restrict__ att_off,
                                               const uint32_t *__restrict__ att_adj,
                                               const uint32_t *__restrict__ perm,
                                               const Node *__restrict__ nodes,
                                               unsigned long long *__restrict__ partial,
                                               const float4 *__restrict__ top_a,
                                               const uint4 *__restrict__ top_b) {
    __shared__ unsigned long long acc[2 * kLimbs];
    if (threadIdx.x < 2 * kLimbs) acc[threadIdx.x] = 0ull;
    __syncthreads();
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const uint32_t b = t < P.n ? perm[t] : P.n;
    const bool live = b < P.n;
    const float4 p = live ? pos[b] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float deg_i = live ? (float)deg[b] + 1.0f : 1.0f;
    const float krd = P.scaling_ratio * deg_i;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const uint32_t N = min(fr->n_tree, P.cap);
    if constexpr (MODE == 0) {
        // the top tree goes to LDS; each lane then walks it and the global array
        __shared__ float4 tA[kTopSlots];
        __shared__ uint4 tB[kTopSlots];
        const WalkForm w = walk_form(P, fr, N);
        const uint32_t K = w.reduced ? min(fr->n_top, kTopSlots) : 0u;
        for (uint32_t q = threadIdx.x; q < K; q += blockDim.x) {
            tA[q] = top_a[q];
            tB[q] = top_b[q];
        }
        __syncthreads();
        if (live && N > 0) {
            if (w.reduced && finite3(p)) {
                if (w.fastdiv)
                    walk_top<true>(tA, tB, K, nodes, N, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
                else
                    walk_top<false>(tA, tB, K, nodes, N, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
            } else {
                (void)walk_exact(nodes, 0u, N, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
            }
        }
    } else {
        const uint64_t wcap = max((uint64_t)N * 4ull, 16ull);
        uint32_t ptr = (live && N > 0) ? 0u : kEnd;
        uint64_t walk = 0;
        uint32_t v = __builtin_amdgcn_readfirstlane(__reduce_min_sync(~0ull, ptr));
        while (v != kEnd) {
            const Node nd = nodes[v];
            if (ptr == v) {
                walk++;
                uint32_t nxt;
                (void)visit_exact(nd, v, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz, nxt);
                if (nxt >= N || walk >= wcap) nxt = kEnd;
                ptr = nxt;
            }
            if (__any(ptr == v + 1u)) v = v + 1u;
            else v = __builtin_amdgcn_readfirstlane(__reduce_min_sync(~0ull, ptr));
        }
    }
    if (live) finish_body(P, b, p, deg_i, fx, fy, fz, pos, old_force, force, att_off, att_adj, acc);
    __syncthreads();
    // the block's limbs go into one of kAccCols columns of the limb-major
    // accumulator; integer adds, so the total is exact in any order
    if (threadIdx.x < 2 * kLimbs && acc[threadIdx.x])
        atomicAdd(&partial[threadIdx.x * kAccCols + blockIdx.x % kAccCols], acc[threadIdx.x]);
}

// K bodies per thread (K consecutive bodies in Morton order), their walks
// interleaved step by step: each step issues the K node loads back to back, so
// K independent pointer chains are in flight per lane instead of one.  Every
// body still performs exactly its own walk in its own order, in the reduced
// form when the launch and all K bodies allow it.
template <int K>
__global__ __launch_bounds__(kTB) void k_force_multi(const Params P, Frame *fr,
                                                     const float4 *__restrict__ pos,
                                                     const float4 *__restrict__ old_force,
                                                     float4 *__restrict__ force,
                                                     const uint32_t *__restrict__ deg,
                                                     const uint32_t *__restrict__ att_off,
                                                     const uint32_t *__restrict__ att_adj,
                                                     const uint32_t *__restrict__ perm,
                                                     const Node *__restrict__ nodes,
                                                     unsigned long long *__restrict__ partial) {
    __shared__ unsigned long long acc[2 * kLimbs];
    if (threadIdx.x < 2 * kLimbs) acc[threadIdx.x] = 0ull;
    __syncthreads();
    const uint32_t t0 = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) * K;
    const uint32_t N = min(fr->n_tree, P.cap);
    uint32_t b[K], idx[K];
    float4 p[K];
    float deg_i[K], krd[K], fx[K], fy[K], fz[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint32_t t = t0 + k;
        b[k] = t < P.n ? perm[t] : P.n;
        const bool live = b[k] < P.n;
        p[k] = live ? pos[b[k]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        deg_i[k] = live ? (float)deg[b[k]] + 1.0f : 1.0f;
        krd[k] = P.scaling_ratio * deg_i[k];
        fx[k] = fy[k] = fz[k] = 0.0f;
        idx[k] = (live && N > 0) ? 0u : kEnd;
    }
    const WalkForm w = walk_form(P, fr, N);
    bool reduced = w.reduced, live[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        live[k] = idx[k] < N;
        reduced = reduced && finite3(p[k]);
    }
    if (reduced) {
        if (w.fastdiv) walk_reduced_multi<K, true>(nodes, N, live, p, b, krd, P.theta2, fx, fy, fz);
        else walk_reduced_multi<K, false>(nodes, N, live, p, b, krd, P.theta2, fx, fy, fz);
    } else {
        // ropes only move forward: at most N steps per walk
        for (uint32_t steps = 0; steps < N; steps++) {
            bool active = false;
#pragma unroll
            for (int k = 0; k < K; k++) active |= idx[k] < N;
            if (!active) break;
            Node nd[K];
#pragma unroll
            for (int k = 0; k < K; k++) nd[k] = nodes[idx[k] < N ? idx[k] : 0u];
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (idx[k] < N) {
                    uint32_t nxt;
                    (void)visit_exact(nd[k], idx[k], p[k].x, p[k].y, p[k].z, b[k], krd[k],
                                      P.theta2, fx[k], fy[k], fz[k], nxt);
                    idx[k] = nxt;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        if (b[k] < P.n)
            finish_body(P, b[k], p[k], deg_i[k], fx[k], fy[k], fz[k], pos, old_force, force, att_off,
                        att_adj, acc);
    __syncthreads();
    if (threadIdx.x < 2 * kLimbs && acc[threadIdx.x])
        atomicAdd(&partial[threadIdx.x * kAccCols + blockIdx.x % kAccCols], acc[threadIdx.x]);
}

// ---- brute-force repulsion (the fa2-brute engine) ----

// Coordinates inside (-2^39, 2^39): every squared distance between two such
// bodies is below 2^82, which with kr in kr_in_fastdiv_range keeps
// krd * mass / max(r2, 1e-4) inside div_in_range's range.
constexpr float kBruteBound = 549755813888.0f;  // 2^39
__device__ __forceinline__ bool brute_bounded(float4 q) {
    return fabsf(q.x) < kBruteBound && fabsf(q.y) < kBruteBound && fabsf(q.z) < kBruteBound;
}

// One repulsion term of force_atlas2.wgsl:82-91, d * cf with d = p - q and
// cf = krd * deg_j / max(|d|^2, 1e-4), deg_j = q.w.
template <bool FASTDIV>
__device__ __forceinline__ void brute_term_of(float4 p, float4 q, float krd, float &tx, float &ty,
                                              float &tz) {
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float r2 = fmaxf(dx * dx + dy * dy + dz * dz, 1.0e-4f);
    const float cf = FASTDIV ? div_in_range(krd * q.w, r2) : krd * q.w / r2;
    tx = dx * cf;
    ty = dy * cf;
    tz = dz * cf;
}
template <bool FASTDIV>
__device__ __forceinline__ void brute_term(float4 p, float4 q, float krd, float &fx, float &fy,
                                           float &fz) {
    float tx, ty, tz;
    brute_term_of<FASTDIV>(p, q, krd, tx, ty, tz);
    fx = fx + tx;
    fy = fy + ty;
    fz = fz + tz;
}

// Lane K of a group of L lanes, in every lane of the group: DPP quad_perm
// [K, K, K, K] for L = 4, row_newbcast:K (one row of 16 lanes) for L = 16.
template <int L, int K>
__device__ __forceinline__ float grp_bcast(float v) {
    static_assert(L == 4 || L == 16, "quads or rows");
    constexpr int ctrl = L == 4 ? K * 0x55 : 0x150 + K;
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, true));
}
// s += t of group lane 0, then 1, ..., L-1, per component: the group's L
// terms in lane order.
template <int L, int K = 0>
__device__ __forceinline__ void grp_add(float &fx, float &fy, float &fz, float tx, float ty,
                                        float tz) {
    fx = fx + grp_bcast<L, K>(tx);
    fy = fy + grp_bcast<L, K>(ty);
    fz = fz + grp_bcast<L, K>(tz);
    if constexpr (K + 1 < L) grp_add<L, K + 1>(fx, fy, fz, tx, ty, tz);
}

// Body j of a tile (w = deg + 1), or a finite pad past the last body.
__device__ __forceinline__ float4 brute_tile_load(const float4 *__restrict__ pos, uint32_t n,
                                                  uint32_t j) {
    return j < n ? pos[j] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

// Repulsion of the brute-force engine (force_atlas2.wgsl:69-91): body b sums
// every other body j in index order, so the f32 sum is the reference's.
// Positions (w = deg + 1, the reference's deg_j) pass through an LDS tile of
// BS bodies that every lane reads in order (broadcast reads).  Fast form: when
// the tile, the body and the scaling ratio keep every quotient in range
// (brute_bounded, Params::fastdiv) the division is div_in_range and the self
// term is not skipped: for a finite body it is (+0) * cf with cf finite, and
// adding +0 leaves a sum that started at +0 unchanged.  Otherwise each term is
// the IEEE quotient and j == b is skipped, as in the reference.
template <int BS>
__global__ __launch_bounds__(BS) void k_force_brute(const Params P,
                                                     const float4 *__restrict__ pos,
                                                     const float4 *__restrict__ old_force,
                                                     float4 *__restrict__ force,
                                                     const uint32_t *__restrict__ deg,
                                                     const uint32_t *__restrict__ att_off,
                                                     const uint32_t *__restrict__ att_adj,
                                                     unsigned long long *__restrict__ partial) {
    __shared__ unsigned long long acc[2 * kLimbs];
    __shared__ float4 tile[BS];
    if (threadIdx.x < 2 * kLimbs) acc[threadIdx.x] = 0ull;
    const uint32_t b = blockIdx.x * BS + threadIdx.x;
    const bool live = b < P.n;
    const float4 p = live ? pos[b] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const float deg_i = live ? (float)deg[b] + 1.0f : 1.0f;
    const float krd = P.scaling_ratio * deg_i;
    const bool own_fast = P.fastdiv != 0u && brute_bounded(p);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    float4 q = brute_tile_load(pos, P.n, threadIdx.x);
    for (uint32_t j0 = 0; j0 < P.n; j0 += BS) {
        __syncthreads();  // the previous tile has been read
        tile[threadIdx.x] = q;
        const bool tile_fast = __syncthreads_and(brute_bounded(q)) != 0;
        q = brute_tile_load(pos, P.n, j0 + BS + threadIdx.x);  // the next tile, in flight meanwhile
        const uint32_t m = min((uint32_t)BS, P.n - j0);
        if (tile_fast && own_fast) {
            uint32_t k = 0;
            for (; k + 4 <= m; k += 4) {
                const float4 q0 = tile[k], q1 = tile[k + 1], q2 = tile[k + 2], q3 = tile[k + 3];
                brute_term<true>(p, q0, krd, fx, fy, fz);
                brute_term<true>(p, q1, krd, fx, fy, fz);
                brute_term<true>(p, q2, krd, fx, fy, fz);
                brute_term<true>(p, q3, krd, fx, fy, fz);
            }
            for (; k < m; k++) brute_term<true>(p, tile[k], krd, fx, fy, fz);
        } else {
            for (uint32_t k = 0; k < m; k++)
                if (j0 + k != b) brute_term<false>(p, tile[k], krd, fx, fy, fz);
        }
    }
    if (live) finish_body(P, b, p, deg_i, fx, fy, fz, pos, old_force, force, att_off, att_adj, acc);
    __syncthreads();
    if (threadIdx.x < 2 * kLimbs && acc[threadIdx.x])
        atomicAdd(&partial[threadIdx.x * kAccCols + blockIdx.x % kAccCols], acc[threadIdx.x]);
}

// The same sums with L = 4 or 16 lanes per body, for graphs too small to fill
// the chip one lane per body: in each step over bodies j0+k .. j0+k+L-1, lane r
// of the group computes the term of body j0+k+r, and all L lanes add the L
// terms to their (identical) running sums in index order through group
// broadcasts -- the f32 additions of the one-lane loop, in its order.  A term
// that is not there (past the tile, or j == b in the IEEE form) is +0, and
// adding +0 leaves a sum that started at +0 unchanged.  Lane 0 of the group
// then finishes the body.  Groups are whole within a wave and every branch
// around a broadcast is uniform over a group (it depends on b, the tile or the
// launch), so the broadcasts read active lanes.
template <int BS, int L>
__global__ __launch_bounds__(BS) void k_force_brute_grp(const Params P,
                                                         const float4 *__restrict__ pos,
                                                         const float4 *__restrict__ old_force,
                                                         float4 *__restrict__ force,
                                                         const uint32_t *__restrict__ deg,
                                                         const uint32_t *__restrict__ att_off,
                                                         const uint32_t *__restrict__ att_adj,
                                                         unsigned long long *__restrict__ partial) {
    static_assert(BS % 64 == 0 && 64 % L == 0, "whole waves, whole groups");
    __shared__ unsigned long long acc[2 * kLimbs];
    __shared__ float4 tile[BS];
    if (threadIdx.x < 2 * kLimbs) acc[threadIdx.x] = 0ull;
    const uint32_t r = threadIdx.x % L;
    const uint32_t b = blockIdx.x * (BS / L) + threadIdx.x / L;
    const bool live = b < P.n;
    const float4 p = live ? pos[b] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const float deg_i = live ? (float)deg[b] + 1.0f : 1.0f;
    const float krd = P.scaling_ratio * deg_i;
    const bool own_fast = P.fastdiv != 0u && brute_bounded(p);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    float4 q = brute_tile_load(pos, P.n, threadIdx.x);
    for (uint32_t j0 = 0; j0 < P.n; j0 += BS) {
        __syncthreads();  // the previous tile has been read
        tile[threadIdx.x] = q;
        const bool fast = (__syncthreads_and(brute_bounded(q)) != 0) && own_fast;
        q = brute_tile_load(pos, P.n, j0 + BS + threadIdx.x);  // the next tile, in flight meanwhile
        const uint32_t m = min((uint32_t)BS, P.n - j0);
        // k + r < BS below: k < m <= BS, and k and BS are multiples of L
        if (fast) {
            uint32_t k = 0;
            for (; k + 2 * L <= m; k += 2 * L) {
                const float4 qa = tile[k + r], qb = tile[k + L + r];
                float ax, ay, az, bx, by, bz;
                brute_term_of<true>(p, qa, krd, ax, ay, az);
                brute_term_of<true>(p, qb, krd, bx, by, bz);
                grp_add<L>(fx, fy, fz, ax, ay, az);
                grp_add<L>(fx, fy, fz, bx, by, bz);
            }
            for (; k < m; k += L) {
                float tx, ty, tz;
                brute_term_of<true>(p, tile[k + r], krd, tx, ty, tz);
                if (k + r >= m) tx = ty = tz = 0.0f;
                grp_add<L>(fx, fy, fz, tx, ty, tz);
            }
        } else {
            for (uint32_t k = 0; k < m; k += L) {
                float tx = 0.0f, ty = 0.0f, tz = 0.0f;
                if (k + r < m && j0 + k + r != b)
                    brute_term_of<false>(p, tile[k + r], krd, tx, ty, tz);
                grp_add<L>(fx, fy, fz, tx, ty, tz);
            }
        }
    }
    if (live && r == 0)
        finish_body(P, b, p, deg_i, fx, fy, fz, pos, old_force, force, att_off, att_adj, acc);
    __syncthreads();
    if (threadIdx.x < 2 * kLimbs && acc[threadIdx.x])
        atomicAdd(&partial[threadIdx.x * kAccCols + blockIdx.x % kAccCols], acc[threadIdx.x]);
}

// Attraction rows longer than kHeavyAtt.  Their terms d*att are independent
// of the walk, so k_heavy_terms computes all of them before the force kernel
// (one thread per entry, every gather in flight at once) into hterm, three
// component arrays over the heavy rows' entries laid end to end (row r owns
// [hoff[r], hoff[r] + hlen[r]), hoff[r] a multiple of 4).  After the force kernel, k_force_heavy adds them to
// the repulsion sum the walk left in force[b]: lanes 0, 1, 2 of one wave per
// row run the x, y, z sums in list order -- the same f32 additions in the same
// order as the per-lane loop -- and lane 0 finishes the body.
__global__ __launch_bounds__(kTB) void k_heavy_terms(const Params P, uint32_t n_ent,
                                                     const uint32_t *__restrict__ hbody,
                                                     const uint32_t *__restrict__ hadj,
                                                     const float4 *__restrict__ pos,
                                                     float *__restrict__ hterm, uint32_t stride) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_ent) return;
    float tx, ty, tz;
    attraction_term(P, pos[hbody[q]], pos[hadj[q]], tx, ty, tz);
    hterm[q] = tx;
    hterm[stride + q] = ty;
    hterm[2 * (size_t)stride + q] = tz;
}

// Per row: the wave copies up to kHeavyStage terms of each component into LDS
// with coalesced 16-byte loads (rows start on 16-byte boundaries in hterm),
// then lanes 0, 1, 2 add them from LDS.  Swing / traction limbs are added
// (integer atomics, so the sum stays exact and order-free) into accumulator
// column blockIdx.x % kAccCols.
__global__ __launch_bounds__(64) void k_force_heavy(const Params P,
                                                   const uint32_t *__restrict__ heavy,
                                                   const uint32_t *__restrict__ hoff,
                                                   const uint32_t *__restrict__ hlen,
                                                   const float *__restrict__ hterm,
                                                   uint32_t stride,
                                                   const float4 *__restrict__ pos,
                                                   const float4 *__restrict__ old_force,
                                                   float4 *__restrict__ force,
                                                   const uint32_t *__restrict__ deg,
                                                   unsigned long long *__restrict__ partial) {
    __shared__ __align__(16) float st[3][kHeavyStage];
    __shared__ unsigned long long acc[2 * kLimbs];
    const uint32_t lane = threadIdx.x;
    const uint32_t b = heavy[blockIdx.x];
    if (lane < 2 * kLimbs) acc[lane] = 0ull;
    const float4 f0 = force[b];
    float run = lane == 0 ? f0.x : lane == 1 ? f0.y : f0.z;
    const uint32_t e0 = hoff[blockIdx.x], len = hlen[blockIdx.x];
    // the next chunk's terms are loaded into registers while lanes 0-2 add the
    // current one (padded entries are loaded, never added)
    constexpr uint32_t V = kHeavyStage / 4 / 64;  // float4 per lane and component
    typedef float v4f __attribute__((ext_vector_type(4)));  // native vectors stay in registers
    v4f nx0[V], nx1[V], nx2[V];
    {
        // chunk 0; the index is clamped to the chunk, so every load is in bounds
        const uint32_t fnv4 = len ? (min(kHeavyStage, len) + 3u) >> 2 : 1u;
        const v4f *s0 = reinterpret_cast<const v4f *>(hterm + e0);
        const v4f *s1 = reinterpret_cast<const v4f *>(hterm + (size_t)stride + e0);
        const v4f *s2 = reinterpret_cast<const v4f *>(hterm + 2 * (size_t)stride + e0);
#pragma unroll
        for (uint32_t k = 0; k < V; k++) {
            const uint32_t v = min(lane + 64 * k, fnv4 - 1u);
            if (len) {
                nx0[k] = s0[v];
                nx1[k] = s1[v];
                nx2[k] = s2[v];
            }
        }
    }
    for (uint32_t c0 = 0; c0 < len; c0 += kHeavyStage) {
        const uint32_t cnt = min(kHeavyStage, len - c0);
        const uint32_t nv4 = (cnt + 3u) >> 2;
#pragma unroll
        for (uint32_t k = 0; k < V; k++) {
            if (lane + 64 * k < nv4) {
                reinterpret_cast<v4f *>(st[0])[lane + 64 * k] = nx0[k];
                reinterpret_cast<v4f *>(st[1])[lane + 64 * k] = nx1[k];
                reinterpret_cast<v4f *>(st[2])[lane + 64 * k] = nx2[k];
            }
        }
        __syncthreads();
        const uint32_t c1 = c0 + kHeavyStage;
        if (c1 < len) {  // the next chunk, in flight while lanes 0-2 add this one
            const uint32_t fnv4 = (min(kHeavyStage, len - c1) + 3u) >> 2;
            const v4f *s0 = reinterpret_cast<const v4f *>(hterm + e0 + c1);
            const v4f *s1 = reinterpret_cast<const v4f *>(hterm + (size_t)stride + e0 + c1);
            const v4f *s2 = reinterpret_cast<const v4f *>(hterm + 2 * (size_t)stride + e0 + c1);
#pragma unroll
            for (uint32_t k = 0; k < V; k++) {
                const uint32_t v = min(lane + 64 * k, fnv4 - 1u);
                nx0[k] = s0[v];
                nx1[k] = s1[v];
                nx2[k] = s2[v];
            }
        }
        if (lane < 3) {
            const float4 *t4 = reinterpret_cast<const float4 *>(st[lane]);
            const uint32_t full = cnt >> 2;
#pragma unroll 8
            for (uint32_t q = 0; q < full; q++) {
                const float4 v = t4[q];
                run = run + v.x;
                run = run + v.y;
                run = run + v.z;
                run = run + v.w;
            }
            for (uint32_t k = full << 2; k < cnt; k++) run = run + st[lane][k];
        }
        __syncthreads();
    }
    const float fx = __shfl(run, 0, 64), fy = __shfl(run, 1, 64), fz = __shfl(run, 2, 64);
    if (lane == 0)
        gravity_and_stats(P, b, pos[b], (float)deg[b] + 1.0f, fx, fy, fz, old_force, force, acc);
    __syncthreads();
    if (lane < 2 * kLimbs && acc[lane])
        atomicAdd(&partial[lane * kAccCols + blockIdx.x % kAccCols], acc[lane]);
}

// Sums the kAccCols accumulator columns per limb (and clears them), then runs
// AdaptiveSpeed::update (fa2_speed.rs:48-100) on one thread.  update_speed = 0
// only records the sums (parity hook).  It also resets what the next tree
// build accumulates into: bounding box, invalid count, depth, top-tree slots.
__global__ __launch_bounds__(kSpeedThreads) void k_speed(const Params P, Frame *fr,
                                                         unsigned long long *__restrict__ partial,
                                                         uint32_t *__restrict__ top_slot,
                                                         uint32_t update_speed) {
    constexpr uint32_t G = kSpeedThreads / (2 * kLimbs);   // threads per limb
    static_assert(kAccCols % G == 0, "columns split evenly over a limb's threads");
    static_assert(64 % G == 0, "a limb's threads sit inside one wave");
    static_assert(kTopSlots <= kSpeedThreads, "one thread per top slot");
    __shared__ unsigned long long red[2 * kLimbs];
    const uint32_t t = threadIdx.x;
    {
        const uint32_t k = t / G, r = t % G;
        unsigned long long *row = partial + (size_t)k * kAccCols;
        unsigned long long s = 0;
#pragma unroll
        for (uint32_t c = r; c < kAccCols; c += G) {
            s += row[c];
            row[c] = 0ull;  // cleared for the next force pass
        }
#pragma unroll
        for (uint32_t o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, (int)o, 64);
        if (r == 0) red[k] = s;
    }
    // the per-build state of the next tree build (k_apply accumulates the box
    // of the positions it writes)
    if (t < kTopSlots) top_slot[t] = kTopAbsent;
    if (t < 3) {
        fr->bbmin[t] = 0xFF800000u;   // f2ord(+inf)
        fr->bbmax[t] = ~0xFF800000u;  // f2ord(-inf)
    }
    if (t == 0) {
        fr->n_invalid = 0;
        fr->max_depth = 0;
        fr->interactions = 0ull;
        fr->n_big = 0;
    }
    __syncthreads();
    if (t != 0) return;
    unsigned long long tot[2 * kLimbs];
#pragma unroll
    for (uint32_t k = 0; k < 2 * kLimbs; k++) tot[k] = red[k];
    exact_normalise(tot);
    exact_normalise(tot + kLimbs);
    const double ts = exact_to_double(tot), tt = exact_to_double(tot + kLimbs);
    fr->sum_swing = ts;
    fr->sum_traction = tt;
    if (!update_speed) return;
    float speed = fr->speed, eff = fr->speed_eff;
    if (!(ts > 0.0) || !(tt > 0.0) || P.n == 0) return;
    const double nn = (double)P.n, jt0 = (double)P.jitter_tolerance;
    double jt = jt0 * fmax(P.min_jt, fmin(10.0, P.est_jt * tt / (nn * nn)));
    const float min_eff = 0.05f;
    if (ts / tt > 2.0) {
        if (eff > min_eff) eff = eff * 0.5f;
        jt = fmax(jt, jt0);
    }
    const double target = jt * (double)eff * tt / ts;
    if (ts > jt * tt) {
        if (eff > min_eff) eff = eff * 0.7f;
    } else if (speed < 1000.0f) {
        eff = eff * 1.3f;
    }
    const double sp = (double)speed;
    speed = (float)(sp + fmin(target - sp, 0.5 * sp));
    fr->speed = speed;
    fr->speed_eff = eff;
}

// fa2_barnes_hut.wgsl:203-226, and the bounding box of the positions it writes
// (finite coordinates, as k_bbox) for the next tree build: per thread over a
// grid-stride loop, then the block, then one atomic per block and axis.
// BBOX = false (the brute-force engine, which builds no tree) skips the box.
template <bool BBOX>
__global__ __launch_bounds__(kTB) void k_apply(const Params P, Frame *fr,
                                               const float4 *__restrict__ pos_in,
                                               float4 *__restrict__ pos_out,
                                               const float4 *__restrict__ force,
                                               float4 *__restrict__ old_force,
                                               const uint32_t *__restrict__ deg,
                                               uint32_t *__restrict__ bbpart) {
    __shared__ float smn[3][kTB / 64], smx[3][kTB / 64];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    const float speed = fr->speed;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x) {
        const float4 f = force[i];
        const float4 o = old_force[i];
        const float deg_i = (float)deg[i] + 1.0f;
        const float swinging = deg_i * len3(o.x - f.x, o.y - f.y, o.z - f.z);
        float factor = speed / (1.0f + sqrtf(speed * swinging));
        const float df = len3(f.x, f.y, f.z);
        if (P.max_displacement > 0.0f && df > 1.0e-9f) factor = fminf(factor, P.max_displacement / df);
        const float k = factor * P.time_step;
        const float4 p = pos_in[i];
        // .w carries the mass for k_emit (one gather there instead of two)
        const float4 q = make_float4(p.x + f.x * k, p.y + f.y * k, p.z + f.z * k, fmaxf(deg_i, 1e-3f));
        pos_out[i] = q;
        old_force[i] = make_float4(f.x, f.y, f.z, 0.0f);
        if constexpr (!BBOX) continue;
        const float v[3] = {q.x + 0.0f, q.y + 0.0f, q.z + 0.0f};  // +0: canonical zero sign
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (finitef(v[c])) {
                mn[c] = fminf(mn[c], v[c]);
                mx[c] = fmaxf(mx[c], v[c]);
            }
        }
    }
    if constexpr (!BBOX) return;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn[c] = fminf(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], o, 64));
        }
        if (lane == 0) {
            smn[c][w] = mn[c];
            smx[c][w] = mx[c];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        float a = INFINITY, b = -INFINITY;
        for (uint32_t q = 0; q < kTB / 64; q++) {
            a = fminf(a, smn[c][q]);
            b = fmaxf(b, smx[c][q]);
        }
        // the block's extremes; k_bbox_final reduces them (one device-scope atomic
        // per block and axis on one line cost about 9 us at 512 blocks)
        bbpart[6 * blockIdx.x + c] = finitef(a) ? f2ord(a) : 0xFF800000u;      // f2ord(+inf)
        bbpart[6 * blockIdx.x + 3 + c] = finitef(b) ? f2ord(b) : ~0xFF800000u; // f2ord(-inf)
    }
}

// The bounding box of the positions k_apply wrote, from its blocks' extremes.
__global__ __launch_bounds__(256) void k_bbox_final(const uint32_t *__restrict__ bbpart, uint32_t nblk,
                                                    Frame *fr) {
    __shared__ uint32_t red[6][4];
    uint32_t v[6] = {0xFF800000u, 0xFF800000u, 0xFF800000u, ~0xFF800000u, ~0xFF800000u, ~0xFF800000u};
    for (uint32_t q = threadIdx.x; q < nblk; q += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[k] = min(v[k], bbpart[6 * q + k]);
            v[3 + k] = max(v[3 + k], bbpart[6 * q + 3 + k]);
        }
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t u = __shfl_xor(v[k], o, 64);
            v[k] = k < 3 ? min(v[k], u) : max(v[k], u);
        }
        if (lane == 0) red[k][w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t r = red[k][0];
        for (uint32_t q = 1; q < 4; q++) r = k < 3 ? min(r, red[k][q]) : max(r, red[k][q]);
        if (k < 3) fr->bbmin[k] = r;
        else fr->bbmax[k - 3] = r;
    }
}

__global__ void k_set_speed(Frame *fr, float speed, float eff) {
    if (threadIdx.x == 0) {
        fr->speed = speed;
        fr->speed_eff = eff;
    }
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------- engine
enum Phase { PH_BBOX, PH_KEYS, PH_SORT, PH_SCAN, PH_EMIT, PH_COM, PH_FORCE, PH_SPEED, PH_APPLY, PH_N };
const char *kPhaseNames[PH_N] = {"bbox", "keys", "sort", "count_scan", "emit",
                                 "com", "force", "speed", "apply"};

// Both ForceAtlas2 engines: Barnes-Hut ("fa2-bh-hip", fa2_bh.rs) and, with
// brute set, the brute-force engine ("fa2-brute-hip", fa2_brute.rs), whose
// iteration is the same minus the tree: k_force_brute for the walk, then the
// same attraction, gravity, stats, speed and apply.
class Fa2BhEngine final : public LayoutEngine {
  public:
    Fa2BhEngine(const jc_fa2bh_settings &s, bool brute) : settings_(s), brute_(brute) {
        desc_.kind = "physics";
        if (brute_) {
            desc_.id = "fa2-brute-hip";
            desc_.display_name = "ForceAtlas2 (brute force, MI355X)";
            desc_.description =
                "O(n^2) repulsion ForceAtlas2 on the GPU: every pair in index order, with the "
                "attraction, gravity and adaptive speed of fa2-brute.";
        } else {
            desc_.id = "fa2-bh-hip";
            desc_.display_name = "ForceAtlas2 (Barnes-Hut, MI355X)";
            desc_.description =
                "O(n log n) ForceAtlas2 with the octree built, walked and integrated on the GPU: "
                "same attraction, gravity, repulsion law and adaptive speed as fa2-bh.";
        }
        desc_.needs_edges = 1;
        desc_.needs_cpu_positions = 1;
        desc_.needs_gpu_positions_buffer = 0;
    }
    ~Fa2BhEngine() override { teardown(); }

    const jc_descriptor &descriptor() const override { return desc_; }

    Status set_params(const char *json, size_t len) override {
        jc_fa2bh_settings s;
        jc_fa2bh_default_settings(&s);
        bool strong = false, linlog = false, overlap = false;
        std::vector<JsonField> f = {
            {JsonField::F32, "gravity", &s.gravity},
            {JsonField::BOOL, "strong_gravity", &strong},
            {JsonField::F32, "scaling_ratio", &s.scaling_ratio},
            {JsonField::F32, "edge_weight_influence", &s.edge_weight_influence},
            {JsonField::F32, "jitter_tolerance", &s.jitter_tolerance},
            {JsonField::BOOL, "lin_log_mode", &linlog},
            {JsonField::BOOL, "prevent_overlap", &overlap},
            {JsonField::F32, "time_step", &s.time_step},
            {JsonField::F32, "max_displacement", &s.max_displacement},
        };
        // Fa2Settings (fa2_brute.rs:63-78) has no theta: the key is ignored there
        if (!brute_) f.push_back({JsonField::F32, "theta", &s.theta});
        int rc = decode_json_object(json, len, f, brute_ ? "fa2-brute settings" : "fa2-bh settings");
        if (rc < 0) return Status{JC_E_INVALID};
        if (rc == 1) return Status::ok();
        s.strong_gravity = strong;
        s.lin_log_mode = linlog;
        s.prevent_overlap = overlap;
        settings_ = s;
        params_dirty_ = true;
        return Status::ok();
    }

    Status init(const CsrView &g, const float *xyz) override {
        int dev = 0;
        JC_TRY(require_device(&dev));
        JC_TRY(validate_csr(g));
        if (g.n && !xyz) return Status::err(JC_E_INVALID, "positions pointer is NULL");
        if (g.n >= 0x7FFFFFFFu) return Status::err(JC_E_INVALID, "too many nodes");
        teardown();
        n_ = g.n;
        // fa2_bh.rs:264-294: degrees, undirected edge list (v<u), attraction order
        std::vector<uint32_t> degs(n_ ? n_ : 1, 0), aoff(n_ + 1, 0), adj;
        {
            std::vector<uint64_t> cnt((size_t)n_ + 1, 0);
            for (uint32_t v = 0; v < n_; v++)
                for (uint64_t e = g.offsets[v]; e < g.offsets[v + 1]; e++) {
                    uint32_t u = g.neighbors[e];
                    if (u == v) continue;
                    degs[v]++;
                    if (v < u) {
                        cnt[v + 1]++;
                        cnt[u + 1]++;
                    }
                }
            for (uint32_t v = 0; v < n_; v++) cnt[v + 1] += cnt[v];
            if (cnt[n_] > 0xFFFFFFFFull) return Status::err(JC_E_INVALID, "too many edges");
            for (uint32_t v = 0; v <= n_; v++) aoff[v] = (uint32_t)cnt[v];
            adj.resize(cnt[n_] ? cnt[n_] : 1);
            for (uint32_t v = 0; v < n_; v++)
                for (uint64_t e = g.offsets[v]; e < g.offsets[v + 1]; e++) {
                    uint32_t u = g.neighbors[e];
                    if (u == v || !(v < u)) continue;
                    adj[cnt[v]++] = u;
                    adj[cnt[u]++] = v;
                }
            n_att_ = aoff[n_];
        }
        std::vector<uint32_t> heavy;
        for (uint32_t v = 0; v < n_; v++)
            if (aoff[v + 1] - aoff[v] > kHeavyAtt) heavy.push_back(v);
        // longest rows first, so the longest wave starts with the launch
        std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t a, uint32_t b) {
            return aoff[a + 1] - aoff[a] > aoff[b + 1] - aoff[b];
        });
        n_heavy_ = (uint32_t)heavy.size();
        // rows laid end to end, each starting on a multiple of 4 entries; the
        // padding entries repeat the row's body as its own neighbour (a zero
        // term that is loaded but never added)
        std::vector<uint32_t> hoff(n_heavy_ + 1, 0), hlen(n_heavy_ ? n_heavy_ : 1, 0), hbody, hadj;
        for (uint32_t r = 0; r < n_heavy_; r++) {
            const uint32_t v = heavy[r];
            hoff[r] = (uint32_t)hbody.size();
            hlen[r] = aoff[v + 1] - aoff[v];
            for (uint32_t e = aoff[v]; e < aoff[v + 1]; e++) {
                hbody.push_back(v);
                hadj.push_back(adj[e]);
            }
            while (hbody.size() & 3u) {
                hbody.push_back(v);
                hadj.push_back(v);
            }
        }
        hoff[n_heavy_] = (uint32_t)hbody.size();
        n_hent_ = (uint32_t)hbody.size();
        hstride_ = (n_hent_ + 3u) & ~3u;
        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        const uint32_t nn = n_ ? n_ : 1;
        JC_TRY(pos_[0].alloc(nn));
        JC_TRY(pos_[1].alloc(nn));
        JC_TRY(old_force_.alloc(nn));
        JC_TRY(force_.alloc(nn));
        JC_TRY(deg_.alloc(nn));
        JC_TRY(att_off_.alloc((size_t)n_ + 1));
        JC_TRY(att_adj_.alloc(adj.size()));
        JC_TRY(keys_.alloc(nn));
        JC_TRY(vals_.alloc(nn));
        JC_TRY(sort_.alloc(n_));
        nb_scan_ = cdiv(nn, kScanTile);
        JC_TRY(loc_.alloc(nn));
        JC_TRY(blk_.alloc(nb_scan_ + 1));
        JC_TRY(blkd_.alloc(nb_scan_ + 1));
        // Internal cells at one depth are disjoint and hold >= 2 bodies each:
        // at most n/2 per depth over 22 depths, plus the n leaves.
        cap_ = nn + (uint64_t)(kLevels + 1) * (nn / 2 + 1) + 8;
        JC_TRY(nodes_.alloc(cap_));
        JC_TRY(node_depth_.alloc((cap_ + 31) & ~15ull));  // whole 16-byte loads in k_com_level
        JC_TRY(cskip_.alloc(cap_));
        JC_TRY(bbpart_.alloc(6 * kApplyBlocks));
        // Cells that do not fit k_com_fused at one depth are disjoint and hold
        // more than kFuseCap nodes each (from depth kTopDepth + 1 for small graphs).
        const uint32_t fuse_from = n_ <= kComSmallMax ? kTopDepth + 1 : kFuseDepth;
        big_max_ = (uint32_t)std::min<uint64_t>((uint64_t)(kLevels + 1 - fuse_from) * (cap_ / kFuseCap + 1),
                                                0xFFFFFFFFull);
        JC_TRY(big_.alloc(big_max_));
        JC_TRY(top_slot_.alloc(kTopSlots));
        JC_TRY(top_a_.alloc(kTopSlots));
        JC_TRY(top_b_.alloc(kTopSlots));
        nb_force_ = cdiv(nn, kTB);
        JC_TRY(partial_.alloc((size_t)kAccCols * 2 * kLimbs));
        JC_TRY(heavy_.alloc(n_heavy_ ? n_heavy_ : 1));
        JC_TRY(hoff_.alloc((size_t)n_heavy_ + 1));
        JC_TRY(hlen_.alloc(n_heavy_ ? n_heavy_ : 1));
        JC_TRY(hbody_.alloc(n_hent_ ? n_hent_ : 1));
        JC_TRY(hadj_.alloc(n_hent_ ? n_hent_ : 1));
        JC_TRY(hterm_.alloc(3 * (size_t)(hstride_ ? hstride_ : 4)));
        if (n_heavy_) {
            JC_HIP(hipMemcpyAsync(heavy_.p, heavy.data(), n_heavy_ * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream_));
            JC_HIP(hipMemcpyAsync(hoff_.p, hoff.data(), hoff.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream_));
            JC_HIP(hipMemcpyAsync(hlen_.p, hlen.data(), n_heavy_ * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream_));
            JC_HIP(hipMemcpyAsync(hbody_.p, hbody.data(), n_hent_ * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream_));
            JC_HIP(hipMemcpyAsync(hadj_.p, hadj.data(), n_hent_ * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream_));
        }
        JC_TRY(frame_.alloc(1));
        mass_.assign(n_, 0.0f);
        for (uint32_t i = 0; i < n_; i++) mass_[i] = std::fmax((float)degs[i] + 1.0f, 1e-3f);
        std::vector<float4> p4(nn, make_float4(0, 0, 0, 0));
        for (uint32_t i = 0; i < n_; i++)
            p4[i] = make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2],
                                mass_[i]);
        JC_HIP(hipMemcpyAsync(pos_[0].p, p4.data(), nn * sizeof(float4), hipMemcpyHostToDevice,
                              stream_));
        JC_HIP(hipMemcpyAsync(pos_[1].p, p4.data(), nn * sizeof(float4), hipMemcpyHostToDevice,
                              stream_));
        JC_HIP(hipMemsetAsync(old_force_.p, 0, nn * sizeof(float4), stream_));
        JC_HIP(hipMemsetAsync(force_.p, 0, nn * sizeof(float4), stream_));
        JC_HIP(hipMemcpyAsync(deg_.p, degs.data(), nn * sizeof(uint32_t), hipMemcpyHostToDevice,
                              stream_));
        JC_HIP(hipMemcpyAsync(att_off_.p, aoff.data(), ((size_t)n_ + 1) * sizeof(uint32_t),
                              hipMemcpyHostToDevice, stream_));
        JC_HIP(hipMemcpyAsync(att_adj_.p, adj.data(), adj.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, stream_));
        JC_HIP(hipMemsetAsync(frame_.p, 0, sizeof(Frame), stream_));
        k_set_speed<<<1, 64, 0, stream_>>>(frame_.p, 1.0f, 1.0f);  // fa2_speed.rs:32-39
        JC_HIP(hipGetLastError());
        JC_HIP(hipStreamSynchronize(stream_));
        cur_ = 0;
        iteration_ = 0;
        params_dirty_ = true;
        ready_ = false;
        inited_ = true;
        return Status::ok();
    }

    Status step(uint32_t n_iters, float *out_xyz) override {
        if (!inited_) return Status::err(JC_E_STATE, std::string(desc_.id) + " step called before successful init");
        if (n_ == 0) return Status::ok();
        refresh_params();
        JC_TRY(ensure_tuned());
        JC_TRY(ensure_ready());
        uint32_t left = n_iters;
        if (left >= 2) {
            JC_TRY(ensure_graph());
            if (cur_ == 1) {  // realign so the 2-iteration graph starts from buffer 0
                JC_TRY(enqueue_iteration(cur_, true, true, nullptr));
                cur_ ^= 1;
                left--;
            }
            while (left >= 2) {
                JC_HIP(hipGraphLaunch(graph_exec_, stream_));
                left -= 2;
            }
        }
        while (left > 0) {
            JC_TRY(enqueue_iteration(cur_, true, true, nullptr));
            cur_ ^= 1;
            left--;
        }
        iteration_ += n_iters;
        if (out_xyz) JC_TRY(read_positions(out_xyz));
        return Status::ok();
    }

    Status sync() override {
        if (stream_) JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }

    Status get_state(float *xyz, float *old_force, float *se2, uint64_t *it) override {
        if (!inited_) return Status::err(JC_E_STATE, "get_state before init");
        if (xyz) JC_TRY(read_positions(xyz));
        if (old_force) JC_TRY(read_vec4(old_force_.p, old_force));
        Frame fr;
        JC_HIP(hipMemcpyAsync(&fr, frame_.p, sizeof(Frame), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        if (se2) {
            se2[0] = fr.speed;
            se2[1] = fr.speed_eff;
        }
        if (it) *it = iteration_;
        return Status::ok();
    }

    Status set_state(const float *xyz, const float *old_force, const float *se2,
                     uint64_t it) override {
        if (!inited_) return Status::err(JC_E_STATE, "set_state before init");
        if (xyz) JC_TRY(write_vec4(pos_[cur_].p, xyz, mass_.data()));
        if (old_force) JC_TRY(write_vec4(old_force_.p, old_force));
        ready_ = false;
        if (se2) {
            k_set_speed<<<1, 64, 0, stream_>>>(frame_.p, se2[0], se2[1]);
            JC_HIP(hipGetLastError());
        }
        iteration_ = it;
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }

    // ---- parity / measurement hooks ----
    Status forces_at(const float *xyz, float *out_force, double *sums2) {
        if (!inited_) return Status::err(JC_E_STATE, "forces_at before init");
        if (n_ == 0) return Status::ok();
        refresh_params();
        DevBuf<float4> tmp;
        JC_TRY(tmp.alloc(n_));
        JC_TRY(write_vec4(tmp.p, xyz, mass_.data()));
        ready_ = false;
        JC_TRY(enqueue_fresh(tmp.p));
        if (!brute_) JC_TRY(enqueue_tree(tmp.p, nullptr));
        JC_TRY(enqueue_force(tmp.p, nullptr));
        if (!brute_)
            k_count_interactions<<<cdiv(n_, kTB), kTB, 0, stream_>>>(params_, frame_.p, tmp.p,
                                                                     deg_.p, vals_.p, nodes_.p);
        k_speed<<<1, kSpeedThreads, 0, stream_>>>(params_, frame_.p, partial_.p, top_slot_.p, 0u);
        JC_HIP(hipGetLastError());
        if (out_force) JC_TRY(read_vec4(force_.p, out_force));
        Frame fr;
        JC_HIP(hipMemcpyAsync(&fr, frame_.p, sizeof(Frame), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        if (sums2) {
            sums2[0] = fr.sum_swing;
            sums2[1] = fr.sum_traction;
        }
        return Status::ok();
    }

    Status tree(uint32_t cap, uint32_t *n_tree, uint32_t *n_valid, uint64_t *keys, uint32_t *perm,
                float *node_a, uint32_t *info, uint32_t *skip, uint8_t *depth, float *root4) {
        if (!inited_) return Status::err(JC_E_STATE, "tree before init");
        Frame fr;
        JC_HIP(hipMemcpyAsync(&fr, frame_.p, sizeof(Frame), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        if (n_tree) *n_tree = fr.n_tree;
        if (n_valid) *n_valid = fr.n_valid;
        if (root4) std::memcpy(root4, fr.root, sizeof fr.root);
        if (n_ == 0) return Status::ok();
        if (keys) JC_TRY(copy_sync(keys, keys_.p, n_ * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
        if (perm) JC_TRY(copy_sync(perm, vals_.p, n_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
        const uint32_t m = std::min<uint32_t>(cap, fr.n_tree);
        if (m) {
            std::vector<Node> h(m);
            JC_TRY(copy_sync(h.data(), nodes_.p, m * sizeof(Node), hipMemcpyDeviceToHost, stream_));
            for (uint32_t q = 0; q < m; q++) {
                if (node_a) std::memcpy(node_a + 4 * (size_t)q, &h[q].a, sizeof(float4));
                // exported form: leaf 0x80000000 | body, cell the bits of (2*half)^2
                if (info) info[q] = h[q].b.z != kCellId ? (kLeaf | h[q].b.z) : h[q].b.y;
                if (skip) skip[q] = h[q].b.w;
            }
            if (depth) JC_TRY(copy_sync(depth, node_depth_.p, m, hipMemcpyDeviceToHost, stream_));
        }
        return Status::ok();
    }

    // Timed iterations: the tree build of each iteration is replayed from a
    // captured graph, and the force kernel is launched between two events on
    // the engine stream, so the events time exactly that kernel; the host reads
    // them back once per kTimedIters iterations.
    Status step_timed(uint32_t n_iters, double *force_ms, double *iter_ms) {
        if (!inited_) return Status::err(JC_E_STATE, "step_timed before init");
        refresh_params();
        if (n_) JC_TRY(ensure_tuned());
        double fsum = 0, isum = 0;
        if (n_ && n_iters) {
            JC_TRY(ensure_ready());
            if (!brute_) JC_TRY(ensure_tree_graphs());
            for (auto &e : tev_)
                if (!e) JC_HIP(hipEventCreate(&e));
            uint32_t done = 0;
            while (done < n_iters) {
                const uint32_t chunk = std::min<uint32_t>(kTimedIters, n_iters - done);
                JC_HIP(hipEventRecord(tev_[2 * kTimedIters], stream_));
                for (uint32_t k = 0; k < chunk; k++) {
                    const uint32_t c = cur_;
                    if (!brute_) JC_HIP(hipGraphLaunch(tree_exec_[c], stream_));
                    JC_HIP(hipEventRecord(tev_[2 * k], stream_));
                    JC_TRY(enqueue_force(pos_[c].p, nullptr));
                    JC_HIP(hipEventRecord(tev_[2 * k + 1], stream_));
                    k_speed<<<1, kSpeedThreads, 0, stream_>>>(params_, frame_.p, partial_.p, top_slot_.p, 1u);
                    enqueue_apply(c);
                    JC_HIP(hipGetLastError());
                    cur_ ^= 1;
                }
                JC_HIP(hipEventRecord(tev_[2 * kTimedIters + 1], stream_));
                JC_HIP(hipEventSynchronize(tev_[2 * kTimedIters + 1]));
                for (uint32_t k = 0; k < chunk; k++) {
                    float a;
                    JC_HIP(hipEventElapsedTime(&a, tev_[2 * k], tev_[2 * k + 1]));
                    fsum += (double)a;
                }
                float c;
                JC_HIP(hipEventElapsedTime(&c, tev_[2 * kTimedIters], tev_[2 * kTimedIters + 1]));
                isum += (double)c;
                done += chunk;
            }
            iteration_ += n_iters;
        }
        if (force_ms) *force_ms = fsum;
        if (iter_ms) *iter_ms = isum;
        return Status::ok();
    }

    Status last_stats(uint64_t *n_tree, uint64_t *interactions, double *sums2) {
        if (!inited_) return Status::err(JC_E_STATE, "last_stats before init");
        Frame fr;
        JC_HIP(hipMemcpyAsync(&fr, frame_.p, sizeof(Frame), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        if (n_tree) *n_tree = fr.n_tree;
        if (interactions) *interactions = fr.interactions;
        if (sums2) {
            sums2[0] = fr.sum_swing;
            sums2[1] = fr.sum_traction;
        }
        return Status::ok();
    }

    Status profile(uint32_t n_iters, double *phase_ms, uint32_t n_phases) {
        if (!inited_) return Status::err(JC_E_STATE, "profile before init");
        refresh_params();
        if (n_) JC_TRY(ensure_tuned());
        if (n_ && n_iters) JC_TRY(ensure_ready());
        std::vector<hipEvent_t> ev(PH_N + 1);
        for (auto &e : ev) JC_HIP(hipEventCreate(&e));
        std::vector<double> acc(PH_N, 0.0);
        for (uint32_t it = 0; it < n_iters; it++) {
            JC_TRY(enqueue_iteration(cur_, true, true, ev.data()));
            cur_ ^= 1;
            iteration_++;
            JC_HIP(hipEventSynchronize(ev[PH_N]));
            for (int k = 0; k < PH_N; k++) {
                float ms;
                JC_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                acc[k] += ms;
            }
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        if (n_iters && n_ && !brute_) {
            // interactions of the last profiled iteration: its tree is still in
            // nodes_, built from the positions that iteration read
            k_count_interactions<<<cdiv(n_, kTB), kTB, 0, stream_>>>(params_, frame_.p,
                                                                     pos_[cur_ ^ 1].p, deg_.p,
                                                                     vals_.p, nodes_.p);
            JC_HIP(hipGetLastError());
            JC_HIP(hipStreamSynchronize(stream_));
        }
        for (uint32_t k = 0; k < n_phases && k < PH_N; k++)
            phase_ms[k] = n_iters ? acc[k] / n_iters : 0.0;
        return Status::ok();
    }

    // mode >= 0 forces a walk (brute: a repulsion kernel); -1 re-runs the timing;
    // *current gets the mode in use.
    Status set_walk_mode(int mode, int *current) {
        if (mode >= 0 && mode < (brute_ ? kBruteModes : kWalkModes)) {
            if (walk_mode_ != mode) drop_graphs();
            walk_mode_ = mode;
        } else if (mode == -1 && inited_ && n_) {
            walk_mode_ = -1;
            drop_graphs();
            refresh_params();
            JC_TRY(ensure_tuned());
        }
        if (current) *current = walk_mode_;
        return Status::ok();
    }

  private:
    static uint32_t chains_of(int mode) { return mode == 2 ? 2u : mode == 3 ? 4u : 1u; }
    // modes 4 and 5 are the lane walk in bigger blocks: fewer, larger Morton
    // ranges share a CU, so its L1 serves more of their common nodes; mode 6
    // is one wave per block, for graphs with fewer waves than CUs
    static uint32_t block_of(int mode) {
        return mode == 4 ? 512u : mode == 5 ? 1024u : mode == 6 ? 64u : kTB;
    }
    uint32_t force_grid(int mode) const {
        return cdiv(cdiv(n_ ? n_ : 1, chains_of(mode)), block_of(mode));
    }
    uint32_t force_grid_now() const { return force_grid(walk_mode_); }

    // Picks the fastest walk on the current positions (all give bit-identical
    // forces).  Engine state is not advanced.
    Status ensure_tuned() {
        if (walk_mode_ >= 0) return Status::ok();
        const int modes = brute_ ? kBruteModes : kWalkModes;
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        // Each timed launch follows a tree build, as in a real iteration (the
        // build leaves the caches the way the walk finds them then; back-to-back
        // walks would favour the variants that reuse a warm L2 most).  Four
        // interleaved rounds over the modes, best round per mode.
        for (int mode = 0; mode < modes; mode++) tune_ms_[mode] = 1e30f;
        for (int round = 0; round < 4; round++)
            for (int mode = 0; mode < modes; mode++) {
                walk_mode_ = mode;
                ready_ = false;
                JC_TRY(enqueue_fresh(pos_[cur_].p));
                if (!brute_) JC_TRY(enqueue_tree(pos_[cur_].p, nullptr));
                JC_HIP(hipEventRecord(a, stream_));
                JC_TRY(enqueue_force(pos_[cur_].p, nullptr));
                JC_HIP(hipEventRecord(b, stream_));
                JC_HIP(hipEventSynchronize(b));
                float ms = 0;
                JC_HIP(hipEventElapsedTime(&ms, a, b));
                tune_ms_[mode] = std::min(tune_ms_[mode], ms);
            }
        int pick = 0;
        for (int mode = 1; mode < modes; mode++)
            if (tune_ms_[mode] < tune_ms_[pick]) pick = mode;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        walk_mode_ = pick;
        drop_graphs();
        return Status::ok();
    }

    void refresh_params() {
        if (!params_dirty_) return;
        const jc_fa2bh_settings &s = settings_;
        params_.n = n_;
        params_.gravity = s.gravity;
        params_.strong_gravity = s.strong_gravity ? 1u : 0u;
        params_.scaling_ratio = s.scaling_ratio;
        params_.jitter_tolerance = s.jitter_tolerance;
        params_.lin_log_mode = s.lin_log_mode ? 1u : 0u;
        params_.time_step = s.time_step;
        const float th = std::fmax(s.theta, 0.0f);  // fa2_bh.rs:716
        params_.theta2 = th * th;
        params_.max_displacement = s.max_displacement;
        params_.ewi_weight = std::pow(std::fmax(1.0f, 1.0e-6f), s.edge_weight_influence);
        params_.est_jt = 0.05 * std::sqrt((double)n_);
        params_.min_jt = std::sqrt(params_.est_jt);
        params_.cap = (uint32_t)cap_;
        params_.fastdiv = kr_in_fastdiv_range(s.scaling_ratio) ? 1u : 0u;
        params_dirty_ = false;
        drop_graphs();  // kernel arguments changed
    }

    Status read_vec4(const float4 *src, float *dst) {
        std::vector<float4> h(n_);
        JC_HIP(hipMemcpyAsync(h.data(), src, n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        for (uint32_t i = 0; i < n_; i++) {
            dst[3 * (size_t)i] = h[i].x;
            dst[3 * (size_t)i + 1] = h[i].y;
            dst[3 * (size_t)i + 2] = h[i].z;
        }
        return Status::ok();
    }
    // xyz (3n) into float4s; w from `w` (n, nullable: 0).  Positions carry the
    // bodies' masses in w.
    Status write_vec4(float4 *dst, const float *src, const float *w = nullptr) {
        std::vector<float4> h(n_);
        for (uint32_t i = 0; i < n_; i++)
            h[i] = make_float4(src[3 * (size_t)i], src[3 * (size_t)i + 1], src[3 * (size_t)i + 2],
                               w ? w[i] : 0.0f);
        JC_HIP(hipMemcpyAsync(dst, h.data(), n_ * sizeof(float4), hipMemcpyHostToDevice, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }
    Status read_positions(float *out) { return read_vec4(pos_[cur_].p, out); }

    // Resets the per-build state and takes the bounding box of pos: needed when
    // pos was not written by k_apply right after k_speed (init, set_state, the
    // parity hooks, the tuner's force passes without k_speed).
    Status enqueue_fresh(const float4 *pos) {
        k_frame_reset<<<1, 1024, 0, stream_>>>(frame_.p, partial_.p, top_slot_.p);
        k_bbox<<<std::min<uint32_t>(cdiv(n_, kTB), 256), kTB, 0, stream_>>>(pos, n_, frame_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }
    Status ensure_ready() {
        if (!ready_) JC_TRY(enqueue_fresh(pos_[cur_].p));
        ready_ = true;
        return Status::ok();
    }
    // k_com_fused strides over the chunks of the tree (at most cap_ nodes)
    uint32_t com_grid() const {
        return std::max<uint32_t>(1, std::min<uint64_t>(cdiv(cap_, kFuseChunk), kComGrid));
    }
    uint32_t apply_grid() const { return std::min<uint32_t>(cdiv(n_ ? n_ : 1, kTB), kApplyBlocks); }
    // k_apply from pos_[c] into pos_[c ^ 1], then the new positions' bounding box
    // (for the next tree build; the brute-force engine has none)
    void enqueue_apply(uint32_t c) {
        if (brute_) {
            k_apply<false><<<apply_grid(), kTB, 0, stream_>>>(params_, frame_.p, pos_[c].p, pos_[c ^ 1].p,
                                                              force_.p, old_force_.p, deg_.p, bbpart_.p);
            return;
        }
        k_apply<true><<<apply_grid(), kTB, 0, stream_>>>(params_, frame_.p, pos_[c].p, pos_[c ^ 1].p,
                                                         force_.p, old_force_.p, deg_.p, bbpart_.p);
        k_bbox_final<<<1, 256, 0, stream_>>>(bbpart_.p, apply_grid(), frame_.p);
    }

    // The tree build of pos; the frame holds pos's bounding box and a reset
    // per-build state (enqueue_fresh, or k_speed + k_apply of the previous
    // iteration).
    Status enqueue_tree(const float4 *pos, hipEvent_t *ev) {
        const uint32_t n = n_;
        if (ev) JC_HIP(hipEventRecord(ev[PH_BBOX], stream_));
        if (ev) JC_HIP(hipEventRecord(ev[PH_KEYS], stream_));
        k_keys<<<cdiv(n, kTB), kTB, 0, stream_>>>(pos, n, frame_.p, keys_.p, vals_.p);
        if (ev) JC_HIP(hipEventRecord(ev[PH_SORT], stream_));
        if (n < kSort24Max) JC_TRY(sort_morton_pairs24(keys_.p, vals_.p, sort_, stream_));
        else JC_TRY(sort_morton_pairs(keys_.p, vals_.p, sort_, stream_));
        if (ev) JC_HIP(hipEventRecord(ev[PH_SCAN], stream_));
        k_count_scan<<<nb_scan_, kTB, 0, stream_>>>(keys_.p, n, frame_.p, loc_.p, blk_.p, blkd_.p);
        k_scan_top<<<1, 1024, 0, stream_>>>(blk_.p, nb_scan_, n, frame_.p, blkd_.p);
        if (ev) JC_HIP(hipEventRecord(ev[PH_EMIT], stream_));
        k_emit<<<cdiv(n, kTB), kTB, 0, stream_>>>(keys_.p, vals_.p, pos, deg_.p, n, frame_.p,
                                                  loc_.p, blk_.p, nodes_.p, node_depth_.p, cap_,
                                                  top_slot_.p, cskip_.p);
        if (ev) JC_HIP(hipEventRecord(ev[PH_COM], stream_));
        if (n <= kComSmallMax) {
            k_com_fused<kTopDepth + 1><<<com_grid(), kFuseThreads, 0, stream_>>>(
                frame_.p, (uint32_t)cap_, nodes_.p, node_depth_.p, cskip_.p, big_.p, big_max_);
            k_com_big<kTopDepth + 1><<<1, 1024, 0, stream_>>>(frame_.p, (uint32_t)cap_, nodes_.p,
                                                              node_depth_.p, big_.p, big_max_);
        } else {
            k_com_fused<kFuseDepth><<<com_grid(), kFuseThreads, 0, stream_>>>(
                frame_.p, (uint32_t)cap_, nodes_.p, node_depth_.p, cskip_.p, big_.p, big_max_);
            k_com_big<kFuseDepth><<<1, 1024, 0, stream_>>>(frame_.p, (uint32_t)cap_, nodes_.p,
                                                           node_depth_.p, big_.p, big_max_);
            const uint32_t gcom = std::min<uint32_t>(cdiv(cap_, kComSpan * (kTB / 64)), 1024);
            for (int d = (int)kFuseDepth - 1; d > kTopDepth; d--)
                k_com_level<<<gcom, kTB, 0, stream_>>>(frame_.p, (uint32_t)d, (uint32_t)cap_,
                                                       nodes_.p, node_depth_.p);
        }
        k_top_build<<<1, 1024, 0, stream_>>>(frame_.p, top_slot_.p, nodes_.p, (uint32_t)cap_,
                                             top_a_.p, top_b_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_force(const float4 *pos, hipEvent_t *ev) {
        if (ev) JC_HIP(hipEventRecord(ev[PH_FORCE], stream_));
        const uint32_t grid = force_grid(walk_mode_);
        if (n_heavy_)
            k_heavy_terms<<<cdiv(n_hent_, kTB), kTB, 0, stream_>>>(params_, n_hent_, hbody_.p,
                                                                  hadj_.p, pos, hterm_.p, hstride_);
        const uint32_t bs = block_of(walk_mode_);
        const int mode = brute_ ? -2 : walk_mode_;
#define JC_FORCE(KERNEL, ...)                                                              \
    KERNEL<<<grid, bs, 0, stream_>>>(params_, frame_.p, pos, old_force_.p, force_.p, deg_.p, \
                                     att_off_.p, att_adj_.p, vals_.p, nodes_.p, partial_.p   \
                                     __VA_ARGS__)
        switch (mode) {
            case -2:
                if (walk_mode_ == 1)
                    k_force_brute_grp<kTB, 4><<<(uint32_t)(((uint64_t)n_ * 4 + kTB - 1) / kTB), kTB,
                                                0, stream_>>>(params_, pos, old_force_.p, force_.p,
                                                              deg_.p, att_off_.p, att_adj_.p,
                                                              partial_.p);
                else if (walk_mode_ == 2)
                    k_force_brute_grp<kTB, 16><<<(uint32_t)(((uint64_t)n_ * 16 + kTB - 1) / kTB),
                                                 kTB, 0, stream_>>>(params_, pos, old_force_.p,
                                                                    force_.p, deg_.p, att_off_.p,
                                                                    att_adj_.p, partial_.p);
                else
                    k_force_brute<kTB><<<cdiv(n_, kTB), kTB, 0, stream_>>>(
                        params_, pos, old_force_.p, force_.p, deg_.p, att_off_.p, att_adj_.p,
                        partial_.p);
                break;
            case 0: JC_FORCE(k_force<0>, , top_a_.p, top_b_.p); break;
            case 2: JC_FORCE(k_force_multi<2>); break;
            case 3: JC_FORCE(k_force_multi<4>); break;
            case 4: JC_FORCE((k_force<0, 512>), , top_a_.p, top_b_.p); break;
            case 5: JC_FORCE((k_force<0, 1024>), , top_a_.p, top_b_.p); break;
            case 6: JC_FORCE((k_force<0, 64>), , top_a_.p, top_b_.p); break;
            default: JC_FORCE(k_force<1>, , top_a_.p, top_b_.p); break;
        }
#undef JC_FORCE
        if (n_heavy_)
            k_force_heavy<<<n_heavy_, 64, 0, stream_>>>(params_, heavy_.p, hoff_.p, hlen_.p,
                                                       hterm_.p, hstride_, pos, old_force_.p,
                                                       force_.p, deg_.p, partial_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    // One full iteration reading pos_[c], writing pos_[c^1].
    Status enqueue_iteration(uint32_t c, bool speed, bool apply, hipEvent_t *ev) {
        if (!brute_) {
            JC_TRY(enqueue_tree(pos_[c].p, ev));
        } else if (ev) {  // no tree: its phases take no time
            for (int k = PH_BBOX; k < PH_FORCE; k++) JC_HIP(hipEventRecord(ev[k], stream_));
        }
        JC_TRY(enqueue_force(pos_[c].p, ev));
        if (ev) JC_HIP(hipEventRecord(ev[PH_SPEED], stream_));
        if (speed)
            k_speed<<<1, kSpeedThreads, 0, stream_>>>(params_, frame_.p, partial_.p, top_slot_.p, 1u);
        if (ev) JC_HIP(hipEventRecord(ev[PH_APPLY], stream_));
        if (apply)
            enqueue_apply(c);
        if (ev) JC_HIP(hipEventRecord(ev[PH_N], stream_));
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status ensure_graph() {
        if (graph_exec_) return Status::ok();
        JC_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        Status s1 = enqueue_iteration(0, true, true, nullptr);
        Status s2 = s1.failed() ? s1 : enqueue_iteration(1, true, true, nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream_, &g);
        if (s2.failed()) return s2;
        if (e != hipSuccess) return Status::err(JC_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        graph_ = g;
        JC_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        return Status::ok();
    }

    // The tree build alone, captured once per ping-pong buffer.
    Status ensure_tree_graphs() {
        for (uint32_t c = 0; c < 2; c++) {
            if (tree_exec_[c]) continue;
            JC_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            Status s = enqueue_tree(pos_[c].p, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(stream_, &g);
            if (s.failed()) return s;
            if (e != hipSuccess)
                return Status::err(JC_E_HIP, std::string("capture: ") + hipGetErrorString(e));
            tree_graph_[c] = g;
            JC_HIP(hipGraphInstantiate(&tree_exec_[c], g, nullptr, nullptr, 0));
        }
        return Status::ok();
    }

    void drop_graphs() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
        for (uint32_t c = 0; c < 2; c++) {
            if (tree_exec_[c]) (void)hipGraphExecDestroy(tree_exec_[c]);
            if (tree_graph_[c]) (void)hipGraphDestroy(tree_graph_[c]);
            tree_exec_[c] = nullptr;
            tree_graph_[c] = nullptr;
        }
    }

    void teardown() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        drop_graphs();
        for (auto &e : tev_)
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        pos_[0].release();
        pos_[1].release();
        old_force_.release();
        force_.release();
        deg_.release();
        att_off_.release();
        att_adj_.release();
        keys_.release();
        vals_.release();
        loc_.release();
        blk_.release();
        nodes_.release();
        node_depth_.release();
        top_slot_.release();
        top_a_.release();
        top_b_.release();
        partial_.release();
        heavy_.release();
        hoff_.release();
        hlen_.release();
        hbody_.release();
        hadj_.release();
        hterm_.release();
        n_heavy_ = 0;
        n_hent_ = 0;
        hstride_ = 0;
        frame_.release();
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
        inited_ = false;
    }

    jc_descriptor desc_{};
    jc_fa2bh_settings settings_;
    bool brute_ = false;  // fa2-brute-hip: no tree, k_force_brute for the repulsion
    Params params_{};
    bool params_dirty_ = true;
    bool inited_ = false;
    // the frame holds the bounding box of pos_[cur_] and a reset per-build state
    bool ready_ = false;
    uint32_t n_ = 0;
    uint64_t n_att_ = 0;
    uint32_t n_heavy_ = 0;          // bodies whose attraction row exceeds kHeavyAtt
    uint32_t n_hent_ = 0;           // entries of those rows
    uint32_t hstride_ = 0;          // n_hent_ rounded up to 4: one component array of hterm_
    DevBuf<uint32_t> heavy_, hoff_, hlen_, hbody_, hadj_;
    DevBuf<float> hterm_;
    uint32_t cur_ = 0;
    uint64_t iteration_ = 0;
    uint64_t cap_ = 0;
    uint32_t nb_scan_ = 0, nb_force_ = 0;
    int walk_mode_ = -1;  // -1: pick at the first step by timing both walks
    float tune_ms_[kWalkModes] = {};
    hipStream_t stream_ = nullptr;
    DevBuf<float4> pos_[2], old_force_, force_;  // positions: w = mass max(deg + 1, 1e-3)
    std::vector<float> mass_;                    // the same masses on the host, for uploads
    DevBuf<uint32_t> deg_, att_off_, att_adj_;
    DevBuf<uint64_t> keys_;
    DevBuf<uint32_t> vals_, loc_, blk_;
    DevBuf<uint32_t> blkd_;       // deepest cell of each count/scan tile
    RadixScratch sort_;
    DevBuf<Node> nodes_;
    DevBuf<uint8_t> node_depth_;
    DevBuf<uint32_t> bbpart_;     // k_apply: each block's bounding-box extremes (6 words)
    DevBuf<uint32_t> cskip_;      // skip rope of each cell (k_emit), read by k_com_fused
    DevBuf<uint32_t> big_;        // cells too large for k_com_fused, summed by k_top_build
    uint32_t big_max_ = 0;
    DevBuf<uint32_t> top_slot_;   // kTopSlots: node of each slot of the top-tree enumeration
    DevBuf<float4> top_a_;        // top tree in pre-order: centre of mass, mass
    DevBuf<uint4> top_b_;         // and size, id, subtree end rank | frontier, byte offset
    DevBuf<unsigned long long> partial_;
    DevBuf<Frame> frame_;
    hipGraph_t graph_ = nullptr, tree_graph_[2] = {};
    hipGraphExec_t graph_exec_ = nullptr, tree_exec_[2] = {};
    hipEvent_t tev_[2 * kTimedIters + 2] = {};
};

}  // namespace

LayoutEngine *make_fa2bh(const jc_fa2bh_settings &s) { return new Fa2BhEngine(s, false); }
LayoutEngine *make_fa2brute(const jc_fa2bh_settings &s) { return new Fa2BhEngine(s, true); }
bool is_fa2bh(const LayoutEngine *e) { return dynamic_cast<const Fa2BhEngine *>(e) != nullptr; }

Status fa2bh_forces_at(LayoutEngine *e, const float *xyz, float *out_force, double *sums2) {
    return static_cast<Fa2BhEngine *>(e)->forces_at(xyz, out_force, sums2);
}
Status fa2bh_tree(LayoutEngine *e, uint32_t cap, uint32_t *n_tree, uint32_t *n_valid,
                  uint64_t *keys, uint32_t *perm, float *node_a, uint32_t *node_info,
                  uint32_t *node_skip, uint8_t *node_depth, float *root4) {
    return static_cast<Fa2BhEngine *>(e)->tree(cap, n_tree, n_valid, keys, perm, node_a, node_info,
                                               node_skip, node_depth, root4);
}
Status fa2bh_step_timed(LayoutEngine *e, uint32_t n_iters, double *force_ms, double *iter_ms) {
    return static_cast<Fa2BhEngine *>(e)->step_timed(n_iters, force_ms, iter_ms);
}
Status fa2bh_last_stats(LayoutEngine *e, uint64_t *n_tree, uint64_t *interactions, double *sums2) {
    return static_cast<Fa2BhEngine *>(e)->last_stats(n_tree, interactions, sums2);
}
Status fa2bh_walk_mode(LayoutEngine *e, int mode, int *current) {
    return static_cast<Fa2BhEngine *>(e)->set_walk_mode(mode, current);
}
Status fa2bh_profile(LayoutEngine *e, uint32_t n_iters, double *phase_ms, uint32_t n_phases) {
    return static_cast<Fa2BhEngine *>(e)->profile(n_iters, phase_ms, n_phases);
}

}  // namespace jc

extern "C" const char *jc_fa2bh_phase_name(uint32_t k) {
    return k < jc::PH_N ? jc::kPhaseNames[k] : "";
}
extern "C" uint32_t jc_fa2bh_phase_count(void) { return jc::PH_N; }
