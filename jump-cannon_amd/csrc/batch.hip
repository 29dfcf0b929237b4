// Batched FA2 + Barnes-Hut: thousands of independent graphs of uneven size laid
// out together, one launch per phase for the whole batch ("fa2-bh-batch-hip").
//
// The reference lays out one graph per engine instance, one dispatch pair per
// step (fa2_bh.rs:546-692; one engine per worker, sim.rs:386-421).  Here all
// graphs form one forest on the device:
//   * bodies are concatenated graph after graph; gid[i] names the graph;
//   * every graph gets its own root box (segmented bbox), its own octant keys,
//     and the sort is graph-major / key-minor (8 key passes + graph-id passes,
//     stable), so each graph's bodies stay in its own segment;
//   * each graph's pre-order octree occupies one contiguous node range whose
//     last rope is the end marker, so a walk never leaves its graph;
//   * swing/traction are summed exactly per graph and every graph runs its own
//     AdaptiveSpeed state (one block per graph).
// Each graph therefore evolves bit-identically to a standalone fa2-bh-hip run
// (and to the oracle), while the dispatch count per iteration is independent
// of the number of graphs.  The iteration is captured in a hipGraph; the force
// kernel's block size is picked at the first step by timing the candidates.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "exact_sum.hpp"
#include "fa2_device.hpp"

struct jc_batch;

namespace jc {
namespace {

constexpr uint32_t kBT = 256;
constexpr uint32_t kBScanItems = 16;
constexpr uint32_t kBScanTile = kBT * kBScanItems;
constexpr uint32_t kBEmitTasks = 1024;    // cell tasks per k_b_emit block held in LDS
constexpr uint32_t kBHeavyAtt = 32;       // attraction rows longer than this go to k_b_force_heavy
constexpr uint32_t kBHeavyStage = 1024;   // heavy-row terms staged in LDS per round
// Centre of mass: cells of depth >= the fusion depth (3, 4 or 5, picked by the
// autotuner) in one launch (com_fused_pass, chunks of 1024 or 2048 nodes), the
// shallower ones one launch per depth.
constexpr uint32_t kBFuseDepthMin = 3;
constexpr uint32_t kBFuseCap = 512;
constexpr uint32_t kBFuseThreads = 512;
constexpr uint32_t kBComGrid = 2048;

struct GFrame {
    uint32_t bbmin[3];
    uint32_t bbmax[3];
    uint32_t n_invalid;
    uint32_t n_tree;
    uint32_t tree_start;
    uint32_t pad;
    float root[4];
    float speed;
    float speed_eff;
    double sum_swing;
    double sum_traction;
};

struct BGlobal {
    uint32_t n_tree_total;
    uint32_t max_depth;
    uint32_t n_big;  // cells too large for k_b_com_fused
};

struct BParams {
    uint32_t n;          // total bodies
    uint32_t n_graphs;
    float gravity;
    uint32_t strong_gravity;
    float scaling_ratio;
    float jitter_tolerance;
    uint32_t lin_log_mode;
    float time_step;
    float theta2;
    float max_displacement;
    float ewi_weight;
    uint32_t cap;
    uint32_t fastdiv;    // scaling ratio inside kr_in_fastdiv_range
};

__global__ void k_b_reset(GFrame *fr, uint32_t G, BGlobal *gl) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) {
        gl->max_depth = 0;
        gl->n_big = 0;
    }
    if (g >= G) return;
    const uint32_t pinf = 0xFF800000u;
    for (int k = 0; k < 3; k++) {
        fr[g].bbmin[k] = pinf;
        fr[g].bbmax[k] = ~pinf;
    }
    fr[g].n_invalid = 0;
}

// Per-graph box over finite coordinates: one block per graph reduces the
// graph's contiguous body range (wave shuffles, then LDS) and writes the box
// with plain stores -- no atomics on the graph's frame, which large graphs
// would otherwise serialise.
__global__ __launch_bounds__(kBT) void k_b_bbox(const float4 *__restrict__ pos,
                                                const uint32_t *__restrict__ base,
                                                const uint32_t *__restrict__ count,
                                                GFrame *fr) {
    __shared__ float red[kBT / 64][6];
    const uint32_t g = blockIdx.x;
    const uint32_t b0 = base[g], ng = count[g];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = threadIdx.x; i < ng; i += kBT) {
        const float4 p = pos[b0 + i];
        const float v[3] = {p.x + 0.0f, p.y + 0.0f, p.z + 0.0f};
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (finitef(v[k])) {
                mn[k] = fminf(mn[k], v[k]);
                mx[k] = fmaxf(mx[k], v[k]);
            }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
        }
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            red[w][k] = mn[k];
            red[w][3 + k] = mx[k];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        float a = red[0][k], c = red[0][3 + k];
        for (uint32_t q = 1; q < kBT / 64; q++) {
            a = fminf(a, red[q][k]);
            c = fmaxf(c, red[q][3 + k]);
        }
        // the single-graph engine's box: +inf / -inf (as f2ord) when no finite value
        fr[g].bbmin[k] = finitef(a) ? f2ord(a) : 0xFF800000u;
        fr[g].bbmax[k] = finitef(c) ? f2ord(c) : ~0xFF800000u;
    }
}

__global__ __launch_bounds__(kBT) void k_b_keys(const float4 *__restrict__ pos,
                                                const uint32_t *__restrict__ gid,
                                                const uint32_t *__restrict__ base, uint32_t n,
                                                GFrame *fr, uint64_t *__restrict__ keys_raw,
                                                uint64_t *__restrict__ keys,
                                                uint32_t *__restrict__ vals, uint32_t ck_p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = gid[i];
    float r4[4];
    compute_root(fr[g].bbmin, fr[g].bbmax, r4);
    if (i == base[g])
        for (int k = 0; k < 4; k++) fr[g].root[k] = r4[k];
    const float4 p = pos[i];
    uint64_t key;
    if (finitef(p.x) && finitef(p.y) && finitef(p.z)) {
        key = octant_key(r4, p.x, p.y, p.z);
    } else {
        key = ~0ull;
        atomicAdd(&fr[g].n_invalid, 1u);
    }
    keys_raw[i] = key;
    // the sort key: graph id above the key's top ck_p bits (sort_segmented_pairs)
    keys[i] = ck_p >= 64 ? key : ((uint64_t)g << ck_p) | (key >> (64 - ck_p));
    vals[i] = i;
}

struct Seg {
    uint32_t g, lo, ve;  // graph, first sorted index, end of its valid bodies
};
__device__ __forceinline__ Seg seg_of(uint32_t i, const uint32_t *gid, const uint32_t *base,
                                      const uint32_t *count, const GFrame *fr) {
    Seg s;
    s.g = gid[i];
    s.lo = base[s.g];
    s.ve = s.lo + count[s.g] - fr[s.g].n_invalid;
    return s;
}

// Per sorted body i, the nodes it writes: its opened cells plus its leaf (0 if
// invalid, 2 for a graph's lone body); then the block-local exclusive scan.
__global__ __launch_bounds__(kBT) void k_b_count_scan(const uint64_t *__restrict__ keys,
                                                      const uint32_t *__restrict__ gid,
                                                      const uint32_t *__restrict__ base,
                                                      const uint32_t *__restrict__ count,
                                                      uint32_t n, const GFrame *fr,
                                                      uint32_t *__restrict__ loc,
                                                      uint32_t *__restrict__ blk_tot,
                                                      uint32_t *__restrict__ blk_depth) {
    __shared__ uint32_t ws[kBT / 64], wd[kBT / 64];
    int md = 0;  // deepest cell: the largest common prefix of two neighbours in one graph
    const uint32_t b0 = blockIdx.x * kBScanTile + threadIdx.x * kBScanItems;
    uint32_t c[kBScanItems];
    uint32_t s = 0;
    // the thread's keys and both neighbours, loaded together; the graph segment
    // is looked up again only when the items cross into the next graph
    uint64_t kk[kBScanItems + 2];
#pragma unroll
    for (uint32_t j = 0; j < kBScanItems + 2; j++) {
        const uint32_t i = b0 + j - 1;   // wraps to 0xFFFFFFFF for b0 = 0: never read
        kk[j] = (b0 + j >= 1 && i < n) ? keys[i] : 0ull;
    }
    Seg sg{0u, 0u, 0u};
    uint32_t seg_end = 0;
#pragma unroll
    for (uint32_t j = 0; j < kBScanItems; j++) {
        const uint32_t i = b0 + j;
        c[j] = 0;
        if (i < n) {
            if (j == 0 || i >= seg_end) {
                sg = seg_of(i, gid, base, count, fr);
                seg_end = base[sg.g] + count[sg.g];
            }
            if (i < sg.ve) {
                if (sg.ve - sg.lo == 1) {
                    c[j] = 2;
                } else {
                    const int lp = i > sg.lo ? lcp_digits(kk[j], kk[j + 1]) : -1;
                    const int li = i + 1 < sg.ve ? lcp_digits(kk[j + 1], kk[j + 2]) : -1;
                    c[j] = (li > lp ? (uint32_t)(li - lp) : 0u) + 1u;
                    md = li > md ? li : md;
                }
            }
        }
        s += c[j];
    }
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
    for (uint32_t q = 0; q < kBT / 64; q++) {
        if (q < w) wp += ws[q];
        tot += ws[q];
        bd = max(bd, wd[q]);
    }
    uint32_t run = incl - s + wp;
#pragma unroll
    for (uint32_t j = 0; j < kBScanItems; j++) {
        if (b0 + j < n) loc[b0 + j] = run;
        run += c[j];
    }
    if (threadIdx.x == 0) {
        blk_tot[blockIdx.x] = tot;
        blk_depth[blockIdx.x] = bd;  // k_b_scan_top takes the maximum
    }
}

__global__ __launch_bounds__(1024) void k_b_scan_top(uint32_t *__restrict__ blk, uint32_t nb,
                                                     BGlobal *gl, const uint32_t *__restrict__ blk_depth) {
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
        gl->n_tree_total = tot;
        gl->max_depth = md;
    }
}

// First node written by sorted body i (= total for i == n).
__device__ __forceinline__ uint32_t node_start(const uint32_t *loc, const uint32_t *blk,
                                               uint32_t i, uint32_t n, const BGlobal *gl) {
    return i < n ? loc[i] + blk[i / kBScanTile] : gl->n_tree_total;
}

__global__ void k_b_graph_tree(const uint32_t *__restrict__ base, const uint32_t *__restrict__ count,
                               uint32_t G, uint32_t n, const uint32_t *__restrict__ loc,
                               const uint32_t *__restrict__ blk, const BGlobal *gl, GFrame *fr) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const uint32_t s = node_start(loc, blk, base[g], n, gl);
    const uint32_t e = node_start(loc, blk, base[g] + count[g], n, gl);
    fr[g].tree_start = s;
    fr[g].n_tree = e - s;
}

// One internal cell of graph sg.g at depth d, opened by sorted body i, one of
// the block's (its half-size: the graph's root halved d times; its skip: the
// first node of the body after the last one sharing the d-digit prefix, inside
// the graph).  The search reads the block's LDS window (cell_end_win): keys wk
// and first nodes wf of the sorted bodies [w0, w0 + wn).
__device__ __forceinline__ void b_emit_cell(const uint64_t *__restrict__ keys, const Seg &sg,
                                            uint32_t i, int d, uint32_t node, uint32_t n,
                                            const GFrame *fr, const BGlobal *gl,
                                            const uint32_t *__restrict__ loc,
                                            const uint32_t *__restrict__ blk,
                                            Node *__restrict__ nodes,
                                            uint8_t *__restrict__ node_depth,
                                            uint32_t *__restrict__ cskip, const uint64_t *wk,
                                            const uint32_t *wf, uint32_t w0, uint32_t wn) {
    float h = fr[sg.g].root[3];
    for (int q = 0; q < d; q++) h = h * 0.5f;
    const uint32_t j = cell_end_win(keys, wk, w0, wn, i, sg.ve, wk[i - w0], d);
    const bool last = j + 1 >= sg.ve;  // the subtree ends with its graph: rope kEnd
    const uint32_t e = last ? sg.ve : j + 1;
    const uint32_t end = e - w0 < wn ? wf[e - w0] : node_start(loc, blk, e, n, gl);
    const float s = h * 2.0f;
    nodes[node].b = cell_b(s * s, last ? kEnd : end);
    cskip[node] = end;  // where the subtree ends in the forest, for k_b_com_fused
    node_depth[node] = (uint8_t)d;
}

// Pre-order emission per graph segment (as k_emit of the single-graph engine):
// leaves by their body's thread, cells through a block-wide task list.
__global__ __launch_bounds__(kBT) void k_b_emit(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ perm,
    const float4 *__restrict__ pos, const uint32_t *__restrict__ deg,
    const uint32_t *__restrict__ gid, const uint32_t *__restrict__ base,
    const uint32_t *__restrict__ count, uint32_t n, const GFrame *fr, BGlobal *gl,
    const uint32_t *__restrict__ loc, const uint32_t *__restrict__ blk,
    Node *__restrict__ nodes, uint8_t *__restrict__ node_depth, uint64_t cap,
    uint32_t *__restrict__ cskip) {
    __shared__ uint32_t ntask;
    __shared__ uint32_t t_node[kBEmitTasks], t_body[kBEmitTasks];
    __shared__ uint8_t t_depth[kBEmitTasks];
    // the block's sorted keys and the next kEmitWin, each with its first node
    __shared__ uint64_t wk[kBT + kEmitWin];
    __shared__ uint32_t wf[kBT + kEmitWin];
    const uint32_t w0 = blockIdx.x * kBT;
    const uint32_t wn = w0 < n ? min(kBT + kEmitWin, n - w0) : 0u;
    if (threadIdx.x == 0) ntask = 0;
    for (uint32_t q = threadIdx.x; q < wn; q += kBT) {
        wk[q] = keys[w0 + q];
        wf[q] = node_start(loc, blk, w0 + q, n, gl);
    }
    __syncthreads();
    const uint32_t i = w0 + threadIdx.x;
    // the neighbours' keys (only read inside the body's own graph segment)
    const uint64_t kown = i < n ? wk[threadIdx.x] : 0ull;
    const uint64_t kprev = (i > 0 && i < n) ? (threadIdx.x > 0 ? wk[threadIdx.x - 1] : keys[i - 1]) : 0ull;
    const uint64_t knext = i + 1 < n ? wk[threadIdx.x + 1] : 0ull;
    if (i < n) {
        const Seg sg = seg_of(i, gid, base, count, fr);
        if (i < sg.ve) {
            const uint64_t k = kown;
            int lp, li;
            if (sg.ve - sg.lo == 1) {
                lp = -1;
                li = 0;
            } else {
                lp = i > sg.lo ? lcp_digits(kprev, k) : -1;
                li = i + 1 < sg.ve ? lcp_digits(k, knext) : -1;
            }
            const uint32_t c = li > lp ? (uint32_t)(li - lp) : 0u;
            const int a0 = lp + 1;
            const uint32_t pos0 = wf[threadIdx.x];
            if ((uint64_t)pos0 + c < cap) {
                if (c) {
                    const uint32_t slot = atomicAdd(&ntask, c);
                    for (uint32_t t = 0; t < c; t++) {
                        if (slot + t < kBEmitTasks) {
                            t_node[slot + t] = pos0 + t;
                            t_body[slot + t] = i;
                            t_depth[slot + t] = (uint8_t)(a0 + (int)t);
                        } else {
                            b_emit_cell(keys, sg, i, a0 + (int)t, pos0 + t, n, fr, gl, loc, blk,
                                        nodes, node_depth, cskip, wk, wf, w0, wn);
                        }
                    }
                }
                const uint32_t leaf = pos0 + c;
                const uint32_t b = perm[i];
                if (b < n) {
                    const float4 p = pos[b];
                    nodes[leaf].a = p;  // .w: the body's mass max(deg + 1, 1e-3) (k_b_apply)
                    nodes[leaf].b = leaf_b(b, (i + 1 < sg.ve) ? leaf + 1 : kEnd);
                    node_depth[leaf] = (uint8_t)(0x80 | ((lp > li ? lp : li) + 1));
                }
            }
        }
    }
    __syncthreads();
    const uint32_t nt = min(ntask, kBEmitTasks);
    for (uint32_t q = threadIdx.x; q < nt; q += blockDim.x) {
        const uint32_t ib = t_body[q];
        const Seg sg = seg_of(ib, gid, base, count, fr);
        b_emit_cell(keys, sg, ib, t_depth[q], t_node[q], n, fr, gl, loc, blk, nodes, node_depth, cskip,
                    wk, wf, w0, wn);
    }
}

// The cells of depth >= Depth over the whole forest: each fitting subtree
// summed in LDS by one workgroup (com_fused_pass), then the ones too large for
// it (com_big_pass).  Every (Depth, Chunk) gives the same bits: each cell is
// summed from its children in octant order by the same operations.
template <uint32_t Depth, uint32_t Chunk>
__global__ __launch_bounds__(kBFuseThreads) void k_b_com_fused(BGlobal *gl, uint32_t cap,
                                                               Node *__restrict__ nodes,
                                                               const uint8_t *__restrict__ node_depth,
                                                               const uint32_t *__restrict__ cskip,
                                                               uint32_t *__restrict__ big, uint32_t big_max) {
    com_fused_pass<Depth, Chunk, kBFuseCap, kBFuseThreads>(
        min(gl->n_tree_total, cap), nodes, node_depth, cskip, big, big_max, &gl->n_big);
}

template <uint32_t Depth>
__global__ __launch_bounds__(1024) void k_b_com_big(const BGlobal *gl, uint32_t cap, Node *__restrict__ nodes,
                                                    const uint8_t *__restrict__ node_depth,
                                                    const uint32_t *__restrict__ big, uint32_t big_max) {
    const uint32_t nbig = min(gl->n_big, big_max);
    if (!nbig) return;
    com_big_pass<Depth>(min(gl->n_tree_total, cap), gl->max_depth, nodes, node_depth, big, nbig);
}

// The cells of depth d over the whole forest (com_level_pass).
__global__ __launch_bounds__(kBT) void k_b_com_level(const BGlobal *gl, uint32_t d, uint32_t cap,
                                                     Node *__restrict__ nodes,
                                                     const uint8_t *__restrict__ node_depth) {
    if (d > gl->max_depth) return;
    __shared__ uint32_t list[kBT / 64][kComSpan];
    com_level_pass<kBT>(nodes, node_depth, min(gl->n_tree_total, cap), d, list);
}

// One attraction term d * att (fa2_barnes_hut.wgsl:158-181), d = q - p.
__device__ __forceinline__ void b_attraction_term(const BParams &P, float4 p, float4 q, float &tx,
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

// Gravity (fa2_barnes_hut.wgsl:183-189), the force write and the per-body
// swing / traction stats (:191-197).
__device__ __forceinline__ void b_gravity_and_stats(const BParams &P, uint32_t b, float4 p,
                                                    float deg_i, float fx, float fy, float fz,
                                                    const float4 *__restrict__ old_force,
                                                    float4 *__restrict__ force,
                                                    float2 *__restrict__ stats) {
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
    stats[b] = make_float2(deg_i * len3(o.x - fx, o.y - fy, o.z - fz),
                           deg_i * 0.5f * len3(o.x + fx, o.y + fy, o.z + fz));
}

// Walk of the body's own graph tree (fa2_barnes_hut.wgsl:108-156), attraction
// over rows of at most kBHeavyAtt entries (four gathers in flight, list order
// kept), gravity and stats.  Longer rows leave the repulsion sum in force[b]
// for k_b_force_heavy.
template <int BS>
__global__ __launch_bounds__(BS) void k_b_force(const BParams P, const GFrame *fr,
                                                const BGlobal *gl,
                                                const uint32_t *__restrict__ gid,
                                                const float4 *__restrict__ pos,
                                                const float4 *__restrict__ old_force,
                                                float4 *__restrict__ force,
                                                const uint32_t *__restrict__ deg,
                                                const uint32_t *__restrict__ att_off,
                                                const uint32_t *__restrict__ att_adj,
                                                const uint32_t *__restrict__ perm,
                                                const Node *__restrict__ nodes,
                                                float2 *__restrict__ stats) {
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const uint32_t b = perm[t];
    if (b >= P.n) return;
    const uint32_t g = gid[b];
    const float4 p = pos[b];
    const float deg_i = (float)deg[b] + 1.0f;
    const float krd = P.scaling_ratio * deg_i;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const uint32_t N = fr[g].n_tree;
    if (N > 0) {
        // the graph's tree is nodes [start, start + N); its last rope is kEnd
        const uint32_t start = fr[g].tree_start, end = start + N;
        const uint32_t total = min(gl->n_tree_total, P.cap);  // uniform bound of the buffer
        if (end <= total && tree_reducible(total, fr[g].root[3], P.theta2) && finite3(p)) {
            if (P.fastdiv)
                walk_reduced<true>(nodes, start, end, total, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
            else
                walk_reduced<false>(nodes, start, end, total, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
        } else {
            (void)walk_exact(nodes, start, min(end, P.cap), p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
        }
    }
    const uint32_t e0 = att_off[b], e1 = att_off[b + 1];
    if (e1 - e0 > kBHeavyAtt) {
        force[b] = make_float4(fx, fy, fz, 0.0f);
        return;
    }
    uint32_t e = e0;
    for (; e + 4 <= e1; e += 4) {
        const uint32_t j0 = att_adj[e], j1 = att_adj[e + 1], j2 = att_adj[e + 2], j3 = att_adj[e + 3];
        const float4 q[4] = {pos[j0], pos[j1], pos[j2], pos[j3]};
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float tx, ty, tz;
            b_attraction_term(P, p, q[u], tx, ty, tz);
            fx = fx + tx;
            fy = fy + ty;
            fz = fz + tz;
        }
    }
    for (; e < e1; e++) {
        float tx, ty, tz;
        b_attraction_term(P, p, pos[att_adj[e]], tx, ty, tz);
        fx = fx + tx;
        fy = fy + ty;
        fz = fz + tz;
    }
    b_gravity_and_stats(P, b, p, deg_i, fx, fy, fz, old_force, force, stats);
}

// Attraction terms of the rows longer than kBHeavyAtt, one thread per entry,
// into three component arrays (rows start on multiples of 4 entries).
__global__ __launch_bounds__(kBT) void k_b_heavy_terms(const BParams P, uint32_t n_ent,
                                                       const uint32_t *__restrict__ hbody,
                                                       const uint32_t *__restrict__ hadj,
                                                       const float4 *__restrict__ pos,
                                                       float *__restrict__ hterm, uint32_t stride) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_ent) return;
    float tx, ty, tz;
    b_attraction_term(P, pos[hbody[q]], pos[hadj[q]], tx, ty, tz);
    hterm[q] = tx;
    hterm[stride + q] = ty;
    hterm[2 * (size_t)stride + q] = tz;
}

// One wave per long row: the terms are staged in LDS with coalesced loads and
// lanes 0, 1, 2 add the x, y, z terms to the repulsion sum in list order.
__global__ __launch_bounds__(64) void k_b_force_heavy(const BParams P,
                                                     const uint32_t *__restrict__ heavy,
                                                     const uint32_t *__restrict__ hoff,
                                                     const uint32_t *__restrict__ hlen,
                                                     const float *__restrict__ hterm,
                                                     uint32_t stride,
                                                     const float4 *__restrict__ pos,
                                                     const float4 *__restrict__ old_force,
                                                     float4 *__restrict__ force,
                                                     const uint32_t *__restrict__ deg,
                                                     float2 *__restrict__ stats) {
    __shared__ __align__(16) float st[3][kBHeavyStage];
    const uint32_t lane = threadIdx.x;
    const uint32_t b = heavy[blockIdx.x];
    const float4 f0 = force[b];
    float run = lane == 0 ? f0.x : lane == 1 ? f0.y : f0.z;
    const uint32_t e0 = hoff[blockIdx.x], len = hlen[blockIdx.x];
    for (uint32_t c0 = 0; c0 < len; c0 += kBHeavyStage) {
        const uint32_t cnt = min(kBHeavyStage, len - c0);
        const uint32_t nv4 = (cnt + 3u) >> 2;
#pragma unroll
        for (uint32_t comp = 0; comp < 3; comp++) {
            const float4 *src = reinterpret_cast<const float4 *>(hterm + (size_t)comp * stride + e0 + c0);
            float4 *dst = reinterpret_cast<float4 *>(st[comp]);
            for (uint32_t v = lane; v < nv4; v += 64) dst[v] = src[v];
        }
        __syncthreads();
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
        b_gravity_and_stats(P, b, pos[b], (float)deg[b] + 1.0f, fx, fy, fz, old_force, force, stats);
}

__global__ __launch_bounds__(256) void k_b_speed(const BParams P, GFrame *fr,
                                                 const uint32_t *__restrict__ base,
                                                 const uint32_t *__restrict__ count,
                                                 const double *__restrict__ est_jt,
                                                 const double *__restrict__ min_jt,
                                                 const float2 *__restrict__ stats) {
    // Each thread sums its share of the graph's stats exactly in registers (no
    // atomics), then the limbs are added across the block.
    __shared__ unsigned long long red[4][2 * kLimbs];
    const uint32_t g = blockIdx.x;
    const uint32_t b0 = base[g], ng = count[g];
    unsigned long long sw[kLimbs] = {}, tr[kLimbs] = {};
    for (uint32_t i = threadIdx.x; i < ng; i += blockDim.x) {
        const float2 s = stats[b0 + i];
        if (finitef(s.x) && finitef(s.y)) {
            exact_reg_add(sw, s.x);
            exact_reg_add(tr, s.y);
        }
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t k = 0; k < kLimbs; k++) {
        unsigned long long a = sw[k], c = tr[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            c += __shfl_xor(c, o, 64);
        }
        if (lane == 0) {
            red[w][k] = a;
            red[w][kLimbs + k] = c;
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    unsigned long long tot[2 * kLimbs];
#pragma unroll
    for (uint32_t k = 0; k < 2 * kLimbs; k++) tot[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
    exact_normalise(tot);
    exact_normalise(tot + kLimbs);
    const double ts = exact_to_double(tot), tt = exact_to_double(tot + kLimbs);
    fr[g].sum_swing = ts;
    fr[g].sum_traction = tt;
    float speed = fr[g].speed, eff = fr[g].speed_eff;
    if (!(ts > 0.0) || !(tt > 0.0) || ng == 0) return;
    const double nn = (double)ng, jt0 = (double)P.jitter_tolerance;
    double jt = jt0 * fmax(min_jt[g], fmin(10.0, est_jt[g] * tt / (nn * nn)));
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
    fr[g].speed = (float)(sp + fmin(target - sp, 0.5 * sp));
    fr[g].speed_eff = eff;
}

__global__ __launch_bounds__(kBT) void k_b_apply(const BParams P, const GFrame *fr,
                                                 const uint32_t *__restrict__ gid,
                                                 const float4 *__restrict__ pos_in,
                                                 float4 *__restrict__ pos_out,
                                                 const float4 *__restrict__ force,
                                                 float4 *__restrict__ old_force,
                                                 const uint32_t *__restrict__ deg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const float speed = fr[gid[i]].speed;
    const float4 f = force[i];
    const float4 o = old_force[i];
    const float deg_i = (float)deg[i] + 1.0f;
    const float swinging = deg_i * len3(o.x - f.x, o.y - f.y, o.z - f.z);
    float factor = speed / (1.0f + sqrtf(speed * swinging));
    const float df = len3(f.x, f.y, f.z);
    if (P.max_displacement > 0.0f && df > 1.0e-9f) factor = fminf(factor, P.max_displacement / df);
    const float k = factor * P.time_step;
    const float4 p = pos_in[i];
    // .w carries the mass for k_b_emit (one gather there instead of two)
    pos_out[i] = make_float4(p.x + f.x * k, p.y + f.y * k, p.z + f.z * k, fmaxf(deg_i, 1e-3f));
    old_force[i] = make_float4(f.x, f.y, f.z, 0.0f);
}

__global__ void k_b_init_speed(GFrame *fr, uint32_t G) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < G) {
        fr[g].speed = 1.0f;
        fr[g].speed_eff = 1.0f;
    }
}

// ---- persistent schedule: one workgroup owns a graph for every iteration ----
//
// The graphs of a batch are independent, so a workgroup can lay one out from
// start to finish with nothing but its own barriers: box, keys, sort, tree,
// centre of mass, force, speed and apply of iteration t, then iteration t + 1,
// all in one launch.  Workgroups stay resident and take graphs from a queue in
// decreasing size (longest first, so the big graphs do not finish last); while
// one workgroup builds its tree, the others on the CU walk theirs.  Every phase
// performs the operations of the phased kernels above on the same values, so
// each graph still evolves bit-identically to its standalone run.
//
// Per graph g the node array holds a static region of pg[g].ncap nodes from
// pg[g].nstart (room for any tree of the graph: n leaves and at most n/2 + 1
// cells per depth), and ropes and byte offsets are relative to the region, so
// the walk's buffer resource spans only the graph's own nodes.

struct PGraph {
    uint32_t base;    // first body
    uint32_t count;   // bodies
    uint32_t nstart;  // first node of the graph's node region
    uint32_t ncap;    // nodes in the region
    uint32_t hstart;  // first of the graph's long attraction rows in prow
    uint32_t hcount;  // long rows of the graph
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) {
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}

// Lanes of the wave holding the same 8-bit digit as this lane (valid lanes only);
// every lane of the wave must be active.
__device__ __forceinline__ uint64_t p_match8(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
        const bool on = (d >> bit) & 1u;
        const uint64_t bl = __ballot(on);
        m &= on ? bl : ~bl;
    }
    return m;
}

// Exclusive prefix sum of one value per thread over the workgroup (thread
// order); *total gets the sum.  Uses s_w[T / 64]; ends with a barrier, so s_w
// may be reused at once.
template <uint32_t T>
__device__ __forceinline__ uint32_t p_excl_scan(uint32_t v, uint32_t *s_w, uint32_t *total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += t;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t wp = 0, tot = 0;
#pragma unroll
    for (uint32_t q = 0; q < T / 64; q++) {
        const uint32_t x = s_w[q];
        wp += q < w ? x : 0u;
        tot += x;
    }
    *total = tot;
    __syncthreads();
    return incl - v + wp;
}

// Maximum of one value per thread over the workgroup (same barrier rules).
template <uint32_t T>
__device__ __forceinline__ uint32_t p_max(uint32_t v, uint32_t *s_w) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
    if (lane == 0) s_w[w] = v;
    __syncthreads();
    uint32_t m = 0;
#pragma unroll
    for (uint32_t q = 0; q < T / 64; q++) m = max(m, s_w[q]);
    __syncthreads();
    return m;
}

// One stable pass of the workgroup's LSD radix sort: items [0, ng) of (ks, vs)
// by the 8-bit digit at bit sh, into (kd, vd).  Wave w owns the contiguous tile
// [w * tile, (w + 1) * tile) and goes through it 64 items per round, so the
// order of (wave, round, lane) is the item order; an item's place is the count
// of smaller digits, plus its digit in earlier waves, earlier rounds and lower
// lanes.  s_hist holds 256 * (T / 64) counters, digit-major.
template <uint32_t T>
__device__ __forceinline__ void p_radix_pass(const uint64_t *__restrict__ ks,
                                             const uint32_t *__restrict__ vs,
                                             uint64_t *__restrict__ kd, uint32_t *__restrict__ vd,
                                             uint32_t ng, uint32_t sh, uint32_t *s_hist,
                                             uint32_t *s_w) {
    constexpr uint32_t NW = T / 64;
    static_assert(256 * NW == 4 * T, "four counters per thread in the scan");
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t tile = (ng + NW - 1) / NW;
    const uint32_t t0 = min(w * tile, ng), t1 = min(t0 + tile, ng);
    const uint32_t rounds = (t1 - t0 + 63) / 64;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t q = threadIdx.x; q < 256 * NW; q += T) s_hist[q] = 0;
    __syncthreads();
    for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t i = t0 + r * 64 + lane;
        const bool valid = i < t1;
        const uint32_t d = valid ? (uint32_t)(ks[i] >> sh) & 0xFFu : 0u;
        const uint64_t peers = p_match8(d, valid);
        if (valid && (peers & below) == 0) s_hist[d * NW + w] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    uint32_t e[4], s = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
        e[j] = s_hist[4 * threadIdx.x + j];
        s += e[j];
    }
    uint32_t tot;
    uint32_t run = p_excl_scan<T>(s, s_w, &tot);
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
        s_hist[4 * threadIdx.x + j] = run;
        run += e[j];
    }
    __syncthreads();
    for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t i = t0 + r * 64 + lane;
        const bool valid = i < t1;
        uint64_t k = 0;
        uint32_t v = 0, d = 0;
        if (valid) {
            k = ks[i];
            v = vs[i];
            d = (uint32_t)(k >> sh) & 0xFFu;
        }
        const uint64_t peers = p_match8(d, valid);
        uint32_t at = 0;
        if (valid) at = s_hist[d * NW + w] + (uint32_t)__popcll(peers & below);
        // every lane has read the cursor before the group's lowest lane moves it
        // past the group
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            kd[at] = k;
            vd[at] = v;
            if ((peers & below) == 0) s_hist[d * NW + w] = at + (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
}

// LDS of the persistent kernel.
template <uint32_t T>
struct PShared {
    static constexpr uint32_t NW = T / 64;
    uint32_t hist[256 * NW];
    float st[NW][3][64];                    // long-row terms, one wave's chunk
    unsigned long long lim[NW][2 * kLimbs];  // swing / traction limbs per wave
    unsigned long long tot[2 * kLimbs];
    uint32_t w[NW];
    float box[NW][6];
    uint64_t kbits[NW][2];
    float speed, eff;
    uint32_t g;
};

template <uint32_t T>
__global__ __launch_bounds__(T) void k_p_run(
    const BParams P, uint32_t n_iters, uint32_t cur0, const PGraph *__restrict__ pg,
    const uint32_t *__restrict__ order, uint32_t *__restrict__ queue, GFrame *fr,
    const double *__restrict__ est_jt, const double *__restrict__ min_jt, float4 *pos0,
    float4 *pos1, float4 *__restrict__ old_force, float4 *__restrict__ force,
    float2 *__restrict__ stats, const uint32_t *__restrict__ deg,
    const uint32_t *__restrict__ att_off, const uint32_t *__restrict__ att_adj,
    const uint32_t *__restrict__ prow, uint64_t *__restrict__ keys_a,
    uint64_t *__restrict__ keys_b, uint32_t *__restrict__ vals_a, uint32_t *__restrict__ vals_b,
    uint32_t *__restrict__ loc, uint32_t *__restrict__ cinfo, Node *__restrict__ nodes_all,
    uint8_t *__restrict__ depth_all) {
    constexpr uint32_t NW = T / 64;
    __shared__ PShared<T> S;
    // Values the whole workgroup shares are taken through readfirstlane
    // (uni/unif), so they live in scalar registers: this kernel holds every
    // phase at once, and vector registers decide its occupancy.
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = uni(tid >> 6);
    for (;;) {
        if (tid == 0) S.g = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t q = uni(S.g);
        __syncthreads();
        if (q >= P.n_graphs) break;
        const uint32_t g = order[q];
        const PGraph pgr = pg[g];
        const uint32_t b0 = pgr.base, ng = pgr.count;
        Node *__restrict__ nodes = nodes_all + pgr.nstart;
        uint8_t *__restrict__ ndep = depth_all + pgr.nstart;
        if (tid == 0) {
            S.speed = fr[g].speed;
            S.eff = fr[g].speed_eff;
        }
        __syncthreads();
        for (uint32_t it = 0; it < n_iters; it++) {
            float4 *pin = ((cur0 + it) & 1u) ? pos1 : pos0;
            float4 *pout = ((cur0 + it) & 1u) ? pos0 : pos1;
            // 1. box over finite coordinates (k_b_bbox), root cube
            float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t i = tid; i < ng; i += T) {
                const float4 p = pin[b0 + i];
                const float v[3] = {p.x + 0.0f, p.y + 0.0f, p.z + 0.0f};
#pragma unroll
                for (int k = 0; k < 3; k++)
                    if (finitef(v[k])) {
                        mn[k] = fminf(mn[k], v[k]);
                        mx[k] = fmaxf(mx[k], v[k]);
                    }
            }
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
                    mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
                }
            if (lane == 0)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    S.box[w][k] = mn[k];
                    S.box[w][3 + k] = mx[k];
                }
            __syncthreads();
            uint32_t bbmin[3], bbmax[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float a = S.box[0][k], c = S.box[0][3 + k];
                for (uint32_t u = 1; u < NW; u++) {
                    a = fminf(a, S.box[u][k]);
                    c = fmaxf(c, S.box[u][3 + k]);
                }
                bbmin[k] = uni(finitef(a) ? f2ord(a) : 0xFF800000u);
                bbmax[k] = uni(finitef(c) ? f2ord(c) : ~0xFF800000u);
            }
            float r4[4];
            compute_root(bbmin, bbmax, r4);
#pragma unroll
            for (int k = 0; k < 4; k++) r4[k] = unif(r4[k]);
            // 2. keys (k_b_keys): invalid bodies get ~0 and sort last
            uint32_t inval = 0;
            uint64_t kand = ~0ull, kor = 0ull;
            for (uint32_t i = tid; i < ng; i += T) {
                const float4 p = pin[b0 + i];
                uint64_t key;
                if (finitef(p.x) && finitef(p.y) && finitef(p.z)) {
                    key = octant_key(r4, p.x, p.y, p.z);
                } else {
                    key = ~0ull;
                    inval++;
                }
                keys_a[b0 + i] = key;
                vals_a[b0 + i] = b0 + i;
                kand &= key;
                kor |= key;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                kand &= __shfl_xor(kand, o, 64);
                kor |= __shfl_xor(kor, o, 64);
                inval += __shfl_xor(inval, o, 64);
            }
            if (lane == 0) {
                S.kbits[w][0] = kand;
                S.kbits[w][1] = kor;
                S.w[w] = inval;
            }
            __syncthreads();
            uint64_t vary = 0;
            uint32_t nvalid = ng;
            for (uint32_t u = 0; u < NW; u++) {
                vary |= S.kbits[u][0] ^ S.kbits[u][1];
                nvalid -= S.w[u];
            }
            vary = uni64(vary);
            nvalid = uni(nvalid);
            __syncthreads();
            // 3. stable sort by key (ties keep body order): the digits that vary
            uint64_t *ks = keys_a + b0, *kt = keys_b + b0;
            uint32_t *vs = vals_a + b0, *vt = vals_b + b0;
            for (uint32_t sh = 0; sh < 64; sh += 8) {
                if (((vary >> sh) & 0xFFull) == 0) continue;
                p_radix_pass<T>(ks, vs, kt, vt, ng, sh, S.hist, S.w);
                uint64_t *tk = ks;
                ks = kt;
                kt = tk;
                uint32_t *tv = vs;
                vs = vt;
                vt = tv;
            }
            // 4. nodes per sorted body (k_b_count_scan): opened cells + its leaf,
            //    the graph's lone body opens the root; exclusive scan in sorted order
            const uint32_t nv = nvalid;
            const uint32_t chunk = (nv + T - 1) / T;
            const uint32_t c0 = min(tid * chunk, nv), c1 = min(c0 + chunk, nv);
            uint32_t s = 0, md = 0;
            for (uint32_t i = c0; i < c1; i++) {
                int lp, li;
                if (nv == 1) {
                    lp = -1;
                    li = 0;
                } else {
                    const uint64_t k = ks[i];
                    lp = i > 0 ? lcp_digits(ks[i - 1], k) : -1;
                    li = i + 1 < nv ? lcp_digits(k, ks[i + 1]) : -1;
                }
                cinfo[b0 + i] = (uint32_t)(lp + 1) | ((uint32_t)(li + 1) << 8);
                s += (li > lp ? (uint32_t)(li - lp) : 0u) + 1u;
                md = max(md, (uint32_t)(li + 1));
            }
            uint32_t ntree;
            uint32_t run = p_excl_scan<T>(s, S.w, &ntree);
            ntree = uni(ntree);
            const uint32_t dmax1 = uni(p_max<T>(md, S.w));  // deepest cell depth + 1 (0: no cell)
            for (uint32_t i = c0; i < c1; i++) {
                loc[b0 + i] = run;
                const uint32_t ci = cinfo[b0 + i];
                const int lp = (int)(ci & 0xFFu) - 1, li = (int)(ci >> 8) - 1;
                run += (li > lp ? (uint32_t)(li - lp) : 0u) + 1u;
            }
            __syncthreads();
            const uint32_t N = min(ntree, pgr.ncap);
            // 5. pre-order emission (k_b_emit): cells opened by body i at depths
            //    lp+1 .. li, then its leaf
            for (uint32_t i = tid; i < nv; i += T) {
                const uint32_t ci = cinfo[b0 + i];
                const int lp = (int)(ci & 0xFFu) - 1, li = (int)(ci >> 8) - 1;
                const uint32_t c = li > lp ? (uint32_t)(li - lp) : 0u;
                const uint32_t pos0 = loc[b0 + i];
                if ((uint64_t)pos0 + c >= pgr.ncap) continue;
                const uint64_t k = ks[i];
                for (uint32_t t = 0; t < c; t++) {
                    const int d = lp + 1 + (int)t;
                    float h = r4[3];
                    for (int u = 0; u < d; u++) h = h * 0.5f;
                    const uint32_t j = cell_end(ks, i, nv, k, d);
                    const bool last = j + 1 >= nv;
                    const uint32_t end = last ? ntree : loc[b0 + j + 1];
                    const float sz = h * 2.0f;
                    nodes[pos0 + t].b = cell_b(sz * sz, last ? kEnd : end);
                    ndep[pos0 + t] = (uint8_t)d;
                }
                const uint32_t leaf = pos0 + c;
                const uint32_t b = vs[i];
                nodes[leaf].a = pin[b];  // .w: the body's mass
                nodes[leaf].b = leaf_b(b, (i + 1 < nv) ? leaf + 1 : kEnd);
                ndep[leaf] = (uint8_t)(0x80 | ((lp > li ? lp : li) + 1));
            }
            __syncthreads();
            // 6. centre of mass, deepest cells first, children in octant order
            for (int d = (int)dmax1 - 1; d >= 0; d--) {
                for (uint32_t i = tid; i < nv; i += T) {
                    const uint32_t ci = cinfo[b0 + i];
                    const int lp = (int)(ci & 0xFFu) - 1, li = (int)(ci >> 8) - 1;
                    if (d > lp && d <= li) {
                        const uint32_t node = loc[b0 + i] + (uint32_t)(d - lp - 1);
                        if (node < N) com_of_cell(nodes, node, N);
                    }
                }
                __syncthreads();
            }
            // 7. force (k_b_force): walk, attraction of the short rows, gravity,
            //    stats; long rows leave the repulsion sum for step 8
            const bool reducible = tree_reducible(N, r4[3], P.theta2);
            for (uint32_t i = tid; i < ng; i += T) {
                const uint32_t b = vs[i];
                const float4 p = pin[b];
                const float deg_i = (float)deg[b] + 1.0f;
                const float krd = P.scaling_ratio * deg_i;
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
                if (N > 0) {
                    if (reducible && finite3(p)) {
                        if (P.fastdiv)
                            walk_reduced<true>(nodes, 0, N, N, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
                        else
                            walk_reduced<false>(nodes, 0, N, N, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
                    } else {
                        (void)walk_exact(nodes, 0, N, p.x, p.y, p.z, b, krd, P.theta2, fx, fy, fz);
                    }
                }
                const uint32_t e0 = att_off[b], e1 = att_off[b + 1];
                if (e1 - e0 > kBHeavyAtt) {
                    force[b] = make_float4(fx, fy, fz, 0.0f);
                    continue;
                }
                uint32_t e = e0;
                for (; e + 4 <= e1; e += 4) {
                    const uint32_t j0 = att_adj[e], j1 = att_adj[e + 1], j2 = att_adj[e + 2],
                                   j3 = att_adj[e + 3];
                    const float4 qv[4] = {pin[j0], pin[j1], pin[j2], pin[j3]};
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        float tx, ty, tz;
                        b_attraction_term(P, p, qv[u], tx, ty, tz);
                        fx = fx + tx;
                        fy = fy + ty;
                        fz = fz + tz;
                    }
                }
                for (; e < e1; e++) {
                    float tx, ty, tz;
                    b_attraction_term(P, p, pin[att_adj[e]], tx, ty, tz);
                    fx = fx + tx;
                    fy = fy + ty;
                    fz = fz + tz;
                }
                b_gravity_and_stats(P, b, p, deg_i, fx, fy, fz, old_force, force, stats);
            }
            __syncthreads();
            // 8. long rows (k_b_heavy_terms + k_b_force_heavy): one wave per row,
            //    64 terms per round into LDS, lanes 0, 1, 2 add x, y, z in list order
            for (uint32_t r = w; r < pgr.hcount; r += NW) {
                const uint32_t b = prow[pgr.hstart + r];
                const float4 p = pin[b];
                const float4 f0 = force[b];
                float acc = lane == 0 ? f0.x : lane == 1 ? f0.y : f0.z;
                const uint32_t e0 = att_off[b], len = att_off[b + 1] - e0;
                for (uint32_t c0r = 0; c0r < len; c0r += 64) {
                    const uint32_t cnt = min(64u, len - c0r);
                    if (lane < cnt) {
                        float tx, ty, tz;
                        b_attraction_term(P, p, pin[att_adj[e0 + c0r + lane]], tx, ty, tz);
                        S.st[w][0][lane] = tx;
                        S.st[w][1][lane] = ty;
                        S.st[w][2][lane] = tz;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (lane < 3)
                        for (uint32_t k = 0; k < cnt; k++) acc = acc + S.st[w][lane][k];
                    __builtin_amdgcn_wave_barrier();  // read before the next round writes
                }
                const float fx = __shfl(acc, 0, 64), fy = __shfl(acc, 1, 64), fz = __shfl(acc, 2, 64);
                if (lane == 0)
                    b_gravity_and_stats(P, b, p, (float)deg[b] + 1.0f, fx, fy, fz, old_force, force, stats);
            }
            __syncthreads();
            // 9. exact swing / traction sums and AdaptiveSpeed (k_b_speed)
            {
                // each wave adds its bodies' values into its own LDS limbs
                // (integer adds: the order does not matter)
                if (lane < 2 * kLimbs) S.lim[w][lane] = 0ull;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (uint32_t i = tid; i < ng; i += T) {
                    const float2 st = stats[b0 + i];
                    if (finitef(st.x) && finitef(st.y)) {
                        exact_acc_add(S.lim[w], st.x);
                        exact_acc_add(S.lim[w] + kLimbs, st.y);
                    }
                }
                __syncthreads();
                if (tid < 2 * kLimbs) {
                    unsigned long long a = 0;
                    for (uint32_t u = 0; u < NW; u++) a += S.lim[u][tid];
                    S.tot[tid] = a;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (tid == 0) {
                    // the limbs stay in LDS: this one thread's arithmetic does not
                    // claim registers from every other phase of the kernel
                    exact_normalise(S.tot);
                    exact_normalise(S.tot + kLimbs);
                    const double ts = exact_to_double(S.tot), tt = exact_to_double(S.tot + kLimbs);
                    fr[g].sum_swing = ts;
                    fr[g].sum_traction = tt;
                    float speed = S.speed, eff = S.eff;
                    if (ts > 0.0 && tt > 0.0 && ng != 0) {
                        const double nn = (double)ng, jt0 = (double)P.jitter_tolerance;
                        double jt = jt0 * fmax(min_jt[g], fmin(10.0, est_jt[g] * tt / (nn * nn)));
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
                        S.speed = (float)(sp + fmin(target - sp, 0.5 * sp));
                        S.eff = eff;
                    }
                }
                __syncthreads();
            }
            // 10. apply (k_b_apply)
            const float speed = unif(S.speed);
            for (uint32_t i = tid; i < ng; i += T) {
                const uint32_t b = b0 + i;
                const float4 f = force[b];
                const float4 o = old_force[b];
                const float deg_i = (float)deg[b] + 1.0f;
                const float swinging = deg_i * len3(o.x - f.x, o.y - f.y, o.z - f.z);
                float factor = speed / (1.0f + sqrtf(speed * swinging));
                const float df = len3(f.x, f.y, f.z);
                if (P.max_displacement > 0.0f && df > 1.0e-9f)
                    factor = fminf(factor, P.max_displacement / df);
                const float k = factor * P.time_step;
                const float4 p = pin[b];
                pout[b] = make_float4(p.x + f.x * k, p.y + f.y * k, p.z + f.z * k, fmaxf(deg_i, 1e-3f));
                old_force[b] = make_float4(f.x, f.y, f.z, 0.0f);
            }
            __syncthreads();
        }
        if (tid == 0) {
            fr[g].speed = S.speed;
            fr[g].speed_eff = S.eff;
        }
    }
}

__global__ void k_p_queue_reset(uint32_t *queue) { *queue = 0; }

inline uint32_t cdivb(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

constexpr uint32_t kSchedAuto = 0, kSchedPhased = 1, kSchedPersistent = 2;
// The persistent schedule (k_p_run) is switched off.  Its first GPU test (24
// mixed graphs, 256-thread workgroups) ended in a memory access fault.  The
// kernel source gives the oracle's bits when its workgroups are run as threads
// on the CPU, under address and undefined-behaviour checks, and no reading of
// the code or of its ISA has shown an out-of-bounds access or a missing wait,
// so the fault's cause is unknown and the kernel is not launched.  The
// autotuner then keeps the phased schedule.
constexpr bool kPersistentEnabled = false;

class Fa2BhBatch {
  public:
    explicit Fa2BhBatch(const jc_fa2bh_settings &s) : settings_(s) {}
    ~Fa2BhBatch() { teardown(); }

    Status init(uint32_t G, const uint32_t *n_nodes, const uint64_t *nnz, const uint32_t *offsets,
                const uint32_t *neighbors, const float *xyz) {
        JC_TRY(require_device(nullptr));
        if (G == 0) return Status::err(JC_E_INVALID, "empty batch");
        teardown();
        G_ = G;
        std::vector<uint32_t> base(G), count(G), gid;
        std::vector<double> est(G), mjt(G);
        uint64_t ntot = 0, off_pos = 0, nb_pos = 0;
        for (uint32_t g = 0; g < G; g++) {
            base[g] = (uint32_t)ntot;
            count[g] = n_nodes[g];
            ntot += n_nodes[g];
            if (ntot >= 0x7FFFFFFFull) return Status::err(JC_E_INVALID, "batch too large");
        }
        n_ = (uint32_t)ntot;
        gid.resize(n_ ? n_ : 1);
        std::vector<uint32_t> degs(n_ ? n_ : 1, 0), aoff((size_t)n_ + 1, 0), adj;
        // per graph: validate, degrees and attraction lists (fa2_bh.rs:264-287)
        std::vector<uint64_t> cnt((size_t)n_ + 1, 0);
        for (uint32_t g = 0; g < G; g++) {
            const uint32_t ng = n_nodes[g];
            const uint32_t *off = offsets + off_pos;
            const uint32_t *nb = neighbors + nb_pos;
            CsrView v{ng, off, nb, nnz[g]};
            Status s = validate_csr(v);
            if (s.failed())
                return Status::err(JC_E_INVALID, "graph " + std::to_string(g) + ": " + last_error());
            est[g] = 0.05 * std::sqrt((double)ng);
            mjt[g] = std::sqrt(est[g]);
            for (uint32_t v2 = 0; v2 < ng; v2++) {
                const uint32_t gv = base[g] + v2;
                gid[gv] = g;
                for (uint64_t e = off[v2]; e < off[v2 + 1]; e++) {
                    const uint32_t u = nb[e];
                    if (u == v2) continue;
                    degs[gv]++;
                    if (v2 < u) {
                        cnt[gv + 1]++;
                        cnt[base[g] + u + 1]++;
                    }
                }
            }
            off_pos += (uint64_t)ng + 1;
            nb_pos += nnz[g];
        }
        for (uint32_t v2 = 0; v2 < n_; v2++) cnt[v2 + 1] += cnt[v2];
        if (cnt[n_] > 0xFFFFFFFFull) return Status::err(JC_E_INVALID, "too many edges");
        for (uint32_t v2 = 0; v2 <= n_; v2++) aoff[v2] = (uint32_t)cnt[v2];
        adj.resize(cnt[n_] ? cnt[n_] : 1);
        off_pos = nb_pos = 0;
        for (uint32_t g = 0; g < G; g++) {
            const uint32_t ng = n_nodes[g];
            const uint32_t *off = offsets + off_pos;
            const uint32_t *nb = neighbors + nb_pos;
            for (uint32_t v2 = 0; v2 < ng; v2++)
                for (uint64_t e = off[v2]; e < off[v2 + 1]; e++) {
                    const uint32_t u = nb[e];
                    if (u == v2 || !(v2 < u)) continue;
                    adj[cnt[base[g] + v2]++] = base[g] + u;
                    adj[cnt[base[g] + u]++] = base[g] + v2;
                }
            off_pos += (uint64_t)ng + 1;
            nb_pos += nnz[g];
        }
        // rows longer than kBHeavyAtt, longest first, laid end to end from
        // multiples of 4 entries (padding: the body as its own neighbour, a zero
        // term loaded but never added)
        std::vector<uint32_t> heavy;
        for (uint32_t v2 = 0; v2 < n_; v2++)
            if (aoff[v2 + 1] - aoff[v2] > kBHeavyAtt) heavy.push_back(v2);
        std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t a, uint32_t b) {
            return aoff[a + 1] - aoff[a] > aoff[b + 1] - aoff[b];
        });
        n_heavy_ = (uint32_t)heavy.size();
        std::vector<uint32_t> hoff(n_heavy_ ? n_heavy_ : 1, 0), hlen(n_heavy_ ? n_heavy_ : 1, 0),
            hbody, hadj;
        for (uint32_t r = 0; r < n_heavy_; r++) {
            const uint32_t v2 = heavy[r];
            hoff[r] = (uint32_t)hbody.size();
            hlen[r] = aoff[v2 + 1] - aoff[v2];
            for (uint32_t e = aoff[v2]; e < aoff[v2 + 1]; e++) {
                hbody.push_back(v2);
                hadj.push_back(adj[e]);
            }
            while (hbody.size() & 3u) {
                hbody.push_back(v2);
                hadj.push_back(v2);
            }
        }
        n_hent_ = (uint32_t)hbody.size();
        hstride_ = n_hent_;
        // persistent schedule: per-graph node regions, long rows graph-major,
        // graphs in decreasing size (stable)
        std::vector<PGraph> pgh(G);
        std::vector<uint32_t> prow;
        uint64_t nacc = 0, ncap_max = 0;
        for (uint32_t g = 0; g < G; g++) {
            const uint64_t nc = (uint64_t)count[g] + (uint64_t)(kLevels + 1) * (count[g] / 2 + 1) + 1;
            pgh[g].base = base[g];
            pgh[g].count = count[g];
            pgh[g].nstart = (uint32_t)std::min<uint64_t>(nacc, 0xFFFFFFFFull);
            pgh[g].ncap = (uint32_t)std::min<uint64_t>(nc, 0xFFFFFFFFull);
            pgh[g].hstart = (uint32_t)prow.size();
            for (uint32_t v2 = base[g]; v2 < base[g] + count[g]; v2++)
                if (aoff[v2 + 1] - aoff[v2] > kBHeavyAtt) prow.push_back(v2);
            pgh[g].hcount = (uint32_t)prow.size() - pgh[g].hstart;
            nacc += nc;
            ncap_max = std::max(ncap_max, nc);
        }
        std::vector<uint32_t> order(G);
        for (uint32_t g = 0; g < G; g++) order[g] = g;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return count[a] > count[b]; });
        // the walk addresses a graph's region with 32-bit byte offsets
        p_ok_ = ncap_max < kMaxNodes32;
        // the sort's composite key: ceil(log2 G) bits of graph id above ck_p_ bits
        // of the raw key, ck_bits_ in all (a multiple of 16); ck_p_ leaves each
        // graph at least about 32 prefixes per body, so runs of equal prefix stay short
        uint32_t gb = 0, cmax = 1;
        while ((1ull << gb) < G) gb++;
        for (uint32_t g = 0; g < G; g++) cmax = std::max(cmax, count[g]);
        uint32_t lc = 0;
        while ((1ull << lc) < cmax) lc++;
        const uint32_t pmin = std::min<uint32_t>(std::max<uint32_t>(16, lc + 5), 63);
        gb_ = gb;
        ck_min_ = std::min<uint32_t>((gb + pmin + 15) / 16 * 16, 64);
        if (want_ck_ && want_ck_ < ck_min_)
            return Status::err(JC_E_INVALID, "sort key width narrower than the batch needs");
        set_ck_bits(want_ck_ ? want_ck_ : ck_min_);

        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        const uint32_t nn = n_ ? n_ : 1;
        JC_TRY(pos_[0].alloc(nn));
        JC_TRY(pos_[1].alloc(nn));
        JC_TRY(old_force_.alloc(nn));
        JC_TRY(force_.alloc(nn));
        JC_TRY(stats_.alloc(nn));
        JC_TRY(deg_.alloc(nn));
        JC_TRY(gid_.alloc(nn));
        JC_TRY(base_.alloc(G));
        JC_TRY(count_.alloc(G));
        JC_TRY(est_.alloc(G));
        JC_TRY(mjt_.alloc(G));
        JC_TRY(att_off_.alloc((size_t)n_ + 1));
        JC_TRY(att_adj_.alloc(adj.size()));
        JC_TRY(keys_raw_.alloc(nn));
        JC_TRY(keys_.alloc(nn));
        JC_TRY(vals_.alloc(nn));
        JC_TRY(sort_.alloc(n_));
        nb_scan_ = cdivb(nn, kBScanTile);
        JC_TRY(loc_.alloc(nn));
        JC_TRY(blk_.alloc(nb_scan_ + 1));
        JC_TRY(blkd_.alloc(nb_scan_ + 1));
        // per graph: n_g leaves + at most n_g/2 cells per depth over 22 depths + 1
        cap_ = 0;
        for (uint32_t g = 0; g < G; g++) cap_ += (uint64_t)count[g] + (uint64_t)(kLevels + 1) * (count[g] / 2 + 1) + 1;
        cap_ += 8;
        if (cap_ >= 0xFFFFFFF0ull) return Status::err(JC_E_INVALID, "batch too large for u32 node ids");
        JC_TRY(nodes_.alloc(cap_));
        JC_TRY(heavy_.alloc(n_heavy_ ? n_heavy_ : 1));
        JC_TRY(hoff_.alloc(n_heavy_ ? n_heavy_ : 1));
        JC_TRY(hlen_.alloc(n_heavy_ ? n_heavy_ : 1));
        JC_TRY(hbody_.alloc(n_hent_ ? n_hent_ : 1));
        JC_TRY(hadj_.alloc(n_hent_ ? n_hent_ : 1));
        JC_TRY(hterm_.alloc(3 * (size_t)(hstride_ ? hstride_ : 4)));
        JC_TRY(node_depth_.alloc((cap_ + 31) & ~15ull));  // whole 16-byte loads in the COM pass
        JC_TRY(cskip_.alloc(cap_));
        // Cells that do not fit k_b_com_fused at one depth are disjoint and hold
        // more than kBFuseCap nodes each.
        big_max_ = (uint32_t)std::min<uint64_t>((uint64_t)(kLevels + 1 - kBFuseDepthMin) * (cap_ / kBFuseCap + 1),
                                                0xFFFFFFFFull);
        JC_TRY(big_.alloc(big_max_));
        JC_TRY(frame_.alloc(G));
        JC_TRY(global_.alloc(1));
        std::vector<float4> p4(nn, make_float4(0, 0, 0, 0));
        for (uint32_t i = 0; i < n_; i++)
            p4[i] = make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2],
                                std::fmax((float)degs[i] + 1.0f, 1e-3f));  // w: the mass
        auto up = [&](void *d, const void *h, size_t bytes) -> Status {
            JC_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream_));
            return Status::ok();
        };
        JC_TRY(up(pos_[0].p, p4.data(), nn * sizeof(float4)));
        JC_TRY(up(pos_[1].p, p4.data(), nn * sizeof(float4)));
        JC_HIP(hipMemsetAsync(old_force_.p, 0, nn * sizeof(float4), stream_));
        JC_TRY(up(deg_.p, degs.data(), nn * 4));
        JC_TRY(up(gid_.p, gid.data(), nn * 4));
        JC_TRY(up(base_.p, base.data(), G * 4));
        JC_TRY(up(count_.p, count.data(), G * 4));
        JC_TRY(up(est_.p, est.data(), G * 8));
        JC_TRY(up(mjt_.p, mjt.data(), G * 8));
        JC_TRY(up(att_off_.p, aoff.data(), ((size_t)n_ + 1) * 4));
        JC_TRY(up(att_adj_.p, adj.data(), adj.size() * 4));
        if (n_heavy_) {
            JC_TRY(up(heavy_.p, heavy.data(), n_heavy_ * 4));
            JC_TRY(up(hoff_.p, hoff.data(), n_heavy_ * 4));
            JC_TRY(up(hlen_.p, hlen.data(), n_heavy_ * 4));
            JC_TRY(up(hbody_.p, hbody.data(), n_hent_ * 4));
            JC_TRY(up(hadj_.p, hadj.data(), n_hent_ * 4));
        }
        JC_TRY(pg_.alloc(G));
        JC_TRY(order_.alloc(G));
        JC_TRY(prow_.alloc(prow.size()));
        JC_TRY(queue_.alloc(1));
        JC_TRY(up(pg_.p, pgh.data(), G * sizeof(PGraph)));
        JC_TRY(up(order_.p, order.data(), G * 4));
        if (!prow.empty()) JC_TRY(up(prow_.p, prow.data(), prow.size() * 4));
        JC_HIP(hipMemsetAsync(frame_.p, 0, G * sizeof(GFrame), stream_));
        JC_HIP(hipMemsetAsync(global_.p, 0, sizeof(BGlobal), stream_));
        k_b_init_speed<<<cdivb(G, 256), 256, 0, stream_>>>(frame_.p, G);
        JC_HIP(hipGetLastError());
        JC_HIP(hipStreamSynchronize(stream_));
        base_h_ = base;
        count_h_ = count;
        cur_ = 0;
        iteration_ = 0;
        inited_ = true;
        params_dirty_ = true;
        return Status::ok();
    }

    Status step(uint32_t n_iters, float *out_xyz) {
        if (!inited_) return Status::err(JC_E_STATE, "batch step before init");
        refresh_params();
        if (n_ && n_iters) {
            if (!tuned_) JC_TRY(autotune());
            if (sched_ == kSchedPersistent) {
                JC_TRY(launch_persistent(n_iters, p_threads_));
                cur_ ^= n_iters & 1u;
                iteration_ += n_iters;
                if (out_xyz) JC_TRY(read_positions(out_xyz));
                return Status::ok();
            }
            uint32_t left = n_iters;
            if (left >= 2) {
                JC_TRY(ensure_graph());
                if (cur_ == 1) {
                    JC_TRY(enqueue_iteration(cur_));
                    cur_ ^= 1;
                    left--;
                }
                while (left >= 2) {
                    JC_HIP(hipGraphLaunch(graph_exec_, stream_));
                    left -= 2;
                }
            }
            while (left) {
                JC_TRY(enqueue_iteration(cur_));
                cur_ ^= 1;
                left--;
            }
            iteration_ += n_iters;
        }
        if (out_xyz) JC_TRY(read_positions(out_xyz));
        return Status::ok();
    }

    Status speeds(float *se) {
        std::vector<GFrame> h(G_);
        JC_HIP(hipMemcpyAsync(h.data(), frame_.p, G_ * sizeof(GFrame), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        for (uint32_t g = 0; g < G_; g++) {
            se[2 * g] = h[g].speed;
            se[2 * g + 1] = h[g].speed_eff;
        }
        return Status::ok();
    }

    Status timed(uint32_t n_iters, double *ms) {
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        if (!tuned_ && n_) JC_TRY(autotune());
        JC_HIP(hipEventRecord(a, stream_));
        Status s = step(n_iters, nullptr);
        JC_HIP(hipEventRecord(b, stream_));
        JC_HIP(hipEventSynchronize(b));
        float t = 0;
        JC_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms) *ms = t;
        return s;
    }

    uint32_t force_block() const { return force_bs_; }

    // mode: kSchedAuto (tuned at the next step), kSchedPhased or
    // kSchedPersistent; threads: the persistent workgroup size (0: tuned).
    Status set_schedule(uint32_t mode, uint32_t threads) {
        if (mode > kSchedPersistent) return Status::err(JC_E_INVALID, "unknown batch schedule");
        if (threads != 0 && threads != 256 && threads != 512 && threads != 1024)
            return Status::err(JC_E_INVALID, "persistent workgroup size must be 0, 256, 512 or 1024");
        if (mode == kSchedPersistent && !kPersistentEnabled)
            return Status::err(JC_E_INVALID,
                               "the persistent batch schedule is switched off: its first GPU run "
                               "made a memory fault whose cause has not been found");
        if (mode == kSchedPersistent && inited_ && !p_ok_)
            return Status::err(JC_E_INVALID, "a graph is too large for the persistent schedule");
        want_sched_ = mode;
        want_threads_ = threads;
        tuned_ = false;
        drop_graph();
        return Status::ok();
    }

    // The phased schedule's knobs, each 0 to leave it to the autotuner: force
    // block size (64 to 1024), centre-of-mass fusion depth (3 to 5) and chunk
    // (1024 or 2048 nodes), composite sort key width in bits (16 to 64, a
    // multiple of 16, no narrower than the one that holds the graph id and a
    // key prefix).  Every setting gives the same bits.
    Status set_tuning(uint32_t fbs, uint32_t depth, uint32_t chunk, uint32_t ck) {
        if (fbs != 0 && fbs != 64 && fbs != 128 && fbs != 256 && fbs != 512 && fbs != 1024)
            return Status::err(JC_E_INVALID, "force block size must be 0, 64, 128, 256, 512 or 1024");
        if (depth != 0 && (depth < 3 || depth > 5))
            return Status::err(JC_E_INVALID, "centre-of-mass fusion depth must be 0, 3, 4 or 5");
        if (chunk != 0 && chunk != 1024 && chunk != 2048)
            return Status::err(JC_E_INVALID, "centre-of-mass chunk must be 0, 1024 or 2048");
        if (ck != 0 && (ck % 16 || ck > 64 || (inited_ && ck < ck_min_)))
            return Status::err(JC_E_INVALID, "sort key width must be 0 or a multiple of 16 up to 64, "
                                             "no narrower than the batch needs");
        want_fbs_ = fbs;
        want_com_depth_ = depth;
        want_com_chunk_ = chunk;
        want_ck_ = ck;
        tuned_ = false;
        drop_graph();
        return Status::ok();
    }

    // mode, persistent workgroup size, force block size, centre-of-mass fusion
    // depth and chunk, sort key width (6 values; mode 0 until the first step)
    void schedule(uint32_t *out) const {
        out[0] = tuned_ ? sched_ : kSchedAuto;
        out[1] = p_threads_;
        out[2] = force_bs_;
        out[3] = com_depth_;
        out[4] = com_chunk_;
        out[5] = ck_bits_;
    }

  private:
    void set_ck_bits(uint32_t bits) {
        ck_bits_ = bits;
        ck_p_ = bits - gb_;
    }

    void refresh_params() {
        if (!params_dirty_) return;
        const jc_fa2bh_settings &s = settings_;
        P_.n = n_;
        P_.n_graphs = G_;
        P_.gravity = s.gravity;
        P_.strong_gravity = s.strong_gravity ? 1u : 0u;
        P_.scaling_ratio = s.scaling_ratio;
        P_.jitter_tolerance = s.jitter_tolerance;
        P_.lin_log_mode = s.lin_log_mode ? 1u : 0u;
        P_.time_step = s.time_step;
        const float th = std::fmax(s.theta, 0.0f);
        P_.theta2 = th * th;
        P_.max_displacement = s.max_displacement;
        P_.ewi_weight = std::pow(std::fmax(1.0f, 1.0e-6f), s.edge_weight_influence);
        P_.cap = (uint32_t)cap_;
        P_.fastdiv = kr_in_fastdiv_range(s.scaling_ratio) ? 1u : 0u;
        params_dirty_ = false;
        drop_graph();
    }

    Status enqueue_tree(const float4 *pos) {
        const uint32_t n = n_;
        k_b_reset<<<cdivb(G_, 256), 256, 0, stream_>>>(frame_.p, G_, global_.p);
        k_b_bbox<<<G_, kBT, 0, stream_>>>(pos, base_.p, count_.p, frame_.p);
        // graph-major, key-minor: radix passes over (graph id, top ck_p_ key bits),
        // then each run of equal composite keys ranked by its raw keys
        k_b_keys<<<cdivb(n, kBT), kBT, 0, stream_>>>(pos, gid_.p, base_.p, n, frame_.p, keys_raw_.p,
                                                     sort_.keys_alt.p, sort_.vals_alt.p, ck_p_);
        JC_TRY(sort_segmented_pairs(keys_.p, vals_.p, keys_raw_.p, sort_, stream_, ck_bits_));
        k_b_count_scan<<<nb_scan_, kBT, 0, stream_>>>(keys_.p, gid_.p, base_.p, count_.p, n,
                                                      frame_.p, loc_.p, blk_.p, blkd_.p);
        k_b_scan_top<<<1, 1024, 0, stream_>>>(blk_.p, nb_scan_, global_.p, blkd_.p);
        k_b_graph_tree<<<cdivb(G_, 256), 256, 0, stream_>>>(base_.p, count_.p, G_, n, loc_.p,
                                                            blk_.p, global_.p, frame_.p);
        k_b_emit<<<cdivb(n, kBT), kBT, 0, stream_>>>(keys_.p, vals_.p, pos, deg_.p, gid_.p,
                                                     base_.p, count_.p, n, frame_.p, global_.p,
                                                     loc_.p, blk_.p, nodes_.p, node_depth_.p,
                                                     cap_, cskip_.p);
        const uint32_t gfuse = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cdivb(cap_, com_chunk_), kBComGrid));
#define JC_BCOM(D, C)                                                                              \
    k_b_com_fused<D, C><<<gfuse, kBFuseThreads, 0, stream_>>>(global_.p, (uint32_t)cap_, nodes_.p, \
                                                              node_depth_.p, cskip_.p, big_.p,    \
                                                              big_max_);                          \
    k_b_com_big<D><<<1, 1024, 0, stream_>>>(global_.p, (uint32_t)cap_, nodes_.p, node_depth_.p,   \
                                            big_.p, big_max_)
        switch (com_depth_ * 10000 + com_chunk_) {
            case 31024: JC_BCOM(3, 1024); break;
            case 32048: JC_BCOM(3, 2048); break;
            case 41024: JC_BCOM(4, 1024); break;
            case 42048: JC_BCOM(4, 2048); break;
            case 51024: JC_BCOM(5, 1024); break;
            case 52048: JC_BCOM(5, 2048); break;
            default: return Status::err(JC_E_STATE, "unknown centre-of-mass variant");
        }
#undef JC_BCOM
        const uint32_t gcom = std::min<uint32_t>(cdivb(cap_, kComSpan * (kBT / 64)), 1024);
        for (int d = (int)com_depth_ - 1; d >= 0; d--)
            k_b_com_level<<<gcom, kBT, 0, stream_>>>(global_.p, (uint32_t)d, (uint32_t)cap_,
                                                     nodes_.p, node_depth_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status launch_force(const float4 *pos, uint32_t bs) {
        const uint32_t grid = cdivb(n_, bs);
        if (n_heavy_)
            k_b_heavy_terms<<<cdivb(n_hent_, kBT), kBT, 0, stream_>>>(P_, n_hent_, hbody_.p, hadj_.p,
                                                                    pos, hterm_.p, hstride_);
#define JC_BF(BS)                                                                           \
    k_b_force<BS><<<grid, BS, 0, stream_>>>(P_, frame_.p, global_.p, gid_.p, pos, old_force_.p, \
                                            force_.p, \
                                            deg_.p, att_off_.p, att_adj_.p, vals_.p, nodes_.p, \
                                            stats_.p)
        switch (bs) {
            case 64: JC_BF(64); break;
            case 128: JC_BF(128); break;
            case 512: JC_BF(512); break;
            case 1024: JC_BF(1024); break;
            default: JC_BF(256); break;
        }
#undef JC_BF
        if (n_heavy_)
            k_b_force_heavy<<<n_heavy_, 64, 0, stream_>>>(P_, heavy_.p, hoff_.p, hlen_.p, hterm_.p,
                                                         hstride_, pos, old_force_.p, force_.p,
                                                         deg_.p, stats_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_iteration(uint32_t c) {
        JC_TRY(enqueue_tree(pos_[c].p));
        JC_TRY(launch_force(pos_[c].p, force_bs_));
        k_b_speed<<<G_, 256, 0, stream_>>>(P_, frame_.p, base_.p, count_.p, est_.p, mjt_.p, stats_.p);
        k_b_apply<<<cdivb(n_, kBT), kBT, 0, stream_>>>(P_, frame_.p, gid_.p, pos_[c].p,
                                                       pos_[c ^ 1].p, force_.p, old_force_.p,
                                                       deg_.p);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status launch_persistent(uint32_t n_iters, uint32_t threads) {
        JC_TRY(ensure_persistent_scratch());
        k_p_queue_reset<<<1, 1, 0, stream_>>>(queue_.p);
#define JC_PR(T)                                                                                 \
    k_p_run<T><<<p_grid(T), T, 0, stream_>>>(                                                    \
        P_, n_iters, cur_, pg_.p, order_.p, queue_.p, frame_.p, est_.p, mjt_.p, pos_[0].p,       \
        pos_[1].p, old_force_.p, force_.p, stats_.p, deg_.p, att_off_.p, att_adj_.p, prow_.p,    \
        p_keys_[0].p, p_keys_[1].p, p_vals_[0].p, p_vals_[1].p, loc_.p, p_cinfo_.p, nodes_.p,    \
        node_depth_.p)
        switch (threads) {
            case 256: JC_PR(256); break;
            case 512: JC_PR(512); break;
            default: JC_PR(1024); break;
        }
#undef JC_PR
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    // Resident workgroups of k_p_run<T>: every CU full, no more than graphs.
    uint32_t p_grid(uint32_t T) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        hipError_t e = hipSuccess;
        switch (T) {
            case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_p_run<256>, 256, 0); break;
            case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_p_run<512>, 512, 0); break;
            default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_p_run<1024>, 1024, 0); break;
        }
        if (e != hipSuccess || per < 1) per = 1;
        if (cus < 1) cus = 1;
        return std::max<uint32_t>(1, std::min<uint32_t>(G_, (uint32_t)cus * (uint32_t)per));
    }

    Status ensure_persistent_scratch() {
        const size_t nn = n_ ? n_ : 1;
        if (p_cinfo_.p && p_cinfo_.n == nn) return Status::ok();
        JC_TRY(p_keys_[0].alloc(nn));
        JC_TRY(p_keys_[1].alloc(nn));
        JC_TRY(p_vals_[0].alloc(nn));
        JC_TRY(p_vals_[1].alloc(nn));
        JC_TRY(p_cinfo_.alloc(nn));
        return Status::ok();
    }

    // Picks the schedule.  Phased: the force kernel's block size is timed on the
    // current positions (tree built once).  Then, unless a schedule was asked
    // for, two whole iterations of the phased schedule and of the persistent
    // one at each workgroup size are timed from the same state (saved before
    // and restored after each candidate, so no state is advanced) and the
    // fastest is kept.
    Status autotune() {
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        auto elapsed = [&]() -> float {
            float t = 0;
            (void)hipEventSynchronize(b);
            (void)hipEventElapsedTime(&t, a, b);
            return t;
        };
        if (want_sched_ != kSchedPersistent) {
            // tree builds from the current positions (no state is advanced):
            // one to warm up, three timed
            auto time_tree = [&](float *ms) -> Status {
                JC_TRY(enqueue_tree(pos_[cur_].p));
                JC_HIP(hipEventRecord(a, stream_));
                for (int r = 0; r < 3; r++) JC_TRY(enqueue_tree(pos_[cur_].p));
                JC_HIP(hipEventRecord(b, stream_));
                *ms = elapsed();
                return Status::ok();
            };
            // 1. centre of mass: fusion depth and chunk
            float best = 1e30f;
            uint32_t bd = want_com_depth_ ? want_com_depth_ : 3, bc = want_com_chunk_ ? want_com_chunk_ : 1024;
            if (!want_com_depth_ || !want_com_chunk_) {
                for (uint32_t d : {3u, 4u, 5u}) {
                    if (want_com_depth_ && d != want_com_depth_) continue;
                    for (uint32_t c : {1024u, 2048u}) {
                        if (want_com_chunk_ && c != want_com_chunk_) continue;
                        com_depth_ = d;
                        com_chunk_ = c;
                        float t = 0;
                        JC_TRY(time_tree(&t));
                        if (t < best) {
                            best = t;
                            bd = d;
                            bc = c;
                        }
                    }
                }
            }
            com_depth_ = bd;
            com_chunk_ = bc;
            // 2. sort: the narrowest composite key (more equal-key runs to rank)
            //    or 16 bits wider (two more radix passes, shorter runs)
            if (!want_ck_) {
                uint32_t bk = ck_min_;
                best = 1e30f;
                for (uint32_t ck = ck_min_; ck <= std::min<uint32_t>(ck_min_ + 16, 64); ck += 16) {
                    set_ck_bits(ck);
                    float t = 0;
                    JC_TRY(time_tree(&t));
                    if (t < best) {
                        best = t;
                        bk = ck;
                    }
                }
                set_ck_bits(bk);
            }
            // 3. force block size, over one tree
            JC_TRY(enqueue_tree(pos_[cur_].p));
            force_bs_ = want_fbs_ ? want_fbs_ : force_bs_;
            best = 1e30f;
            for (uint32_t bs : {64u, 128u, 256u, 512u, 1024u}) {
                if (want_fbs_) break;
                JC_TRY(launch_force(pos_[cur_].p, bs));  // warm
                JC_HIP(hipEventRecord(a, stream_));
                for (int r = 0; r < 3; r++) JC_TRY(launch_force(pos_[cur_].p, bs));
                JC_HIP(hipEventRecord(b, stream_));
                const float t = elapsed();
                if (t < best) {
                    best = t;
                    force_bs_ = bs;
                }
            }
        }
        if (want_sched_ == kSchedPersistent && !p_ok_)
            return Status::err(JC_E_INVALID, "a graph is too large for the persistent schedule");
        sched_ = want_sched_ == kSchedAuto ? kSchedPhased : want_sched_;
        if (!p_ok_) sched_ = kSchedPhased;
        p_threads_ = want_threads_ ? want_threads_ : 1024u;
        const bool try_p = kPersistentEnabled && p_ok_ && want_sched_ != kSchedPhased;
        const bool pick_t = want_threads_ == 0;
        if (try_p && (want_sched_ == kSchedAuto || pick_t)) {
            DevBuf<float4> sp, so;
            DevBuf<GFrame> sf;
            JC_TRY(sp.alloc(n_));
            JC_TRY(so.alloc(n_));
            JC_TRY(sf.alloc(G_));
            const size_t pb = (size_t)n_ * sizeof(float4), fb = (size_t)G_ * sizeof(GFrame);
            JC_HIP(hipMemcpyAsync(sp.p, pos_[cur_].p, pb, hipMemcpyDeviceToDevice, stream_));
            JC_HIP(hipMemcpyAsync(so.p, old_force_.p, pb, hipMemcpyDeviceToDevice, stream_));
            JC_HIP(hipMemcpyAsync(sf.p, frame_.p, fb, hipMemcpyDeviceToDevice, stream_));
            auto restore = [&]() -> Status {
                JC_HIP(hipMemcpyAsync(pos_[cur_].p, sp.p, pb, hipMemcpyDeviceToDevice, stream_));
                JC_HIP(hipMemcpyAsync(old_force_.p, so.p, pb, hipMemcpyDeviceToDevice, stream_));
                JC_HIP(hipMemcpyAsync(frame_.p, sf.p, fb, hipMemcpyDeviceToDevice, stream_));
                return Status::ok();
            };
            // two iterations from the saved state, run twice (the first warms up)
            auto time_one = [&](uint32_t mode, uint32_t T, float *ms) -> Status {
                for (int r = 0; r < 2; r++) {
                    JC_TRY(restore());
                    JC_HIP(hipEventRecord(a, stream_));
                    if (mode == kSchedPersistent) {
                        JC_TRY(launch_persistent(2, T));
                    } else {
                        JC_TRY(enqueue_iteration(cur_));
                        JC_TRY(enqueue_iteration(cur_ ^ 1u));
                    }
                    JC_HIP(hipEventRecord(b, stream_));
                    *ms = elapsed();
                }
                return Status::ok();
            };
            float best = 1e30f;
            if (want_sched_ == kSchedAuto) JC_TRY(time_one(kSchedPhased, 0, &best));
            for (uint32_t T : {256u, 512u, 1024u}) {
                if (!pick_t && T != want_threads_) continue;
                float t = 0;
                JC_TRY(time_one(kSchedPersistent, T, &t));
                if (t < best) {
                    best = t;
                    sched_ = kSchedPersistent;
                    p_threads_ = T;
                }
            }
            JC_TRY(restore());
            JC_HIP(hipStreamSynchronize(stream_));
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        tuned_ = true;
        drop_graph();
        return Status::ok();
    }

    Status ensure_graph() {
        if (graph_exec_) return Status::ok();
        JC_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        Status s1 = enqueue_iteration(0);
        Status s2 = s1.failed() ? s1 : enqueue_iteration(1);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream_, &g);
        if (s2.failed()) return s2;
        if (e != hipSuccess) return Status::err(JC_E_HIP, hipGetErrorString(e));
        graph_ = g;
        JC_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        return Status::ok();
    }

    void drop_graph() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
    }

    Status read_positions(float *out) {
        std::vector<float4> h(n_);
        if (n_) {
            JC_HIP(hipMemcpyAsync(h.data(), pos_[cur_].p, n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
        }
        for (uint32_t i = 0; i < n_; i++) {
            out[3 * (size_t)i] = h[i].x;
            out[3 * (size_t)i + 1] = h[i].y;
            out[3 * (size_t)i + 2] = h[i].z;
        }
        return Status::ok();
    }

    void teardown() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        drop_graph();
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
        inited_ = false;
        tuned_ = false;
    }

    jc_fa2bh_settings settings_;
    BParams P_{};
    // schedule: the autotuner's pick (sched_, p_threads_) or the caller's
    uint32_t want_sched_ = kSchedAuto, want_threads_ = 0, sched_ = kSchedPhased, p_threads_ = 1024;
    bool p_ok_ = false;
    DevBuf<PGraph> pg_;
    DevBuf<uint32_t> order_, prow_, queue_, p_vals_[2], p_cinfo_;
    DevBuf<uint64_t> p_keys_[2];
    bool params_dirty_ = true, inited_ = false, tuned_ = false;
    uint32_t G_ = 0, n_ = 0, cur_ = 0, nb_scan_ = 0, ck_bits_ = 16, ck_p_ = 16, force_bs_ = 256;
    // tuned knobs (want_*: the caller's choice, 0 = tuned) and the composite
    // key's graph-id bits and narrowest width
    uint32_t com_depth_ = 3, com_chunk_ = 1024, gb_ = 0, ck_min_ = 16;
    uint32_t want_fbs_ = 0, want_com_depth_ = 0, want_com_chunk_ = 0, want_ck_ = 0;
    uint64_t cap_ = 0, iteration_ = 0;
    std::vector<uint32_t> base_h_, count_h_;
    hipStream_t stream_ = nullptr;
    DevBuf<float4> pos_[2], old_force_, force_;
    DevBuf<float2> stats_;
    DevBuf<uint32_t> deg_, gid_, base_, count_, att_off_, att_adj_, vals_, loc_, blk_;
    DevBuf<uint32_t> blkd_;       // deepest cell of each count/scan tile
    DevBuf<uint32_t> cskip_;      // end of each cell's subtree (k_b_emit), read by k_b_com_fused
    DevBuf<uint32_t> big_;        // cells too large for k_b_com_fused, summed by k_b_com_big
    uint32_t big_max_ = 0;
    DevBuf<double> est_, mjt_;
    DevBuf<uint64_t> keys_raw_, keys_;
    RadixScratch sort_;
    DevBuf<Node> nodes_;
    uint32_t n_heavy_ = 0, n_hent_ = 0, hstride_ = 0;   // rows longer than kBHeavyAtt
    DevBuf<uint32_t> heavy_, hoff_, hlen_, hbody_, hadj_;
    DevBuf<float> hterm_;
    DevBuf<uint8_t> node_depth_;
    DevBuf<GFrame> frame_;
    DevBuf<BGlobal> global_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace jc

struct jc_batch {
    jc::Fa2BhBatch impl;
    explicit jc_batch(const jc_fa2bh_settings &s) : impl(s) {}
};

extern "C" {

int jc_fa2bh_batch_create(const jc_fa2bh_settings *s, jc_batch **out) {
    if (!out) {
        jc::set_error("NULL output handle");
        return JC_E_INVALID;
    }
    jc_fa2bh_settings d;
    jc_fa2bh_default_settings(&d);
    *out = new jc_batch(s ? *s : d);
    return JC_OK;
}

int jc_fa2bh_batch_init(jc_batch *b, uint32_t n_graphs, const uint32_t *n_nodes,
                        const uint64_t *nnz, const uint32_t *offsets, const uint32_t *neighbors,
                        const float *xyz) {
    if (!b) {
        jc::set_error("NULL batch");
        return JC_E_INVALID;
    }
    jc::Status s = b->impl.init(n_graphs, n_nodes, nnz, offsets, neighbors, xyz);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_fa2bh_batch_step(jc_batch *b, uint32_t n_iters, float *out_xyz) {
    if (!b) {
        jc::set_error("NULL batch");
        return JC_E_INVALID;
    }
    jc::Status s = b->impl.step(n_iters, out_xyz);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_fa2bh_batch_speeds(jc_batch *b, float *speed_eff) {
    if (!b) {
        jc::set_error("NULL batch");
        return JC_E_INVALID;
    }
    return b->impl.speeds(speed_eff).code;
}

int jc_fa2bh_batch_step_timed(jc_batch *b, uint32_t n_iters, double *gpu_ms, uint32_t *force_block) {
    if (!b) {
        jc::set_error("NULL batch");
        return JC_E_INVALID;
    }
    jc::Status s = b->impl.timed(n_iters, gpu_ms);
    if (force_block) *force_block = b->impl.force_block();
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_fa2bh_batch_set_schedule(jc_batch *b, uint32_t mode, uint32_t threads) {
    if (!b) {
        jc::set_error("NULL batch");
        return JC_E_INVALID;
    }
    jc::Status s = b->impl.set_schedule(mode, threads);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_fa2bh_batch_set_tuning(jc_batch *b, uint32_t force_block, uint32_t com_depth,
                              uint32_t com_chunk, uint32_t sort_key_bits) {
    if (!b) {
        jc::set_error("NULL batch");
        return JC_E_INVALID;
    }
    jc::Status s = b->impl.set_tuning(force_block, com_depth, com_chunk, sort_key_bits);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_fa2bh_batch_schedule(jc_batch *b, uint32_t *out) {
    if (!b || !out) {
        jc::set_error(!b ? "NULL batch" : "NULL output");
        return JC_E_INVALID;
    }
    b->impl.schedule(out);
    return JC_OK;
}

void jc_fa2bh_batch_destroy(jc_batch *b) { delete b; }

}  // extern "C"
