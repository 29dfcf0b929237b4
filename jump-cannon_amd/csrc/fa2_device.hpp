// Device helpers shared by the FA2+BH kernels (single graph and batched).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace jc {

constexpr int kLevels = 21;            // octree key depth (3 bits per level)
constexpr uint32_t kLeaf = 0x80000000u;  // node_info flag of a leaf (exported tree form)
constexpr uint32_t kEnd = 0xFFFFFFFFu;   // end of the rope
constexpr uint32_t kCellId = 0xFFFFFFFFu;   // Node::b.z of a cell: never a live body
constexpr uint32_t kLeafSize = 0xFF800000u;  // Node::b.y of a leaf: bits of -inf
// Byte offsets of nodes fit 32 bits below this many nodes (the fast walk's addressing).
constexpr uint32_t kMaxNodes32 = 1u << 27;
// Word 3 of a gfx9-family buffer resource for raw (untyped) dword loads.
constexpr int kBufferRsrcWord3 = 0x00020000;

// One octree node, 32 bytes: one aligned record per visit.
//   a   = centre of mass xyz, mass
//   b.x = skip rope as a byte offset (b.w * 32; 0xFFFFFFFF at the end or when
//         b.w * 32 would not fit 32 bits)
//   b.y = cell: bits of the opening size (2*half)^2; leaf: bits of -inf
//   b.z = leaf: its body; cell: kCellId
//   b.w = skip rope (first node after the subtree, kEnd at the end)
struct __align__(32) Node {
    float4 a;
    uint4 b;
};

__device__ __forceinline__ uint32_t skip_bytes(uint32_t skip) {
    return skip < kMaxNodes32 ? skip << 5 : 0xFFFFFFFFu;
}
__device__ __forceinline__ uint4 cell_b(float s2, uint32_t skip) {
    return make_uint4(skip_bytes(skip), __float_as_uint(s2), kCellId, skip);
}
__device__ __forceinline__ uint4 leaf_b(uint32_t body, uint32_t skip) {
    return make_uint4(skip_bytes(skip), kLeafSize, body, skip);
}

// Correctly rounded a / b in 8 instructions: the reciprocal refinement and the
// two residual corrections of the compiler's IEEE f32 division, without its
// operand scaling (v_div_scale) and special-case fixup (v_div_fixup).  Both are
// identities when a and b are positive normals, a >= 2^-100, b < 2^126,
// a / b is a normal and the exponents differ by less than 96; there this returns
// the bits of a / b.  Callers guarantee that range.
__device__ __forceinline__ float div_in_range(float a, float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    const float y = __builtin_fmaf(e, y0, y0);
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-b, q0, a);
    const float q1 = __builtin_fmaf(r0, y, q0);
    const float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y, q1);
}

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(u);
}
__device__ __forceinline__ bool finitef(float x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool finite3(float4 p) {
    return finitef(p.x) && finitef(p.y) && finitef(p.z);
}
__device__ __forceinline__ float len3(float x, float y, float z) {
    return sqrtf(x * x + y * y + z * z);
}

// Root cube of fa2_bh.rs:768-799 from the reduced box (order-preserving u32).
__device__ __forceinline__ void compute_root(const uint32_t *bbmin, const uint32_t *bbmax, float *r4) {
    float mn[3], mx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        mn[k] = ord2f(bbmin[k]);
        mx[k] = ord2f(bbmax[k]);
    }
    if (!finitef(mn[0])) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            mn[k] = -1.0f;
            mx[k] = 1.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) r4[k] = 0.5f * (mn[k] + mx[k]);
    float half = fmaxf(fmaxf(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]) * 0.5f;
    if (!finitef(half) || half <= 0.0f) half = 1.0f;
    r4[3] = half * 1.05f;
}

__device__ __forceinline__ int lcp_digits(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    if (!x) return kLevels;
    return (__clzll((long long)x) - 1) / 3;
}

// Last sorted index j >= i whose first d octal digits equal those of key k.
__device__ __forceinline__ uint32_t cell_end(const uint64_t *keys, uint32_t i, uint32_t nv,
                                             uint64_t k, int d) {
    if (d == 0) return nv - 1;
    const int sh = 3 * (kLevels - d);
    const uint64_t pre = k >> sh;
    uint32_t lo = i, step = 1, hi;
    for (;;) {
        uint32_t probe = lo + step;
        if (probe >= nv || (keys[probe] >> sh) != pre) {
            hi = probe < nv ? probe : nv;
            break;
        }
        lo = probe;
        step <<= 1;
    }
    while (hi - lo > 1) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if ((keys[mid] >> sh) == pre) lo = mid;
        else hi = mid;
    }
    return lo;
}

// cell_end with the sorted keys [w0, w0 + wn) staged in LDS (wk): a probe that
// falls in that window reads LDS, any other reads global memory.  The probes
// and their order are cell_end's, so the answer is the same.  An emit block
// stages its own keys and the next kEmitWin ones, which hold the ends of most
// of its cells, so their searches leave global memory.
constexpr uint32_t kEmitWin = 256;
__device__ __forceinline__ uint32_t cell_end_win(const uint64_t *__restrict__ keys,
                                                 const uint64_t *wk, uint32_t w0, uint32_t wn,
                                                 uint32_t i, uint32_t nv, uint64_t k, int d) {
    if (d == 0) return nv - 1;
    const int sh = 3 * (kLevels - d);
    const uint64_t pre = k >> sh;
    uint32_t lo = i, step = 1, hi;
    for (;;) {
        const uint32_t probe = lo + step;
        if (probe >= nv) {
            hi = nv;
            break;
        }
        const uint64_t kp = probe - w0 < wn ? wk[probe - w0] : keys[probe];
        if ((kp >> sh) != pre) {
            hi = probe;
            break;
        }
        lo = probe;
        step <<= 1;
    }
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        const uint64_t km = mid - w0 < wn ? wk[mid - w0] : keys[mid];
        if ((km >> sh) == pre) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  This bijection gives XCD x the contiguous run
// of logical blocks [x*G/8, (x+1)*G/8): with bodies in Morton order, each L2
// then holds one region of the tree instead of all of it.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G) {
    constexpr uint32_t kXcd = 8;
    const uint32_t q = G / kXcd, r = G % kXcd;
    const uint32_t x = b % kXcd, slot = b / kXcd;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// One visit of the walk in the reference form (fa2_barnes_hut.wgsl:108-156):
// leaf: skip self, take any other body; cell: accept when (2*half)^2 <
// theta^2 * r^2 and r^2 > 0, else descend to the next node in pre-order.
// Returns whether the node was taken; *next is the node visited after it.
__device__ __forceinline__ bool visit_exact(const Node &nd, uint32_t idx, float px, float py,
                                            float pz, uint32_t b, float krd, float theta2,
                                            float &fx, float &fy, float &fz, uint32_t &next) {
    const float dx = px - nd.a.x, dy = py - nd.a.y, dz = pz - nd.a.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const bool leaf = nd.b.z != kCellId;
    bool take;
    if (leaf) take = nd.b.z != b && nd.a.w > 0.0f;
    else take = nd.a.w > 0.0f && r2 > 0.0f && __uint_as_float(nd.b.y) < theta2 * r2;
    if (take) {
        const float r2c = fmaxf(r2, 1.0e-4f);
        const float cf = krd * nd.a.w / r2c;
        fx = fx + dx * cf;
        fy = fy + dy * cf;
        fz = fz + dz * cf;
    }
    next = (leaf || take) ? nd.b.w : idx + 1u;
    return take;
}

// The walk of one body in the reference form, nodes [start, end) (end: one past
// the tree's last node, where the ropes also stop).  Returns the nodes taken.
__device__ __forceinline__ uint32_t walk_exact(const Node *__restrict__ nodes, uint32_t start,
                                               uint32_t end, float px, float py, float pz,
                                               uint32_t b, float krd, float theta2, float &fx,
                                               float &fy, float &fz) {
    uint32_t taken = 0, idx = start;
    // ropes only move forward, so a walk visits at most end - start nodes
    for (uint32_t steps = 0; idx < end && steps < end - start; steps++) {
        uint32_t next;
        taken += visit_exact(nodes[idx], idx, px, py, pz, b, krd, theta2, fx, fy, fz, next);
        idx = next;
    }
    return taken;
}

// Host-side range check of the scaling ratio for div_in_range in the repulsion
// quotient kr*deg*mass / max(r2, 1e-4): with masses in [1, 2^34], degrees below
// 2^31 and squared distances below 2^84 (tree_reducible), kr in [2^-10, 2^10]
// keeps numerator, denominator and quotient inside div_in_range's range.
inline bool kr_in_fastdiv_range(float kr) { return kr >= 0x1p-10f && kr <= 0x1p10f; }

// Where walk_reduced may stand in for walk_exact: node byte offsets fit 32
// bits, the root half-size is below 2^40 (so every squared distance between
// finite bodies of the tree is finite and below 2^84), theta^2 is finite, and
// the walking body is finite.
constexpr float kReducedRootHalf = 1099511627776.0f;  // 2^40
__device__ __forceinline__ bool tree_reducible(uint32_t end, float root_half, float theta2) {
    return end < kMaxNodes32 && root_half < kReducedRootHalf && finitef(theta2);
}

// The same walk in reduced form: same nodes in the same order, same f32
// operations, same sums.  Every mass is >= 1 (leaves: deg+1; cells: sums of
// those) and every squared distance r2 is finite, so the reference tests become
//     take = size < theta^2 * r2  &&  id != b
// with size = (2*half)^2 >= 0 for a cell (size < theta^2*r2 then implies
// r2 > 0) and -inf for a leaf (theta^2*r2 is finite or +inf, never NaN), and
// id = the leaf's body, kCellId for a cell.  A leaf's skip rope is the next
// node, so next = take ? skip : this + 1 covers both kinds; the ropes are byte
// offsets, which saves the index shift per step.  r2 >= +0 is never -0 or NaN,
// so max(r2, 1e-4) is the unsigned maximum of the bit patterns.  FASTDIV uses
// div_in_range: the caller has checked that krd * mass / max(r2, 1e-4) stays in
// its range.  Interactions are not counted here (walk_exact counts them).
// buf_nodes bounds the buffer resource: it must be wave-uniform (a resource
// lives in scalar registers; a per-lane one costs a loop over the distinct
// values around every load), so a caller whose lanes walk different trees
// passes the size of the whole node array.
template <bool FASTDIV>
__device__ __forceinline__ void walk_reduced(const Node *__restrict__ nodes, uint32_t start,
                                             uint32_t end, uint32_t buf_nodes, float px,
                                             float py, float pz, uint32_t b, float krd,
                                             float theta2, float &fx, float &fy, float &fz) {
    // Both halves of a node come in as one 16-byte buffer load each (a plain
    // pointer load is split by the compiler into the fields each branch uses).
    const uint32_t end_b = end << 5;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<Node *>(nodes), (short)0,
                                          (int)(buf_nodes << 5), kBufferRsrcWord3);
    uint32_t off = start << 5;
    for (uint32_t steps = 0; off < end_b && steps < end - start; steps++) {
        const auto av = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        const auto bv = __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16u, 0, 0);
        const float4 a = make_float4(__uint_as_float(av[0]), __uint_as_float(av[1]),
                                     __uint_as_float(av[2]), __uint_as_float(av[3]));
        const uint32_t skip_b = bv[0], size = bv[1], id = bv[2];
        const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool take = __uint_as_float(size) < theta2 * r2 && id != b;
        const uint32_t next = take ? skip_b : off + 32u;
        if (take) {
            const float r2c = __uint_as_float(max(__float_as_uint(r2), 0x38D1B717u));  // 1e-4f
            const float cf = FASTDIV ? div_in_range(krd * a.w, r2c) : krd * a.w / r2c;
            fx = fx + dx * cf;
            fy = fy + dy * cf;
            fz = fz + dz * cf;
        }
        off = next;
    }
}

// K reduced walks interleaved step by step in one lane (K independent
// pointer chains in flight): each walk is walk_reduced's, nodes [0, end).  A
// finished walk parks at end * 32, past the buffer, where its loads return
// zeros and its visits are skipped.
template <int K, bool FASTDIV>
__device__ __forceinline__ void walk_reduced_multi(const Node *__restrict__ nodes, uint32_t end,
                                                   const bool (&live)[K], const float4 (&p)[K],
                                                   const uint32_t (&b)[K], const float (&krd)[K],
                                                   float theta2, float (&fx)[K], float (&fy)[K],
                                                   float (&fz)[K]) {
    const uint32_t end_b = end << 5;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<Node *>(nodes), (short)0, (int)end_b,
                                          kBufferRsrcWord3);
    uint32_t off[K];
#pragma unroll
    for (int k = 0; k < K; k++) off[k] = live[k] ? 0u : end_b;
    for (uint32_t steps = 0; steps < end; steps++) {
        bool active = false;
#pragma unroll
        for (int k = 0; k < K; k++) active |= off[k] < end_b;
        if (!active) break;
        uint4 av[K], bv[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, off[k], 0, 0);
            const auto y = __builtin_amdgcn_raw_buffer_load_b128(rs, off[k] + 16u, 0, 0);
            av[k] = make_uint4(x[0], x[1], x[2], x[3]);
            bv[k] = make_uint4(y[0], y[1], y[2], y[3]);
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (off[k] < end_b) {
                const float dx = p[k].x - __uint_as_float(av[k].x);
                const float dy = p[k].y - __uint_as_float(av[k].y);
                const float dz = p[k].z - __uint_as_float(av[k].z);
                const float r2 = dx * dx + dy * dy + dz * dz;
                const bool take = __uint_as_float(bv[k].y) < theta2 * r2 && bv[k].z != b[k];
                const uint32_t next = take ? bv[k].x : off[k] + 32u;
                if (take) {
                    const float r2c = __uint_as_float(max(__float_as_uint(r2), 0x38D1B717u));
                    const float m = krd[k] * __uint_as_float(av[k].w);
                    const float cf = FASTDIV ? div_in_range(m, r2c) : m / r2c;
                    fx[k] = fx[k] + dx * cf;
                    fy[k] = fy[k] + dy * cf;
                    fz[k] = fz[k] + dz * cf;
                }
                off[k] = next;
            }
        }
    }
}

// ---- top of the tree in LDS ----
// The nodes of depth <= kTopDepth (at most kTopSlots of them) are visited by
// nearly every walk.  They are gathered once per tree build into a compact
// pre-order array ("top tree") that each force block copies into LDS, so those
// visits read LDS instead of going through the vector-memory address path.
// A node's place in the gathering is its slot in the pre-order enumeration of
// the complete octree of depth kTopDepth: slot(root) = 0 and a child with
// octant o of a node at depth j-1 sits at slot(parent) + 1 + o * top_sub(j),
// where top_sub(j) is the size of a complete subtree rooted at depth j.
constexpr int kTopDepth = 3;
constexpr uint32_t kTopSlots = 585;  // 1 + 8 + 64 + 512
constexpr uint32_t kTopAbsent = 0xFFFFFFFFu;
// Closed form (8^(kTopDepth - j + 1) - 1) / 7, not a recursion: a recursive
// call in device code is not inlined, and the kernels calling it would get a
// dynamic stack (scratch set up at every dispatch).
__host__ __device__ constexpr uint32_t top_sub(int j) {
    return j > kTopDepth ? 0u : ((1u << (3 * (kTopDepth - j + 1))) - 1u) / 7u;
}
static_assert(top_sub(kTopDepth) == 1u && top_sub(kTopDepth - 1) == 9u, "1 + 8 top_sub(j + 1)");
static_assert(top_sub(0) == kTopSlots, "complete octree of depth kTopDepth");
// Slot of the depth-d (d <= kTopDepth) prefix of a key.
__device__ __forceinline__ uint32_t top_slot_of(uint64_t key, int d) {
    uint32_t s = 0;
    for (int j = 1; j <= d; j++) s += 1u + ((uint32_t)(key >> (3 * (kLevels - j))) & 7u) * top_sub(j);
    return s;
}
// Depth of a slot of the enumeration.
__device__ __forceinline__ int top_slot_depth(uint32_t s) {
    int d = 0;
    while (s != 0u && d < kTopDepth) {
        d++;
        s = (s - 1u) % top_sub(d);
    }
    return d;
}
// Top-tree record b: x = opening size bits, y = id (as Node::b.y, b.z),
// z = rank of the first top node after the subtree | 0x80000000 when the node
// is a cell of depth kTopDepth (its children live only in the global array),
// w = the node's byte offset in the global array.
constexpr uint32_t kTopFrontier = 0x80000000u;

// The walk of one body with the top tree read from LDS (tA, tB: K records)
// and everything below it from the global array (nodes [0, N)); the node
// sequence, the operations and the sums are walk_reduced's.  In the top tree
// the walk moves by rank: a taken node or a leaf goes to its subtree's end
// rank, an opened cell above the frontier to rank + 1, an opened frontier cell
// walks its subtree in the global array (from the cell's offset + 32 up to the
// next top node's offset) and then resumes at its subtree's end rank.  Callers
// guarantee what walk_reduced needs (tree_reducible, finite body).
template <bool FASTDIV>
__device__ __forceinline__ void walk_top(const float4 *tA, const uint4 *tB, uint32_t K,
                                         const Node *__restrict__ nodes, uint32_t N, float px,
                                         float py, float pz, uint32_t b, float krd, float theta2,
                                         float &fx, float &fy, float &fz) {
    const uint32_t end_b = N << 5;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<Node *>(nodes), (short)0, (int)end_b,
                                          kBufferRsrcWord3);
    bool in_top = true;
    uint32_t u = 0, off = 0, sub_end = 0;
    for (uint32_t steps = 0; steps < N + K; steps++) {
        float4 a;
        uint32_t size, id, link;
        if (in_top) {
            if (u >= K) break;
            a = tA[u];
            const uint4 q = tB[u];
            size = q.x;
            id = q.y;
            link = q.z;
        } else {
            // Below the top tree only the rope and the size are read (8 bytes,
            // not 12): the body's own leaf is then taken like any other leaf.
            // It holds the body's own position, so its offset is zero in every
            // component and its term dx * cf is +0; adding +0 leaves every
            // partial sum unchanged (a sum that starts at +0 is never -0 in
            // round-to-nearest), so the forces keep their bits.
            const auto av = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
            const auto bv = __builtin_amdgcn_raw_buffer_load_b64(rs, off + 16u, 0, 0);
            a = make_float4(__uint_as_float(av[0]), __uint_as_float(av[1]), __uint_as_float(av[2]),
                            __uint_as_float(av[3]));
            size = bv[1];
            id = kCellId;
            link = bv[0];
        }
        const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool take = __uint_as_float(size) < theta2 * r2 && id != b;
        if (take) {
            const float r2c = __uint_as_float(max(__float_as_uint(r2), 0x38D1B717u));  // 1e-4f
            const float cf = FASTDIV ? div_in_range(krd * a.w, r2c) : krd * a.w / r2c;
            fx = fx + dx * cf;
            fy = fy + dy * cf;
            fz = fz + dz * cf;
        }
        if (in_top) {
            const uint32_t ts = link & ~kTopFrontier;
            if (take) {
                u = ts;
            } else if (link & kTopFrontier) {
                off = tB[u].w + 32u;
                sub_end = ts < K ? tB[ts].w : end_b;
                u = ts;
                in_top = false;
            } else {
                u = u + 1u;
            }
        } else {
            const uint32_t next = take ? link : off + 32u;
            if (next >= sub_end) in_top = true;
            else off = next;
        }
    }
}

// ---- centre of mass, one depth per launch ----
// One cell's centre of mass from its children along the rope, in octant
// (pre-order) order: m = sum m_c; com = sum(com_c * m_c) / m (fa2_bh.rs:944-965).
__device__ __forceinline__ void com_of_cell(Node *__restrict__ nodes, uint32_t p, uint32_t N) {
    const uint32_t end = nodes[p].b.w;
    float tm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (uint32_t ch = p + 1; ch < N && ch != end; ch = nodes[ch].b.w) {
        const float4 a = nodes[ch].a;
        tm = tm + a.w;
        cx = cx + a.x * a.w;
        cy = cy + a.y * a.w;
        cz = cz + a.z * a.w;
    }
    if (tm > 0.0f) {
        cx = cx / tm;
        cy = cy / tm;
        cz = cz / tm;
    }
    nodes[p].a = make_float4(cx, cy, cz, tm);
}

// All cells of depth d among nodes [0, N) (node_depth: cell depth, 0x80 | depth
// for a leaf).  Each lane reads 16 depth bytes with one load, the wave gathers
// the matching node indices into its LDS list (an exclusive scan of the lanes'
// match counts), and the matches are then shared out over the lanes, so a
// lane that found several cells does not chase all their ropes alone.
// node_depth must hold whole 16-byte groups past N (callers allocate rounded up).
constexpr uint32_t kComSpan = 64 * 16;  // nodes scanned per wave and round
template <uint32_t BS>
__device__ __forceinline__ void com_level_pass(Node *__restrict__ nodes,
                                               const uint8_t *__restrict__ node_depth,
                                               uint32_t N, uint32_t d,
                                               uint32_t (*list)[kComSpan]) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t nw = gridDim.x * (BS / 64);
    for (uint32_t base = (blockIdx.x * (BS / 64) + w) * kComSpan; base < N; base += nw * kComSpan) {
        const uint32_t p0 = base + lane * 16;
        uint32_t m = 0;
        if (p0 < N) {
            const uint4 dv = *reinterpret_cast<const uint4 *>(node_depth + p0);
            const uint32_t wd[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
            for (uint32_t j = 0; j < 16; j++)
                if (((wd[j >> 2] >> (8 * (j & 3))) & 0xFFu) == d && p0 + j < N) m |= 1u << j;
        }
        const uint32_t c = (uint32_t)__popc(m);
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if ((int)lane >= o) incl += t;
        }
        const uint32_t total = __shfl(incl, 63, 64);
        uint32_t at = incl - c;
        while (m) {
            list[w][at++] = p0 + (uint32_t)(__ffs(m) - 1);
            m &= m - 1u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t q = lane; q < total; q += 64) com_of_cell(nodes, list[w][q], N);
        __builtin_amdgcn_wave_barrier();  // the list is read before the next round writes it
    }
}

// ---- centre of mass of the deep cells, one launch ----
// The cells of depth >= Depth.  Such a cell "fits" when its subtree, the
// pre-order range [x, end(x)), holds at most Cap nodes; cskip[x] is end(x), the
// first node after the subtree (kEnd: the end of the node array).  A fitting
// cell with no fitting ancestor is a root: one workgroup sums the root's whole
// subtree in LDS, deepest level first with a barrier between levels, each cell
// from its children in octant (pre-order) order: m = sum m_c;
// com = sum(com_c * m_c) / m (fa2_bh.rs:944-965), the same operations as
// com_of_cell, so the bits are those of one launch per depth.
//
// Ownership: the workgroup of the chunk [c0, c0 + Chunk) takes the roots that
// start in it.  A fitting ancestor a of x satisfies x - Cap <= a < x and
// end(a) > x, so x has one iff the prefix maximum of the fitting cells' ends
// over the window [x - Cap, x) passes x.  The roots of a chunk span fewer than
// Chunk + Cap nodes.  Between two roots lie only leaves, shallower cells and
// cells that do not fit; every fitting cell in the span is inside a root.
// Cells that do not fit (fewer than one per Cap nodes and depth) are appended
// to `big` for com_big_pass.
__device__ __forceinline__ uint32_t fit_end(const uint32_t *__restrict__ cskip, uint32_t x,
                                            uint32_t N) {
    const uint32_t sk = cskip[x];
    return sk < N ? sk : N;
}

template <uint32_t Depth, uint32_t Chunk, uint32_t Cap, uint32_t Threads>
__device__ __forceinline__ void com_fused_pass(uint32_t N, Node *__restrict__ nodes,
                                               const uint8_t *__restrict__ node_depth,
                                               const uint32_t *__restrict__ cskip,
                                               uint32_t *__restrict__ big, uint32_t big_max,
                                               uint32_t *n_big) {
    constexpr uint32_t Span = Chunk + Cap;
    constexpr uint32_t Per = Span / Threads;
    static_assert(Depth >= 1, "the level loop counts down to Depth");
    static_assert(Span % Threads == 0, "whole segments per thread");
    static_assert(Span <= 65535, "16-bit ropes in LDS");
    __shared__ float4 ca[Span];  // the window's skips first, then {com, mass} of the span
    __shared__ uint32_t meta[Span];  // depth byte << 16 | rope relative to the span (fitting cells)
    __shared__ uint32_t wmax[Threads / 64];
    __shared__ uint32_t s_lo, s_hi, s_dmax;
    uint32_t *win = reinterpret_cast<uint32_t *>(ca);
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (uint32_t c0 = blockIdx.x * Chunk; c0 < N; c0 += gridDim.x * Chunk) {
        const uint32_t c1 = min(c0 + Chunk, N);
        const uint32_t lo = c0 > Cap ? c0 - Cap : 0u;
        const uint32_t W = c1 - lo;
        if (t == 0) {
            s_lo = 0xFFFFFFFFu;
            s_hi = 0;
            s_dmax = 0;
        }
        // 1. the window's values: the skip of a fitting cell, 0 elsewhere; the
        //    chunk's cells that do not fit join the big list
        for (uint32_t q = t; q < W; q += Threads) {
            const uint32_t x = lo + q;
            const uint32_t d = node_depth[x];
            uint32_t v = 0;
            if (d < 0x80u && d >= Depth) {
                const uint32_t end = fit_end(cskip, x, N);
                if (end - x <= Cap) {
                    v = end;
                } else if (x >= c0) {
                    const uint32_t slot = atomicAdd(n_big, 1u);
                    if (slot < big_max) big[slot] = x;
                }
            }
            win[q] = v;
        }
        __syncthreads();
        // 2. exclusive prefix maximum over the window (a segment of Per
        //    entries per thread, then across the lanes and the waves); a fitting
        //    cell of the chunk whose prefix maximum does not pass it is a root
        uint32_t seg = 0;
        for (uint32_t j = 0; j < Per; j++) {
            const uint32_t q = t * Per + j;
            if (q < W) seg = max(seg, win[q]);
        }
        uint32_t incl = seg;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, 64);
            if ((int)lane >= o) incl = max(incl, u);
        }
        if (lane == 63) wmax[w] = incl;
        __syncthreads();
        uint32_t run = __shfl_up(incl, 1, 64);
        if (lane == 0) run = 0;
        for (uint32_t q = 0; q < w; q++) run = max(run, wmax[q]);
        uint32_t rlo = 0xFFFFFFFFu, rhi = 0;
        for (uint32_t j = 0; j < Per; j++) {
            const uint32_t q = t * Per + j;
            if (q < W) {
                const uint32_t v = win[q], x = lo + q;
                if (v && x >= c0 && run <= x) {
                    rlo = min(rlo, x);
                    rhi = max(rhi, v);
                }
                run = max(run, v);
            }
        }
        if (rhi) {
            atomicMin(&s_lo, rlo);
            atomicMax(&s_hi, rhi);
        }
        __syncthreads();
        const uint32_t r0 = s_lo, r1 = s_hi;
        if (r1 == 0) {  // no root starts in this chunk (the same answer for every thread)
            __syncthreads();
            continue;
        }
        const uint32_t R = r1 - r0;  // < Span
        // 3. the span: depths, ropes of the fitting cells relative to r0, {com, mass}
        uint32_t dmax = 0;
        for (uint32_t q = t; q < R; q += Threads) {
            const uint32_t y = r0 + q;
            const uint32_t d = node_depth[y];
            uint32_t r = 0;
            if (d < 0x80u && d >= Depth) {
                const uint32_t end = fit_end(cskip, y, N);
                if (end - y <= Cap) {
                    r = end - r0;
                    dmax = max(dmax, d);
                }
            }
            meta[q] = d << 16 | r;
            ca[q] = nodes[y].a;
        }
        if (dmax) atomicMax(&s_dmax, dmax);
        __syncthreads();
        // 4. deepest level first; a fitting cell's children are fitting cells or leaves
        for (uint32_t d = s_dmax; d >= Depth; d--) {
            for (uint32_t q = t; q < R; q += Threads) {
                const uint32_t mq = meta[q];
                if ((mq >> 16) != d || !(mq & 0xFFFFu)) continue;
                const uint32_t end = mq & 0xFFFFu;
                float tm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
                for (uint32_t ch = q + 1; ch < end;) {
                    const float4 a = ca[ch];
                    const uint32_t m = meta[ch];
                    tm = tm + a.w;
                    cx = cx + a.x * a.w;
                    cy = cy + a.y * a.w;
                    cz = cz + a.z * a.w;
                    const uint32_t nx = (m & 0x800000u) ? ch + 1 : (m & 0xFFFFu);
                    ch = nx > ch ? nx : end;  // ropes only move forward
                }
                if (tm > 0.0f) {
                    cx = cx / tm;
                    cy = cy / tm;
                    cz = cz / tm;
                }
                ca[q] = make_float4(cx, cy, cz, tm);
            }
            __syncthreads();
        }
        for (uint32_t q = t; q < R; q += Threads)
            if (meta[q] & 0xFFFFu) nodes[r0 + q].a = ca[q];
        __syncthreads();  // the LDS is reused by the next chunk
    }
}

// The cells of depth >= Depth that com_fused_pass left (subtrees larger than its
// Cap), deepest first, in one workgroup; their children are final: roots summed
// by com_fused_pass, leaves, or big cells of a deeper level done here.
template <uint32_t Depth>
__device__ __forceinline__ void com_big_pass(uint32_t N, uint32_t max_depth, Node *__restrict__ nodes,
                                             const uint8_t *__restrict__ node_depth,
                                             const uint32_t *__restrict__ big, uint32_t nbig) {
    static_assert(Depth >= 1, "the level loop counts down to Depth");
    for (uint32_t d = min(max_depth, (uint32_t)kLevels); d >= Depth; d--) {
        for (uint32_t q = threadIdx.x; q < nbig; q += blockDim.x) {
            const uint32_t x = big[q];
            if (node_depth[x] == d) com_of_cell(nodes, x, N);
        }
        __syncthreads();
    }
}

// Octant path of p through the reference cell chain for kLevels levels: digit
// (x>=cx) | (y>=cy)<<1 | (z>=cz)<<2, child half = half*0.5, child centre =
// centre +- child half (fa2_bh.rs:841-861).
__device__ __forceinline__ uint64_t octant_key(const float *r4, float px, float py, float pz) {
    float cx = r4[0], cy = r4[1], cz = r4[2], h = r4[3];
    uint64_t key = 0;
#pragma unroll
    for (int l = 0; l < kLevels; l++) {
        const uint32_t o = (px >= cx ? 1u : 0u) | (py >= cy ? 2u : 0u) | (pz >= cz ? 4u : 0u);
        key = (key << 3) | o;
        h = h * 0.5f;
        cx = cx + ((o & 1u) ? h : -h);
        cy = cy + ((o & 2u) ? h : -h);
        cz = cz + ((o & 4u) ? h : -h);
    }
    return key;
}

}  // namespace jc
