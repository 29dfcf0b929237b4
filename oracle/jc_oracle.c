/*
 * jc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see jc_oracle.h).
 *
 * Restatement of the reference hot path in plain C.  Reference paths below are
 * relative to the ocasazza/jump-cannon repository root.
 *
 * What pins each part (tests under tests/ exercise every item):
 *   SplitMix64 / next_f32 ........ closed form of stability.rs:11-26, known first
 *                                   outputs of the published SplitMix64 for seed 0.
 *   synthetic_scale_free (BA) .... invariants (symmetry, sortedness, no self loops,
 *                                   m edges per new node, degree sum) + golden CSR
 *                                   hashes written from this restatement.
 *   SBM / RGG generators ......... new in this build (absent in the reference);
 *                                   pinned by brute-force recount of the defining
 *                                   predicate on small sizes.
 *   csr_from_pairs ............... closed form (stability.rs:86-110).
 *   reference insertion octree ... fa2_bh.rs:761-1045 restated; pinned by the
 *                                   reference's own test (fa2_bh.rs:1149-1188: leaf
 *                                   mass = sum of masses, theta=0 walk == brute force).
 *   Morton pre-order octree ...... own contract; pinned bit-for-bit against the
 *                                   insertion octree on non-degenerate inputs.
 *   f32 walk / attraction /
 *   gravity / stats / apply ...... fa2_barnes_hut.wgsl:96-226 and the CPU replica
 *                                   fa2_bh.rs:1096-1147; pinned by the FP64 brute force
 *                                   (theta=0) and two-body closed forms.
 *   speed controller ............. fa2_speed.rs:48-100; pinned by the reference's
 *                                   known answers (fa2_speed.rs:103-147).
 *   exact swing/traction sum ..... pinned against an independent exact rational sum
 *                                   (Python fractions) on random inputs.
 *   PageRank pull/push/f64 ....... pagerank.wgsl:30-46, pagerank.rs:298-331,
 *                                   graph-metrics/src/pagerank.rs:4-45; pinned by the
 *                                   closed-form star and ring (uniform 1/n).
 *   BFS / pivots / anneal / sweep  sgd_stress.rs:143-175, 324-394; sgd_stress.wgsl:34-62;
 *                                   pinned by path BFS [0,1,2,3,4], the zero-stress
 *                                   path fixed point and a brute max-min recount.
 *   metrics ...................... metrics.rs:31-164; pinned by collinear-path zero
 *                                   stress, scale invariance and hand-computed CV.
 *   BFS distances ................ analytics/bfs.rs:44-66; pinned by the closed
 *                                   forms of analytics tests (path |i-src|, star,
 *                                   grid Manhattan, unreachable node) and by scipy's
 *                                   unweighted shortest paths on random graphs.
 *   connected components / LCC ... analytics/connected_components.rs:45-66; pinned by
 *                                   disjoint cliques, path, isolated nodes, two cycles
 *                                   and scipy's connected_components on random graphs.
 *   SpMV serial / f16 / hybrid ... spmv.rs:68-99, spmv.wgsl:27-41, spmv_f16.wgsl:27-45,
 *                                   spmv_hybrid.wgsl:44-84; pinned by the identity and
 *                                   the [5,6,13] matrix of the reference's tests, exact
 *                                   integer sums against scipy's sparse product, a
 *                                   hand-worked row whose serial and tree orders round
 *                                   differently, and numpy's float16 conversion.
 */
#include "jc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define JCO_LEVELS 21

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------- */
/* SplitMix64 (stability.rs:11-26; identical copy at sgd_stress.rs:104-125)   */
/* ------------------------------------------------------------------------- */
uint64_t jco_next_u64(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* stability.rs:22-25: (x >> 40) as f32 / 2^24 as f32 */
float jco_next_f32(uint64_t *state) {
    return (float)(jco_next_u64(state) >> 40) / (float)(1ull << 24);
}

static double next_f64(uint64_t *state) {
    return (double)(jco_next_u64(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* Positions uniform in [-half, half]^3: (2u - 1) * half per coordinate, x,y,z
 * drawn in node order (the "2*next_f32()-1" seed convention). */
void jco_uniform_cube(uint32_t n, float half, uint64_t seed, float *xyz) {
    uint64_t s = seed;
    for (uint64_t i = 0; i < 3ull * n; i++) {
        float u = jco_next_f32(&s);
        xyz[i] = (2.0f * u - 1.0f) * half;
    }
}

/* stability.rs:30-50 ball_seed */
void jco_ball_seed(uint32_t n, uint64_t seed, float *xyz) {
    uint64_t s = seed;
    float radius = sqrtf((float)n) * 5.0f;
    for (uint32_t i = 0; i < n; i++) {
        for (;;) {
            float x = 2.0f * jco_next_f32(&s) - 1.0f;
            float y = 2.0f * jco_next_f32(&s) - 1.0f;
            float z = 2.0f * jco_next_f32(&s) - 1.0f;
            if (x * x + y * y + z * z <= 1.0f) {
                xyz[3 * i] = x * radius;
                xyz[3 * i + 1] = y * radius;
                xyz[3 * i + 2] = z * radius;
                break;
            }
        }
    }
}

/* stability.rs:57-82 synthetic_scale_free, pairs only (CSR via csr_from_pairs).
 * Returns the pair count; pairs[2k], pairs[2k+1] = (v, t). */
uint64_t jco_ba_pairs(uint32_t n_in, uint32_t m_in, uint64_t seed, uint32_t *pairs, uint64_t cap) {
    uint64_t s = seed;
    uint32_t n = n_in < 2 ? 2 : n_in;
    uint32_t m = m_in < 1 ? 1 : m_in;
    uint64_t tcap = 4 + 2ull * (uint64_t)(n - 2) * m;
    uint32_t *targets = (uint32_t *)malloc(tcap * sizeof(uint32_t));
    uint32_t *picked = (uint32_t *)malloc((size_t)m * sizeof(uint32_t));
    uint64_t tlen = 4, np = 0;
    targets[0] = 0; targets[1] = 1; targets[2] = 1; targets[3] = 0;
    if (np < cap) { pairs[0] = 0; pairs[1] = 1; }
    np = 1;
    for (uint32_t v = 2; v < n; v++) {
        uint32_t want = m < v ? m : v, got = 0, guard = 0;
        while (got < want && guard < 64) {
            guard++;
            uint32_t t = targets[jco_next_u64(&s) % tlen];
            int dup = 0;
            for (uint32_t q = 0; q < got; q++) dup |= (picked[q] == t);
            if (t != v && !dup) picked[got++] = t;
        }
        for (uint32_t q = 0; q < got; q++) {
            if (np < cap) { pairs[2 * np] = v; pairs[2 * np + 1] = picked[q]; }
            np++;
            targets[tlen++] = v;
            targets[tlen++] = picked[q];
        }
    }
    free(targets);
    free(picked);
    return np;
}

/* Stochastic block model (new in this build; the reference has none).
 * Block of node v: floor(v*B/n).  Block pairs (a<=b) in order; inside each the
 * candidate pairs are enumerated lexicographically and sampled by geometric
 * skipping: idx += 1 + floor(log1p(-u)/log1p(-p)) with u = 53-bit uniform. */
uint64_t jco_sbm_pairs(uint32_t n, uint32_t blocks, double p_in, double p_out, uint64_t seed,
                       uint32_t *pairs, uint64_t cap) {
    uint64_t s = seed, np = 0;
    if (blocks < 1) blocks = 1;
    for (uint32_t a = 0; a < blocks; a++) {
        uint64_t a0 = ((uint64_t)a * n + blocks - 1) / blocks, a1 = ((uint64_t)(a + 1) * n + blocks - 1) / blocks;
        for (uint32_t b = a; b < blocks; b++) {
            uint64_t b0 = ((uint64_t)b * n + blocks - 1) / blocks, b1 = ((uint64_t)(b + 1) * n + blocks - 1) / blocks;
            double p = (a == b) ? p_in : p_out;
            uint64_t sa = a1 - a0, sb = b1 - b0;
            uint64_t count = (a == b) ? (sa * (sa ? sa - 1 : 0)) / 2 : sa * sb;
            if (p <= 0.0 || count == 0) continue;
            double lq = log1p(-p);
            int64_t idx = -1;
            /* row cursor for the triangle case */
            uint64_t row = 0, row_start = 0;
            for (;;) {
                if (p >= 1.0) idx += 1;
                else {
                    double u = next_f64(&s);
                    double k = floor(log1p(-u) / lq);
                    if (k > 1e18) break;
                    idx += 1 + (int64_t)k;
                }
                if ((uint64_t)idx >= count) break;
                uint32_t i, j;
                if (a == b) {
                    while ((uint64_t)idx >= row_start + (sa - 1 - row)) { row_start += sa - 1 - row; row++; }
                    i = (uint32_t)(a0 + row);
                    j = (uint32_t)(a0 + row + 1 + ((uint64_t)idx - row_start));
                } else {
                    i = (uint32_t)(a0 + (uint64_t)idx / sb);
                    j = (uint32_t)(b0 + (uint64_t)idx % sb);
                }
                if (np < cap) { pairs[2 * np] = i; pairs[2 * np + 1] = j; }
                np++;
            }
        }
    }
    return np;
}

/* Random geometric graph in the unit square (new in this build): x,y drawn with
 * next_f32 in node order; edge iff dx^2+dy^2 < r^2 evaluated in f64. */
uint64_t jco_rgg_pairs(uint32_t n, double radius, uint64_t seed, float *xy, uint32_t *pairs,
                       uint64_t cap) {
    uint64_t s = seed, np = 0;
    for (uint32_t i = 0; i < n; i++) {
        xy[2 * i] = jco_next_f32(&s);
        xy[2 * i + 1] = jco_next_f32(&s);
    }
    uint32_t g = (uint32_t)(1.0 / radius);
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    uint64_t cells = (uint64_t)g * g;
    uint32_t *cnt = (uint32_t *)calloc(cells + 1, sizeof(uint32_t));
    uint32_t *idx = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *cell_of = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) {
        uint32_t cx = (uint32_t)(xy[2 * i] * g), cy = (uint32_t)(xy[2 * i + 1] * g);
        if (cx >= g) cx = g - 1;
        if (cy >= g) cy = g - 1;
        cell_of[i] = cy * g + cx;
        cnt[cell_of[i] + 1]++;
    }
    for (uint64_t c = 0; c < cells; c++) cnt[c + 1] += cnt[c];
    uint32_t *fill = (uint32_t *)malloc((cells + 1) * sizeof(uint32_t));
    memcpy(fill, cnt, (cells + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) idx[fill[cell_of[i]]++] = i;
    double r2 = radius * radius;
    for (uint32_t i = 0; i < n; i++) {
        int cx = (int)(cell_of[i] % g), cy = (int)(cell_of[i] / g);
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int x = cx + dx, y = cy + dy;
                if (x < 0 || y < 0 || x >= (int)g || y >= (int)g) continue;
                uint32_t c = (uint32_t)y * g + (uint32_t)x;
                for (uint32_t q = cnt[c]; q < cnt[c + 1]; q++) {
                    uint32_t j = idx[q];
                    if (j <= i) continue;
                    double ddx = (double)xy[2 * i] - (double)xy[2 * j];
                    double ddy = (double)xy[2 * i + 1] - (double)xy[2 * j + 1];
                    if (ddx * ddx + ddy * ddy < r2) {
                        if (np < cap) { pairs[2 * np] = i; pairs[2 * np + 1] = j; }
                        np++;
                    }
                }
            }
    }
    free(cnt); free(idx); free(cell_of); free(fill);
    return np;
}

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* stability.rs:86-110 csr_from_pairs: symmetric, sorted, deduplicated; self
 * loops and out-of-range ids dropped.  Returns nnz. */
uint64_t jco_csr_from_pairs(uint32_t n, const uint32_t *pairs, uint64_t npairs, uint32_t *offsets,
                            uint32_t *nbrs) {
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    for (uint64_t k = 0; k < npairs; k++) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a == b || a >= n || b >= n) continue;
        cnt[a + 1]++;
        cnt[b + 1]++;
    }
    for (uint32_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
    uint64_t total = cnt[n];
    uint32_t *tmp = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
    uint64_t *fill = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    memcpy(fill, cnt, ((size_t)n + 1) * sizeof(uint64_t));
    for (uint64_t k = 0; k < npairs; k++) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a == b || a >= n || b >= n) continue;
        tmp[fill[a]++] = b;
        tmp[fill[b]++] = a;
    }
    uint64_t out = 0;
    offsets[0] = 0;
    for (uint32_t v = 0; v < n; v++) {
        uint64_t s0 = cnt[v], s1 = cnt[v + 1];
        qsort(tmp + s0, s1 - s0, sizeof(uint32_t), cmp_u32);
        for (uint64_t q = s0; q < s1; q++) {
            if (q > s0 && tmp[q] == tmp[q - 1]) continue;
            nbrs[out++] = tmp[q];
        }
        offsets[v + 1] = (uint32_t)out;
    }
    free(cnt); free(tmp); free(fill);
    return out;
}

/* Largest connected component (ties: the one holding the smallest node id).
 * keep_map[v] = new id (increasing in v) or 0xFFFFFFFF.  Returns its size. */
uint32_t jco_largest_component(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                               uint32_t *keep_map) {
    uint32_t *comp = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    uint32_t *queue = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    for (uint32_t v = 0; v < n; v++) comp[v] = 0xFFFFFFFFu;
    uint32_t best = 0xFFFFFFFFu, best_size = 0, nc = 0;
    for (uint32_t s = 0; s < n; s++) {
        if (comp[s] != 0xFFFFFFFFu) continue;
        uint32_t head = 0, tail = 0;
        queue[tail++] = s;
        comp[s] = nc;
        while (head < tail) {
            uint32_t v = queue[head++];
            for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
                uint32_t u = nbrs[e];
                if (comp[u] == 0xFFFFFFFFu) { comp[u] = nc; queue[tail++] = u; }
            }
        }
        if (tail > best_size) { best_size = tail; best = nc; }
        nc++;
    }
    uint32_t next = 0;
    for (uint32_t v = 0; v < n; v++) keep_map[v] = (comp[v] == best) ? next++ : 0xFFFFFFFFu;
    free(comp); free(queue);
    return best_size;
}

/* ------------------------------------------------------------------------- */
/* FA2 preparation: fa2_bh.rs:264-294.                                        */
/* degrees[v] = non-self CSR entries of row v; undirected edge list (v<u) in   */
/* row order; attraction adjacency = for node i, the other endpoint of every   */
/* edge touching i, in edge-list order (the order fa2_barnes_hut.wgsl:158-181  */
/* meets them).  Returns the adjacency length (2 * n_edges).                   */
/* ------------------------------------------------------------------------- */
uint64_t jco_fa2_prepare(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                         uint32_t *degrees, uint32_t *att_off, uint32_t *att_adj) {
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < n; v++) {
        degrees[v] = 0;
        for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
            uint32_t u = nbrs[e];
            if (u == v) continue;
            degrees[v]++;
            if (v < u) { cnt[v + 1]++; cnt[u + 1]++; }
        }
    }
    for (uint32_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
    for (uint32_t v = 0; v <= n; v++) att_off[v] = (uint32_t)cnt[v];
    for (uint32_t v = 0; v < n; v++) {
        for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
            uint32_t u = nbrs[e];
            if (u == v || !(v < u)) continue;
            att_adj[cnt[v]++] = u;
            att_adj[cnt[u]++] = v;
        }
    }
    uint64_t total = att_off[n];
    free(cnt);
    return total;
}

/* ------------------------------------------------------------------------- */
/* Octree root box: fa2_bh.rs:768-799.                                        */
/* ------------------------------------------------------------------------- */
static void root_box(uint32_t n, const float *xyz, float *root4) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            float v = xyz[3 * i + k];
            if (!isfinite(v)) continue;
            if (v < mn[k]) mn[k] = v;
            if (v > mx[k]) mx[k] = v;
        }
    if (!isfinite(mn[0])) {
        for (int k = 0; k < 3; k++) { mn[k] = -1.0f; mx[k] = 1.0f; }
    }
    for (int k = 0; k < 3; k++) root4[k] = 0.5f * (mn[k] + mx[k]);
    float half = fmaxf(fmaxf(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]) * 0.5f;
    if (!isfinite(half) || half <= 0.0f) half = 1.0f;
    half *= 1.05f;
    root4[3] = half;
}

static inline int body_finite(const float *p) {
    return isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]);
}

/* Octant digit of p in the cell centred at c: x>=cx -> 1, y -> 2, z -> 4
 * (fa2_bh.rs:841-853). */
static inline uint32_t octant(const float *c, const float *p) {
    uint32_t o = 0;
    if (p[0] >= c[0]) o |= 1;
    if (p[1] >= c[1]) o |= 2;
    if (p[2] >= c[2]) o |= 4;
    return o;
}

/* Child cell: h = parent.half * 0.5; c = parent.c +- h (fa2_bh.rs:855-861). */
static inline void child_cell(float *c, float *h, uint32_t o) {
    float hh = *h * 0.5f;
    c[0] = c[0] + ((o & 1) ? hh : -hh);
    c[1] = c[1] + ((o & 2) ? hh : -hh);
    c[2] = c[2] + ((o & 4) ? hh : -hh);
    *h = hh;
}

/* Morton key: the octant path of the body through the reference's own cell
 * chain for JCO_LEVELS levels (3 bits per level, root digit most significant,
 * bit 63 clear).  Non-finite bodies get the all-ones key. */
static uint64_t morton_key(const float *root4, const float *p) {
    if (!body_finite(p)) return ~0ull;
    float c[3] = {root4[0], root4[1], root4[2]}, h = root4[3];
    uint64_t key = 0;
    for (int l = 0; l < JCO_LEVELS; l++) {
        uint32_t o = octant(c, p);
        key = (key << 3) | o;
        child_cell(c, &h, o);
    }
    return key;
}

void jco_morton_keys(uint32_t n, const float *xyz, uint64_t *keys, float *root4) {
    root_box(n, xyz, root4);
    for (uint32_t i = 0; i < n; i++) keys[i] = morton_key(root4, xyz + 3 * i);
}

/* Internal-node COM in child (octant) order: fa2_bh.rs:944-965.  Children of
 * internal node p in a pre-order array are p+1, skip[p+1], ... up to skip[p].
 * Reverse pre-order visits every child before its parent. */
static void canonical_com(uint32_t N, float *node_a, const uint32_t *body, const uint32_t *skip) {
    for (int64_t p = (int64_t)N - 1; p >= 0; p--) {
        if (body[p] != JCO_INTERNAL) continue;
        float tm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        uint32_t end = skip[p];
        for (uint32_t ch = (uint32_t)p + 1; ch < N && ch != end; ch = skip[ch]) {
            const float *a = node_a + 4 * (size_t)ch;
            tm += a[3];
            cx += a[0] * a[3];
            cy += a[1] * a[3];
            cz += a[2] * a[3];
        }
        if (tm > 0.0f) { cx /= tm; cy /= tm; cz /= tm; }
        float *o = node_a + 4 * (size_t)p;
        o[0] = cx; o[1] = cy; o[2] = cz; o[3] = tm;
    }
}

/* ------------------------------------------------------------------------- */
/* Reference insertion octree: fa2_bh.rs:761-1045, then relabelled to         */
/* pre-order (the rope walk order) for comparison.                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    float ps[4];
    float cm[4];
    uint32_t body, next, skip, nchild, depth;
    uint32_t child[8];
} ins_node;

typedef struct {
    ins_node *v;
    uint32_t len;
} ins_tree;

static uint32_t ins_push(ins_tree *t, const float *c, float h, const float *cm, uint32_t body,
                         uint32_t depth) {
    ins_node *nd = &t->v[t->len];
    nd->ps[0] = c[0]; nd->ps[1] = c[1]; nd->ps[2] = c[2]; nd->ps[3] = h;
    for (int k = 0; k < 4; k++) nd->cm[k] = cm ? cm[k] : 0.0f;
    nd->body = body; nd->next = JCO_END; nd->skip = JCO_END; nd->nchild = 0; nd->depth = depth;
    for (int k = 0; k < 8; k++) nd->child[k] = JCO_END;
    return t->len++;
}

/* fa2_bh.rs:863-923 insert_body: returns 0 ok, -1 overflow */
static int ins_insert(ins_tree *t, uint32_t root, const float *p, uint32_t body, float bm,
                      uint32_t max_nodes) {
    uint32_t node = root;
    for (int depth = 0; depth < 32; depth++) {
        float c[4];
        memcpy(c, t->v[node].ps, sizeof c);
        uint32_t o = octant(c, p);
        uint32_t ch = t->v[node].child[o];
        if (ch == JCO_END) {
            if (t->len >= max_nodes) return -1;
            float cc[3] = {c[0], c[1], c[2]}, hh = c[3];
            child_cell(cc, &hh, o);
            float cm[4] = {p[0], p[1], p[2], bm};
            uint32_t ni = ins_push(t, cc, hh, cm, body, t->v[node].depth + 1);
            t->v[node].child[o] = ni;
            t->v[node].nchild += 1;
            return 0;
        }
        if (t->v[ch].body == JCO_INTERNAL) { node = ch; continue; }
        float prev[4];
        memcpy(prev, t->v[ch].cm, sizeof prev);
        uint32_t prev_body = t->v[ch].body;
        for (int k = 0; k < 4; k++) t->v[ch].cm[k] = 0.0f;
        t->v[ch].body = JCO_INTERNAL;
        t->v[ch].nchild = 0;
        if (ins_insert(t, ch, prev, prev_body, fmaxf(prev[3], 1e-3f), max_nodes) != 0) return -1;
        node = ch;
    }
    return 0;
}

uint32_t jco_octree_ref(uint32_t n, const float *xyz, const float *mass, uint32_t max_nodes,
                        float *node_a, float *half, uint32_t *body, uint32_t *skip,
                        uint32_t *depth, float *root4, float *centre) {
    if (n == 0 || max_nodes == 0) return 0;
    ins_tree t;
    t.v = (ins_node *)malloc((size_t)max_nodes * sizeof(ins_node));
    t.len = 0;
    float r4[4];
    root_box(n, xyz, r4);
    if (root4) memcpy(root4, r4, sizeof r4);
    ins_push(&t, r4, r4[3], NULL, JCO_INTERNAL, 0);
    for (uint32_t i = 0; i < n; i++) {
        const float *p = xyz + 3 * (size_t)i;
        float bm = fmaxf(mass ? mass[i] : 1.0f, 1e-3f);
        if (!body_finite(p)) continue;
        if (ins_insert(&t, 0, p, i, bm, max_nodes) != 0) break;
    }
    /* pre-order relabel: octant order DFS == the rope order of assign_ropes
     * (fa2_bh.rs:969-1045). */
    uint32_t N = t.len;
    uint32_t *order = (uint32_t *)malloc((size_t)N * sizeof(uint32_t));
    uint32_t *newid = (uint32_t *)malloc((size_t)N * sizeof(uint32_t));
    uint32_t *stack = (uint32_t *)malloc(((size_t)N + 8) * sizeof(uint32_t));
    uint32_t sp = 0, cnt = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t x = stack[--sp];
        newid[x] = cnt;
        order[cnt++] = x;
        for (int k = 7; k >= 0; k--)
            if (t.v[x].child[k] != JCO_END) stack[sp++] = t.v[x].child[k];
    }
    /* skip[x] = first node after x's subtree in pre-order = x + subtree size */
    uint32_t *size = (uint32_t *)malloc((size_t)N * sizeof(uint32_t));
    for (int64_t q = (int64_t)cnt - 1; q >= 0; q--) {
        uint32_t x = order[q], s = 1;
        for (int k = 0; k < 8; k++)
            if (t.v[x].child[k] != JCO_END) s += size[newid[t.v[x].child[k]]];
        size[q] = s;
    }
    for (uint32_t q = 0; q < cnt; q++) {
        const ins_node *nd = &t.v[order[q]];
        for (int k = 0; k < 4; k++) node_a[4 * (size_t)q + k] = nd->cm[k];
        half[q] = nd->ps[3];
        body[q] = nd->body;
        skip[q] = (q + size[q] < cnt) ? q + size[q] : JCO_END;
        if (depth) depth[q] = nd->depth;
        if (centre) for (int k = 0; k < 3; k++) centre[3 * (size_t)q + k] = nd->ps[k];
    }
    canonical_com(cnt, node_a, body, skip);
    free(order); free(newid); free(stack); free(size); free(t.v);
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* Morton-built pre-order octree (the GPU contract).                          */
/* Sorted keys k_0..k_{nv-1}; l_i = common leading octal digits of k_i,k_{i+1} */
/* (-1 at the ends).  Cells at depths l_{i-1}+1..l_i open at body i, then its  */
/* leaf at depth max(l_{i-1},l_i)+1.  A lone body still hangs below the root.  */
/* ------------------------------------------------------------------------- */
typedef struct { uint64_t k; uint32_t i; } kv;
static int cmp_kv(const void *a, const void *b) {
    const kv *x = (const kv *)a, *y = (const kv *)b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    return (x->i > y->i) - (x->i < y->i);
}

static inline int lcp_digits(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    if (!x) return JCO_LEVELS;
    return (__builtin_clzll(x) - 1) / 3;
}

uint32_t jco_octree_morton(uint32_t n, const float *xyz, const float *mass, float *node_a,
                           float *half, uint32_t *body, uint32_t *skip, uint32_t *depth,
                           uint64_t *keys_sorted, uint32_t *perm, uint32_t *n_valid,
                           float *root4) {
    float r4[4];
    kv *a = (kv *)malloc(((size_t)n + 1) * sizeof(kv));
    root_box(n, xyz, r4);
    if (root4) memcpy(root4, r4, sizeof r4);
    uint32_t nv = 0;
    for (uint32_t i = 0; i < n; i++) {
        a[i].k = morton_key(r4, xyz + 3 * (size_t)i);
        a[i].i = i;
        nv += (a[i].k != ~0ull);
    }
    qsort(a, n, sizeof(kv), cmp_kv);
    for (uint32_t i = 0; i < n; i++) {
        if (keys_sorted) keys_sorted[i] = a[i].k;
        if (perm) perm[i] = a[i].i;
    }
    if (n_valid) *n_valid = nv;
    if (nv == 0) { free(a); return 0; }
    int *l = (int *)malloc((size_t)nv * sizeof(int));
    uint32_t *off = (uint32_t *)malloc((size_t)nv * sizeof(uint32_t));
    uint32_t *c = (uint32_t *)malloc((size_t)nv * sizeof(uint32_t));
    for (uint32_t i = 0; i < nv; i++) l[i] = (i + 1 < nv) ? lcp_digits(a[i].k, a[i + 1].k) : -1;
    uint32_t tot = 0;
    for (uint32_t i = 0; i < nv; i++) {
        int lp = i ? l[i - 1] : -1;
        int ci = l[i] - lp;
        c[i] = (nv == 1) ? 1u : (ci > 0 ? (uint32_t)ci : 0u);
        off[i] = tot;
        tot += c[i];
    }
    uint32_t N = tot + nv;
    float hd[JCO_LEVELS + 2];
    hd[0] = r4[3];
    for (int d = 1; d < JCO_LEVELS + 2; d++) hd[d] = hd[d - 1] * 0.5f;
    for (uint32_t i = 0; i < nv; i++) {
        int lp = i ? l[i - 1] : -1;
        int a0 = (nv == 1) ? 0 : lp + 1;
        uint32_t leafpos = off[i] + i + c[i];
        for (uint32_t t = 0; t < c[i]; t++) {
            uint32_t pos = off[i] + i + t;
            int d = a0 + (int)t;
            uint32_t j = i;
            while (j + 1 < nv && l[j] >= d) j++;
            half[pos] = hd[d];
            body[pos] = JCO_INTERNAL;
            skip[pos] = (j + 1 < nv) ? off[j] + j + c[j] + 1 : JCO_END;
            if (depth) depth[pos] = (uint32_t)d;
            for (int k = 0; k < 4; k++) node_a[4 * (size_t)pos + k] = 0.0f;
        }
        int ld = (nv == 1) ? 1 : ((lp > l[i] ? lp : l[i]) + 1);
        const float *p = xyz + 3 * (size_t)a[i].i;
        node_a[4 * (size_t)leafpos + 0] = p[0];
        node_a[4 * (size_t)leafpos + 1] = p[1];
        node_a[4 * (size_t)leafpos + 2] = p[2];
        node_a[4 * (size_t)leafpos + 3] = fmaxf(mass ? mass[a[i].i] : 1.0f, 1e-3f);
        half[leafpos] = hd[ld];
        body[leafpos] = a[i].i;
        skip[leafpos] = (i + 1 < nv) ? leafpos + 1 : JCO_END;
        if (depth) depth[leafpos] = (uint32_t)ld;
    }
    canonical_com(N, node_a, body, skip);
    free(a); free(l); free(off); free(c);
    return N;
}

/* ------------------------------------------------------------------------- */
/* FA2 force pass: fa2_barnes_hut.wgsl:96-198 (walk as in the CPU replica     */
/* fa2_bh.rs:1096-1147), attraction :158-181, gravity :183-189, stats :191-197 */
/* ------------------------------------------------------------------------- */
static inline float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

/* Attraction (fa2_barnes_hut.wgsl:158-181), gravity (:183-189) and the force
 * and stats writes (:191-197) of body i after its repulsion sum (fx, fy, fz).
 * force_atlas2.wgsl:93-135 is the same tail for the brute-force engine. */
static void finish_one(uint32_t i, float deg_i, float fx, float fy, float fz, const float *xyz,
                       const float *old_force, const uint32_t *att_off, const uint32_t *att_adj,
                       const jco_fa2_settings *s, float *force, float *stats) {
    const float *p = xyz + 3 * (size_t)i;
    /* pow(max(w,1e-6), ewi) with every edge weight 1.0 (fa2_bh.rs:281, fa2_brute.rs:239). */
    float w = powf(fmaxf(1.0f, 1.0e-6f), s->edge_weight_influence);
    for (uint32_t e = att_off[i]; e < att_off[i + 1]; e++) {
        const float *q = xyz + 3 * (size_t)att_adj[e];
        float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
        float att = w;
        if (s->lin_log_mode) {
            float r = fmaxf(len3(dx, dy, dz), 1.0e-3f);
            att = w * logf(1.0f + r) / r;
        }
        fx = fx + dx * att; fy = fy + dy * att; fz = fz + dz * att;
    }
    float r0 = fmaxf(len3(p[0], p[1], p[2]), 1.0e-3f);
    float gd = s->gravity * deg_i;
    if (s->strong_gravity) {
        fx = fx - p[0] * gd; fy = fy - p[1] * gd; fz = fz - p[2] * gd;
    } else {
        fx = fx - (p[0] / r0) * gd; fy = fy - (p[1] / r0) * gd; fz = fz - (p[2] / r0) * gd;
    }
    const float *o = old_force ? old_force + 3 * (size_t)i : NULL;
    float ox = o ? o[0] : 0.0f, oy = o ? o[1] : 0.0f, oz = o ? o[2] : 0.0f;
    force[3 * (size_t)i] = fx; force[3 * (size_t)i + 1] = fy; force[3 * (size_t)i + 2] = fz;
    if (stats) {
        stats[2 * (size_t)i] = deg_i * len3(ox - fx, oy - fy, oz - fz);
        stats[2 * (size_t)i + 1] = deg_i * 0.5f * len3(ox + fx, oy + fy, oz + fz);
    }
}

static void force_one(uint32_t i, const float *xyz, const float *old_force, const uint32_t *degrees,
                      const uint32_t *att_off, const uint32_t *att_adj, uint32_t N,
                      const float *node_a, const float *half, const uint32_t *body,
                      const uint32_t *skip, const jco_fa2_settings *s, float *force, float *stats) {
    const float *p = xyz + 3 * (size_t)i;
    float deg_i = (float)degrees[i] + 1.0f;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    float theta = fmaxf(s->theta, 0.0f);
    float theta2 = theta * theta;
    if (N > 0) {
        uint32_t idx = 0, walk = 0, cap = N * 4u > 16u ? N * 4u : 16u;
        while (idx < N && walk < cap) {
            walk++;
            const float *a = node_a + 4 * (size_t)idx;
            float mass_n = a[3];
            float sz = half[idx] * 2.0f;
            float dx = p[0] - a[0], dy = p[1] - a[1], dz = p[2] - a[2];
            float r2 = dx * dx + dy * dy + dz * dz;
            if (body[idx] != JCO_INTERNAL) {
                if (body[idx] != i && mass_n > 0.0f) {
                    float r2c = fmaxf(r2, 1.0e-4f);
                    float cf = s->scaling_ratio * deg_i * mass_n / r2c;
                    fx = fx + dx * cf; fy = fy + dy * cf; fz = fz + dz * cf;
                }
                idx = skip[idx];
                continue;
            }
            if (mass_n > 0.0f && r2 > 0.0f && (sz * sz) < (theta2 * r2)) {
                float r2c = fmaxf(r2, 1.0e-4f);
                float cf = s->scaling_ratio * deg_i * mass_n / r2c;
                fx = fx + dx * cf; fy = fy + dy * cf; fz = fz + dz * cf;
                idx = skip[idx];
            } else {
                idx = idx + 1;
            }
        }
    }
    finish_one(i, deg_i, fx, fy, fz, xyz, old_force, att_off, att_adj, s, force, stats);
}

/* fa2_force of the brute-force engine (force_atlas2.wgsl:69-135): repulsion
 * from every other body j in index order, coefficient
 * scaling_ratio * deg_i * deg_j / max(|d|^2, 1e-4) along d = p_i - p_j
 * (:82-91), then the shared attraction / gravity / stats tail. */
static void brute_one(uint32_t i, uint32_t n, const float *xyz, const float *old_force,
                      const uint32_t *degrees, const uint32_t *att_off, const uint32_t *att_adj,
                      const jco_fa2_settings *s, float *force, float *stats) {
    const float *p = xyz + 3 * (size_t)i;
    float deg_i = (float)degrees[i] + 1.0f;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (uint32_t j = 0; j < n; j++) {
        if (j == i) continue;
        const float *q = xyz + 3 * (size_t)j;
        float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        float r2 = fmaxf(dx * dx + dy * dy + dz * dz, 1.0e-4f);
        float deg_j = (float)degrees[j] + 1.0f;
        float cf = s->scaling_ratio * deg_i * deg_j / r2;
        fx = fx + dx * cf; fy = fy + dy * cf; fz = fz + dz * cf;
    }
    finish_one(i, deg_i, fx, fy, fz, xyz, old_force, att_off, att_adj, s, force, stats);
}

void jco_fa2_brute_forces(uint32_t n, const float *xyz, const float *old_force,
                          const uint32_t *degrees, const uint32_t *att_off,
                          const uint32_t *att_adj, const jco_fa2_settings *s, float *force,
                          float *stats, int n_threads) {
    (void)n_threads;
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t i = 0; i < (int64_t)n; i++)
        brute_one((uint32_t)i, n, xyz, old_force, degrees, att_off, att_adj, s, force, stats);
}

void jco_fa2_forces(uint32_t n, const float *xyz, const float *old_force, const uint32_t *degrees,
                    const uint32_t *att_off, const uint32_t *att_adj, uint32_t n_tree,
                    const float *node_a, const float *half, const uint32_t *body,
                    const uint32_t *skip, const jco_fa2_settings *s, float *force, float *stats,
                    int n_threads) {
    (void)n_threads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t i = 0; i < (int64_t)n; i++)
        force_one((uint32_t)i, xyz, old_force, degrees, att_off, att_adj, n_tree, node_a, half, body,
                  skip, s, force, stats);
}

/* FP64 companion: same tree, same acceptance decisions as the f32 walk (taken
 * from the f32 test), contributions in f64 with f64 cell centres of mass.
 * flips counts nodes whose f64 acceptance test would have decided otherwise. */
void jco_fa2_forces_f64(uint32_t n, const float *xyz, const uint32_t *degrees,
                        const uint32_t *att_off, const uint32_t *att_adj, uint32_t N,
                        const float *node_a, const float *half, const uint32_t *body,
                        const uint32_t *skip, const jco_fa2_settings *s, double *force,
                        uint64_t *flips, int n_threads) {
    double *com64 = (double *)malloc(((size_t)N + 1) * 4 * sizeof(double));
    for (int64_t p = (int64_t)N - 1; p >= 0; p--) {
        double *o = com64 + 4 * p;
        if (body[p] != JCO_INTERNAL) {
            const float *q = xyz + 3 * (size_t)body[p];
            o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = node_a[4 * p + 3];
            continue;
        }
        double tm = 0, cx = 0, cy = 0, cz = 0;
        for (uint32_t ch = (uint32_t)p + 1; ch < N && ch != skip[p]; ch = skip[ch]) {
            const double *c = com64 + 4 * (size_t)ch;
            tm += c[3]; cx += c[0] * c[3]; cy += c[1] * c[3]; cz += c[2] * c[3];
        }
        if (tm > 0) { cx /= tm; cy /= tm; cz /= tm; }
        o[0] = cx; o[1] = cy; o[2] = cz; o[3] = tm;
    }
    float theta = fmaxf(s->theta, 0.0f);
    float theta2 = theta * theta;
    double th2d = (double)theta * (double)theta;
    uint64_t nflip = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(n_threads > 0 ? n_threads : 1) reduction(+ : nflip)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        uint32_t i = (uint32_t)ii;
        const float *p = xyz + 3 * (size_t)i;
        double px = p[0], py = p[1], pz = p[2];
        double deg_i = (double)degrees[i] + 1.0;
        double fx = 0, fy = 0, fz = 0;
        if (N > 0) {
            uint32_t idx = 0, walk = 0, cap = N * 4u > 16u ? N * 4u : 16u;
            while (idx < N && walk < cap) {
                walk++;
                const float *a = node_a + 4 * (size_t)idx;
                const double *a64 = com64 + 4 * (size_t)idx;
                float sz = half[idx] * 2.0f;
                float dx = p[0] - a[0], dy = p[1] - a[1], dz = p[2] - a[2];
                float r2 = dx * dx + dy * dy + dz * dz;
                double ex = px - a64[0], ey = py - a64[1], ez = pz - a64[2];
                double r2d = ex * ex + ey * ey + ez * ez;
                int take;
                if (body[idx] != JCO_INTERNAL) {
                    if (body[idx] != i && a[3] > 0.0f) {
                        double r2c = r2d > 1e-4 ? r2d : 1e-4;
                        double cf = (double)s->scaling_ratio * deg_i * a64[3] / r2c;
                        fx += ex * cf; fy += ey * cf; fz += ez * cf;
                    }
                    idx = skip[idx];
                    continue;
                }
                take = (a[3] > 0.0f && r2 > 0.0f && (sz * sz) < (theta2 * r2));
                int take64 = (a64[3] > 0 && r2d > 0 && (2.0 * half[idx]) * (2.0 * half[idx]) < th2d * r2d);
                nflip += (take != take64);
                if (take) {
                    double r2c = r2d > 1e-4 ? r2d : 1e-4;
                    double cf = (double)s->scaling_ratio * deg_i * a64[3] / r2c;
                    fx += ex * cf; fy += ey * cf; fz += ez * cf;
                    idx = skip[idx];
                } else {
                    idx = idx + 1;
                }
            }
        }
        for (uint32_t e = att_off[i]; e < att_off[i + 1]; e++) {
            const float *q = xyz + 3 * (size_t)att_adj[e];
            double dx = (double)q[0] - px, dy = (double)q[1] - py, dz = (double)q[2] - pz;
            double att = 1.0;
            if (s->lin_log_mode) {
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r < 1e-3) r = 1e-3;
                att = log(1.0 + r) / r;
            }
            fx += dx * att; fy += dy * att; fz += dz * att;
        }
        double r0 = sqrt(px * px + py * py + pz * pz);
        if (r0 < 1e-3) r0 = 1e-3;
        double gd = (double)s->gravity * deg_i;
        if (s->strong_gravity) { fx -= px * gd; fy -= py * gd; fz -= pz * gd; }
        else { fx -= px / r0 * gd; fy -= py / r0 * gd; fz -= pz / r0 * gd; }
        force[3 * (size_t)i] = fx; force[3 * (size_t)i + 1] = fy; force[3 * (size_t)i + 2] = fz;
    }
    if (flips) *flips = nflip;
    free(com64);
}

/* Brute-force FP64 forces: force_atlas2.wgsl:82-91 repulsion over all j != i
 * with (deg_i+1)(deg_j+1), plus the same attraction and gravity. */
void jco_fa2_brute_f64(uint32_t n, const float *xyz, const uint32_t *degrees,
                       const uint32_t *att_off, const uint32_t *att_adj,
                       const jco_fa2_settings *s, double *force) {
    for (uint32_t i = 0; i < n; i++) {
        double px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
        double di = (double)degrees[i] + 1.0, fx = 0, fy = 0, fz = 0;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) continue;
            double dx = px - xyz[3 * j], dy = py - xyz[3 * j + 1], dz = pz - xyz[3 * j + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < 1e-4) r2 = 1e-4;
            double cf = (double)s->scaling_ratio * di * ((double)degrees[j] + 1.0) / r2;
            fx += dx * cf; fy += dy * cf; fz += dz * cf;
        }
        for (uint32_t e = att_off[i]; e < att_off[i + 1]; e++) {
            const float *q = xyz + 3 * (size_t)att_adj[e];
            double dx = q[0] - px, dy = q[1] - py, dz = q[2] - pz, att = 1.0;
            if (s->lin_log_mode) {
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r < 1e-3) r = 1e-3;
                att = log(1.0 + r) / r;
            }
            fx += dx * att; fy += dy * att; fz += dz * att;
        }
        double r0 = sqrt(px * px + py * py + pz * pz);
        if (r0 < 1e-3) r0 = 1e-3;
        double gd = (double)s->gravity * di;
        if (s->strong_gravity) { fx -= px * gd; fy -= py * gd; fz -= pz * gd; }
        else { fx -= px / r0 * gd; fy -= py / r0 * gd; fz -= pz / r0 * gd; }
        force[3 * i] = fx; force[3 * i + 1] = fy; force[3 * i + 2] = fz;
    }
}

/* ------------------------------------------------------------------------- */
/* Exact sum of non-negative finite f32 values: a 320-bit fixed-point          */
/* accumulator (LSB = 2^-149) in ten 32-bit limbs held in u64 (carry room),    */
/* rounded once to the nearest f64 (ties to even).  Order-independent, so the  */
/* GPU's block-parallel sum reproduces it bit for bit.                         */
/* ------------------------------------------------------------------------- */
#define JCO_LIMBS 10
static void acc_add(uint64_t *limb, float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    u &= 0x7FFFFFFFu;
    uint32_t e = u >> 23, m = u & 0x7FFFFFu;
    if (e) m |= 0x800000u;
    uint32_t sh = (e ? e : 1) - 1;
    uint64_t t = (uint64_t)m << (sh & 31);
    limb[sh >> 5] += t & 0xFFFFFFFFull;
    limb[(sh >> 5) + 1] += t >> 32;
}

static double acc_to_double(const uint64_t *in) {
    uint64_t l[JCO_LIMBS];
    memcpy(l, in, sizeof l);
    for (int k = 0; k + 1 < JCO_LIMBS; k++) { l[k + 1] += l[k] >> 32; l[k] &= 0xFFFFFFFFull; }
    int top = -1;
    for (int k = JCO_LIMBS - 1; k >= 0; k--)
        if (l[k]) { top = 32 * k + 63 - __builtin_clzll(l[k]); break; }
    if (top < 0) return 0.0;
#define BIT(pos) ((pos) < 0 ? 0u : (uint32_t)((l[(pos) >> 5] >> ((pos) & 31)) & 1u))
    uint64_t mant = 0;
    int lo = top - 52;
    for (int b = top; b >= (lo > 0 ? lo : 0); b--) mant = (mant << 1) | BIT(b);
    if (lo <= 0) return ldexp((double)mant, -149 + (lo < 0 ? 0 : lo));
    uint32_t rb = BIT(lo - 1);
    uint32_t sticky = 0;
    for (int b = lo - 2; b >= 0 && !sticky; b--) sticky |= BIT(b);
#undef BIT
    if (rb && (sticky || (mant & 1))) mant++;
    return ldexp((double)mant, lo - 149);
}

/* (p - q) * 2^-149 rounded to f64, for carry-normalised p >= q. */
static double acc_diff_to_double(const uint64_t *pin, const uint64_t *qin) {
    uint64_t p[JCO_LIMBS], q[JCO_LIMBS];
    memcpy(p, pin, sizeof p);
    memcpy(q, qin, sizeof q);
    for (int k = 0; k + 1 < JCO_LIMBS; k++) {
        p[k + 1] += p[k] >> 32; p[k] &= 0xFFFFFFFFull;
        q[k + 1] += q[k] >> 32; q[k] &= 0xFFFFFFFFull;
    }
    int64_t borrow = 0;
    for (int k = 0; k < JCO_LIMBS; k++) {
        int64_t d = (int64_t)p[k] - (int64_t)q[k] - borrow;
        borrow = d < 0;
        p[k] = (uint64_t)(d + (borrow ? ((int64_t)1 << 32) : 0));
    }
    return acc_to_double(p);
}

void jco_exact_sum2(uint32_t n, const float *stats, double *out2) {
    uint64_t a[JCO_LIMBS + 1] = {0}, b[JCO_LIMBS + 1] = {0};
    for (uint32_t i = 0; i < n; i++) {
        float s = stats[2 * i], t = stats[2 * i + 1];
        if (isfinite(s) && isfinite(t)) { acc_add(a, s); acc_add(b, t); }
    }
    out2[0] = acc_to_double(a);
    out2[1] = acc_to_double(b);
}

/* fa2_speed.rs:48-100 AdaptiveSpeed::update.  speed_eff2 = {speed, eff}.
 * exact_sum = 0: f64 sums in node order (the reference); 1: exact sums. */
float jco_speed_update(float *se, uint32_t n, float jitter_tolerance, const float *stats,
                       int exact_sum, double *sums_out) {
    double ts = 0.0, tt = 0.0;
    if (exact_sum) {
        double o[2];
        jco_exact_sum2(n, stats, o);
        ts = o[0]; tt = o[1];
    } else {
        for (uint32_t i = 0; i < n; i++) {
            float swg = stats[2 * i], tra = stats[2 * i + 1];
            if (isfinite(swg) && isfinite(tra)) { ts += (double)swg; tt += (double)tra; }
        }
    }
    if (sums_out) { sums_out[0] = ts; sums_out[1] = tt; }
    if (!(ts > 0.0) || !(tt > 0.0) || n == 0) return se[0];
    double nn = (double)n, jt0 = (double)jitter_tolerance;
    double est = 0.05 * sqrt(nn);
    double min_jt = sqrt(est), max_jt = 10.0;
    double jt = jt0 * fmax(min_jt, fmin(max_jt, est * tt / (nn * nn)));
    const float min_eff = 0.05f;
    if (ts / tt > 2.0) {
        if (se[1] > min_eff) se[1] *= 0.5f;
        jt = fmax(jt, jt0);
    }
    double target = jt * (double)se[1] * tt / ts;
    if (ts > jt * tt) {
        if (se[1] > min_eff) se[1] *= 0.7f;
    } else if (se[0] < 1000.0f) {
        se[1] *= 1.3f;
    }
    double sp = (double)se[0];
    se[0] = (float)(sp + fmin(target - sp, 0.5 * sp));
    return se[0];
}

/* fa2_barnes_hut.wgsl:203-226 fa2_apply (ping-pong). */
void jco_fa2_apply(uint32_t n, const float *xyz_in, float *xyz_out, const float *force,
                   float *old_force, const uint32_t *degrees, const jco_fa2_settings *s,
                   float speed, int n_threads) {
    (void)n_threads;
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        size_t i = (size_t)ii;
        float fx = force[3 * i], fy = force[3 * i + 1], fz = force[3 * i + 2];
        float ox = old_force[3 * i], oy = old_force[3 * i + 1], oz = old_force[3 * i + 2];
        float deg_i = (float)degrees[i] + 1.0f;
        float swinging = deg_i * len3(ox - fx, oy - fy, oz - fz);
        float factor = speed / (1.0f + sqrtf(speed * swinging));
        float df = len3(fx, fy, fz);
        if (s->max_displacement > 0.0f && df > 1.0e-9f) factor = fminf(factor, s->max_displacement / df);
        float k = factor * s->time_step;
        xyz_out[3 * i] = xyz_in[3 * i] + fx * k;
        xyz_out[3 * i + 1] = xyz_in[3 * i + 1] + fy * k;
        xyz_out[3 * i + 2] = xyz_in[3 * i + 2] + fz * k;
        old_force[3 * i] = fx; old_force[3 * i + 1] = fy; old_force[3 * i + 2] = fz;
    }
}

/* Whole iterations (fa2_bh.rs:546-692): tree from the step-start positions,
 * force pass, speed update, apply.  tree_kind 0 = reference insertion octree
 * (capacity 2n+8, serial as in fa2_bh.rs:556), 1 = Morton pre-order octree,
 * 2 = no tree: the brute-force engine's iteration (fa2_brute.rs:454-565).
 * timings (optional, 3 doubles): seconds in tree build, force pass, rest. */
int jco_fa2_run(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float *xyz,
                float *old_force, float *se, uint32_t iters, const jco_fa2_settings *s,
                int tree_kind, int exact_sum, int n_threads, double *timings) {
    if (n == 0) return 0;
    uint32_t *deg = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *att_off = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    uint64_t nnz = offsets[n];
    uint32_t *att_adj = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    jco_fa2_prepare(n, offsets, nbrs, deg, att_off, att_adj);
    float *mass = (float *)malloc((size_t)n * sizeof(float));
    for (uint32_t i = 0; i < n; i++) mass[i] = (float)deg[i] + 1.0f;
    uint32_t cap = tree_kind == 0 ? 2 * n + 8 : tree_kind == 1 ? (JCO_LEVELS + 2) * n + 8 : 1;
    float *node_a = (float *)malloc((size_t)cap * 4 * sizeof(float));
    float *half = (float *)malloc((size_t)cap * sizeof(float));
    uint32_t *body = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
    uint32_t *skip = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
    float *force = (float *)malloc((size_t)n * 3 * sizeof(float));
    float *stats = (float *)malloc((size_t)n * 2 * sizeof(float));
    float *out = (float *)malloc((size_t)n * 3 * sizeof(float));
    double t_tree = 0, t_force = 0, t_rest = 0;
    for (uint32_t it = 0; it < iters; it++) {
        double t0 = now_s();
        uint32_t N = tree_kind == 0
                         ? jco_octree_ref(n, xyz, mass, cap, node_a, half, body, skip, NULL, NULL, NULL)
                     : tree_kind == 1
                         ? jco_octree_morton(n, xyz, mass, node_a, half, body, skip, NULL, NULL,
                                             NULL, NULL, NULL)
                         : 0u;
        double t1 = now_s();
        if (tree_kind == 2)
            jco_fa2_brute_forces(n, xyz, old_force, deg, att_off, att_adj, s, force, stats,
                                 n_threads);
        else
            jco_fa2_forces(n, xyz, old_force, deg, att_off, att_adj, N, node_a, half, body, skip,
                           s, force, stats, n_threads);
        double t2 = now_s();
        jco_speed_update(se, n, s->jitter_tolerance, stats, exact_sum, NULL);
        jco_fa2_apply(n, xyz, out, force, old_force, deg, s, se[0], n_threads);
        memcpy(xyz, out, (size_t)n * 3 * sizeof(float));
        double t3 = now_s();
        t_tree += t1 - t0; t_force += t2 - t1; t_rest += t3 - t2;
    }
    if (timings) { timings[0] = t_tree; timings[1] = t_force; timings[2] = t_rest; }
    free(deg); free(att_off); free(att_adj); free(mass); free(node_a); free(half); free(body);
    free(skip); free(force); free(stats); free(out);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* SpMV                                                                        */
/* ------------------------------------------------------------------------- */
/* cpu_spmv (spmv.rs:88-99) = gpu_spmv's thread per row (spmv.wgsl:27-41): acc
 * from +0, acc = acc + w[e] * x[nbr[e]] over the row in CSR order. */
void jco_spmv(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, const float *w,
              const float *x, float *y) {
    for (uint32_t v = 0; v < n; v++) {
        float acc = 0.0f;
        for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) acc = acc + w[e] * x[nbrs[e]];
        y[v] = acc;
    }
}

/* spmv_hybrid.wgsl:44-84: 64 lanes per row; lane l from +0 adds the entries
 * start+l, start+l+64, ...; then for stride 32, 16, ..., 1 every lane below the
 * stride adds partials[l + stride]; y = partials[0]. */
void jco_spmv_hybrid(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, const float *w,
                     const float *x, float *y) {
    float part[64];
    for (uint32_t v = 0; v < n; v++) {
        for (uint32_t l = 0; l < 64; l++) {
            float acc = 0.0f;
            for (uint32_t e = offsets[v] + l; e < offsets[v + 1]; e += 64)
                acc = acc + w[e] * x[nbrs[e]];
            part[l] = acc;
        }
        for (uint32_t stride = 32; stride > 0; stride >>= 1)
            for (uint32_t l = 0; l < stride; l++) part[l] = part[l] + part[l + stride];
        y[v] = part[0];
    }
}

/* IEEE binary16 from binary32, round to nearest, ties to even: what
 * f16::from_f32 gives pack_f16_pairs (spmv.rs:68-86).  Overflow goes to inf
 * (the rounding carry out of the largest exponent), results below half the
 * smallest subnormal to zero, NaN stays a quiet NaN with the top payload bits. */
uint16_t jco_f32_to_f16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u, ex = (u >> 23) & 0xFFu, man = u & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u | (man >> 13) : 0u));
    const int e = (int)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        const uint32_t m = man | 0x800000u;
        const uint32_t shift = (uint32_t)(14 - e);
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)h;
}

/* unpack2x16float's decode: binary16 to binary32, exact. */
float jco_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, ex = (h >> 10) & 0x1Fu;
    uint32_t man = h & 0x3FFu, u;
    if (ex == 0) {
        if (man == 0) {
            u = sign;
        } else {
            int k = 0;
            while (!(man & 0x400u)) {
                man <<= 1;
                k++;
            }
            u = sign | ((uint32_t)(113 - k) << 23) | ((man & 0x3FFu) << 13);
        }
    } else if (ex == 0x1Fu) {
        u = sign | 0x7F800000u | (man << 13);
    } else {
        u = sign | ((ex + 112u) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* gpu_spmv_f16 (spmv.rs:293-469, spmv_f16.wgsl:27-45): weights and x stored as
 * f16, each product f32(w) * f32(x) and the row sum in f32, in CSR order. */
void jco_spmv_f16(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, const float *w,
                  const float *x, float *y) {
    for (uint32_t v = 0; v < n; v++) {
        float acc = 0.0f;
        for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
            const float a = jco_f16_to_f32(jco_f32_to_f16(w[e]));
            const float b = jco_f16_to_f32(jco_f32_to_f16(x[nbrs[e]]));
            acc = acc + a * b;
        }
        y[v] = acc;
    }
}

/* ------------------------------------------------------------------------- */
/* PageRank                                                                    */
/* ------------------------------------------------------------------------- */
/* Pull form of pagerank.wgsl:30-46 with the host set-up of pagerank.rs:92-124;
 * stops when the L1 norm of (r_{t+1} - r_t) -- summed exactly and rounded once
 * to f64 (the stopping rule is this build's addition; the reference runs a
 * fixed count) -- is < tol, or after max_iters.  Returns -1 on a dangling node
 * (pagerank.rs:98-104). */
int jco_pagerank_pull(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float damping,
                      uint32_t max_iters, double tol_l1, float *ranks, uint32_t *iters_done,
                      double *resid, int n_threads) {
    const int nt = n_threads > 0 ? n_threads : 1;
    if (n == 0) { if (iters_done) *iters_done = 0; return 0; }
    float *inv = (float *)malloc((size_t)n * sizeof(float));
    for (uint32_t v = 0; v < n; v++) {
        uint32_t d = offsets[v + 1] - offsets[v];
        if (d == 0) { free(inv); return -1; }
        inv[v] = 1.0f / (float)d;
    }
    float inv_n = 1.0f / (float)n;
    float teleport = (1.0f - damping) * inv_n;
    float *a = (float *)malloc((size_t)n * sizeof(float)), *b = (float *)malloc((size_t)n * sizeof(float));
    for (uint32_t v = 0; v < n; v++) a[v] = inv_n;
    uint32_t it = 0;
    double r = 0.0;
    while (it < max_iters) {
        /* rows are independent; the residual's limbs add exactly in any order */
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
        for (int64_t vv = 0; vv < (int64_t)n; vv++) {
            uint32_t v = (uint32_t)vv;
            float acc = 0.0f;
            for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
                uint32_t u = nbrs[e];
                acc = acc + a[u] * inv[u];
            }
            b[v] = teleport + damping * acc;
        }
        uint64_t P[JCO_LIMBS + 1] = {0}, Q[JCO_LIMBS + 1] = {0};
#pragma omp parallel num_threads(nt)
        {
            uint64_t p[JCO_LIMBS + 1] = {0}, q[JCO_LIMBS + 1] = {0};
#pragma omp for schedule(static)
            for (int64_t vv = 0; vv < (int64_t)n; vv++) {
                acc_add(p, fmaxf(a[vv], b[vv]));
                acc_add(q, fminf(a[vv], b[vv]));
            }
#pragma omp critical
            for (int k = 0; k <= JCO_LIMBS; k++) { P[k] += p[k]; Q[k] += q[k]; }
        }
        r = acc_diff_to_double(P, Q);
        float *t = a; a = b; b = t;
        it++;
        if (r < tol_l1) break;
    }
    memcpy(ranks, a, (size_t)n * sizeof(float));
    if (iters_done) *iters_done = it;
    if (resid) *resid = r;
    free(inv); free(a); free(b);
    return 0;
}

/* pagerank.rs:298-331 cpu_pagerank (push, dangling mass spread uniformly). */
void jco_pagerank_push(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float damping,
                       uint32_t iters, float *ranks) {
    if (n == 0) return;
    float inv_n = 1.0f / (float)n;
    float teleport = (1.0f - damping) * inv_n;
    float *next = (float *)malloc((size_t)n * sizeof(float));
    for (uint32_t v = 0; v < n; v++) ranks[v] = inv_n;
    for (uint32_t it = 0; it < iters; it++) {
        for (uint32_t v = 0; v < n; v++) next[v] = 0.0f;
        float dangling = 0.0f;
        for (uint32_t v = 0; v < n; v++) {
            float od = (float)(offsets[v + 1] - offsets[v]);
            if (od == 0.0f) { dangling += ranks[v]; continue; }
            float share = ranks[v] / od;
            for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) next[nbrs[e]] += share;
        }
        float ds = damping * dangling * inv_n;
        for (uint32_t v = 0; v < n; v++) ranks[v] = teleport + ds + damping * next[v];
    }
    free(next);
}

/* graph-metrics/src/pagerank.rs:4-45 compute_pagerank in f64 (out-neighbours =
 * CSR rows). */
void jco_pagerank_f64(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, double damping,
                      uint32_t iters, double *ranks) {
    if (n == 0) return;
    double *nx = (double *)malloc((size_t)n * sizeof(double));
    for (uint32_t v = 0; v < n; v++) ranks[v] = 1.0 / n;
    for (uint32_t it = 0; it < iters; it++) {
        for (uint32_t v = 0; v < n; v++) nx[v] = (1.0 - damping) / n;
        for (uint32_t v = 0; v < n; v++) {
            uint32_t d = offsets[v + 1] - offsets[v];
            if (d == 0) {
                double share = damping * ranks[v] / n;
                for (uint32_t j = 0; j < n; j++) nx[j] += share;
            } else {
                double share = damping * ranks[v] / d;
                for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) nx[nbrs[e]] += share;
            }
        }
        memcpy(ranks, nx, (size_t)n * sizeof(double));
    }
    free(nx);
}

/* ------------------------------------------------------------------------- */
/* SGD sparse stress                                                          */
/* ------------------------------------------------------------------------- */
/* sgd_stress.rs:372-394 */
void jco_bfs(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, uint32_t src,
             uint32_t *dist) {
    uint32_t *q = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    for (uint32_t v = 0; v < n; v++) dist[v] = 0xFFFFFFFFu;
    /* analytics/bfs.rs:48-50: a source out of range reaches nothing */
    if (n == 0 || src >= n) { free(q); return; }
    uint32_t h = 0, t = 0;
    dist[src] = 0;
    q[t++] = src;
    while (h < t) {
        uint32_t v = q[h++];
        for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
            uint32_t u = nbrs[e];
            if (dist[u] == 0xFFFFFFFFu) { dist[u] = dist[v] + 1; q[t++] = u; }
        }
    }
    free(q);
}

/* analytics/connected_components.rs:45-66 cpu_connected_components: union-find
 * over every CSR entry taken as an undirected edge, path halving, the larger
 * root hooked under the smaller, so labels[v] = smallest node id of v's
 * component. */
static uint32_t jco_uf_find(uint32_t *parent, uint32_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

void jco_connected_components(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                              uint32_t *labels) {
    for (uint32_t v = 0; v < n; v++) labels[v] = v;
    for (uint32_t v = 0; v < n; v++)
        for (uint32_t e = offsets[v]; e < offsets[v + 1]; e++) {
            uint32_t rv = jco_uf_find(labels, v), ru = jco_uf_find(labels, nbrs[e]);
            if (rv != ru) {
                uint32_t lo = rv < ru ? rv : ru, hi = rv < ru ? ru : rv;
                labels[hi] = lo;
            }
        }
    for (uint32_t v = 0; v < n; v++) labels[v] = jco_uf_find(labels, v);
}

/* sgd_stress.rs:143-175 select_pivots (rng = SplitMix64(seed), first pivot
 * rng.below(n), then max-min).  dist[i*k+t] = hop distance as u16 with 0 for
 * self / unreachable (the sentinel of sgd_stress_gpu.rs:195-209).  Returns k,
 * or 0xFFFFFFFF when a finite distance exceeds 65535. */
uint32_t jco_select_pivots(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                           uint32_t n_pivots, uint64_t seed, uint32_t *pivots, uint16_t *dist) {
    if (n == 0) return 0;
    uint32_t k = n_pivots < 1 ? 1 : (n_pivots > n ? n : n_pivots);
    uint64_t s = seed;
    uint32_t *mind = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *d = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    for (uint32_t v = 0; v < n; v++) mind[v] = 0xFFFFFFFFu;
    uint32_t next = (uint32_t)(jco_next_u64(&s) % n);
    int bad = 0;
    for (uint32_t t = 0; t < k; t++) {
        jco_bfs(n, offsets, nbrs, next, d);
        uint32_t best_node = next, best = 0;
        for (uint32_t v = 0; v < n; v++) {
            uint32_t dv = d[v];
            if (dv != 0xFFFFFFFFu && dv < mind[v]) mind[v] = dv;
            uint32_t cover = mind[v];
            if (cover != 0xFFFFFFFFu && cover > best) { best = cover; best_node = v; }
            if (dv != 0xFFFFFFFFu && dv > 65535u) bad = 1;
            dist[(size_t)v * k + t] = (dv == 0xFFFFFFFFu || dv == 0) ? 0 : (uint16_t)dv;
        }
        pivots[t] = next;
        next = best_node;
    }
    free(mind); free(d);
    return bad ? 0xFFFFFFFFu : k;
}

/* sgd_stress.rs:324-334 anneal_eta */
float jco_anneal_eta(uint64_t sweep, float eta_max, float eta_min, uint32_t n_anneal) {
    if (n_anneal == 0) return eta_min;
    float t = (float)sweep / (float)n_anneal;
    if (t >= 1.0f) return eta_min;
    float ratio = fmaxf(eta_min / eta_max, 1e-9f);
    return eta_max * powf(ratio, t);
}

/* sgd_stress.wgsl:34-62 sgd_step (Jacobi across nodes, Gauss-Seidel over the
 * node's own k pivots).  Positions interleaved xyz. */
void jco_sgd_sweep(uint32_t n, uint32_t k, const uint32_t *pivots, const uint16_t *dist,
                   const float *pos_in, float *pos_out, float eta, int n_threads) {
    (void)n_threads;
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        uint32_t i = (uint32_t)ii;
        float px = pos_in[3 * (size_t)i], py = pos_in[3 * (size_t)i + 1], pz = pos_in[3 * (size_t)i + 2];
        for (uint32_t t = 0; t < k; t++) {
            uint32_t j = pivots[t];
            if (j == i) continue;
            uint16_t du = dist[(size_t)i * k + t];
            if (du == 0) continue;
            float d = (float)du;
            const float *q = pos_in + 3 * (size_t)j;
            float dx = px - q[0], dy = py - q[1], dz = pz - q[2];
            float mag = fmaxf(len3(dx, dy, dz), 1e-6f);
            float w = 1.0f / (d * d);
            float mu = fminf(1.0f, w * eta);
            float r = mu * (mag - d) / mag;
            px = px - r * dx; py = py - r * dy; pz = pz - r * dz;
        }
        pos_out[3 * (size_t)i] = px; pos_out[3 * (size_t)i + 1] = py; pos_out[3 * (size_t)i + 2] = pz;
    }
}

/* ------------------------------------------------------------------------- */
/* Metrics: metrics.rs:31-59 scale_normalized_stress, :150-164 edge_length_cv  */
/* ------------------------------------------------------------------------- */
static inline double euclid(const float *x, uint32_t i, uint32_t j) {
    double dx = (double)(x[3 * (size_t)i] - x[3 * (size_t)j]);
    double dy = (double)(x[3 * (size_t)i + 1] - x[3 * (size_t)j + 1]);
    double dz = (double)(x[3 * (size_t)i + 2] - x[3 * (size_t)j + 2]);
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* sgd_pair_update (sgd_stress.rs:340-366): w = 1/d^2, mu = min(1, w*eta),
 * r = 0.5*mu*(|D| - d)/|D| with |D| >= 1e-6; i moves by -r*D, j by +r*D. */
static void sgd_pair_update(float *pos, uint32_t i, uint32_t j, float d, float eta) {
    float xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
    float xj = pos[3 * (size_t)j], yj = pos[3 * (size_t)j + 1], zj = pos[3 * (size_t)j + 2];
    float dx = xi - xj, dy = yi - yj, dz = zi - zj;
    float mag = fmaxf(sqrtf(dx * dx + dy * dy + dz * dz), 1e-6f);
    float w = 1.0f / (d * d);
    float mu = fminf(w * eta, 1.0f);
    float r = 0.5f * mu * (mag - d) / mag;
    float rx = r * dx, ry = r * dy, rz = r * dz;
    pos[3 * (size_t)i] = xi - rx; pos[3 * (size_t)i + 1] = yi - ry; pos[3 * (size_t)i + 2] = zi - rz;
    pos[3 * (size_t)j] = xj + rx; pos[3 * (size_t)j + 1] = yj + ry; pos[3 * (size_t)j + 2] = zj + rz;
}

/* The reference CPU engine SgdStressEngine (sgd_stress.rs:251-316).  init:
 * SplitMix64(seed) picks the pivots (select_pivots draws once, :152), and the
 * same stream then samples the pairs.  step: sweeps_per_step (at least 1)
 * sweeps, each with eta = anneal_eta(sweep) and n samples i = below(n),
 * t = below(k), j = pivots[t] (:299-301); i == j and unreachable or zero
 * distances are skipped (:302-310); sgd_pair_update moves both endpoints in
 * place.  Runs `steps` steps on xyz (3n) in place.  Returns k, or 0xFFFFFFFF
 * when a hop distance exceeds the u16 table. */
uint32_t jco_sgd_cpu_run(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float *xyz,
                         uint32_t n_pivots, uint32_t sweeps_per_step, float eta_max, float eta_min,
                         uint32_t n_anneal, uint64_t seed, uint32_t steps) {
    if (n == 0) return 0;
    uint32_t k = n_pivots < 1 ? 1 : (n_pivots > n ? n : n_pivots);
    uint32_t *piv = (uint32_t *)malloc((size_t)k * sizeof(uint32_t));
    uint16_t *dist = (uint16_t *)malloc((size_t)n * k * sizeof(uint16_t));
    uint32_t got = jco_select_pivots(n, offsets, nbrs, n_pivots, seed, piv, dist);
    if (got == 0xFFFFFFFFu) { free(piv); free(dist); return got; }
    uint64_t s = seed;
    (void)jco_next_u64(&s);  /* the draw select_pivots made */
    uint64_t sweep = 0;
    uint32_t per = sweeps_per_step ? sweeps_per_step : 1;
    for (uint32_t st = 0; st < steps; st++) {
        for (uint32_t q = 0; q < per; q++) {
            float eta = jco_anneal_eta(sweep, eta_max, eta_min, n_anneal);
            for (uint32_t c = 0; c < n; c++) {
                uint32_t i = (uint32_t)(jco_next_u64(&s) % n);
                uint32_t t = (uint32_t)(jco_next_u64(&s) % k);
                uint32_t j = piv[t];
                if (i == j) continue;
                uint16_t d = dist[(size_t)i * k + t];
                if (d == 0) continue;
                sgd_pair_update(xyz, i, j, (float)d, eta);
            }
            sweep++;
        }
    }
    free(piv); free(dist);
    return k;
}
double jco_stress_normalized(const float *xyz, const uint32_t *ti, const uint32_t *tj,
                             const float *td, uint64_t n_terms) {
    double num = 0, den = 0, wd2 = 0;
    for (uint64_t q = 0; q < n_terms; q++) {
        if (td[q] <= 0.0f) continue;
        double d = td[q], w = 1.0 / (d * d), big = euclid(xyz, ti[q], tj[q]);
        num += w * d * big; den += w * big * big; wd2 += w * d * d;
    }
    if (den <= 0 || wd2 <= 0) return 0.0;
    double alpha = num / den, s = 0;
    for (uint64_t q = 0; q < n_terms; q++) {
        if (td[q] <= 0.0f) continue;
        double d = td[q], w = 1.0 / (d * d), r = alpha * euclid(xyz, ti[q], tj[q]) - d;
        s += w * r * r;
    }
    return (double)(float)(s / wd2);
}

double jco_edge_length_cv(const float *xyz, const uint32_t *ei, const uint32_t *ej,
                          uint64_t n_edges) {
    if (n_edges < 2) return 0.0;
    double sum = 0;
    for (uint64_t q = 0; q < n_edges; q++) sum += euclid(xyz, ei[q], ej[q]);
    double mean = sum / (double)n_edges;
    if (mean <= 1e-12) return 0.0;
    double var = 0;
    for (uint64_t q = 0; q < n_edges; q++) {
        double l = euclid(xyz, ei[q], ej[q]) - mean;
        var += l * l;
    }
    var /= (double)n_edges;
    return (double)(float)(sqrt(var) / mean);
}

/* ------------------------------------------------------------------------- */
/* Multilevel cascade: crates/graph-layouts/src/layout/coarsen.rs and          */
/* crates/graph-compute/src/engines/multilevel.rs                            */
/* ------------------------------------------------------------------------- */
/* coarsen.rs:96-160 contract_one_level.  edges = n_pairs (s,t) pairs.  Writes
 * parent[n] and the contracted pairs (first occurrences, original orientation,
 * self loops dropped) into out_edges (capacity n_pairs pairs); returns the
 * number of super-nodes, *out_pairs the number of contracted pairs. */
uint32_t jco_contract_level(uint32_t n, const uint32_t *edges, uint64_t n_pairs, uint32_t *parent,
                            uint32_t *out_edges, uint64_t *out_pairs) {
    for (uint32_t v = 0; v < n; v++) parent[v] = 0xFFFFFFFFu;
    uint32_t next_super = 0;
    for (uint64_t i = 0; i < n_pairs; i++) {
        uint32_t s = edges[2 * i], t = edges[2 * i + 1];
        if (s == t || s >= n || t >= n) continue;
        if (parent[s] == 0xFFFFFFFFu && parent[t] == 0xFFFFFFFFu) {
            parent[s] = next_super;
            parent[t] = next_super;
            next_super++;
        }
    }
    for (uint32_t v = 0; v < n; v++)
        if (parent[v] == 0xFFFFFFFFu) parent[v] = next_super++;
    /* seen-set: open addressing over (min << 32 | max) */
    uint64_t cap = 16;
    while (cap < 2 * n_pairs + 1) cap <<= 1;
    uint64_t *slot = (uint64_t *)malloc(cap * sizeof(uint64_t));
    for (uint64_t q = 0; q < cap; q++) slot[q] = ~0ull;
    uint64_t m = 0;
    for (uint64_t j = 0; j < n_pairs; j++) {
        uint32_t es = edges[2 * j], et = edges[2 * j + 1];
        if (es >= n || et >= n) continue;
        uint32_t a = parent[es], b = parent[et];
        if (a == b) continue;
        uint64_t key = a < b ? (((uint64_t)a << 32) | b) : (((uint64_t)b << 32) | a);
        uint64_t h = (key ^ (key >> 29)) * 0xBF58476D1CE4E5B9ull & (cap - 1);
        int fresh = 1;
        while (slot[h] != ~0ull) {
            if (slot[h] == key) { fresh = 0; break; }
            h = (h + 1) & (cap - 1);
        }
        if (!fresh) continue;
        slot[h] = key;
        out_edges[2 * m] = a;
        out_edges[2 * m + 1] = b;
        m++;
    }
    free(slot);
    *out_pairs = m;
    return next_super;
}

/* multilevel.rs:486-511 fold_up, one level: coarse[p] = f32 mean of its
 * children (sums in child order, then one division by the count). */
void jco_fold_up_level(uint32_t n_fine, const float *fine, const uint32_t *parent_map, uint32_t n_coarse,
                       float *coarse) {
    uint32_t *cnt = (uint32_t *)calloc(n_coarse ? n_coarse : 1, sizeof(uint32_t));
    for (uint64_t q = 0; q < 3 * (uint64_t)n_coarse; q++) coarse[q] = 0.0f;
    for (uint32_t c = 0; c < n_fine; c++) {
        uint32_t p = parent_map[c];
        if (p >= n_coarse) continue;
        coarse[3 * (size_t)p] += fine[3 * (size_t)c];
        coarse[3 * (size_t)p + 1] += fine[3 * (size_t)c + 1];
        coarse[3 * (size_t)p + 2] += fine[3 * (size_t)c + 2];
        cnt[p]++;
    }
    for (uint32_t p = 0; p < n_coarse; p++) {
        float k = (float)(cnt[p] ? cnt[p] : 1u);
        coarse[3 * (size_t)p] /= k;
        coarse[3 * (size_t)p + 1] /= k;
        coarse[3 * (size_t)p + 2] /= k;
    }
    free(cnt);
}

/* coarsen.rs:166-193 prolong: Vigna's xorshift64* (shifts 12, 25, 27,
 * multiplier 0x2545F4914F6CDD1D) from seed + 0x9E3779B97F4A7C15; u = the top
 * 53 bits as a double in [0,1), rounded to f32, times 2 minus 1. */
void jco_prolong(const float *parent_pos, const uint32_t *parent_map, uint32_t n_child, float jitter,
                 uint32_t seed, float *out) {
    uint64_t s = (uint64_t)seed + 0x9E3779B97F4A7C15ull;
    for (uint32_t i = 0; i < n_child; i++) {
        size_t b = 3 * (size_t)parent_map[i];
        for (int q = 0; q < 3; q++) {
            s ^= s >> 12;
            s ^= s << 25;
            s ^= s >> 27;
            uint64_t v = s * 0x2545F4914F6CDD1Dull;
            float u = (float)((double)(v >> 11) / 9007199254740992.0);
            u = u * 2.0f - 1.0f;
            out[3 * (size_t)i + q] = parent_pos[b + q] + u * jitter;
        }
    }
}

/* multilevel.rs:229-272 level_csr: rows filled in edge order. */
void jco_level_csr(uint32_t n, const uint32_t *edges, uint64_t n_pairs, uint32_t *offsets, uint32_t *nbrs) {
    for (uint32_t v = 0; v <= n; v++) offsets[v] = 0;
    for (uint64_t i = 0; i < n_pairs; i++) {
        uint32_t s = edges[2 * i], t = edges[2 * i + 1];
        if (s < n && t < n && s != t) { offsets[s + 1]++; offsets[t + 1]++; }
    }
    for (uint32_t v = 0; v < n; v++) offsets[v + 1] += offsets[v];
    uint32_t *cur = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    memcpy(cur, offsets, ((size_t)n + 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i < n_pairs; i++) {
        uint32_t s = edges[2 * i], t = edges[2 * i + 1];
        if (s < n && t < n && s != t) { nbrs[cur[s]++] = t; nbrs[cur[t]++] = s; }
    }
    free(cur);
}

