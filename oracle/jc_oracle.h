/*
 * jc_oracle.h -- CPU oracle for the layout / graph-compute hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference algorithms
 * (ocasazza/jump-cannon, Rust + WGSL) in plain C so that the HIP kernels can be
 * checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or as the timed CPU
 * baseline -- never as the product path.
 *
 * Every function cites the reference file:line it follows.  Built with
 * -ffp-contract=off and no fast-math so f32 host semantics match Rust's (Rust
 * never contracts a*b+c into an FMA).
 *
 * Parity status per function is recorded in the header comment of
 * jc_oracle.c ("pinned by" / "parity unpinned").
 */
#ifndef JC_ORACLE_H
#define JC_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JCO_END 0xFFFFFFFFu
#define JCO_INTERNAL 0xFFFFFFFFu

/* ---- SplitMix64 + fixtures (crates/graph-compute/src/stability.rs:9-110) ---- */
uint64_t jco_next_u64(uint64_t *state);
float jco_next_f32(uint64_t *state);
void jco_uniform_cube(uint32_t n, float half, uint64_t seed, float *xyz);
void jco_ball_seed(uint32_t n, uint64_t seed, float *xyz);
uint64_t jco_ba_pairs(uint32_t n, uint32_t m, uint64_t seed, uint32_t *pairs, uint64_t cap);
uint64_t jco_sbm_pairs(uint32_t n, uint32_t blocks, double p_in, double p_out, uint64_t seed,
                       uint32_t *pairs, uint64_t cap);
uint64_t jco_rgg_pairs(uint32_t n, double radius, uint64_t seed, float *xy, uint32_t *pairs,
                       uint64_t cap);
uint64_t jco_csr_from_pairs(uint32_t n, const uint32_t *pairs, uint64_t npairs, uint32_t *offsets,
                            uint32_t *nbrs);
uint32_t jco_largest_component(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                               uint32_t *keep_map);

/* ---- FA2 preparation (fa2_bh.rs:264-294) ---- */
uint64_t jco_fa2_prepare(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                         uint32_t *degrees, uint32_t *att_off, uint32_t *att_adj);

/* ---- Octrees in canonical pre-order form ----
 * node_a[4*k] = com.x, com.y, com.z, mass; half[k] = cell half extent;
 * body[k] = body index or JCO_INTERNAL; skip[k] = pre-order rope or JCO_END;
 * depth[k].  For an internal node with children, next == k+1. */
/* centre[3k] (nullable) = the cell's centre (for a leaf: the cell the body
 * was placed in).  Insertion stops at the first body that does not fit in
 * max_nodes and the partial tree is returned (fa2_bh.rs:813-819, 881-883). */
uint32_t jco_octree_ref(uint32_t n, const float *xyz, const float *mass, uint32_t max_nodes,
                        float *node_a, float *half, uint32_t *body, uint32_t *skip,
                        uint32_t *depth, float *root4, float *centre);
uint32_t jco_octree_morton(uint32_t n, const float *xyz, const float *mass, float *node_a,
                           float *half, uint32_t *body, uint32_t *skip, uint32_t *depth,
                           uint64_t *keys_sorted, uint32_t *perm, uint32_t *n_valid,
                           float *root4);
void jco_morton_keys(uint32_t n, const float *xyz, uint64_t *keys, float *root4);

/* ---- FA2 force / apply / speed ---- */
typedef struct {
    float gravity;
    uint32_t strong_gravity;
    float scaling_ratio;
    float edge_weight_influence;
    float jitter_tolerance;
    uint32_t lin_log_mode;
    uint32_t prevent_overlap;
    float time_step;
    float theta;
    float max_displacement;
} jco_fa2_settings;

void jco_fa2_forces(uint32_t n, const float *xyz, const float *old_force, const uint32_t *degrees,
                    const uint32_t *att_off, const uint32_t *att_adj, uint32_t n_tree,
                    const float *node_a, const float *half, const uint32_t *body,
                    const uint32_t *skip, const jco_fa2_settings *s, float *force, float *stats,
                    int n_threads);
void jco_fa2_forces_f64(uint32_t n, const float *xyz, const uint32_t *degrees,
                        const uint32_t *att_off, const uint32_t *att_adj, uint32_t n_tree,
                        const float *node_a, const float *half, const uint32_t *body,
                        const uint32_t *skip, const jco_fa2_settings *s, double *force,
                        uint64_t *flips, int n_threads);
/* f32 brute-force force pass (force_atlas2.wgsl:69-135): repulsion from every
 * other body in index order, then the same attraction, gravity and stats. */
void jco_fa2_brute_forces(uint32_t n, const float *xyz, const float *old_force,
                          const uint32_t *degrees, const uint32_t *att_off,
                          const uint32_t *att_adj, const jco_fa2_settings *s, float *force,
                          float *stats, int n_threads);
void jco_fa2_brute_f64(uint32_t n, const float *xyz, const uint32_t *degrees,
                       const uint32_t *att_off, const uint32_t *att_adj,
                       const jco_fa2_settings *s, double *force);
void jco_exact_sum2(uint32_t n, const float *stats, double *out2);
float jco_speed_update(float *speed_eff2, uint32_t n, float jitter_tolerance, const float *stats,
                       int exact_sum, double *sums_out);
void jco_fa2_apply(uint32_t n, const float *xyz_in, float *xyz_out, const float *force,
                   float *old_force, const uint32_t *degrees, const jco_fa2_settings *s,
                   float speed, int n_threads);
int jco_fa2_run(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float *xyz,
                float *old_force, float *speed_eff2, uint32_t iters, const jco_fa2_settings *s,
                int tree_kind, int exact_sum, int n_threads, double *timings);

/* ---- PageRank (analytics/pagerank.rs:67-331, pagerank.wgsl:30-46) ---- */
int jco_pagerank_pull(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float damping,
                      uint32_t max_iters, double tol_l1, float *ranks, uint32_t *iters_done,
                      double *resid, int n_threads);
void jco_pagerank_push(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float damping,
                       uint32_t iters, float *ranks);
void jco_pagerank_f64(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, double damping,
                      uint32_t iters, double *ranks);

/* ---- SpMV (analytics/spmv.rs, spmv.wgsl, spmv_f16.wgsl, spmv_hybrid.wgsl) ---- */
void jco_spmv(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, const float *w,
              const float *x, float *y);
void jco_spmv_hybrid(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, const float *w,
                     const float *x, float *y);
uint16_t jco_f32_to_f16(float f);
float jco_f16_to_f32(uint16_t h);
void jco_spmv_f16(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, const float *w,
                  const float *x, float *y);

/* ---- SGD sparse stress (engines/sgd_stress.rs, sgd_stress.wgsl) ---- */
void jco_bfs(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, uint32_t src,
             uint32_t *dist);
void jco_connected_components(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                              uint32_t *labels);
uint32_t jco_select_pivots(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs,
                           uint32_t n_pivots, uint64_t seed, uint32_t *pivots, uint16_t *dist);
float jco_anneal_eta(uint64_t sweep, float eta_max, float eta_min, uint32_t n_anneal);
void jco_sgd_sweep(uint32_t n, uint32_t k, const uint32_t *pivots, const uint16_t *dist,
                   const float *pos_in, float *pos_out, float eta, int n_threads);

/* ---- metrics (crates/graph-layouts/src/metrics.rs:17-164) ---- */
uint32_t jco_sgd_cpu_run(uint32_t n, const uint32_t *offsets, const uint32_t *nbrs, float *xyz,
                         uint32_t n_pivots, uint32_t sweeps_per_step, float eta_max, float eta_min,
                         uint32_t n_anneal, uint64_t seed, uint32_t steps);
double jco_stress_normalized(const float *xyz, const uint32_t *ti, const uint32_t *tj,
                             const float *td, uint64_t n_terms);
double jco_edge_length_cv(const float *xyz, const uint32_t *ei, const uint32_t *ej,
                          uint64_t n_edges);

/* multilevel cascade (coarsen.rs, multilevel.rs) */
uint32_t jco_contract_level(uint32_t n, const uint32_t *edges, uint64_t n_pairs, uint32_t *parent,
                            uint32_t *out_edges, uint64_t *out_pairs);
void jco_fold_up_level(uint32_t n_fine, const float *fine, const uint32_t *parent_map, uint32_t n_coarse,
                       float *coarse);
void jco_prolong(const float *parent_pos, const uint32_t *parent_map, uint32_t n_child, float jitter,
                 uint32_t seed, float *out);
void jco_level_csr(uint32_t n, const uint32_t *edges, uint64_t n_pairs, uint32_t *offsets, uint32_t *nbrs);

#ifdef __cplusplus
}
#endif
#endif
