/*
 * jc.h -- C ABI of libjc_hip.so, the MI355X (gfx950) layout / graph-compute
 * engines.  Every entry point is plain C: pointers and sizes, no C++ or torch
 * types, no exception crosses it.  Reference paths are relative to the root of
 * ocasazza/jump-cannon.
 *
 * Conventions
 *   - Return 0 (JC_OK) on success, a negative JC_E_* code otherwise; the
 *     message of the last failure on this thread is jc_last_error().
 *   - Callers own every input and output buffer; inputs are copied at init.
 *   - One engine handle is driven by one thread at a time; separate handles are
 *     independent.  Each engine owns one HIP stream and its device buffers.
 *   - Positions are interleaved x,y,z f32, length 3*n, as at the reference
 *     LayoutEngine boundary.
 */
#ifndef JC_H
#define JC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    JC_OK = 0,
    JC_E_NODEVICE = -1, /* no HIP device: the reference's "requires a wgpu device" Err */
    JC_E_INVALID = -2,  /* malformed input (CSR, positions, settings JSON) */
    JC_E_OOM = -3,      /* device allocation failed */
    JC_E_HIP = -4,      /* a HIP runtime call failed */
    JC_E_STATE = -5,    /* call order violated (e.g. step before init) */
    JC_E_DANGLING = -6  /* PageRank input has a degree-0 node (pagerank.rs:92-106) */
};

typedef struct jc_engine jc_engine;

/* Field-for-field Fa2BhSettings (crates/graph-compute/src/engines/fa2_bh.rs:85-125);
 * bools as u32.  prevent_overlap is accepted and unused, as in the reference. */
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
} jc_fa2bh_settings;

/* Field-for-field Fa2Settings of the brute-force engine
 * (crates/graph-compute/src/engines/fa2_brute.rs:63-93); bools as u32. */
typedef struct {
    float gravity;
    uint32_t strong_gravity;
    float scaling_ratio;
    float edge_weight_influence;
    float jitter_tolerance;
    uint32_t lin_log_mode;
    uint32_t prevent_overlap;
    float time_step;
    float max_displacement;
} jc_fa2_settings;

/* Field-for-field SgdStressSettings (crates/graph-compute/src/engines/sgd_stress.rs:59-98).
 * eta_max <= 0 selects the s_gd2 start 1/w_min = (max pivot distance)^2. */
typedef struct {
    uint32_t n_pivots;
    uint32_t sweeps_per_step;
    float eta_max;
    float eta_min;
    uint32_t n_anneal_steps;
    uint32_t reserved;
    uint64_t seed;
} jc_sgd_settings;

/* LayoutDescriptor (crates/graph-layouts/src/layout/layout_trait.rs:35-42). */
typedef struct {
    const char *id;
    const char *kind;
    const char *display_name;
    const char *description;
    uint32_t needs_edges;
    uint32_t needs_cpu_positions;
    uint32_t needs_gpu_positions_buffer;
} jc_descriptor;

/* Thread-local message of the last failing call (empty if none). */
const char *jc_last_error(void);
/* Number of visible HIP devices (0 when there is none). */
int jc_device_count(void);

/* ---- engines: the LayoutEngine trait (crates/graph-compute/src/engines/mod.rs:383-444) ---- */

/* Fa2BhSettings::default() (fa2_bh.rs:109-125). */
void jc_fa2bh_default_settings(jc_fa2bh_settings *out);
/* Fa2BhEngine::new (fa2_bh.rs:179-186), registry id "fa2-bh-hip". */
int jc_fa2bh_create(const jc_fa2bh_settings *settings, jc_engine **out);
/* Fa2Settings::default() (fa2_brute.rs:80-93). */
void jc_fa2_default_settings(jc_fa2_settings *out);
/* Fa2BruteEngine::new (fa2_brute.rs:135-141), registry id "fa2-brute-hip": O(n^2)
 * repulsion in index order (force_atlas2.wgsl:69-91), then the attraction,
 * gravity, AdaptiveSpeed and apply of fa2-bh-hip.  The fa2bh parity and
 * measurement hooks below accept it (it has no octree: jc_fa2bh_tree reports
 * none, and the walk mode is ignored). */
int jc_fa2brute_create(const jc_fa2_settings *settings, jc_engine **out);
/* SgdStressSettings::default() (sgd_stress.rs:84-98). */
void jc_sgd_default_settings(jc_sgd_settings *out);
/* SgdStressGpuEngine::new (sgd_stress_gpu.rs:96-120), registry id "sgd-stress-hip". */
int jc_sgd_create(const jc_sgd_settings *settings, jc_engine **out);

/* LayoutEngine::descriptor (mod.rs:393). */
int jc_engine_descriptor(const jc_engine *e, jc_descriptor *out);
/* LayoutEngine::set_params (mod.rs:400-402; fa2_bh.rs:212-220): serde-compatible JSON
 * object; "null" keeps the current settings; a missing key takes its default; an
 * unknown key is ignored; a wrongly typed value is JC_E_INVALID. */
int jc_engine_set_params_json(jc_engine *e, const char *json, size_t len);
/* LayoutEngine::init (mod.rs:407-412; fa2_bh.rs:222-544): CSR offsets[n+1],
 * neighbors[nnz], xyz[3n].  JC_E_NODEVICE when no GPU is visible. */
int jc_engine_init(jc_engine *e, uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                   uint64_t nnz, const float *xyz);
/* LayoutEngine::reinit (mod.rs:422-431): same contract as init. */
int jc_engine_reinit(jc_engine *e, uint32_t n, const uint32_t *offsets,
                     const uint32_t *neighbors, uint64_t nnz, const float *xyz);
/* LayoutEngine::step (mod.rs:435; fa2_bh.rs:546-692) run n_iters times on the device.
 * out_xyz (3n) receives the final positions, or NULL for no read-back (the
 * positions stay resident in HBM).  The call returns once the work is queued
 * when out_xyz is NULL; jc_engine_sync waits for it. */
int jc_engine_step(jc_engine *e, uint32_t n_iters, float *out_xyz);
int jc_engine_sync(jc_engine *e);
/* Checkpoint / resume: positions, F_{t-1} (3n each), {speed, speed_efficiency}
 * (the AdaptiveSpeed state, fa2_speed.rs:25-39) and the iteration count. */
int jc_engine_get_state(jc_engine *e, float *xyz, float *old_force, float *speed_eff2,
                        uint64_t *iteration);
int jc_engine_set_state(jc_engine *e, const float *xyz, const float *old_force,
                        const float *speed_eff2, uint64_t iteration);
/* LayoutEngine::is_halted (mod.rs:442-444): always 0 for these engines. */
int jc_engine_is_halted(const jc_engine *e);
void jc_engine_destroy(jc_engine *e);

/* ---- FA2+BH parity and measurement hooks (no counterpart in the reference) ---- */

/* One force pass at positions xyz (tree build + walk + attraction + gravity)
 * against the engine's current F_{t-1}; writes F (3n) and the exact
 * {sum swing, sum traction}.  Engine state is left unchanged. */
int jc_fa2bh_forces_at(jc_engine *e, const float *xyz, float *out_force, double *out_sums2);
/* Copies the octree built by the last tree build: sorted Morton keys and
 * permutation (n each), node arrays (cap each): node_a = com.xyz,mass (4 f32);
 * node_info = internal: bits of (2*half)^2, leaf: 0x80000000|body;
 * node_skip = pre-order rope (0xFFFFFFFF = end); node_depth = depth, 0x80 set
 * on leaves.  root4 = centre xyz, half.  Any output pointer may be NULL. */
int jc_fa2bh_tree(jc_engine *e, uint32_t cap, uint32_t *n_tree, uint32_t *n_valid, uint64_t *keys,
                  uint32_t *perm, float *node_a, uint32_t *node_info, uint32_t *node_skip,
                  uint8_t *node_depth, float *root4);
/* Runs n_iters iterations (the octree build replayed from a captured hipGraph)
 * with HIP events around the force kernel of every iteration, on the engine
 * stream.  Returns the summed force-kernel time and the summed whole-iteration
 * time (ms). */
int jc_fa2bh_step_timed(jc_engine *e, uint32_t n_iters, double *force_ms_sum,
                        double *iter_ms_sum);
/* Octree size and the exact {sum swing, sum traction} of the last iteration, and
 * the number of interactions summed by the walk (accepted cells plus leaf bodies)
 * in the last iteration run by jc_fa2bh_profile or jc_fa2bh_forces_at (plain
 * steps do not count them: the count is a separate walk, 0 after a plain step). */
int jc_fa2bh_last_stats(jc_engine *e, uint64_t *n_tree, uint64_t *interactions, double *sums2);
/* Octree walk of the force kernel: 0 = every lane follows its own rope (the
 * top three octree levels read from LDS), 1 = wave-union walk (uniform node
 * loads), 2 / 3 = each lane interleaves the walks of 2 / 4 Morton-adjacent
 * bodies, 4 / 5 / 6 = mode 0 in 512- / 1024- / 64-thread blocks.  All give bit-identical
 * forces; by default the engine times them at its first step and keeps the
 * fastest.  mode 0..6 forces one, -1 re-times; *current receives the mode in use.
 * fa2-brute-hip has three repulsion kernels, timed the same way: 0 = one lane
 * per body, 1 / 2 = 4 / 16 lanes per body (for graphs too small to fill the
 * chip one lane per body). */
int jc_fa2bh_walk_mode(jc_engine *e, int mode, int *current);
/* Eager (uncaptured) iterations with events around every phase; phase_ms[k]
 * receives the mean ms of phase k over n_iters.  Phase names: jc_fa2bh_phase_name. */
int jc_fa2bh_profile(jc_engine *e, uint32_t n_iters, double *phase_ms, uint32_t n_phases);
const char *jc_fa2bh_phase_name(uint32_t k);
uint32_t jc_fa2bh_phase_count(void);

/* ---- SGD stress hooks ---- */
/* Pivots chosen at init (sgd_stress.rs:143-175), their hop distances in the
 * reference's node-major layout dist[i*k + t] (0 = self/unreachable), and the
 * eta_max in force (settings.eta_max, or max distance^2 when eta_max <= 0).
 * Call with pivots = dist = NULL first to learn k. */
int jc_sgd_pivots(jc_engine *e, uint32_t *k, uint32_t *pivots, uint16_t *dist, float *eta_max);
/* Runs n_steps steps (LayoutEngine::step without read-back) between HIP events
 * on the engine stream and returns their device time in ms. */
int jc_sgd_time(jc_engine *e, uint32_t n_steps, double *gpu_ms);

/* ---- layout quality: scale-normalized stress over pivot terms ----
 * scale_normalized_stress (crates/graph-layouts/src/metrics.rs:31-59) over the
 * terms (i, pivot t, hop distance d > 0) of a pivot distance table in the
 * node-major layout dist[i*k + t] (as jc_sgd_pivots returns it), f64 sums. */
typedef struct jc_stress_plan jc_stress_plan;
int jc_stress_plan_create(uint32_t n, uint32_t k, const uint32_t *pivots, const uint16_t *dist,
                          jc_stress_plan **out);
int jc_stress_plan_eval(jc_stress_plan *p, const float *xyz, double *out_stress);
void jc_stress_plan_destroy(jc_stress_plan *p);

/* ---- PageRank (crates/graph-compute/src/analytics/pagerank.rs:67-287) ----
 * Pull power iteration: rank_{t+1}[v] = (1-d)/n + d * sum_{u in N(v)} rank_t[u]/deg(u),
 * rank_0 = 1/n, stopping when the f64 L1 norm of rank_{t+1}-rank_t < tol_l1 or
 * after max_iters.  JC_E_DANGLING when a node has no neighbours. */
int jc_pagerank(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                float damping, uint32_t max_iters, double tol_l1, float *out_ranks,
                uint32_t *out_iters, double *out_residual_l1);
/* The same computation split into set-up (CSR and 1/deg uploaded once) and runs
 * over the resident graph; out_gpu_ms (nullable) is the device time of the run
 * measured with HIP events.  plan_time runs n_iters iterations with the
 * stopping rule disabled and returns the mean device ms per iteration. */
typedef struct jc_pagerank_plan jc_pagerank_plan;
int jc_pagerank_plan_create(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                            uint64_t nnz, float damping, jc_pagerank_plan **out);
int jc_pagerank_plan_run(jc_pagerank_plan *plan, uint32_t max_iters, double tol_l1,
                         float *out_ranks, uint32_t *out_iters, double *out_residual_l1,
                         double *out_gpu_ms);
int jc_pagerank_plan_time(jc_pagerank_plan *plan, uint32_t n_iters, double *ms_per_iter);
void jc_pagerank_plan_destroy(jc_pagerank_plan *plan);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): jc_pagerank over a CSR the caller already holds in device
 * memory, with no device allocation.  d_offsets (n+1), d_neighbors (nnz) and
 * d_out_ranks (n) are device pointers; stream is a hipStream_t (NULL: the null
 * stream); out_iters / out_residual_l1 are host pointers (nullable).  Same
 * iteration, stopping rule and result bits as jc_pagerank; the CSR is checked
 * on the device (JC_E_INVALID, JC_E_DANGLING as jc_pagerank).  The workspace
 * (1/deg, two rank and two x vectors, the residual accumulator and state) is
 * one static device array of about 20.0 MiB per device, held until the process
 * ends; n above JC_PAGERANK_NOALLOC_MAX_N is JC_E_INVALID.  Every call on a
 * device shares that device's workspace: calls that overlap on one device
 * race and must be serialised by the caller. */
#define JC_PAGERANK_NOALLOC_MAX_N (1u << 20)
int jc_pagerank_noalloc(uint32_t n, const uint32_t *d_offsets, const uint32_t *d_neighbors, uint64_t nnz,
                        float damping, uint32_t max_iters, double tol_l1, float *d_out_ranks,
                        uint32_t *out_iters, double *out_residual_l1, void *stream);

/* ---- SpMV (crates/graph-compute/src/analytics/spmv.rs) ----
 * y = A x for a weighted CSR matrix (WeightedCsr, spmv.rs:26-32: n rows, offsets
 * [n+1], neighbors and weights [nnz]); x and y hold n floats.  n == 0 returns
 * JC_OK and touches nothing (Ok(Vec::new())).  Rows are summed in f32:
 *   jc_spmv         gpu_spmv (spmv.rs:103-287): each row in CSR order, as cpu_spmv
 *                   (:88-99) does;
 *   jc_spmv_f16     gpu_spmv_f16 (:293-469): weights and x rounded to f16
 *                   (nearest-even, as pack_f16_pairs :68-86), products and row sums in
 *                   f32, in CSR order;
 *   jc_spmv_hybrid  gpu_spmv_hybrid (:491-676): lane l of 64 sums entries start+l,
 *                   start+l+64, ...; a tree over the 64 partials (stride 32 .. 1)
 *                   gives the row (shaders/spmv_hybrid.wgsl).
 * JC_E_INVALID for a malformed CSR (the reference indexes out of bounds) or a NULL
 * weights / x / y pointer. */
int jc_spmv(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, const float *weights,
            uint64_t nnz, const float *x, float *y);
int jc_spmv_f16(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                const float *weights, uint64_t nnz, const float *x, float *y);
int jc_spmv_hybrid(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                   const float *weights, uint64_t nnz, const float *x, float *y);
/* The matrix uploaded once (kind: JC_SPMV_F32 / _F16 / _HYBRID, the three functions
 * above) and products over it; out_gpu_ms (nullable) is the device time of the
 * product.  plan_time runs n_iters products of the x of the last run and returns
 * the mean device ms per product. */
enum { JC_SPMV_F32 = 0, JC_SPMV_F16 = 1, JC_SPMV_HYBRID = 2 };
typedef struct jc_spmv_plan jc_spmv_plan;
int jc_spmv_plan_create(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                        const float *weights, uint64_t nnz, int kind, jc_spmv_plan **out);
int jc_spmv_plan_run(jc_spmv_plan *plan, const float *x, float *y, double *out_gpu_ms);
int jc_spmv_plan_time(jc_spmv_plan *plan, uint32_t n_iters, double *ms_per_product);
void jc_spmv_plan_destroy(jc_spmv_plan *plan);

/* ---- multilevel wrapper ----
 * Replaces MultilevelEngine (crates/graph-compute/src/engines/multilevel.rs:196-480),
 * registry id "multilevel-hip": coarsen (greedy matching, coarsen.rs:53-160) ->
 * solve the coarsest level -> prolong (coarsen.rs:166-193) -> refine, with
 * "fa2-brute-hip" (the default, as multilevel.rs:176 has fa2-brute), "fa2-bh-hip"
 * or "sgd-stress-hip" as the inner engine.  Settings
 * through jc_engine_set_params_json, the MultilevelSettings keys
 * (multilevel.rs:132-165): inner, inner_params (forwarded to the inner engine
 * at init), max_levels, target_size, sweeps_per_level, sweep_schedule
 * ("uniform" | "linear" | "geometric"), spring_len, seed.  An unknown inner id
 * or a multilevel inner is JC_E_INVALID.  jc_engine_step(e, k, out) runs k
 * ticks; out receives the current level's 3 * n_level coordinates (n_level
 * <= n, from jc_multilevel_state).  get_state / set_state are JC_E_INVALID
 * (the reference wrapper has no checkpoint state). */
int jc_multilevel_create(jc_engine **out);
/* The cascade position: current level (0 = the input graph), number of levels,
 * the current level's node count and the ticks taken at it.  Each nullable. */
int jc_multilevel_state(jc_engine *e, uint32_t *level, uint32_t *n_levels, uint32_t *n_level_nodes,
                        uint32_t *sweeps_done);

/* ---- graph traversal ---- */
#define JC_UNREACHABLE 0xFFFFFFFFu
/* Replaces gpu_bfs (crates/graph-compute/src/analytics/bfs.rs:71-300): hop
 * distances from `source` over the CSR, in node order, JC_UNREACHABLE for nodes
 * it does not reach (bfs.rs:20-22).  n == 0 writes nothing (bfs.rs:82-84);
 * source >= n is JC_E_INVALID (bfs.rs:91-93); JC_E_NODEVICE without a GPU. */
int jc_bfs(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
           uint32_t source, uint32_t *out_dist);
/* The same search with the CSR uploaded once for many sources.  out_dist
 * (nullable) receives the distances; out_gpu_ms (nullable) the device time of
 * the search measured with HIP events. */
typedef struct jc_bfs_plan jc_bfs_plan;
int jc_bfs_plan_create(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                       jc_bfs_plan **out);
int jc_bfs_plan_run(jc_bfs_plan *plan, uint32_t source, uint32_t *out_dist, double *out_gpu_ms);
void jc_bfs_plan_destroy(jc_bfs_plan *plan);
/* Replaces gpu_connected_components
 * (crates/graph-compute/src/analytics/connected_components.rs:74): out_labels[v]
 * = smallest node index of v's component, every CSR entry taken as an
 * undirected edge (the labels of cpu_connected_components, :45-66). */
int jc_connected_components(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                            uint64_t nnz, uint32_t *out_labels);
/* Largest component, the pre-processing of the stress workloads (no GPU
 * counterpart in the reference; its CPU form is the union-find above plus a
 * size count).  Ties go to the component holding the smallest node id.
 * keep_map[v] (nullable, n entries) = new id of v, increasing in v, or
 * JC_UNREACHABLE; *out_size = its node count.  With sub_offsets (n+1 entries)
 * and sub_neighbors (nnz entries) non-NULL the induced CSR with renumbered ids
 * is written too, *out_sub_nnz its entry count (each nullable). */
int jc_largest_component(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, uint64_t nnz,
                         uint32_t *keep_map, uint32_t *out_size, uint32_t *sub_offsets,
                         uint32_t *sub_neighbors, uint64_t *out_sub_nnz);

/* ---- batched FA2+BH: many independent graphs per launch ----
 * No reference counterpart (the reference runs one engine per graph,
 * sim.rs:386-421); each graph follows exactly the fa2-bh-hip semantics with its
 * own AdaptiveSpeed state.  Inputs are concatenated graph after graph:
 * offsets = the n_g+1 local offsets of each graph, neighbors = local ids,
 * xyz = 3*sum(n_g) positions. */
typedef struct jc_batch jc_batch;
int jc_fa2bh_batch_create(const jc_fa2bh_settings *settings, jc_batch **out);
int jc_fa2bh_batch_init(jc_batch *b, uint32_t n_graphs, const uint32_t *n_nodes,
                        const uint64_t *nnz, const uint32_t *offsets, const uint32_t *neighbors,
                        const float *xyz);
int jc_fa2bh_batch_step(jc_batch *b, uint32_t n_iters, float *out_xyz);
/* {speed, speed_efficiency} per graph (2 * n_graphs floats). */
int jc_fa2bh_batch_speeds(jc_batch *b, float *speed_eff);
/* n_iters iterations between HIP events; also reports the force-kernel block
 * size the autotuner picked. */
int jc_fa2bh_batch_step_timed(jc_batch *b, uint32_t n_iters, double *gpu_ms,
                              uint32_t *force_block);
/* Schedule of the batch (no reference counterpart).  mode 0: picked by the
 * autotuner at the next step (default); 1: phased, one launch per phase over the
 * whole forest, iterations replayed from a hipGraph; 2: persistent, one launch
 * for all iterations of a step, each resident workgroup laying out whole graphs
 * taken largest first from a queue.  threads: the persistent workgroup size,
 * 256, 512 or 1024, or 0 for the autotuner's pick.  Either schedule gives the
 * same bits.  Mode 2 is JC_E_INVALID for a graph whose node region would need
 * 2^27 nodes or more.  Mode 2 is currently switched off and returns JC_E_INVALID
 * (its first GPU run made a memory fault whose cause has not been found); the
 * autotuner then keeps mode 1. */
int jc_fa2bh_batch_set_schedule(jc_batch *b, uint32_t mode, uint32_t threads);
/* Knobs of the phased schedule (no reference counterpart), each 0 to leave it
 * to the autotuner, which times the candidates on the live positions at the
 * next step: force block size (64, 128, 256, 512, 1024), centre-of-mass fusion
 * depth (3, 4, 5: cells at least that deep summed in one launch) and chunk
 * (1024 or 2048 nodes per workgroup), composite sort key width (16 to 64 bits,
 * a multiple of 16, no narrower than the width the batch needs for its graph
 * ids and key prefixes).  Every setting gives the same bits; invalid values are
 * JC_E_INVALID. */
int jc_fa2bh_batch_set_tuning(jc_batch *b, uint32_t force_block, uint32_t com_depth,
                              uint32_t com_chunk, uint32_t sort_key_bits);
/* The schedule in use, out[6]: mode (0 until the first step has tuned it),
 * persistent workgroup size, force block size, centre-of-mass fusion depth and
 * chunk, sort key width. */
int jc_fa2bh_batch_schedule(jc_batch *b, uint32_t *out);
void jc_fa2bh_batch_destroy(jc_batch *b);

/* ---- learned warm start: 3-layer message-passing network, hidden 128 ----
 * No reference counterpart (the reference seeds at random or from a multilevel
 * solve, engines/multilevel.rs:358-470).  BF16 weights/activations on MFMA
 * with FP32 accumulation; FP32 master weights trained with Adam on the GPU.
 * The CSR must be symmetric.  Parameter layout (floats): for l = 0..2,
 * W_l (K_l x 128, K = 32, 256, 256) then b_l (128); then W_h (128 x 3), b_h (3). */
typedef struct jc_mpnn jc_mpnn;
uint64_t jc_mpnn_param_count(void);
int jc_mpnn_create(uint64_t seed, jc_mpnn **out);
void jc_mpnn_destroy(jc_mpnn *m);
int jc_mpnn_get_params(jc_mpnn *m, float *out);
int jc_mpnn_set_params(jc_mpnn *m, const float *in);
int jc_mpnn_get_grads(jc_mpnn *m, float *out);
/* Adam's first and second moments (param_count floats each) and its step
 * count: a training run saves them to stop and resumes from them exactly. */
int jc_mpnn_get_adam(jc_mpnn *m, float *m1, float *m2, uint64_t *step);
int jc_mpnn_set_adam(jc_mpnn *m, const float *m1, const float *m2, uint64_t step);
/* Binds a graph and computes its input features (random-walk return
 * probabilities from `walks` walks per node, degree terms, noise: 3n values
 * in [-1,1] or NULL to draw them from feat_seed). */
int jc_mpnn_set_graph(jc_mpnn *m, uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                      uint64_t nnz, uint64_t feat_seed, uint32_t walks, const float *noise);
int jc_mpnn_features(jc_mpnn *m, float *out /* n x 16 */);
int jc_mpnn_activations(jc_mpnn *m, int layer, float *out /* n x 128 */);
/* Predicted coordinates (3n) and the device time of the forward pass. */
int jc_mpnn_forward(jc_mpnn *m, float *out_xyz, double *gpu_ms);
/* The same prediction centred at the origin and scaled to RMS radius target_rms
 * on the device (mean and RMS summed in f64 in a fixed order: deterministic);
 * gpu_ms covers the forward pass and the rescale. */
int jc_mpnn_forward_scaled(jc_mpnn *m, double target_rms, float *out_xyz, double *gpu_ms);
/* Target layout (3n) whose pair distances the loss matches: graph edges plus
 * pairs with n_anchor seeded anchor nodes. */
int jc_mpnn_set_target(jc_mpnn *m, const float *xyz, uint32_t n_anchor, uint64_t seed);
/* Forward + loss + backward; the gradient is left in the network (get_grads). */
int jc_mpnn_loss_grad(jc_mpnn *m, double *loss);
int jc_mpnn_adam(jc_mpnn *m, float lr);
int jc_mpnn_train(jc_mpnn *m, uint32_t steps, float lr, double *last_loss);

/* ---- fixtures (crates/graph-compute/src/stability.rs:9-110) ---- */
/* synthetic_scale_free(n, m, seed) as undirected pairs; returns the pair count
 * (pairs written only while count <= cap). */
uint64_t jc_gen_ba_pairs(uint32_t n, uint32_t m, uint64_t seed, uint32_t *pairs, uint64_t cap);
uint64_t jc_gen_sbm_pairs(uint32_t n, uint32_t blocks, double p_in, double p_out, uint64_t seed,
                          uint32_t *pairs, uint64_t cap);
uint64_t jc_gen_rgg_pairs(uint32_t n, double radius, uint64_t seed, float *xy, uint32_t *pairs,
                          uint64_t cap);
/* csr_from_pairs: symmetric, sorted, deduplicated rows; returns nnz
 * (nbrs needs room for 2*npairs). */
uint64_t jc_csr_from_pairs(uint32_t n, const uint32_t *pairs, uint64_t npairs, uint32_t *offsets,
                           uint32_t *nbrs);
/* Positions uniform in [-half, half]^3 ((2u-1)*half per coordinate). */
void jc_gen_uniform_cube(uint32_t n, float half, uint64_t seed, float *xyz);
/* ball_seed(n, seed) (stability.rs:30-50): the renderer's seed, a uniform random
 * ball of radius sqrt(n)*5 by rejection sampling (xyz: 3n). */
void jc_gen_ball_seed(uint32_t n, uint64_t seed, float *xyz);
/* ring_seed(n) (sim.rs:371-384): the worker's deterministic seed for a graph
 * loaded without positions, 3n draws of a 32-bit LCG mapped to [-1, 1] and
 * scaled by sqrt(n)*5 (xyz: 3n). */
void jc_gen_ring_seed(uint32_t n, float *xyz);

/* ---- layout telemetry: StepStats (crates/graph-compute/src/stability.rs:128-186) ---- */
typedef struct {
    float max_abs;      /* max |coordinate| over finite nodes: the divergence detector */
    float p50_radius;   /* median |pos| */
    float p99_radius;   /* 99th-percentile |pos| */
    float mean_disp;    /* mean per-node displacement from prev (f64 sum over finite nodes / n) */
    float p50_disp;     /* median displacement */
    uint64_t nonfinite; /* nodes with a NaN/Inf coordinate */
} jc_step_stats_t;
/* StepStats::measure(prev, cur) over n nodes (interleaved xyz); host-side, no device. */
int jc_step_stats(uint32_t n, const float *prev, const float *cur, jc_step_stats_t *out);

/* ---- binary CSR wire format (CsrGraph::from_bin_bytes / to_bin,
 * crates/graph-compute/src/sim.rs:52-128; the graph-api /graph/csr.bin payload
 * and the LoadGraph gRPC body) ----
 * Little-endian [u32 n_nodes][u32 n_edges][u32 x (n_nodes+1) offsets]
 * [u32 x n_edges neighbors].  Like the reference parser, the contents are not
 * validated here (engine init validates).  No device is needed. */
/* Header and length check: JC_E_INVALID with the reference's messages
 * ("CSR truncated: ..." / "CSR length mismatch: ..."). */
int jc_csr_bin_header(const uint8_t *bytes, size_t len, uint32_t *n_nodes, uint32_t *n_edges);
/* Copies the arrays out (offsets: n_nodes+1, neighbors: n_edges u32s). */
int jc_csr_from_bin(const uint8_t *bytes, size_t len, uint32_t *offsets, uint32_t *neighbors);
/* Size in bytes of the encoding of a graph (8 + 4(n_nodes+1) + 4 n_edges). */
uint64_t jc_csr_bin_size(uint32_t n_nodes, uint64_t n_edges);
/* Encodes into out[0..cap); JC_E_INVALID when cap is too small or n_edges > 2^32-1. */
int jc_csr_to_bin(uint32_t n_nodes, const uint32_t *offsets, const uint32_t *neighbors,
                  uint64_t n_edges, uint8_t *out, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
