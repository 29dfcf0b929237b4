// C++ mirror of the reference LayoutEngine trait
// (crates/graph-compute/src/engines/mod.rs:383-444) that the C ABI wraps.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "jc_internal.hpp"

namespace jc {

// CsrShard (mod.rs:318-334) for the unsharded case: the whole CSR graph.
struct CsrView {
    uint32_t n = 0;
    const uint32_t *offsets = nullptr;  // n + 1
    const uint32_t *neighbors = nullptr;  // nnz
    uint64_t nnz = 0;
};

// Rejects what the reference would index out of bounds on: offsets must start
// at 0, never decrease, end at nnz, and every neighbour id must be < n.
Status validate_csr(const CsrView &g);

class LayoutEngine {
  public:
    virtual ~LayoutEngine() = default;
    virtual const jc_descriptor &descriptor() const = 0;
    virtual Status set_params(const char *json, size_t len) = 0;
    virtual Status init(const CsrView &g, const float *xyz) = 0;
    virtual Status reinit(const CsrView &g, const float *xyz) { return init(g, xyz); }
    virtual Status step(uint32_t n_iters, float *out_xyz) = 0;
    virtual Status sync() = 0;
    virtual Status get_state(float *xyz, float *old_force, float *speed_eff2,
                             uint64_t *iteration) = 0;
    virtual Status set_state(const float *xyz, const float *old_force, const float *speed_eff2,
                             uint64_t iteration) = 0;
    virtual bool is_halted() const { return false; }
};

// ---- minimal serde-style JSON object decoding (host.cpp) ----
struct JsonField {
    // STR: a JSON string into std::string; RAW: any JSON value, kept as its
    // text in std::string (settings forwarded to another decoder)
    enum Kind { F32, BOOL, U32, U64, STR, RAW } kind;
    const char *name;
    void *dst;
};
// Decodes `json` into the fields.  Returns 1 when json is `null` (keep the
// current settings), 0 on success, <0 with jc_last_error set on a malformed or
// wrongly typed input.  Fields absent from the object are left untouched, so
// callers pre-fill defaults; unknown keys are ignored.
int decode_json_object(const char *json, size_t len, const std::vector<JsonField> &fields,
                       const char *what);

// traverse.hip: max-min pivot selection on the device (sgd_stress.rs:143-175).
// dist8_dev[((t/8) * n + v) * 8 + t%8] = hop distance of v from pivot t as u16
// (0 for the pivot itself and for unreachable nodes; the caller zeroes the
// buffer, ceil(k/8) * n * 8 entries); pivots_dev[k]; *dmax = largest such
// distance.  JC_E_INVALID when a finite distance exceeds 65535.
Status select_pivots_device(const CsrView &g, uint32_t k, uint32_t first, hipStream_t stream,
                            uint32_t *pivots_dev, uint16_t *dist8_dev, std::vector<uint32_t> &h_pivots,
                            uint32_t *dmax);

LayoutEngine *make_fa2bh(const jc_fa2bh_settings &s);
// the brute-force engine (fa2_brute.rs); theta is not used
LayoutEngine *make_fa2brute(const jc_fa2bh_settings &s);
LayoutEngine *make_sgd(const jc_sgd_settings &s);
LayoutEngine *make_multilevel();
// multilevel.cpp: the cascade position of a multilevel-hip engine.
Status multilevel_state(LayoutEngine *e, uint32_t *level, uint32_t *n_levels, uint32_t *n_level_nodes,
                        uint32_t *sweeps_done);

// FA2+BH specific hooks (fa2bh.hip)
Status fa2bh_forces_at(LayoutEngine *e, const float *xyz, float *out_force, double *sums2);
Status fa2bh_tree(LayoutEngine *e, uint32_t cap, uint32_t *n_tree, uint32_t *n_valid,
                  uint64_t *keys, uint32_t *perm, float *node_a, uint32_t *node_info,
                  uint32_t *node_skip, uint8_t *node_depth, float *root4);
Status fa2bh_step_timed(LayoutEngine *e, uint32_t n_iters, double *force_ms, double *iter_ms);
Status fa2bh_last_stats(LayoutEngine *e, uint64_t *n_tree, uint64_t *interactions, double *sums2);
Status fa2bh_walk_mode(LayoutEngine *e, int mode, int *current);
Status fa2bh_profile(LayoutEngine *e, uint32_t n_iters, double *phase_ms, uint32_t n_phases);
bool is_fa2bh(const LayoutEngine *e);

// SGD stress hooks (sgd.hip)
Status sgd_pivots(LayoutEngine *e, uint32_t *k, uint32_t *pivots, uint16_t *dist, float *eta_max);
Status sgd_time(LayoutEngine *e, uint32_t n, double *ms);

// Pivot-term stress evaluator (sgd.hip)
class StressPlan;
StressPlan *make_stress_plan();
Status stress_plan_create(StressPlan *p, uint32_t n, uint32_t k, const uint32_t *pivots,
                          const uint16_t *dist);
Status stress_plan_eval(StressPlan *p, const float *xyz, double *out);
void free_stress_plan(StressPlan *p);

}  // namespace jc
