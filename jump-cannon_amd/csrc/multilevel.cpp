// Multilevel wrapper engine (registry id "multilevel-hip").
//
// Reference: crates/graph-compute/src/engines/multilevel.rs (the cascade state
// machine over an inner LayoutEngine: coarsen -> solve the coarsest level ->
// prolong -> refine, one inner step per tick) and
// crates/graph-layouts/src/layout/coarsen.rs (greedy maximal-matching
// contraction, child-average fold-up, prolongation with xorshift64* jitter).
//
// The inner solvers are this library's HIP engines ("fa2-brute-hip", the
// default, "fa2-bh-hip", "sgd-stress-hip"), rebound to every level through reinit.  The cascade itself
// is host code, as in the reference: the matching is greedy in edge-list order
// (first touch wins), a sequential rule whose result fixes every coarse graph,
// so it stays on the CPU where it is computed exactly as the reference does;
// it runs once per init and is linear in the edge count.
//
// step(k, out) runs k ticks.  Within a level the inner engine runs as many
// ticks as remain in the level's budget in one call (its trajectory does not
// depend on how the iterations are grouped); positions come back to the host
// only where the reference needs them: at a level change (prolongation) and
// for `out`.  `out` receives the current level's 3 * n_level coordinates, as
// the reference's StepOutput carries the level's positions during the descent.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"

namespace jc {
namespace {

enum class Schedule { Uniform, Linear, Geometric };

// MultilevelSettings::default() (multilevel.rs:167-180); the inner default is
// the brute-force FA2 engine, as there (Fa2BruteEngine::ID, :176).
struct MlSettings {
    std::string inner = "fa2-brute-hip";
    std::string inner_params = "null";
    uint64_t max_levels = 6;
    uint64_t target_size = 500;
    uint32_t sweeps_per_level = 30;
    Schedule schedule = Schedule::Linear;
    float spring_len = 30.0f;
    uint32_t seed = 0x5EED;
};

uint32_t sat_mul(uint32_t a, uint64_t b) {
    const uint64_t p = (uint64_t)a * b;
    return p > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)p;
}

// SweepSchedule::sweeps_for_level (multilevel.rs:101-126): never below 1.
uint32_t sweeps_for_level(Schedule s, uint64_t level, uint32_t base) {
    base = std::max<uint32_t>(base, 1);
    uint32_t r = base;
    if (s == Schedule::Linear) r = sat_mul(base, 1 + level);
    else if (s == Schedule::Geometric) r = sat_mul(base, 1ull << std::min<uint64_t>(level, 16));
    return std::max<uint32_t>(r, 1);
}

struct Level {
    uint32_t n = 0;
    std::vector<uint32_t> edges;       // flat [s, t, ...]
    std::vector<uint32_t> parent_map;  // finer level's node -> this level's node
};

// Insert-if-absent set of 64-bit keys (open addressing, linear probing).
class KeySet {
  public:
    explicit KeySet(size_t expect) {
        size_t cap = 16;
        while (cap < 2 * expect + 1) cap <<= 1;
        slots_.assign(cap, kEmpty);
        mask_ = cap - 1;
    }
    bool insert(uint64_t k) {
        size_t h = (size_t)((k * 0x9E3779B97F4A7C15ull) >> 17) & mask_;
        while (slots_[h] != kEmpty) {
            if (slots_[h] == k) return false;
            h = (h + 1) & mask_;
        }
        slots_[h] = k;
        return true;
    }

  private:
    static constexpr uint64_t kEmpty = ~0ull;  // never a key: both halves < 2^32 - 1
    std::vector<uint64_t> slots_;
    size_t mask_ = 0;
};

// contract_one_level (coarsen.rs:96-160): greedy matching in edge order, the
// unmatched nodes numbered after the pairs in node order, contracted edges in
// first-occurrence order with self loops and repeats dropped.
Level contract_one_level(uint32_t n, const std::vector<uint32_t> &edges) {
    Level out;
    out.parent_map.assign(n, 0xFFFFFFFFu);
    std::vector<uint32_t> &parent = out.parent_map;
    uint32_t next_super = 0;
    for (size_t i = 0; i + 1 < edges.size(); i += 2) {
        const uint32_t s = edges[i], t = edges[i + 1];
        if (s == t || s >= n || t >= n) continue;
        if (parent[s] == 0xFFFFFFFFu && parent[t] == 0xFFFFFFFFu) {
            parent[s] = next_super;
            parent[t] = next_super;
            next_super++;
        }
    }
    for (uint32_t &p : parent)
        if (p == 0xFFFFFFFFu) p = next_super++;
    out.n = next_super;
    KeySet seen(edges.size() / 2);
    out.edges.reserve(edges.size() / 2);
    for (size_t j = 0; j + 1 < edges.size(); j += 2) {
        const uint32_t es = edges[j], et = edges[j + 1];
        if (es >= n || et >= n) continue;
        const uint32_t s = parent[es], t = parent[et];
        if (s == t) continue;
        const uint32_t a = std::min(s, t), b = std::max(s, t);
        if (seen.insert(((uint64_t)a << 32) | b)) {
            out.edges.push_back(s);
            out.edges.push_back(t);
        }
    }
    return out;
}

// coarsen (coarsen.rs:53-86): stop at target_size nodes, at a coarsening ratio
// below 1.6, or at max_levels levels.
std::vector<Level> coarsen(uint32_t n, const std::vector<uint32_t> &edges, uint64_t max_levels,
                           uint64_t target_size) {
    std::vector<Level> levels;
    Level l0;
    l0.n = n;
    l0.edges = edges;
    levels.push_back(std::move(l0));
    for (uint64_t it = 1; it < max_levels; it++) {
        const Level &cur = levels.back();
        if (cur.n <= target_size) break;
        Level next = contract_one_level(cur.n, cur.edges);
        const float ratio = (float)cur.n / (float)std::max<uint32_t>(next.n, 1);
        if (ratio < 1.6f && next.n == cur.n) break;
        const uint32_t next_n = next.n;
        levels.push_back(std::move(next));
        if (next_n <= target_size || ratio < 1.6f) break;
    }
    return levels;
}

// fold_up (multilevel.rs:483-512): each coarse node starts at the f32 average of
// its children, level by level.
std::vector<float> fold_up(const std::vector<Level> &levels, const float *fine, uint32_t n0) {
    std::vector<float> cur(fine, fine + 3 * (size_t)n0);
    for (size_t l = 1; l < levels.size(); l++) {
        const std::vector<uint32_t> &pm = levels[l].parent_map;
        const uint32_t cn = levels[l].n;
        std::vector<float> sums(3 * (size_t)cn, 0.0f);
        std::vector<uint32_t> counts(cn, 0);
        for (size_t child = 0; child < pm.size(); child++) {
            const uint32_t p = pm[child];
            if (p >= cn || child * 3 + 2 >= cur.size()) continue;
            sums[3 * (size_t)p] += cur[child * 3];
            sums[3 * (size_t)p + 1] += cur[child * 3 + 1];
            sums[3 * (size_t)p + 2] += cur[child * 3 + 2];
            counts[p]++;
        }
        for (uint32_t p = 0; p < cn; p++) {
            const float c = (float)std::max<uint32_t>(counts[p], 1);
            sums[3 * (size_t)p] /= c;
            sums[3 * (size_t)p + 1] /= c;
            sums[3 * (size_t)p + 2] /= c;
        }
        cur.swap(sums);
    }
    return cur;
}

// prolong (coarsen.rs:166-193): child = parent + u * jitter per coordinate, u
// uniform in [-1, 1) from xorshift64* seeded with seed + golden-ratio increment.
std::vector<float> prolong(const std::vector<float> &parent_pos, const std::vector<uint32_t> &parent_map,
                           uint32_t n_child, float jitter, uint32_t seed) {
    std::vector<float> out(3 * (size_t)n_child, 0.0f);
    uint64_t s = (uint64_t)seed + 0x9E3779B97F4A7C15ull;
    auto next = [&s]() -> float {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        const uint64_t v = s * 0x2545F4914F6CDD1Dull;
        return (float)((double)(v >> 11) / 9007199254740992.0) * 2.0f - 1.0f;
    };
    for (uint32_t i = 0; i < n_child; i++) {
        const size_t b = 3 * (size_t)parent_map[i];
        for (int q = 0; q < 3; q++) {
            const float u = next();
            out[3 * (size_t)i + q] = parent_pos[b + q] + u * jitter;
        }
    }
    return out;
}

// level_csr (multilevel.rs:229-272): symmetric CSR from the flat edge list, rows
// filled in edge order, self loops and out-of-range ids dropped.
void level_csr(uint32_t n, const std::vector<uint32_t> &edges, std::vector<uint32_t> &off,
               std::vector<uint32_t> &nbr) {
    std::vector<uint32_t> degree(n, 0);
    for (size_t i = 0; i + 1 < edges.size(); i += 2) {
        const uint32_t s = edges[i], t = edges[i + 1];
        if (s < n && t < n && s != t) {
            degree[s]++;
            degree[t]++;
        }
    }
    off.assign((size_t)n + 1, 0);
    for (uint32_t v = 0; v < n; v++) off[v + 1] = off[v] + degree[v];
    nbr.assign(off[n], 0);
    std::vector<uint32_t> cursor(off.begin(), off.end() - 1);
    for (size_t j = 0; j + 1 < edges.size(); j += 2) {
        const uint32_t s = edges[j], t = edges[j + 1];
        if (s < n && t < n && s != t) {
            nbr[cursor[s]++] = t;
            nbr[cursor[t]++] = s;
        }
    }
}

bool is_inner_id(const std::string &id) {
    return id == "fa2-brute-hip" || id == "fa2-bh-hip" || id == "sgd-stress-hip";
}

class MultilevelEngine final : public LayoutEngine {
  public:
    MultilevelEngine() {
        desc_.id = "multilevel-hip";
        desc_.kind = "physics";
        desc_.display_name = "Multilevel (multiscale wrapper, MI355X)";
        desc_.description =
            "Coarsen -> solve coarsest -> prolong -> refine, wrapping fa2-brute-hip, fa2-bh-hip or "
            "sgd-stress-hip (Walshaw/FM3/sfdp). Select the inner solver via `inner`.";
        desc_.needs_edges = 1;
        desc_.needs_cpu_positions = 1;
        desc_.needs_gpu_positions_buffer = 0;
    }

    const jc_descriptor &descriptor() const override { return desc_; }

    // multilevel.rs:307-331: serde decode with every missing field at its
    // default, then the inner id checked.
    Status set_params(const char *json, size_t len) override {
        MlSettings s;
        std::string schedule = "linear";
        std::vector<JsonField> f = {
            {JsonField::STR, "inner", &s.inner},
            {JsonField::RAW, "inner_params", &s.inner_params},
            {JsonField::U64, "max_levels", &s.max_levels},
            {JsonField::U64, "target_size", &s.target_size},
            {JsonField::U32, "sweeps_per_level", &s.sweeps_per_level},
            {JsonField::STR, "sweep_schedule", &schedule},
            {JsonField::F32, "spring_len", &s.spring_len},
            {JsonField::U32, "seed", &s.seed},
        };
        const int rc = decode_json_object(json, len, f, "multilevel settings");
        if (rc < 0) return Status{JC_E_INVALID};
        if (rc == 1) return Status::ok();
        if (schedule == "uniform") s.schedule = Schedule::Uniform;
        else if (schedule == "linear") s.schedule = Schedule::Linear;
        else if (schedule == "geometric") s.schedule = Schedule::Geometric;
        else
            return Status::err(JC_E_INVALID, "decode multilevel settings: unknown variant `" + schedule +
                                                 "`, expected one of `uniform`, `linear`, `geometric`");
        if (s.inner == "multilevel" || s.inner == "multilevel-hip")
            return Status::err(JC_E_INVALID, "multilevel cannot wrap itself");
        if (!is_inner_id(s.inner))
            return Status::err(JC_E_INVALID, "unknown inner engine \"" + s.inner + "\" for multilevel wrapper");
        settings_ = s;
        return Status::ok();
    }

    // multilevel.rs:333-383
    Status init(const CsrView &g, const float *xyz) override {
        JC_TRY(require_device(nullptr));
        JC_TRY(validate_csr(g));
        if (g.n && !xyz) return Status::err(JC_E_INVALID, "positions pointer is NULL");
        inited_ = false;
        std::vector<uint32_t> edges;
        edges.reserve(g.nnz);
        for (uint32_t v = 0; v < g.n; v++)
            for (uint64_t e = g.offsets[v]; e < g.offsets[v + 1]; e++)
                if (v < g.neighbors[e]) {
                    edges.push_back(v);
                    edges.push_back(g.neighbors[e]);
                }
        levels_ = coarsen(g.n, edges, settings_.max_levels, settings_.target_size);
        level_ = (uint32_t)levels_.size() - 1;
        sweeps_done_ = 0;
        positions_ = fold_up(levels_, xyz, g.n);
        pos_current_ = true;
        JC_TRY(build_inner());
        JC_TRY(bind_inner());
        inited_ = true;
        return Status::ok();
    }

    // multilevel.rs:385-468, k ticks at a time
    Status step(uint32_t n_iters, float *out_xyz) override {
        if (!inited_) return Status::err(JC_E_STATE, "multilevel step called before successful init");
        uint32_t left = n_iters;
        while (left > 0) {
            if (level_ == 0) {
                JC_TRY(inner_->step(left, nullptr));
                sweeps_done_ += left;
                left = 0;
                pos_current_ = false;
                break;
            }
            const uint32_t budget = sweeps_for_level(settings_.schedule, level_, settings_.sweeps_per_level);
            const uint32_t chunk = std::min<uint32_t>(left, budget - sweeps_done_);
            const uint32_t nl = levels_[level_].n;
            positions_.resize(3 * (size_t)nl);
            JC_TRY(inner_->step(chunk, nl ? positions_.data() : nullptr));
            sweeps_done_ += chunk;
            left -= chunk;
            if (sweeps_done_ >= budget) {
                const uint32_t child = level_ - 1;
                const float jitter = settings_.spring_len * 0.5f;
                const uint32_t seed = settings_.seed + child + 1;
                positions_ = prolong(positions_, levels_[level_].parent_map, levels_[child].n, jitter, seed);
                level_ = child;
                sweeps_done_ = 0;
                JC_TRY(bind_inner());
            }
            pos_current_ = true;
        }
        if (out_xyz) {
            const uint32_t nl = levels_[level_].n;
            if (pos_current_) std::copy(positions_.begin(), positions_.begin() + 3 * (size_t)nl, out_xyz);
            else if (nl) {
                // level 0: the inner engine holds the positions
                JC_TRY(inner_->get_state(out_xyz, nullptr, nullptr, nullptr));
            }
        }
        return Status::ok();
    }

    Status sync() override { return inner_ ? inner_->sync() : Status::ok(); }

    Status get_state(float *, float *, float *, uint64_t *) override {
        return Status::err(JC_E_INVALID, "multilevel-hip keeps no checkpoint state (the reference wrapper has none)");
    }
    Status set_state(const float *, const float *, const float *, uint64_t) override {
        return Status::err(JC_E_INVALID, "multilevel-hip keeps no checkpoint state (the reference wrapper has none)");
    }

    // multilevel.rs:470-479
    bool is_halted() const override { return inited_ && level_ == 0 && inner_ && inner_->is_halted(); }

    Status state(uint32_t *level, uint32_t *n_levels, uint32_t *n_level_nodes, uint32_t *sweeps) const {
        if (!inited_) return Status::err(JC_E_STATE, "multilevel state before init");
        if (level) *level = level_;
        if (n_levels) *n_levels = (uint32_t)levels_.size();
        if (n_level_nodes) *n_level_nodes = levels_[level_].n;
        if (sweeps) *sweeps = sweeps_done_;
        return Status::ok();
    }

  private:
    // multilevel.rs:274-283: one inner instance per init, settings forwarded.
    Status build_inner() {
        if (settings_.inner == "fa2-bh-hip" || settings_.inner == "fa2-brute-hip") {
            jc_fa2bh_settings d;
            jc_fa2bh_default_settings(&d);
            inner_.reset(settings_.inner == "fa2-bh-hip" ? make_fa2bh(d) : make_fa2brute(d));
        } else {
            jc_sgd_settings d;
            jc_sgd_default_settings(&d);
            inner_.reset(make_sgd(d));
        }
        return inner_->set_params(settings_.inner_params.data(), settings_.inner_params.size());
    }

    // multilevel.rs:296-314: the level's CSR and positions into the same inner
    // engine through reinit.
    Status bind_inner() {
        const Level &lv = levels_[level_];
        level_csr(lv.n, lv.edges, off_, nbr_);
        CsrView g;
        g.n = lv.n;
        g.offsets = off_.data();
        g.neighbors = nbr_.empty() ? nullptr : nbr_.data();
        g.nnz = nbr_.size();
        return inner_->reinit(g, positions_.data());
    }

    jc_descriptor desc_{};
    MlSettings settings_;
    std::unique_ptr<LayoutEngine> inner_;
    std::vector<Level> levels_;
    std::vector<float> positions_;
    std::vector<uint32_t> off_, nbr_;
    uint32_t level_ = 0, sweeps_done_ = 0;
    bool inited_ = false;
    bool pos_current_ = false;  // positions_ holds the current level's positions
};

}  // namespace

LayoutEngine *make_multilevel() { return new MultilevelEngine(); }

Status multilevel_state(LayoutEngine *e, uint32_t *level, uint32_t *n_levels, uint32_t *n_level_nodes,
                        uint32_t *sweeps_done) {
    auto *m = dynamic_cast<MultilevelEngine *>(e);
    if (!m) return Status::err(JC_E_INVALID, "engine is not multilevel-hip");
    return m->state(level, n_levels, n_level_nodes, sweeps_done);
}

}  // namespace jc
