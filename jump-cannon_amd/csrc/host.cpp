// Host side of libjc_hip.so: the C ABI over the LayoutEngine mirror, error
// plumbing, CSR validation, serde-style settings decoding and the seeded
// fixture generators (crates/graph-compute/src/stability.rs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"

struct jc_engine {
    jc::LayoutEngine *impl;
};

namespace jc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

Status require_device(int *dev_out) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return Status::err(JC_E_NODEVICE, "fa2-bh-hip requires a HIP device but none is available");
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return Status::err(JC_E_HIP, hipGetErrorString(e));
    if (dev_out) *dev_out = dev;
    return Status::ok();
}

Status validate_csr(const CsrView &g) {
    if (g.n == 0) return Status::ok();
    if (!g.offsets) return Status::err(JC_E_INVALID, "CSR offsets pointer is NULL");
    if (g.offsets[0] != 0) return Status::err(JC_E_INVALID, "CSR offsets[0] != 0");
    for (uint32_t v = 0; v < g.n; v++)
        if (g.offsets[v + 1] < g.offsets[v])
            return Status::err(JC_E_INVALID, "CSR offsets decrease at node " + std::to_string(v));
    if ((uint64_t)g.offsets[g.n] != g.nnz)
        return Status::err(JC_E_INVALID, "CSR offsets[n] = " + std::to_string(g.offsets[g.n]) +
                                             " != nnz = " + std::to_string(g.nnz));
    if (g.nnz && !g.neighbors) return Status::err(JC_E_INVALID, "CSR neighbors pointer is NULL");
    for (uint64_t e = 0; e < g.nnz; e++)
        if (g.neighbors[e] >= g.n)
            return Status::err(JC_E_INVALID, "CSR neighbor " + std::to_string(g.neighbors[e]) +
                                                 " out of range at entry " + std::to_string(e));
    return Status::ok();
}

// ------------------------------------------------------------------ JSON
namespace {
struct JsonCursor {
    const char *s;
    size_t n, i = 0;
    void ws() {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
    }
    bool eat(char c) {
        ws();
        if (i < n && s[i] == c) {
            i++;
            return true;
        }
        return false;
    }
    bool lit(const char *w) {
        ws();
        size_t L = std::strlen(w);
        if (i + L <= n && std::strncmp(s + i, w, L) == 0) {
            i += L;
            return true;
        }
        return false;
    }
};

enum class JKind { Null, Bool, Num, Str, Arr, Obj, Bad };
struct JVal {
    JKind kind = JKind::Bad;
    bool b = false;
    double num = 0;
    bool integral = false;
    bool negative = false;
    uint64_t u = 0;
};

bool parse_string(JsonCursor &c, std::string *out) {
    c.ws();
    if (c.i >= c.n || c.s[c.i] != '"') return false;
    c.i++;
    std::string r;
    while (c.i < c.n && c.s[c.i] != '"') {
        if (c.s[c.i] == '\\') {
            c.i++;
            if (c.i >= c.n) return false;
            char e = c.s[c.i];
            if (e == 'u') {
                if (c.i + 4 >= c.n) return false;
                c.i += 4;
                r.push_back('?');
            } else {
                r.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
            }
            c.i++;
        } else {
            r.push_back(c.s[c.i++]);
        }
    }
    if (c.i >= c.n) return false;
    c.i++;
    if (out) *out = r;
    return true;
}

bool parse_value(JsonCursor &c, JVal *v, int depth);

bool skip_container(JsonCursor &c, char open, char close, int depth) {
    if (depth > 64) return false;
    if (!c.eat(open)) return false;
    if (c.eat(close)) return true;
    for (;;) {
        if (open == '{') {
            if (!parse_string(c, nullptr) || !c.eat(':')) return false;
        }
        JVal tmp;
        if (!parse_value(c, &tmp, depth + 1)) return false;
        if (c.eat(',')) continue;
        return c.eat(close);
    }
}

bool parse_value(JsonCursor &c, JVal *v, int depth) {
    c.ws();
    if (c.i >= c.n) return false;
    char ch = c.s[c.i];
    if (ch == 'n') {
        v->kind = JKind::Null;
        return c.lit("null");
    }
    if (ch == 't') {
        v->kind = JKind::Bool;
        v->b = true;
        return c.lit("true");
    }
    if (ch == 'f') {
        v->kind = JKind::Bool;
        v->b = false;
        return c.lit("false");
    }
    if (ch == '"') {
        v->kind = JKind::Str;
        return parse_string(c, nullptr);
    }
    if (ch == '[') {
        v->kind = JKind::Arr;
        return skip_container(c, '[', ']', depth);
    }
    if (ch == '{') {
        v->kind = JKind::Obj;
        return skip_container(c, '{', '}', depth);
    }
    size_t st = c.i;
    bool integral = true;
    if (c.i < c.n && (c.s[c.i] == '-' || c.s[c.i] == '+')) c.i++;
    while (c.i < c.n && (std::isdigit((unsigned char)c.s[c.i]) || c.s[c.i] == '.' ||
                         c.s[c.i] == 'e' || c.s[c.i] == 'E' || c.s[c.i] == '-' || c.s[c.i] == '+')) {
        if (c.s[c.i] == '.' || c.s[c.i] == 'e' || c.s[c.i] == 'E') integral = false;
        c.i++;
    }
    if (c.i == st) return false;
    std::string tok(c.s + st, c.i - st);
    char *end = nullptr;
    v->num = std::strtod(tok.c_str(), &end);
    if (!end || *end) return false;
    v->kind = JKind::Num;
    v->integral = integral;
    v->negative = tok[0] == '-';
    if (integral && !v->negative) v->u = std::strtoull(tok.c_str(), nullptr, 10);
    return true;
}
}  // namespace

int decode_json_object(const char *json, size_t len, const std::vector<JsonField> &fields,
                       const char *what) {
    JsonCursor c{json ? json : "", json ? len : 0};
    c.ws();
    if (c.i >= c.n) {
        set_error(std::string("decode ") + what + ": empty input");
        return -1;
    }
    if (c.lit("null")) {
        c.ws();
        if (c.i != c.n) {
            set_error(std::string("decode ") + what + ": trailing characters");
            return -1;
        }
        return 1;
    }
    if (!c.eat('{')) {
        set_error(std::string("decode ") + what + ": invalid type, expected struct");
        return -1;
    }
    if (!c.eat('}')) {
        for (;;) {
            std::string key;
            if (!parse_string(c, &key) || !c.eat(':')) {
                set_error(std::string("decode ") + what + ": malformed object");
                return -1;
            }
            JVal v;
            c.ws();
            const size_t v0 = c.i;
            std::string str;
            bool is_str = false;
            if (c.i < c.n && c.s[c.i] == '"') {
                is_str = true;
                v.kind = JKind::Str;
                if (!parse_string(c, &str)) {
                    set_error(std::string("decode ") + what + ": malformed value for `" + key + "`");
                    return -1;
                }
            } else if (!parse_value(c, &v, 0)) {
                set_error(std::string("decode ") + what + ": malformed value for `" + key + "`");
                return -1;
            }
            const std::string raw(c.s + v0, c.i - v0);
            for (const JsonField &f : fields) {
                if (key != f.name) continue;
                bool ok = false;
                switch (f.kind) {
                    case JsonField::F32:
                        ok = v.kind == JKind::Num;
                        if (ok) *static_cast<float *>(f.dst) = (float)v.num;
                        break;
                    case JsonField::BOOL:
                        ok = v.kind == JKind::Bool;
                        if (ok) *static_cast<bool *>(f.dst) = v.b;
                        break;
                    case JsonField::U32:
                        ok = v.kind == JKind::Num && v.integral && !v.negative && v.u <= 0xFFFFFFFFull;
                        if (ok) *static_cast<uint32_t *>(f.dst) = (uint32_t)v.u;
                        break;
                    case JsonField::U64:
                        ok = v.kind == JKind::Num && v.integral && !v.negative;
                        if (ok) *static_cast<uint64_t *>(f.dst) = v.u;
                        break;
                    case JsonField::STR:
                        ok = is_str;
                        if (ok) *static_cast<std::string *>(f.dst) = str;
                        break;
                    case JsonField::RAW:
                        ok = true;
                        *static_cast<std::string *>(f.dst) = raw;
                        break;
                }
                if (!ok) {
                    set_error(std::string("decode ") + what + ": invalid type for field `" + key + "`");
                    return -1;
                }
            }
            if (c.eat(',')) continue;
            if (c.eat('}')) break;
            set_error(std::string("decode ") + what + ": expected `,` or `}`");
            return -1;
        }
    }
    c.ws();
    if (c.i != c.n) {
        set_error(std::string("decode ") + what + ": trailing characters");
        return -1;
    }
    return 0;
}

}  // namespace jc

// ====================================================================== C ABI
using jc::Status;

namespace {
int finish(Status s) {
    if (!s.failed()) jc::set_error("");
    return s.code;
}
int bad_handle() {
    jc::set_error("NULL engine handle");
    return JC_E_INVALID;
}
}  // namespace

extern "C" {

const char *jc_last_error(void) { return jc::last_error(); }

int jc_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

void jc_fa2bh_default_settings(jc_fa2bh_settings *o) {
    if (!o) return;
    o->gravity = 1.0f;
    o->strong_gravity = 0;
    o->scaling_ratio = 2.0f;
    o->edge_weight_influence = 1.0f;
    o->jitter_tolerance = 1.0f;
    o->lin_log_mode = 0;
    o->prevent_overlap = 0;
    o->time_step = 1.0f;
    o->theta = 0.7f;
    o->max_displacement = 10.0f;
}

void jc_fa2_default_settings(jc_fa2_settings *o) {
    if (!o) return;
    o->gravity = 1.0f;
    o->strong_gravity = 0;
    o->scaling_ratio = 2.0f;
    o->edge_weight_influence = 1.0f;
    o->jitter_tolerance = 1.0f;
    o->lin_log_mode = 0;
    o->prevent_overlap = 0;
    o->time_step = 1.0f;
    o->max_displacement = 10.0f;
}

void jc_sgd_default_settings(jc_sgd_settings *o) {
    if (!o) return;
    o->n_pivots = 50;
    o->sweeps_per_step = 1;
    o->eta_max = 1.0f;
    o->eta_min = 0.01f;
    o->n_anneal_steps = 30;
    o->reserved = 0;
    o->seed = 0x5EEDC0DE12345678ull;
}

int jc_fa2bh_create(const jc_fa2bh_settings *s, jc_engine **out) {
    if (!out) return bad_handle();
    jc_fa2bh_settings d;
    jc_fa2bh_default_settings(&d);
    auto *h = new (std::nothrow) jc_engine{jc::make_fa2bh(s ? *s : d)};
    if (!h) return JC_E_OOM;
    *out = h;
    return finish(Status::ok());
}

int jc_fa2brute_create(const jc_fa2_settings *s, jc_engine **out) {
    if (!out) return bad_handle();
    jc_fa2_settings d;
    jc_fa2_default_settings(&d);
    const jc_fa2_settings &v = s ? *s : d;
    jc_fa2bh_settings b;
    jc_fa2bh_default_settings(&b);
    b.gravity = v.gravity;
    b.strong_gravity = v.strong_gravity;
    b.scaling_ratio = v.scaling_ratio;
    b.edge_weight_influence = v.edge_weight_influence;
    b.jitter_tolerance = v.jitter_tolerance;
    b.lin_log_mode = v.lin_log_mode;
    b.prevent_overlap = v.prevent_overlap;
    b.time_step = v.time_step;
    b.max_displacement = v.max_displacement;
    auto *h = new (std::nothrow) jc_engine{jc::make_fa2brute(b)};
    if (!h) return JC_E_OOM;
    *out = h;
    return finish(Status::ok());
}

int jc_sgd_create(const jc_sgd_settings *s, jc_engine **out) {
    if (!out) return bad_handle();
    jc_sgd_settings d;
    jc_sgd_default_settings(&d);
    jc::LayoutEngine *impl = jc::make_sgd(s ? *s : d);
    if (!impl) return JC_E_OOM;
    *out = new jc_engine{impl};
    return finish(Status::ok());
}

int jc_multilevel_create(jc_engine **out) {
    if (!out) return bad_handle();
    *out = new jc_engine{jc::make_multilevel()};
    return finish(Status::ok());
}

int jc_multilevel_state(jc_engine *e, uint32_t *level, uint32_t *n_levels, uint32_t *n_level_nodes,
                        uint32_t *sweeps_done) {
    if (!e) return bad_handle();
    return finish(jc::multilevel_state(e->impl, level, n_levels, n_level_nodes, sweeps_done));
}

int jc_engine_descriptor(const jc_engine *e, jc_descriptor *out) {
    if (!e || !out) return bad_handle();
    *out = e->impl->descriptor();
    return JC_OK;
}

int jc_engine_set_params_json(jc_engine *e, const char *json, size_t len) {
    if (!e) return bad_handle();
    return finish(e->impl->set_params(json, len));
}

int jc_engine_init(jc_engine *e, uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                   uint64_t nnz, const float *xyz) {
    if (!e) return bad_handle();
    jc::CsrView g{n, offsets, neighbors, nnz};
    return finish(e->impl->init(g, xyz));
}

int jc_engine_reinit(jc_engine *e, uint32_t n, const uint32_t *offsets,
                     const uint32_t *neighbors, uint64_t nnz, const float *xyz) {
    if (!e) return bad_handle();
    jc::CsrView g{n, offsets, neighbors, nnz};
    return finish(e->impl->reinit(g, xyz));
}

int jc_engine_step(jc_engine *e, uint32_t n_iters, float *out_xyz) {
    if (!e) return bad_handle();
    return finish(e->impl->step(n_iters, out_xyz));
}

int jc_engine_sync(jc_engine *e) {
    if (!e) return bad_handle();
    return finish(e->impl->sync());
}

int jc_engine_get_state(jc_engine *e, float *xyz, float *old_force, float *se2, uint64_t *it) {
    if (!e) return bad_handle();
    return finish(e->impl->get_state(xyz, old_force, se2, it));
}

int jc_engine_set_state(jc_engine *e, const float *xyz, const float *old_force, const float *se2,
                        uint64_t it) {
    if (!e) return bad_handle();
    return finish(e->impl->set_state(xyz, old_force, se2, it));
}

int jc_engine_is_halted(const jc_engine *e) { return (e && e->impl->is_halted()) ? 1 : 0; }

void jc_engine_destroy(jc_engine *e) {
    if (!e) return;
    delete e->impl;
    delete e;
}

static int need_fa2(jc_engine *e) {
    if (!e) return bad_handle();
    if (!jc::is_fa2bh(e->impl)) {
        jc::set_error("engine is not fa2-bh-hip or fa2-brute-hip");
        return JC_E_INVALID;
    }
    return JC_OK;
}

int jc_fa2bh_forces_at(jc_engine *e, const float *xyz, float *out_force, double *sums2) {
    if (int rc = need_fa2(e)) return rc;
    return finish(jc::fa2bh_forces_at(e->impl, xyz, out_force, sums2));
}

int jc_fa2bh_tree(jc_engine *e, uint32_t cap, uint32_t *n_tree, uint32_t *n_valid, uint64_t *keys,
                  uint32_t *perm, float *node_a, uint32_t *node_info, uint32_t *node_skip,
                  uint8_t *node_depth, float *root4) {
    if (int rc = need_fa2(e)) return rc;
    return finish(jc::fa2bh_tree(e->impl, cap, n_tree, n_valid, keys, perm, node_a, node_info,
                                 node_skip, node_depth, root4));
}

int jc_fa2bh_step_timed(jc_engine *e, uint32_t n_iters, double *force_ms, double *iter_ms) {
    if (int rc = need_fa2(e)) return rc;
    return finish(jc::fa2bh_step_timed(e->impl, n_iters, force_ms, iter_ms));
}

int jc_fa2bh_last_stats(jc_engine *e, uint64_t *n_tree, uint64_t *interactions, double *sums2) {
    if (int rc = need_fa2(e)) return rc;
    return finish(jc::fa2bh_last_stats(e->impl, n_tree, interactions, sums2));
}

int jc_fa2bh_walk_mode(jc_engine *e, int mode, int *current) {
    if (int rc = need_fa2(e)) return rc;
    return finish(jc::fa2bh_walk_mode(e->impl, mode, current));
}

int jc_fa2bh_profile(jc_engine *e, uint32_t n_iters, double *phase_ms, uint32_t n_phases) {
    if (int rc = need_fa2(e)) return rc;
    return finish(jc::fa2bh_profile(e->impl, n_iters, phase_ms, n_phases));
}

int jc_sgd_pivots(jc_engine *e, uint32_t *k, uint32_t *pivots, uint16_t *dist, float *eta_max) {
    if (!e) return bad_handle();
    return finish(jc::sgd_pivots(e->impl, k, pivots, dist, eta_max));
}

int jc_sgd_time(jc_engine *e, uint32_t n_steps, double *gpu_ms) {
    if (!e) return bad_handle();
    return finish(jc::sgd_time(e->impl, n_steps, gpu_ms));
}

struct jc_stress_plan {
    jc::StressPlan *impl;
};

int jc_stress_plan_create(uint32_t n, uint32_t k, const uint32_t *pivots, const uint16_t *dist,
                          jc_stress_plan **out) {
    if (!out) return bad_handle();
    auto *p = new jc_stress_plan{jc::make_stress_plan()};
    jc::Status s = jc::stress_plan_create(p->impl, n, k, pivots, dist);
    if (s.failed()) {
        jc::free_stress_plan(p->impl);
        delete p;
        return s.code;
    }
    *out = p;
    return finish(s);
}

int jc_stress_plan_eval(jc_stress_plan *p, const float *xyz, double *out) {
    if (!p) return bad_handle();
    return finish(jc::stress_plan_eval(p->impl, xyz, out));
}

void jc_stress_plan_destroy(jc_stress_plan *p) {
    if (!p) return;
    jc::free_stress_plan(p->impl);
    delete p;
}

// ------------------------------------------------------------------ fixtures
static inline uint64_t sm_next(uint64_t &s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline float sm_f32(uint64_t &s) { return (float)(sm_next(s) >> 40) / (float)(1ull << 24); }
static inline double sm_f64(uint64_t &s) {
    return (double)(sm_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

// stability.rs:57-82 synthetic_scale_free: each new node v attaches min(m, v)
// distinct endpoints drawn from the running endpoint list (<= 64 draws).
uint64_t jc_gen_ba_pairs(uint32_t n_in, uint32_t m_in, uint64_t seed, uint32_t *pairs,
                         uint64_t cap) {
    uint64_t s = seed;
    const uint32_t n = std::max<uint32_t>(n_in, 2), m = std::max<uint32_t>(m_in, 1);
    std::vector<uint32_t> targets = {0, 1, 1, 0};
    targets.reserve(4 + 2ull * (n - 2) * m);
    std::vector<uint32_t> picked;
    picked.reserve(m);
    uint64_t np = 0;
    auto emit = [&](uint32_t a, uint32_t b) {
        if (np < cap) {
            pairs[2 * np] = a;
            pairs[2 * np + 1] = b;
        }
        np++;
    };
    emit(0, 1);
    for (uint32_t v = 2; v < n; v++) {
        picked.clear();
        const uint32_t want = std::min(m, v);
        uint32_t guard = 0;
        while (picked.size() < want && guard < 64) {
            guard++;
            uint32_t t = targets[sm_next(s) % targets.size()];
            if (t != v && std::find(picked.begin(), picked.end(), t) == picked.end())
                picked.push_back(t);
        }
        for (uint32_t t : picked) {
            emit(v, t);
            targets.push_back(v);
            targets.push_back(t);
        }
    }
    return np;
}

// Stochastic block model, sampled by geometric skipping over the candidate
// pairs of every block pair (a <= b) in lexicographic order.
uint64_t jc_gen_sbm_pairs(uint32_t n, uint32_t blocks, double p_in, double p_out, uint64_t seed,
                          uint32_t *pairs, uint64_t cap) {
    uint64_t s = seed, np = 0;
    if (blocks < 1) blocks = 1;
    auto start = [&](uint64_t b) { return (b * n + blocks - 1) / blocks; };
    for (uint32_t a = 0; a < blocks; a++) {
        const uint64_t a0 = start(a), a1 = start(a + 1), sa = a1 - a0;
        for (uint32_t b = a; b < blocks; b++) {
            const uint64_t b0 = start(b), b1 = start(b + 1), sb = b1 - b0;
            const double p = (a == b) ? p_in : p_out;
            const uint64_t count = (a == b) ? (sa * (sa ? sa - 1 : 0)) / 2 : sa * sb;
            if (p <= 0.0 || count == 0) continue;
            const double lq = std::log1p(-p);
            int64_t idx = -1;
            uint64_t row = 0, row_start = 0;
            for (;;) {
                if (p >= 1.0) idx += 1;
                else {
                    const double u = sm_f64(s);
                    const double k = std::floor(std::log1p(-u) / lq);
                    if (k > 1e18) break;
                    idx += 1 + (int64_t)k;
                }
                if ((uint64_t)idx >= count) break;
                uint32_t i, j;
                if (a == b) {
                    while ((uint64_t)idx >= row_start + (sa - 1 - row)) {
                        row_start += sa - 1 - row;
                        row++;
                    }
                    i = (uint32_t)(a0 + row);
                    j = (uint32_t)(a0 + row + 1 + ((uint64_t)idx - row_start));
                } else {
                    i = (uint32_t)(a0 + (uint64_t)idx / sb);
                    j = (uint32_t)(b0 + (uint64_t)idx % sb);
                }
                if (np < cap) {
                    pairs[2 * np] = i;
                    pairs[2 * np + 1] = j;
                }
                np++;
            }
        }
    }
    return np;
}

// Random geometric graph in the unit square: x,y = next_f32 in node order,
// edge iff (dx^2 + dy^2) < r^2 in f64; grid buckets of side ~r.
uint64_t jc_gen_rgg_pairs(uint32_t n, double radius, uint64_t seed, float *xy, uint32_t *pairs,
                          uint64_t cap) {
    uint64_t s = seed, np = 0;
    for (uint32_t i = 0; i < n; i++) {
        xy[2 * (size_t)i] = sm_f32(s);
        xy[2 * (size_t)i + 1] = sm_f32(s);
    }
    uint32_t g = (uint32_t)(1.0 / radius);
    g = std::min<uint32_t>(std::max<uint32_t>(g, 1), 4096);
    const uint64_t cells = (uint64_t)g * g;
    std::vector<uint32_t> cnt(cells + 1, 0), idx(n), cell_of(n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t cx = (uint32_t)(xy[2 * (size_t)i] * g), cy = (uint32_t)(xy[2 * (size_t)i + 1] * g);
        cx = std::min(cx, g - 1);
        cy = std::min(cy, g - 1);
        cell_of[i] = cy * g + cx;
        cnt[cell_of[i] + 1]++;
    }
    for (uint64_t c = 0; c < cells; c++) cnt[c + 1] += cnt[c];
    std::vector<uint32_t> fill(cnt);
    for (uint32_t i = 0; i < n; i++) idx[fill[cell_of[i]]++] = i;
    const double r2 = radius * radius;
    for (uint32_t i = 0; i < n; i++) {
        const int cx = (int)(cell_of[i] % g), cy = (int)(cell_of[i] / g);
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int x = cx + dx, y = cy + dy;
                if (x < 0 || y < 0 || x >= (int)g || y >= (int)g) continue;
                const uint32_t c = (uint32_t)y * g + (uint32_t)x;
                for (uint32_t q = cnt[c]; q < cnt[c + 1]; q++) {
                    const uint32_t j = idx[q];
                    if (j <= i) continue;
                    const double ddx = (double)xy[2 * (size_t)i] - (double)xy[2 * (size_t)j];
                    const double ddy = (double)xy[2 * (size_t)i + 1] - (double)xy[2 * (size_t)j + 1];
                    if (ddx * ddx + ddy * ddy < r2) {
                        if (np < cap) {
                            pairs[2 * np] = i;
                            pairs[2 * np + 1] = j;
                        }
                        np++;
                    }
                }
            }
    }
    return np;
}

// stability.rs:86-110 csr_from_pairs.
uint64_t jc_csr_from_pairs(uint32_t n, const uint32_t *pairs, uint64_t npairs, uint32_t *offsets,
                           uint32_t *nbrs) {
    std::vector<uint64_t> cnt((size_t)n + 1, 0);
    for (uint64_t k = 0; k < npairs; k++) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a == b || a >= n || b >= n) continue;
        cnt[a + 1]++;
        cnt[b + 1]++;
    }
    for (uint32_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
    std::vector<uint32_t> tmp(cnt[n] ? cnt[n] : 1);
    std::vector<uint64_t> fill(cnt);
    for (uint64_t k = 0; k < npairs; k++) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a == b || a >= n || b >= n) continue;
        tmp[fill[a]++] = b;
        tmp[fill[b]++] = a;
    }
    uint64_t out = 0;
    offsets[0] = 0;
    for (uint32_t v = 0; v < n; v++) {
        std::sort(tmp.begin() + cnt[v], tmp.begin() + cnt[v + 1]);
        for (uint64_t q = cnt[v]; q < cnt[v + 1]; q++) {
            if (q > cnt[v] && tmp[q] == tmp[q - 1]) continue;
            nbrs[out++] = tmp[q];
        }
        offsets[v + 1] = (uint32_t)out;
    }
    return out;
}

void jc_gen_uniform_cube(uint32_t n, float half, uint64_t seed, float *xyz) {
    uint64_t s = seed;
    for (uint64_t i = 0; i < 3ull * n; i++) {
        const float u = sm_f32(s);
        xyz[i] = (2.0f * u - 1.0f) * half;
    }
}

// ball_seed (stability.rs:30-50): for each node, rejection-sample the unit ball
// from three next_f32 draws (2u - 1 each, accepted when x^2 + y^2 + z^2 <= 1),
// scaled by sqrt(n) * 5.
void jc_gen_ball_seed(uint32_t n, uint64_t seed, float *xyz) {
    uint64_t s = seed;
    const float radius = std::sqrt((float)n) * 5.0f;
    for (uint32_t i = 0; i < n; i++) {
        for (;;) {
            const float x = 2.0f * sm_f32(s) - 1.0f;
            const float y = 2.0f * sm_f32(s) - 1.0f;
            const float z = 2.0f * sm_f32(s) - 1.0f;
            if (x * x + y * y + z * z <= 1.0f) {
                xyz[3 * (size_t)i] = x * radius;
                xyz[3 * (size_t)i + 1] = y * radius;
                xyz[3 * (size_t)i + 2] = z * radius;
                break;
            }
        }
    }
}

// The gRPC worker's deterministic seed when a graph arrives without positions
// (ring_seed, sim.rs:371-384): a 32-bit LCG (s = s * 1664525 + 1013904223 from
// 0x9E3779B1), each draw mapped to f32 (s / 2^32) * 2 - 1 and scaled by
// sqrt(n) * 5, filling x, y, z of each node in turn.
void jc_gen_ring_seed(uint32_t n, float *xyz) {
    uint32_t s = 0x9E3779B1u;
    const float radius = std::sqrt((float)(n > 1 ? n : 1)) * 5.0f;
    const float umax = (float)0xFFFFFFFFu;  // rounds to 2^32, as the f32 cast of u32::MAX does
    for (uint64_t k = 0; k < 3 * (uint64_t)n; k++) {
        s = s * 1664525u + 1013904223u;
        xyz[k] = ((float)s / umax * 2.0f - 1.0f) * radius;
    }
}

// StepStats::measure (stability.rs:128-186).  Non-finite nodes are counted
// and left out of every statistic; radii and displacements are sorted (NaN
// cannot occur among the finite nodes) and picked at index
// floor((len - 1) * q); the mean displacement divides the f64 sum of the
// finite nodes' displacements by n.
int jc_step_stats(uint32_t n, const float *prev, const float *cur, jc_step_stats_t *out) {
    if (!out || (n && (!prev || !cur))) {
        jc::set_error("jc_step_stats: NULL pointer");
        return JC_E_INVALID;
    }
    float max_abs = 0.0f;
    double sum_disp = 0.0;
    uint64_t nonfinite = 0;
    std::vector<float> radii, disps;
    radii.reserve(n);
    disps.reserve(n);
    for (uint32_t i = 0; i < n; i++) {
        const float x = cur[3 * (size_t)i], y = cur[3 * (size_t)i + 1], z = cur[3 * (size_t)i + 2];
        if (!(std::isfinite(x) && std::isfinite(y) && std::isfinite(z))) {
            nonfinite++;
            continue;
        }
        max_abs = std::fmax(std::fmax(std::fmax(max_abs, std::fabs(x)), std::fabs(y)), std::fabs(z));
        radii.push_back(std::sqrt(x * x + y * y + z * z));
        const float dx = x - prev[3 * (size_t)i], dy = y - prev[3 * (size_t)i + 1], dz = z - prev[3 * (size_t)i + 2];
        const double d = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
        disps.push_back((float)d);
        sum_disp += d;
    }
    std::sort(radii.begin(), radii.end());
    std::sort(disps.begin(), disps.end());
    auto pick = [](const std::vector<float> &v, double q) -> float {
        return v.empty() ? 0.0f : v[(size_t)((double)(v.size() - 1) * q)];
    };
    out->max_abs = max_abs;
    out->p50_radius = pick(radii, 0.5);
    out->p99_radius = pick(radii, 0.99);
    out->mean_disp = n ? (float)(sum_disp / (double)n) : 0.0f;
    out->p50_disp = pick(disps, 0.5);
    out->nonfinite = nonfinite;
    jc::set_error("");
    return JC_OK;
}

namespace {
inline uint32_t rd_le32(const uint8_t *b) {
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}
inline void wr_le32(uint8_t *b, uint32_t v) {
    b[0] = (uint8_t)v;
    b[1] = (uint8_t)(v >> 8);
    b[2] = (uint8_t)(v >> 16);
    b[3] = (uint8_t)(v >> 24);
}
}  // namespace

// sim.rs:72-104 (from_bin_bytes): header, then an exact length check.
int jc_csr_bin_header(const uint8_t *bytes, size_t len, uint32_t *n_nodes, uint32_t *n_edges) {
    if (!bytes && len) return finish(jc::Status::err(JC_E_INVALID, "CSR bytes pointer is NULL"));
    if (len < 8)
        return finish(jc::Status::err(JC_E_INVALID, "CSR truncated: " + std::to_string(len) +
                                                        " bytes < 8-byte header"));
    const uint32_t n = rd_le32(bytes), m = rd_le32(bytes + 4);
    const uint64_t expected = 8ull + 4ull * ((uint64_t)n + 1) + 4ull * (uint64_t)m;
    if ((uint64_t)len != expected)
        return finish(jc::Status::err(
            JC_E_INVALID, "CSR length mismatch: got " + std::to_string(len) + " bytes, expected " +
                              std::to_string(expected) + " (n_nodes=" + std::to_string(n) +
                              ", n_edges=" + std::to_string(m) + ")"));
    if (n_nodes) *n_nodes = n;
    if (n_edges) *n_edges = m;
    return finish(jc::Status::ok());
}

int jc_csr_from_bin(const uint8_t *bytes, size_t len, uint32_t *offsets, uint32_t *neighbors) {
    uint32_t n = 0, m = 0;
    const int rc = jc_csr_bin_header(bytes, len, &n, &m);
    if (rc) return rc;
    if ((!offsets) || (m && !neighbors))
        return finish(jc::Status::err(JC_E_INVALID, "CSR output pointer is NULL"));
    const uint8_t *p = bytes + 8;
    for (uint64_t i = 0; i <= n; i++, p += 4) offsets[i] = rd_le32(p);
    for (uint64_t i = 0; i < m; i++, p += 4) neighbors[i] = rd_le32(p);
    return finish(jc::Status::ok());
}

uint64_t jc_csr_bin_size(uint32_t n_nodes, uint64_t n_edges) {
    return 8ull + 4ull * ((uint64_t)n_nodes + 1) + 4ull * n_edges;
}

// sim.rs:118-127 (to_bin).
int jc_csr_to_bin(uint32_t n_nodes, const uint32_t *offsets, const uint32_t *neighbors,
                  uint64_t n_edges, uint8_t *out, uint64_t cap) {
    if (n_edges > 0xFFFFFFFFull)
        return finish(jc::Status::err(JC_E_INVALID, "CSR has more than 2^32-1 edges"));
    if (!offsets || (n_edges && !neighbors) || !out)
        return finish(jc::Status::err(JC_E_INVALID, "CSR pointer is NULL"));
    const uint64_t need = jc_csr_bin_size(n_nodes, n_edges);
    if (cap < need)
        return finish(jc::Status::err(JC_E_INVALID, "CSR output buffer too small: " +
                                                        std::to_string(cap) + " < " +
                                                        std::to_string(need)));
    wr_le32(out, n_nodes);
    wr_le32(out + 4, (uint32_t)n_edges);
    uint8_t *p = out + 8;
    for (uint64_t i = 0; i <= n_nodes; i++, p += 4) wr_le32(p, offsets[i]);
    for (uint64_t i = 0; i < n_edges; i++, p += 4) wr_le32(p, neighbors[i]);
    return finish(jc::Status::ok());
}

}  // extern "C"
