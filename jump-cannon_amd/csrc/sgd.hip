// Sparse-pivot stress by SGD on gfx950 (registry id "sgd-stress-hip").
//
// Reference: crates/graph-compute/src/engines/sgd_stress_gpu.rs (host pivot
// BFS, dense f32 distance matrix, one Jacobi kernel per sweep) and
// shaders/sgd_stress.wgsl:34-62.  Same arithmetic per term:
//   w = 1/d^2, mu = min(1, w*eta), r = mu*(|p-q| - d)/|p-q|, p -= r*(p-q)
// applied in pivot order to the node's own running position (Gauss-Seidel over
// its pivots, Jacobi across nodes: q comes from the sweep's input positions).
//
// MI355X layout: hop distances are u16 (0 = skip) in chunks of eight pivots per
// node, dist8[c][i] = the distances of node i to pivots 8c..8c+7 (16 bytes), so
// lane i of a wave reads one 16-byte word per eight terms, coalesced across the
// wave; the k pivot positions and ids of the sweep are staged in LDS once per
// block; w and mu depend only on d and eta, so they are tabulated per sweep in
// LDS (values identical to the per-term expressions).
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace jc {
namespace {

constexpr uint32_t kSgdTB = 256;
constexpr uint32_t kMaxTable = 8192;  // distances above this use the inline formula

__device__ __forceinline__ float len3f(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// Correctly rounded sqrt(s) for 2^-96 <= s < inf: the hardware square root
// (within one ulp) and the two neighbouring candidates, chosen by the sign of
// the residual s - c*r, which is the compiler's own IEEE sequence without the
// operand scaling for tiny inputs and the class fix-up for 0, inf and NaN.
__device__ __forceinline__ float sqrt_normal(float s) {
    const float r = __builtin_amdgcn_sqrtf(s);
    const float dn = __uint_as_float(__float_as_uint(r) - 1u);
    const float up = __uint_as_float(__float_as_uint(r) + 1u);
    float out = r;
    if (__builtin_fmaf(-dn, r, s) <= 0.0f) out = dn;
    if (__builtin_fmaf(-up, r, s) > 0.0f) out = up;
    return out;
}

// One sweep (sgd_stress.wgsl:34-62).  dist8[c * n + i] holds node i's u16 hop
// distances to pivots 8c .. 8c+7 (0: skip; the padding past k is 0), so one
// 16-byte load serves eight terms and the next chunk's load is in flight while
// the current eight are applied.  LDS holds the pivot positions and ids (when
// they fit) and the mu table; the terms run in pivot order as in the reference.
template <bool LDS_PIV>
__global__ __launch_bounds__(kSgdTB) void k_sgd_sweep(uint32_t n, uint32_t k,
                                                      const uint32_t *__restrict__ pivots,
                                                      const uint4 *__restrict__ dist8,
                                                      const float4 *__restrict__ pos_in,
                                                      float4 *__restrict__ pos_out, float eta,
                                                      uint32_t dmax) {
    extern __shared__ float4 smem[];
    float4 *piv = smem;                                                            // k positions
    uint32_t *pid = reinterpret_cast<uint32_t *>(smem + (LDS_PIV ? k : 0));       // k ids
    float *mu_tab = reinterpret_cast<float *>(pid + (LDS_PIV ? k : 0));            // dmax+1 entries
    const bool tab = dmax <= kMaxTable;
    if (LDS_PIV)
        for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) {
            pid[t] = pivots[t];
            piv[t] = pos_in[pivots[t]];
        }
    if (tab)
        for (uint32_t d = threadIdx.x; d <= dmax; d += blockDim.x) {
            const float df = (float)d;
            const float w = 1.0f / (df * df);
            mu_tab[d] = fminf(1.0f, w * eta);
        }
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p0 = pos_in[i];
    float px = p0.x, py = p0.y, pz = p0.z;
    const uint32_t kc = (k + 7) >> 3;
    uint4 cur = kc ? dist8[i] : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t c = 0; c < kc; c++) {
        const uint4 nxt = c + 1 < kc ? dist8[(size_t)(c + 1) * n + i] : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
        // Branch-free over the chunk, so the eight terms' LDS reads issue
        // together: a skipped term (distance 0: self, unreachable, padding; or
        // the node itself as pivot) is computed and then not taken, by select,
        // which leaves px/py/pz bit for bit unchanged.
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) {
            const uint32_t du = (wd[j >> 1] >> ((j & 1u) * 16u)) & 0xFFFFu;
            const uint32_t t = min(c * 8 + j, k - 1);
            const uint32_t pv = LDS_PIV ? pid[t] : pivots[t];
            const bool take = du != 0u && pv != i;
            const float d = (float)du;
            const float4 q = LDS_PIV ? piv[t] : pos_in[pv];
            const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
            const float ss = dx * dx + dy * dy + dz * dz;
            // wave-uniform choice: the short sequence when every lane's sum is
            // in its range (always, for finite layouts not collapsed to a point)
            const bool plain = ss >= 0x1p-96f && ss < __builtin_inff();
            const float root = __all(plain) ? sqrt_normal(ss) : sqrtf(ss);
            const float mag = fmaxf(root, 1e-6f);
            float mu;
            if (tab) mu = mu_tab[du];
            else {
                const float w = 1.0f / (d * d);
                mu = fminf(1.0f, w * eta);
            }
            const float r = mu * (mag - d) / mag;
            const float nx = px - r * dx, ny = py - r * dy, nz = pz - r * dz;
            px = take ? nx : px;
            py = take ? ny : py;
            pz = take ? nz : pz;
        }
        cur = nxt;
    }
    pos_out[i] = make_float4(px, py, pz, 0.0f);
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Scale-normalized stress over pivot terms (crates/graph-layouts/src/metrics.rs:31-59):
// terms (i, pivot t, d = hop distance > 0), w = 1/d^2.  Pass 0 accumulates
// num = sum w d D, den = sum w D^2, wd2 = sum w d^2 (D = layout distance);
// pass 1 accumulates sum w (alpha D - d)^2.  Per-block f64 partials in a fixed
// order, summed in block order by the host.
__global__ __launch_bounds__(256) void k_stress(uint32_t n, uint32_t k, const uint32_t *__restrict__ pivots,
                                                const uint16_t *__restrict__ dist,
                                                const float4 *__restrict__ pos, double alpha,
                                                int pass, double *__restrict__ part) {
    __shared__ double red[3][256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0, b = 0, c = 0;
    if (i < n) {
        const float4 p = pos[i];
        for (uint32_t t = 0; t < k; t++) {
            const uint32_t du = dist[(size_t)t * n + i];
            if (!du) continue;
            const float4 q = pos[pivots[t]];
            const double dx = (double)(p.x - q.x), dy = (double)(p.y - q.y), dz = (double)(p.z - q.z);
            const double D = sqrt(dx * dx + dy * dy + dz * dz);
            const double d = (double)du, w = 1.0 / (d * d);
            if (pass == 0) {
                a += w * d * D;
                b += w * D * D;
                c += w * d * d;
            } else {
                const double r = alpha * D - d;
                a += w * r * r;
            }
        }
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) part[3 * (size_t)blockIdx.x + q] = red[q][0];
}

// sgd_stress.rs:104-125 SplitMix64 (below(n) = next % n)
struct Rng {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// sgd_stress.rs:324-334 anneal_eta (host f32 powf, as the reference's host code)
float anneal_eta(uint64_t sweep, float eta_max, float eta_min, uint32_t n_anneal) {
    if (n_anneal == 0) return eta_min;
    const float t = (float)sweep / (float)n_anneal;
    if (t >= 1.0f) return eta_min;
    const float ratio = std::fmax(eta_min / eta_max, 1e-9f);
    return eta_max * powf(ratio, t);
}

class SgdStressEngine final : public LayoutEngine {
  public:
    explicit SgdStressEngine(const jc_sgd_settings &s) : settings_(s) {
        desc_.id = "sgd-stress-hip";
        desc_.kind = "physics";
        desc_.display_name = "SGD stress (pivot, MI355X)";
        desc_.description =
            "Sparse-pivot stress majorization by SGD (s_gd2 step, Ortmann max-min pivots): one "
            "Jacobi pivot update per node per sweep on the GPU, u16 hop distances.";
        desc_.needs_edges = 1;
        desc_.needs_cpu_positions = 1;
        desc_.needs_gpu_positions_buffer = 0;
    }
    ~SgdStressEngine() override { teardown(); }

    const jc_descriptor &descriptor() const override { return desc_; }

    Status set_params(const char *json, size_t len) override {
        jc_sgd_settings s;
        jc_sgd_default_settings(&s);
        std::vector<JsonField> f = {
            {JsonField::U32, "n_pivots", &s.n_pivots},
            {JsonField::U32, "sweeps_per_step", &s.sweeps_per_step},
            {JsonField::F32, "eta_max", &s.eta_max},
            {JsonField::F32, "eta_min", &s.eta_min},
            {JsonField::U32, "n_anneal_steps", &s.n_anneal_steps},
            {JsonField::U64, "seed", &s.seed},
        };
        int rc = decode_json_object(json, len, f, "sgd-stress settings");
        if (rc < 0) return Status{JC_E_INVALID};
        if (rc == 1) return Status::ok();
        settings_ = s;
        return Status::ok();
    }

    Status init(const CsrView &g, const float *xyz) override {
        JC_TRY(require_device(nullptr));
        JC_TRY(validate_csr(g));
        if (g.n && !xyz) return Status::err(JC_E_INVALID, "positions pointer is NULL");
        teardown();
        n_ = g.n;
        sweep_ = 0;
        // sgd_stress.rs:143-175 select_pivots: first pivot rng.below(n), then
        // max-min, every BFS row on the device (traverse.hip)
        dmax_ = 0;
        k_ = 0;
        std::vector<uint32_t> pivots;
        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        const uint32_t nn = n_ ? n_ : 1;
        JC_TRY(pos_[0].alloc(nn));
        JC_TRY(pos_[1].alloc(nn));
        if (n_) {
            k_ = std::min<uint32_t>(std::max<uint32_t>(settings_.n_pivots, 1), n_);
            Rng rng{settings_.seed};
            const uint32_t first = (uint32_t)(rng.next() % n_);
            JC_TRY(pivots_.alloc(k_));
            const size_t chunks = (k_ + 7) / 8;
            JC_TRY(dist_.alloc(chunks * n_ * 8));
            JC_HIP(hipMemsetAsync(dist_.p, 0, chunks * n_ * 8 * sizeof(uint16_t), stream_));
            JC_TRY(select_pivots_device(g, k_, first, stream_, pivots_.p, dist_.p, pivots, &dmax_));
        } else {
            JC_TRY(pivots_.alloc(1));
            JC_TRY(dist_.alloc(1));
        }
        eta_max_ = settings_.eta_max > 0.0f ? settings_.eta_max : (float)dmax_ * (float)dmax_;
        std::vector<float4> p4(nn, make_float4(0, 0, 0, 0));
        for (uint32_t i = 0; i < n_; i++)
            p4[i] = make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0.0f);
        JC_HIP(hipMemcpyAsync(pos_[0].p, p4.data(), nn * sizeof(float4), hipMemcpyHostToDevice, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        h_pivots_ = pivots;
        cur_ = 0;
        inited_ = true;
        return Status::ok();
    }

    // sgd_stress_gpu.rs:368-449: one eta per step, sweeps_per_step sweeps.
    Status step(uint32_t n_iters, float *out_xyz) override {
        if (!inited_) return Status::err(JC_E_STATE, "sgd-stress-hip step before init");
        if (n_ == 0 || k_ == 0) return Status::ok();
        const uint32_t sweeps = std::max<uint32_t>(settings_.sweeps_per_step, 1);
        const size_t tab = dmax_ <= kMaxTable ? (size_t)(dmax_ + 1) * sizeof(float) : 0;
        const size_t piv_bytes = (size_t)k_ * (sizeof(float4) + sizeof(uint32_t));
        const bool lds_piv = piv_bytes + tab <= 60 * 1024;
        const size_t shmem = (lds_piv ? piv_bytes : 0) + tab;
        const uint4 *d8 = reinterpret_cast<const uint4 *>(dist_.p);
        for (uint32_t it = 0; it < n_iters; it++) {
            const float eta = anneal_eta(sweep_, eta_max_, settings_.eta_min, settings_.n_anneal_steps);
            for (uint32_t s = 0; s < sweeps; s++) {
                if (lds_piv)
                    k_sgd_sweep<true><<<cdiv(n_, kSgdTB), kSgdTB, shmem, stream_>>>(
                        n_, k_, pivots_.p, d8, pos_[cur_].p, pos_[cur_ ^ 1].p, eta, dmax_);
                else
                    k_sgd_sweep<false><<<cdiv(n_, kSgdTB), kSgdTB, shmem, stream_>>>(
                        n_, k_, pivots_.p, d8, pos_[cur_].p, pos_[cur_ ^ 1].p, eta, dmax_);
                cur_ ^= 1;
            }
            sweep_ += sweeps;
        }
        JC_HIP(hipGetLastError());
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
        if (old_force) std::fill(old_force, old_force + 3 * (size_t)n_, 0.0f);
        if (se2) {
            se2[0] = anneal_eta(sweep_, eta_max_, settings_.eta_min, settings_.n_anneal_steps);
            se2[1] = (float)sweep_;
        }
        if (it) *it = iteration_;
        return Status::ok();
    }

    Status set_state(const float *xyz, const float *, const float *se2, uint64_t it) override {
        if (!inited_) return Status::err(JC_E_STATE, "set_state before init");
        if (xyz) {
            std::vector<float4> h(n_);
            for (uint32_t i = 0; i < n_; i++)
                h[i] = make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0.0f);
            JC_TRY(copy_sync(pos_[cur_].p, h.data(), n_ * sizeof(float4), hipMemcpyHostToDevice, stream_));
        }
        if (se2) sweep_ = (uint64_t)se2[1];
        iteration_ = it;
        return Status::ok();
    }

    uint32_t n_pivots() const { return k_; }
    const std::vector<uint32_t> &pivots() const { return h_pivots_; }
    Status read_dist(uint16_t *out) {
        // back to the reference's node-major layout dist[i*k + t]
        const size_t chunks = (k_ + 7) / 8;
        std::vector<uint16_t> c8(chunks * n_ * 8);
        if (!c8.empty())
            JC_TRY(copy_sync(c8.data(), dist_.p, c8.size() * 2, hipMemcpyDeviceToHost, stream_));
        for (uint32_t t = 0; t < k_; t++)
            for (uint32_t i = 0; i < n_; i++)
                out[(size_t)i * k_ + t] = c8[((size_t)(t >> 3) * n_ + i) * 8 + (t & 7)];
        return Status::ok();
    }
    float eta_max() const { return eta_max_; }

    Status time_sweeps(uint32_t n, double *ms) {
        if (!inited_) return Status::err(JC_E_STATE, "time before init");
        hipEvent_t a, b;
        JC_HIP(hipEventCreate(&a));
        JC_HIP(hipEventCreate(&b));
        JC_HIP(hipEventRecord(a, stream_));
        Status s = step(n, nullptr);
        JC_HIP(hipEventRecord(b, stream_));
        JC_HIP(hipEventSynchronize(b));
        float t = 0;
        JC_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms) *ms = t;
        return s;
    }

  private:
    Status read_positions(float *out) {
        std::vector<float4> h(n_);
        JC_HIP(hipMemcpyAsync(h.data(), pos_[cur_].p, n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        for (uint32_t i = 0; i < n_; i++) {
            out[3 * (size_t)i] = h[i].x;
            out[3 * (size_t)i + 1] = h[i].y;
            out[3 * (size_t)i + 2] = h[i].z;
        }
        return Status::ok();
    }

    void teardown() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        pos_[0].release();
        pos_[1].release();
        pivots_.release();
        dist_.release();
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
        inited_ = false;
    }

    jc_descriptor desc_{};
    jc_sgd_settings settings_;
    bool inited_ = false;
    uint32_t n_ = 0, k_ = 0, cur_ = 0, dmax_ = 0;
    uint64_t sweep_ = 0, iteration_ = 0;
    float eta_max_ = 1.0f;
    hipStream_t stream_ = nullptr;
    DevBuf<float4> pos_[2];
    DevBuf<uint32_t> pivots_;
    DevBuf<uint16_t> dist_;
    std::vector<uint32_t> h_pivots_;
};

}  // namespace

class StressPlan {
  public:
    ~StressPlan() {
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    Status create(uint32_t n, uint32_t k, const uint32_t *pivots, const uint16_t *dist_node_major) {
        JC_TRY(require_device(nullptr));
        n_ = n;
        k_ = k;
        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        std::vector<uint16_t> pm((size_t)k * n);
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t t = 0; t < k; t++) pm[(size_t)t * n + i] = dist_node_major[(size_t)i * k + t];
        for (uint32_t t = 0; t < k; t++)
            if (pivots[t] >= n) return Status::err(JC_E_INVALID, "pivot out of range");
        JC_TRY(piv_.alloc(k ? k : 1));
        JC_TRY(dist_.alloc(pm.size()));
        JC_TRY(pos_.alloc(n ? n : 1));
        JC_TRY(part_.alloc(3 * (size_t)cdiv(n ? n : 1, 256)));
        if (k) JC_TRY(copy_sync(piv_.p, pivots, k * 4, hipMemcpyHostToDevice, stream_));
        if (!pm.empty()) JC_TRY(copy_sync(dist_.p, pm.data(), pm.size() * 2, hipMemcpyHostToDevice, stream_));
        return Status::ok();
    }
    Status eval(const float *xyz, double *out) {
        if (!n_ || !k_) {
            *out = 0.0;
            return Status::ok();
        }
        std::vector<float4> h(n_);
        for (uint32_t i = 0; i < n_; i++)
            h[i] = make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0.0f);
        JC_HIP(hipMemcpyAsync(pos_.p, h.data(), n_ * sizeof(float4), hipMemcpyHostToDevice, stream_));
        const uint32_t nb = cdiv(n_, 256);
        std::vector<double> part(3 * (size_t)nb);
        auto run = [&](double alpha, int pass, double *acc) -> Status {
            k_stress<<<nb, 256, 0, stream_>>>(n_, k_, piv_.p, dist_.p, pos_.p, alpha, pass, part_.p);
            JC_HIP(hipGetLastError());
            JC_HIP(hipMemcpyAsync(part.data(), part_.p, part.size() * 8, hipMemcpyDeviceToHost, stream_));
            JC_HIP(hipStreamSynchronize(stream_));
            acc[0] = acc[1] = acc[2] = 0.0;
            for (uint32_t b = 0; b < nb; b++)
                for (int q = 0; q < 3; q++) acc[q] += part[3 * (size_t)b + q];
            return Status::ok();
        };
        double s0[3], s1[3];
        JC_TRY(run(0.0, 0, s0));
        if (s0[1] <= 0.0 || s0[2] <= 0.0) {
            *out = 0.0;
            return Status::ok();
        }
        const double alpha = s0[0] / s0[1];
        JC_TRY(run(alpha, 1, s1));
        *out = s1[0] / s0[2];
        return Status::ok();
    }

  private:
    uint32_t n_ = 0, k_ = 0;
    hipStream_t stream_ = nullptr;
    DevBuf<uint32_t> piv_;
    DevBuf<uint16_t> dist_;
    DevBuf<float4> pos_;
    DevBuf<double> part_;
};

LayoutEngine *make_sgd(const jc_sgd_settings &s) { return new SgdStressEngine(s); }

Status sgd_pivots(LayoutEngine *e, uint32_t *k, uint32_t *pivots, uint16_t *dist, float *eta_max) {
    auto *s = dynamic_cast<SgdStressEngine *>(e);
    if (!s) return Status::err(JC_E_INVALID, "engine is not sgd-stress-hip");
    if (k) *k = s->n_pivots();
    if (pivots) std::copy(s->pivots().begin(), s->pivots().end(), pivots);
    if (eta_max) *eta_max = s->eta_max();
    if (dist) JC_TRY(s->read_dist(dist));
    return Status::ok();
}

StressPlan *make_stress_plan() { return new StressPlan(); }
Status stress_plan_create(StressPlan *p, uint32_t n, uint32_t k, const uint32_t *pivots,
                          const uint16_t *dist) {
    return p->create(n, k, pivots, dist);
}
Status stress_plan_eval(StressPlan *p, const float *xyz, double *out) { return p->eval(xyz, out); }
void free_stress_plan(StressPlan *p) { delete p; }

Status sgd_time(LayoutEngine *e, uint32_t n, double *ms) {
    auto *s = dynamic_cast<SgdStressEngine *>(e);
    if (!s) return Status::err(JC_E_INVALID, "engine is not sgd-stress-hip");
    return s->time_sweeps(n, ms);
}

}  // namespace jc
