// Weighted sparse matrix-vector product y = A x over a CSR matrix on gfx950.
//
// Reference: crates/graph-compute/src/analytics/spmv.rs
//   cpu_spmv        :88-99    y[v] = sum over row v in CSR order of w[e] * x[nbr[e]]
//   gpu_spmv        :103-287  the same per-row serial sum (shaders/spmv.wgsl:27-41)
//   gpu_spmv_f16    :293-469  weights and x stored as f16 (pack_f16_pairs :68-86,
//                             round-to-nearest-even), products and sum in f32
//                             (shaders/spmv_f16.wgsl:27-45)
//   gpu_spmv_hybrid :491-676  one 64-lane workgroup per row: lane l sums the
//                             entries start+l, start+l+64, ... and a tree over the
//                             64 partials (stride 32, 16, ..., 1) gives y[v]
//                             (shaders/spmv_hybrid.wgsl:44-84)
//
// Each variant here performs the reference's f32 operations in the reference's
// order, so its result has the reference algorithm's bits:
//   serial (f32, f16): one launch (k_spmv_rows).  Rows of at most kLight
//     entries go 64 to a wave; the wave reads their index and weight runs
//     coalesced, forms every product in parallel into LDS, then each lane adds
//     its own row's products in CSR order.  Longer rows take one wave each: the
//     products of 1024 entries per round land in LDS and one lane adds them in
//     order.
//   hybrid: the reference's lane-strided partials and tree as cross-lane
//     shuffles that pair the same partials; rows of at most S entries share a
//     wave 64 / S to a wave (S a power of two), longer rows take a wave each.
#include <algorithm>
#include <vector>

#include "engine.hpp"

struct jc_spmv_plan;

namespace jc {
namespace {

constexpr uint32_t kSpTB = 256;
constexpr uint32_t kLight = 32;       // rows with more entries take a wave each
constexpr uint32_t kLongChunk = 1024;  // products per long-row round: 16 per lane

inline uint32_t cdiv_u32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Storage of the weights and x: f32, or f16 bits (the f16 variant).
template <typename T>
__device__ __forceinline__ float load_val(const T *p, uint32_t i);
template <>
__device__ __forceinline__ float load_val<float>(const float *p, uint32_t i) {
    return p[i];
}
template <>
__device__ __forceinline__ float load_val<uint16_t>(const uint16_t *p, uint32_t i) {
    return (float)__builtin_bit_cast(_Float16, p[i]);
}

// f32 -> f16 with round-to-nearest-even (the hardware conversion in the default
// rounding mode), as pack_f16_pairs does on the host with f16::from_f32.
__global__ __launch_bounds__(kSpTB) void k_to_f16(const float *__restrict__ src, uint64_t n,
                                                  uint16_t *__restrict__ dst) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = __builtin_bit_cast(uint16_t, (_Float16)src[i]);
}

// One product over all rows in one launch.  The first long_blocks blocks take
// the long rows (more than kLight entries), one row per wave, longest first;
// the rest take the light rows, 64 consecutive rows per wave, whose entries
// are contiguous in the light CSR (offl / nbrl / wl, where long rows are
// empty).  Both halves run side by side: a long row's serial sum is a latency
// chain that leaves most of the machine to the light rows.
//   light: the wave reads its index and weight runs coalesced, forms every
//     product into its LDS segment, and each lane adds its own row in order;
//   long: each round forms up to kLongChunk products into the segment (all
//     loads of a lane in flight together), then lane 0 adds them in order.
template <typename T>
__global__ __launch_bounds__(kSpTB) void k_spmv_rows(
    uint32_t n, uint32_t long_blocks, uint32_t n_long, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr, const T *__restrict__ wt,
    const uint32_t *__restrict__ offl, const uint32_t *__restrict__ nbrl,
    const T *__restrict__ wl, const T *__restrict__ x, float *__restrict__ y) {
    constexpr uint32_t W = kSpTB / 64, SEG = 64 * kLight;
    static_assert(kLongChunk <= SEG, "a long-row round fits the wave's segment");
    __shared__ __align__(16) float prod[W][SEG];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (blockIdx.x < long_blocks) {
        const uint32_t li = blockIdx.x * W + w;
        if (li >= n_long) return;
        const uint32_t v = rows[li];
        const uint32_t s = off[v], e = off[v + 1];
        float acc = 0.0f;
        for (uint32_t c = s; c < e; c += kLongChunk) {
            const uint32_t len = min(kLongChunk, e - c);
            constexpr uint32_t R = kLongChunk / 64;
            uint32_t j[R];
            float a[R], b[R];
#pragma unroll
            for (uint32_t r = 0; r < R; r++) {
                const uint32_t k = lane + 64u * r;
                j[r] = k < len ? nbr[c + k] : 0u;
                a[r] = k < len ? load_val(wt, c + k) : 0.0f;
            }
#pragma unroll
            for (uint32_t r = 0; r < R; r++) b[r] = load_val(x, j[r]);
#pragma unroll
            for (uint32_t r = 0; r < R; r++) {
                const uint32_t k = lane + 64u * r;
                if (k < len) prod[w][k] = a[r] * b[r];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) {
                uint32_t k = 0;
                for (; k + 4 <= len; k += 4) {
                    const float4 q = *reinterpret_cast<const float4 *>(&prod[w][k]);
                    acc = acc + q.x;
                    acc = acc + q.y;
                    acc = acc + q.z;
                    acc = acc + q.w;
                }
                for (; k < len; k++) acc = acc + prod[w][k];
            }
            // lane 0 has read the round before the next one overwrites it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane == 0) y[v] = acc;
        return;
    }
    const uint32_t r0 = ((blockIdx.x - long_blocks) * W + w) * 64;
    if (r0 >= n) return;
    const uint32_t v = r0 + lane;
    const uint32_t seg0 = offl[r0], seg1 = offl[min(r0 + 64u, n)];
    for (uint32_t base = seg0; base < seg1; base += 8 * 64) {
        uint32_t u[8];
        float a[8], b[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) {
            const uint32_t k = base + lane + 64u * q;
            u[q] = k < seg1 ? nbrl[k] : 0u;
            a[q] = k < seg1 ? load_val(wl, k) : 0.0f;
        }
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) b[q] = load_val(x, u[q]);
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) {
            const uint32_t k = base + lane + 64u * q;
            if (k < seg1) prod[w][k - seg0] = a[q] * b[q];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (v < n && off[v + 1] - off[v] <= kLight) {
        const uint32_t s = offl[v] - seg0, e = offl[v + 1] - seg0;
        float acc = 0.0f;
        for (uint32_t k = s; k < e; k++) acc = acc + prod[w][k];
        y[v] = acc;
    }
}

// gpu_spmv_hybrid's order (spmv_hybrid.wgsl:44-84): lane l of 64 adds
// w[e] * x[nbr[e]] for e = start + l, start + l + 64, ... from +0; then for
// stride 32, 16, ..., 1 the lanes below the stride add the partial `stride`
// lanes up; lane 0 has y[v].
// A row of L <= S entries (S a power of two <= 64) gives the same bits on S
// lanes with the strides S/2 ... 1 only: every partial past the row is +0, no
// partial is -0 (each starts as +0 + product), so the reference's stages with
// stride >= S add +0 to every partial that matters and change nothing.  Rows
// are therefore grouped by S on the host: a wave takes 64 / S rows of one
// class, and rows longer than 64 take a whole wave.
constexpr uint32_t kHybridClasses = 8;  // S = 1, 2, 4, ..., 64; then rows > 64
struct HybridClasses {
    uint32_t block_begin[kHybridClasses + 1];
    uint32_t list_begin[kHybridClasses];
    uint32_t count[kHybridClasses];
};
__device__ __forceinline__ uint32_t hybrid_width(uint32_t c) { return c < 7 ? 1u << c : 64u; }

__global__ __launch_bounds__(kSpTB) void k_spmv_hybrid(const HybridClasses hc,
                                                       const uint32_t *__restrict__ rows,
                                                       const uint32_t *__restrict__ off,
                                                       const uint32_t *__restrict__ nbr,
                                                       const float *__restrict__ wt,
                                                       const float *__restrict__ x,
                                                       float *__restrict__ y) {
    uint32_t c = 0;
#pragma unroll
    for (uint32_t k = 1; k < kHybridClasses; k++)
        if (blockIdx.x >= hc.block_begin[k]) c = k;
    const uint32_t S = hybrid_width(c);
    const uint32_t lane = threadIdx.x & 63, sub = lane & (S - 1);
    const uint32_t wave = (blockIdx.x - hc.block_begin[c]) * (kSpTB / 64) + (threadIdx.x >> 6);
    const uint32_t slot = wave * (64u / S) + lane / S;
    const bool live = slot < hc.count[c];
    uint32_t v = 0, s = 0, e = 0;
    if (live) {
        v = rows[hc.list_begin[c] + slot];
        s = off[v];
        e = off[v + 1];
    }
    float acc = 0.0f;
    uint32_t k = s + sub;
    if (c == kHybridClasses - 1) {
        // four strided entries per round, their loads in flight together; the
        // additions stay in the lane's order
        for (; k + 3u * 64u < e; k += 4u * 64u) {
            const uint32_t j0 = nbr[k], j1 = nbr[k + 64], j2 = nbr[k + 128], j3 = nbr[k + 192];
            const float w0 = wt[k], w1 = wt[k + 64], w2 = wt[k + 128], w3 = wt[k + 192];
            const float x0 = x[j0], x1 = x[j1], x2 = x[j2], x3 = x[j3];
            acc = acc + w0 * x0;
            acc = acc + w1 * x1;
            acc = acc + w2 * x2;
            acc = acc + w3 * x3;
        }
        for (; k < e; k += 64u) acc = acc + wt[k] * x[nbr[k]];
    } else if (k < e) {
        acc = acc + wt[k] * x[nbr[k]];
    }
    for (uint32_t stride = S >> 1; stride > 0; stride >>= 1) {
        const float up = __shfl_down(acc, stride, (int)S);
        if (sub < stride) acc = acc + up;
    }
    if (live && sub == 0) y[v] = acc;
}

}  // namespace

class SpmvPlan {
  public:
    Status create(uint32_t n, const uint32_t *off, const uint32_t *nbr, const float *wt,
                  uint64_t nnz, int kind) {
        if (kind != JC_SPMV_F32 && kind != JC_SPMV_F16 && kind != JC_SPMV_HYBRID)
            return Status::err(JC_E_INVALID, "unknown SpMV kind " + std::to_string(kind));
        JC_TRY(require_device(nullptr));
        CsrView g{n, off, nbr, nnz};
        JC_TRY(validate_csr(g));
        if (nnz && !wt) return Status::err(JC_E_INVALID, "weights pointer is NULL");
        kind_ = kind;
        n_ = n;
        if (n == 0) return Status::ok();
        JC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        JC_TRY(off_.alloc((size_t)n + 1));
        JC_TRY(nbr_.alloc(nnz));
        JC_TRY(y_.alloc(n));
        JC_HIP(hipMemcpyAsync(off_.p, off, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
        if (nnz) JC_HIP(hipMemcpyAsync(nbr_.p, nbr, nnz * 4, hipMemcpyHostToDevice, stream_));
        // f32 weights and x: also the staging area of the f16 conversion
        JC_TRY(w32_.alloc(nnz));
        JC_TRY(x32_.alloc(n));
        if (nnz) JC_HIP(hipMemcpyAsync(w32_.p, wt, nnz * 4, hipMemcpyHostToDevice, stream_));
        if (kind == JC_SPMV_HYBRID) {
            // rows by class: S = the smallest power of two >= the row length
            // (1 for empty rows) up to 64, then the longer rows, longest first
            std::vector<uint32_t> cls[kHybridClasses];
            for (uint32_t v = 0; v < n; v++) {
                const uint32_t d = off[v + 1] - off[v];
                uint32_t c = 0;
                while (c < kHybridClasses - 1 && (1u << c) < d) c++;
                cls[c].push_back(v);
            }
            std::stable_sort(cls[kHybridClasses - 1].begin(), cls[kHybridClasses - 1].end(),
                             [&](uint32_t a, uint32_t b) {
                                 return off[a + 1] - off[a] > off[b + 1] - off[b];
                             });
            std::vector<uint32_t> order;
            order.reserve(n);
            uint32_t blocks = 0;
            for (uint32_t c = 0; c < kHybridClasses; c++) {
                const uint32_t width = c < 7 ? 1u << c : 64u;
                hc_.block_begin[c] = blocks;
                hc_.list_begin[c] = (uint32_t)order.size();
                hc_.count[c] = (uint32_t)cls[c].size();
                blocks += cdiv_u32(cls[c].size(), (kSpTB / 64) * (64u / width));
                order.insert(order.end(), cls[c].begin(), cls[c].end());
            }
            hc_.block_begin[kHybridClasses] = blocks;
            JC_TRY(rows_.alloc(order.size()));
            JC_HIP(hipMemcpyAsync(rows_.p, order.data(), order.size() * 4, hipMemcpyHostToDevice,
                                  stream_));
            JC_HIP(hipStreamSynchronize(stream_));
            return Status::ok();
        }
        // light / long split of the rows (the long rows longest first: they
        // are the critical path of a product)
        std::vector<uint32_t> rows, offl((size_t)n + 1, 0), nbrl;
        std::vector<float> wl;
        for (uint32_t v = 0; v < n; v++) {
            const uint32_t d = off[v + 1] - off[v];
            if (d > kLight) {
                rows.push_back(v);
            } else {
                nbrl.insert(nbrl.end(), nbr + off[v], nbr + off[v + 1]);
                wl.insert(wl.end(), wt + off[v], wt + off[v + 1]);
            }
            offl[v + 1] = (uint32_t)nbrl.size();
        }
        std::stable_sort(rows.begin(), rows.end(), [&](uint32_t a, uint32_t b) {
            return off[a + 1] - off[a] > off[b + 1] - off[b];
        });
        n_long_ = (uint32_t)rows.size();
        JC_TRY(offl_.alloc((size_t)n + 1));
        JC_TRY(nbrl_.alloc(nbrl.size()));
        JC_TRY(rows_.alloc(rows.size()));
        JC_TRY(wl32_.alloc(wl.size()));
        JC_HIP(hipMemcpyAsync(offl_.p, offl.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
        if (!nbrl.empty())
            JC_HIP(hipMemcpyAsync(nbrl_.p, nbrl.data(), nbrl.size() * 4, hipMemcpyHostToDevice, stream_));
        if (!rows.empty())
            JC_HIP(hipMemcpyAsync(rows_.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, stream_));
        if (!wl.empty())
            JC_HIP(hipMemcpyAsync(wl32_.p, wl.data(), wl.size() * 4, hipMemcpyHostToDevice, stream_));
        if (kind == JC_SPMV_F16) {
            // the f16 variant keeps only f16 weights on the device
            JC_TRY(w16_.alloc(nnz));
            JC_TRY(wl16_.alloc(wl.size()));
            JC_TRY(x16_.alloc(n));
            JC_TRY(convert(w32_.p, nnz, w16_.p));
            JC_TRY(convert(wl32_.p, wl.size(), wl16_.p));
            JC_HIP(hipStreamSynchronize(stream_));
            w32_.release();
            wl32_.release();
        }
        JC_HIP(hipStreamSynchronize(stream_));
        return Status::ok();
    }

    ~SpmvPlan() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (auto &e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    // y = A x (x: n host floats; y: n host floats).
    Status run(const float *x, float *y, double *ms) {
        if (ms) *ms = 0.0;
        if (n_ == 0) return Status::ok();
        if (!x || !y) return Status::err(JC_E_INVALID, "x or y pointer is NULL");
        JC_TRY(ensure_ready());
        JC_HIP(hipMemcpyAsync(x32_.p, x, (size_t)n_ * 4, hipMemcpyHostToDevice, stream_));
        JC_HIP(hipEventRecord(ev_[0], stream_));
        JC_TRY(enqueue_product());
        JC_HIP(hipEventRecord(ev_[1], stream_));
        JC_HIP(hipMemcpyAsync(y, y_.p, (size_t)n_ * 4, hipMemcpyDeviceToHost, stream_));
        JC_HIP(hipStreamSynchronize(stream_));
        float t = 0.0f;
        JC_HIP(hipEventElapsedTime(&t, ev_[0], ev_[1]));
        if (ms) *ms = t;
        return Status::ok();
    }

    // Device time per product over n_iters back-to-back products of the x
    // resident from the last run (zeros before any run).
    Status time_products(uint32_t n_iters, double *ms_per) {
        if (ms_per) *ms_per = 0.0;
        if (n_ == 0 || n_iters == 0) return Status::ok();
        JC_TRY(ensure_ready());
        JC_TRY(enqueue_product());  // warm-up
        JC_HIP(hipEventRecord(ev_[0], stream_));
        for (uint32_t k = 0; k < n_iters; k++) JC_TRY(enqueue_product());
        JC_HIP(hipEventRecord(ev_[1], stream_));
        JC_HIP(hipEventSynchronize(ev_[1]));
        float t = 0.0f;
        JC_HIP(hipEventElapsedTime(&t, ev_[0], ev_[1]));
        if (ms_per) *ms_per = t / n_iters;
        return Status::ok();
    }

  private:
    // Timing events, and x zeroed before its first use (plan_time may run
    // before any product has uploaded an x).
    Status ensure_ready() {
        for (auto &e : ev_)
            if (!e) JC_HIP(hipEventCreate(&e));
        if (!x_ready_) {
            JC_HIP(hipMemsetAsync(x32_.p, 0, (size_t)n_ * 4, stream_));
            x_ready_ = true;
        }
        return Status::ok();
    }

    Status convert(const float *src, uint64_t count, uint16_t *dst) {
        if (!count) return Status::ok();
        const uint32_t grid = std::min<uint32_t>(cdiv_u32(count, kSpTB), 4096);
        k_to_f16<<<grid, kSpTB, 0, stream_>>>(src, count, dst);
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    Status enqueue_product() {
        if (kind_ == JC_SPMV_HYBRID) {
            if (hc_.block_begin[kHybridClasses])
                k_spmv_hybrid<<<hc_.block_begin[kHybridClasses], kSpTB, 0, stream_>>>(
                    hc_, rows_.p, off_.p, nbr_.p, w32_.p, x32_.p, y_.p);
        } else {
            const uint32_t lb = cdiv_u32(n_long_, kSpTB / 64);
            const uint32_t grid = lb + cdiv_u32(n_, kSpTB);
            if (kind_ == JC_SPMV_F16) {
                JC_TRY(convert(x32_.p, n_, x16_.p));
                k_spmv_rows<uint16_t><<<grid, kSpTB, 0, stream_>>>(
                    n_, lb, n_long_, rows_.p, off_.p, nbr_.p, w16_.p, offl_.p, nbrl_.p, wl16_.p,
                    x16_.p, y_.p);
            } else {
                k_spmv_rows<float><<<grid, kSpTB, 0, stream_>>>(
                    n_, lb, n_long_, rows_.p, off_.p, nbr_.p, w32_.p, offl_.p, nbrl_.p, wl32_.p,
                    x32_.p, y_.p);
            }
        }
        JC_HIP(hipGetLastError());
        return Status::ok();
    }

    HybridClasses hc_{};
    int kind_ = JC_SPMV_F32;
    uint32_t n_ = 0, n_long_ = 0;
    bool x_ready_ = false;
    hipStream_t stream_ = nullptr;
    DevBuf<uint32_t> off_, nbr_, offl_, nbrl_, rows_;
    DevBuf<float> w32_, wl32_, x32_, y_;
    DevBuf<uint16_t> w16_, wl16_, x16_;
    hipEvent_t ev_[2] = {};
};

}  // namespace jc

struct jc_spmv_plan {
    jc::SpmvPlan impl;
};

extern "C" {

int jc_spmv_plan_create(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                        const float *weights, uint64_t nnz, int kind, jc_spmv_plan **out) {
    if (!out) {
        jc::set_error("NULL output handle");
        return JC_E_INVALID;
    }
    auto *p = new jc_spmv_plan();
    jc::Status s = p->impl.create(n, offsets, neighbors, weights, nnz, kind);
    if (s.failed()) {
        delete p;
        return s.code;
    }
    *out = p;
    jc::set_error("");
    return JC_OK;
}

int jc_spmv_plan_run(jc_spmv_plan *p, const float *x, float *y, double *out_gpu_ms) {
    if (!p) {
        jc::set_error("NULL plan");
        return JC_E_INVALID;
    }
    jc::Status s = p->impl.run(x, y, out_gpu_ms);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

int jc_spmv_plan_time(jc_spmv_plan *p, uint32_t n_iters, double *ms_per_product) {
    if (!p) {
        jc::set_error("NULL plan");
        return JC_E_INVALID;
    }
    jc::Status s = p->impl.time_products(n_iters, ms_per_product);
    if (!s.failed()) jc::set_error("");
    return s.code;
}

void jc_spmv_plan_destroy(jc_spmv_plan *p) { delete p; }

static int spmv_once(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                     const float *weights, uint64_t nnz, const float *x, float *y, int kind) {
    jc_spmv_plan *p = nullptr;
    int rc = jc_spmv_plan_create(n, offsets, neighbors, weights, nnz, kind, &p);
    if (rc) return rc;
    rc = jc_spmv_plan_run(p, x, y, nullptr);
    jc_spmv_plan_destroy(p);
    return rc;
}

int jc_spmv(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors, const float *weights,
            uint64_t nnz, const float *x, float *y) {
    return spmv_once(n, offsets, neighbors, weights, nnz, x, y, JC_SPMV_F32);
}

int jc_spmv_f16(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                const float *weights, uint64_t nnz, const float *x, float *y) {
    return spmv_once(n, offsets, neighbors, weights, nnz, x, y, JC_SPMV_F16);
}

int jc_spmv_hybrid(uint32_t n, const uint32_t *offsets, const uint32_t *neighbors,
                   const float *weights, uint64_t nnz, const float *x, float *y) {
    return spmv_once(n, offsets, neighbors, weights, nnz, x, y, JC_SPMV_HYBRID);
}

}  // extern "C"
