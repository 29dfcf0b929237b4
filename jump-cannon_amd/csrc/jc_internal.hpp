// Internal helpers shared by the HIP translation units of libjc_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/jc.h"

namespace jc {

// Thread-local last error message returned by jc_last_error().
void set_error(const std::string &msg);
const char *last_error();

struct Status {
    int code = JC_OK;
    static Status ok() { return {}; }
    static Status err(int c, const std::string &m) {
        set_error(m);
        Status s;
        s.code = c;
        return s;
    }
    bool failed() const { return code != JC_OK; }
};

#define JC_HIP(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::jc::Status::err(JC_E_HIP, std::string(#expr " failed: ") +             \
                                                   hipGetErrorString(e_));                  \
    } while (0)

#define JC_TRY(expr)                                                                        \
    do {                                                                                    \
        ::jc::Status s_ = (expr);                                                           \
        if (s_.failed()) return s_;                                                         \
    } while (0)

// Returns JC_E_NODEVICE (the reference's "requires a GPU" error) when no HIP
// device is visible; selects device `dev` otherwise.
Status require_device(int *dev_out);

// A copy queued on `st` behind the work already there, then a wait for the
// stream.  Engines run on non-blocking streams, which a plain hipMemcpy on the
// null stream does not wait for.
inline Status copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
    if (!bytes) return Status::ok();
    JC_HIP(hipMemcpyAsync(dst, src, bytes, kind, st));
    JC_HIP(hipStreamSynchronize(st));
    return Status::ok();
}

// Owning device allocation.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        cap = 0;
    }
    Status alloc(size_t count) {
        release();
        size_t bytes = (count ? count : 1) * sizeof(T);
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return Status::err(JC_E_OOM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                             " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
        cap = count ? count : 1;
        return Status::ok();
    }
    // Keeps the allocation when it already holds `count` elements.
    Status fit(size_t count) {
        if (p && count <= cap) {
            n = count;
            return Status::ok();
        }
        return alloc(count);
    }
};

// ---- device-wide primitives (prims.hip) ----
// Stable LSD radix sort of (key,val) pairs by key bits [begin_bit, end_bit)
// (8-bit digits).  `n` is fixed for a given scratch set so the launches can be
// captured into a hipGraph.
struct RadixScratch {
    DevBuf<uint64_t> keys_alt;
    DevBuf<uint32_t> vals_alt;
    DevBuf<uint32_t> hist;   // 256 * n_blocks, digit-major
    DevBuf<uint32_t> dbase;  // 256 digit bases
    DevBuf<uint32_t> long_runs, long_count;  // sort_morton_pairs: runs longer than kShortRun
    uint32_t n = 0, n_blocks = 0;
    Status alloc(uint32_t n);
};
Status radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, RadixScratch &s, hipStream_t st,
                            uint32_t begin_bit = 0, uint32_t end_bit = 64);
// The same result as radix_sort_pairs_u64 over all 64 bits (stable, ties in
// input order), in four radix passes over the upper halves plus a sort of each
// run of equal upper halves by the lower half.  Fast when such runs are short,
// as they are for octree keys; exact for any input.
Status sort_morton_pairs(uint64_t *keys, uint32_t *vals, RadixScratch &s, hipStream_t st);
// The same result for fewer than kSort24Max keys: three radix passes over bits
// [40, 64), then each key ranked within its run of equal upper 24 bits.
// Octree keys leave runs of a few keys (at most 16 in the 1M-node workload).
// Up to kSmallSort keys are ranked by counting instead (two launches).
Status sort_morton_pairs24(uint64_t *keys, uint32_t *vals, RadixScratch &s, hipStream_t st);
constexpr uint32_t kSort24Max = 1u << 24;  // long-run ranks fit the 24 bits above the low 40
constexpr uint32_t kSmallSort = 8192;      // sort_morton_pairs24: sort by ranking up to here
constexpr uint32_t kRankRun = 32;          // sort_morton_pairs24: runs ranked in place by each key
// Keys grouped by segment (many graphs): the caller writes a composite key ck
// (segment id above the top bits of the raw key) into s.keys_alt and the
// element indices into s.vals_alt; radix passes over ck's low ck_bits (a
// multiple of 16) leave them grouped, then each run of equal ck is ordered by
// (raw[index], index).  keys/vals receive the raw keys and indices in order.
Status sort_segmented_pairs(uint64_t *keys, uint32_t *vals, const uint64_t *raw, RadixScratch &s,
                            hipStream_t st, uint32_t ck_bits);

constexpr uint32_t kShortRun = 16;          // sort_morton_pairs: runs sorted by one thread
constexpr uint32_t kLdsRun = 4096;          // longer runs up to this size sort in LDS
constexpr uint32_t kLongRunThreads = 1024;
constexpr uint32_t kLongRunBlocks = 32;
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortItems = 8;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;

}  // namespace jc
