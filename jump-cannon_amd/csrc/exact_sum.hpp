// Exact, order-independent sums of finite f32 values on the device.
//
// A value v = m * 2^(e-150) (m < 2^24) is added as an integer in units of 2^-149
// to a 320-bit accumulator held as ten 32-bit limbs in u64 (the upper halves
// absorb carries), so parallel blocks can add their partial limbs in any order
// and the final rounding to f64 happens once.  The CPU oracle implements the
// same arithmetic, which makes the f64 results bit-identical.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace jc {

constexpr uint32_t kLimbs = 10;

// Adds |v| (v finite) to acc[0..kLimbs) with LDS/global u64 atomics.
__device__ __forceinline__ void exact_acc_add(unsigned long long *acc, float v) {
    const uint32_t u = __float_as_uint(v) & 0x7FFFFFFFu;
    if (!u) return;
    const uint32_t e = u >> 23;
    uint32_t m = u & 0x7FFFFFu;
    if (e) m |= 0x800000u;
    const uint32_t sh = (e ? e : 1u) - 1u;
    const uint64_t t = (uint64_t)m << (sh & 31u);
    atomicAdd(&acc[sh >> 5], (unsigned long long)(t & 0xFFFFFFFFull));
    if (t >> 32) atomicAdd(&acc[(sh >> 5) + 1], (unsigned long long)(t >> 32));
}

// Adds |v| (v finite) to a thread-private accumulator held in registers: the
// same two 32-bit pieces exact_acc_add adds, placed by comparisons against
// constant limb indices (no dynamic indexing, no atomics).
__device__ __forceinline__ void exact_reg_add(unsigned long long (&acc)[kLimbs], float v) {
    const uint32_t u = __float_as_uint(v) & 0x7FFFFFFFu;
    const uint32_t e = u >> 23;
    uint32_t m = u & 0x7FFFFFu;
    if (e) m |= 0x800000u;
    const uint32_t sh = (e ? e : 1u) - 1u;
    const uint64_t t = (uint64_t)m << (sh & 31u);
    const uint32_t k = sh >> 5;
    const unsigned long long lo = t & 0xFFFFFFFFull, hi = t >> 32;
#pragma unroll
    for (uint32_t j = 0; j < kLimbs; j++) {
        acc[j] += (j == k) ? lo : 0ull;
        acc[j] += (j == k + 1) ? hi : 0ull;
    }
}

// Carry-normalises l[0..kLimbs) in place (each limb < 2^32 afterwards).
__device__ __forceinline__ void exact_normalise(unsigned long long *l) {
    for (uint32_t k = 0; k + 1 < kLimbs; k++) {
        l[k + 1] += l[k] >> 32;
        l[k] &= 0xFFFFFFFFull;
    }
}

// Nearest f64 (ties to even) of a normalised non-negative limb integer * 2^-149.
// Word-level: the three limbs from the top nonzero one form a 96-bit window
// H:L; the 53-bit mantissa, round bit and sticky bit come from shifts of the
// window plus an OR of the limbs below it.  All limb indices are compile-time
// constants after unrolling, so the limbs stay in registers.
__host__ __device__ inline double exact_to_double(const unsigned long long *l) {
    int tk = -1;
#pragma unroll
    for (int k = (int)kLimbs - 1; k >= 0; k--)
        if (tk < 0 && l[k] != 0ull) tk = k;
    if (tk < 0) return 0.0;
    unsigned long long hi = 0, mid = 0, lo = 0;
    bool below = false;
#pragma unroll
    for (int k = 0; k < (int)kLimbs; k++) {
        if (k == tk) hi = l[k];
        if (k == tk - 1) mid = l[k];
        if (k == tk - 2) lo = l[k];
        if (k < tk - 2) below |= l[k] != 0ull;
    }
    const unsigned long long L = (mid << 32) | lo;      // window bits 0..63; hi is bits 64..95
    const int h = 63 - __builtin_clzll(hi);                // top bit of hi (0..31)
    const int sh = 12 + h;                                 // window bits dropped below the mantissa
    unsigned long long mant = (hi << (52 - h)) | (L >> sh);
    const unsigned long long rb = (L >> (sh - 1)) & 1ull;
    const bool sticky = below || (L & ((1ull << (sh - 1)) - 1ull)) != 0ull;
    if (rb && (sticky || (mant & 1ull))) mant++;
    return ldexp((double)mant, sh + 32 * (tk - 2) - 149);
}

// p - q for normalised limb integers with p >= q, result normalised in p.
__device__ inline void exact_sub(unsigned long long *p, const unsigned long long *q) {
    long long borrow = 0;
    for (uint32_t k = 0; k < kLimbs; k++) {
        long long d = (long long)p[k] - (long long)q[k] - borrow;
        borrow = d < 0 ? 1 : 0;
        p[k] = (unsigned long long)(d + (borrow ? (1ll << 32) : 0));
    }
}

}  // namespace jc
