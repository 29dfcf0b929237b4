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

// Carry-normalises l[0..kLimbs) in place (each limb < 2^32 afterwards).
__device__ __forceinline__ void exact_normalise(unsigned long long *l) {
    for (uint32_t k = 0; k + 1 < kLimbs; k++) {
        l[k + 1] += l[k] >> 32;
        l[k] &= 0xFFFFFFFFull;
    }
}

// Nearest f64 (ties to even) of a normalised non-negative limb integer * 2^-149.
__device__ inline double exact_to_double(const unsigned long long *l) {
    int top = -1;
    for (int k = (int)kLimbs - 1; k >= 0; k--)
        if (l[k]) {
            top = 32 * k + 63 - __clzll((long long)l[k]);
            break;
        }
    if (top < 0) return 0.0;
    auto bit = [&](int q) -> uint32_t {
        return q < 0 ? 0u : (uint32_t)((l[q >> 5] >> (q & 31)) & 1ull);
    };
    unsigned long long mant = 0;
    const int lo = top - 52;
    for (int q = top; q >= (lo > 0 ? lo : 0); q--) mant = (mant << 1) | bit(q);
    if (lo <= 0) return ldexp((double)mant, -149);
    const uint32_t rb = bit(lo - 1);
    uint32_t sticky = 0;
    for (int q = lo - 2; q >= 0 && !sticky; q--) sticky |= bit(q);
    if (rb && (sticky || (mant & 1ull))) mant++;
    return ldexp((double)mant, lo - 149);
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
