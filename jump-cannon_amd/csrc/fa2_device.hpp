// Device helpers shared by the FA2+BH kernels (single graph and batched).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace jc {

constexpr int kLevels = 21;            // octree key depth (3 bits per level)
constexpr uint32_t kLeaf = 0x80000000u;  // node_info flag of a leaf
constexpr uint32_t kEnd = 0xFFFFFFFFu;   // end of the rope

// One octree node, 32 bytes: one aligned record per visit.
//   a = centre of mass xyz, mass
//   b.x = internal: bits of (2*half)^2; leaf: 0x80000000 | body
//   b.y = skip rope (first node after the subtree, kEnd at the end)
struct __align__(32) Node {
    float4 a;
    uint4 b;
};

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(u);
}
__device__ __forceinline__ bool finitef(float x) { return __builtin_isfinite(x); }
__device__ __forceinline__ float len3(float x, float y, float z) {
    return sqrtf(x * x + y * y + z * z);
}

// Root cube of fa2_bh.rs:768-799 from the reduced box (order-preserving u32).
__device__ __forceinline__ void compute_root(const uint32_t *bbmin, const uint32_t *bbmax, float *r4) {
    float mn[3], mx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        mn[k] = ord2f(bbmin[k]);
        mx[k] = ord2f(bbmax[k]);
    }
    if (!finitef(mn[0])) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            mn[k] = -1.0f;
            mx[k] = 1.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) r4[k] = 0.5f * (mn[k] + mx[k]);
    float half = fmaxf(fmaxf(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]) * 0.5f;
    if (!finitef(half) || half <= 0.0f) half = 1.0f;
    r4[3] = half * 1.05f;
}

__device__ __forceinline__ int lcp_digits(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    if (!x) return kLevels;
    return (__clzll((long long)x) - 1) / 3;
}

// Last sorted index j >= i whose first d octal digits equal those of key k.
__device__ __forceinline__ uint32_t cell_end(const uint64_t *keys, uint32_t i, uint32_t nv,
                                             uint64_t k, int d) {
    if (d == 0) return nv - 1;
    const int sh = 3 * (kLevels - d);
    const uint64_t pre = k >> sh;
    uint32_t lo = i, step = 1, hi;
    for (;;) {
        uint32_t probe = lo + step;
        if (probe >= nv || (keys[probe] >> sh) != pre) {
            hi = probe < nv ? probe : nv;
            break;
        }
        lo = probe;
        step <<= 1;
    }
    while (hi - lo > 1) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if ((keys[mid] >> sh) == pre) lo = mid;
        else hi = mid;
    }
    return lo;
}


// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  This bijection gives XCD x the contiguous run
// of logical blocks [x*G/8, (x+1)*G/8): with bodies in Morton order, each L2
// then holds one region of the tree instead of all of it.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G) {
    constexpr uint32_t kXcd = 8;
    const uint32_t q = G / kXcd, r = G % kXcd;
    const uint32_t x = b % kXcd, slot = b / kXcd;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// Octant path of p through the reference cell chain for kLevels levels: digit
// (x>=cx) | (y>=cy)<<1 | (z>=cz)<<2, child half = half*0.5, child centre =
// centre +- child half (fa2_bh.rs:841-861).
__device__ __forceinline__ uint64_t octant_key(const float *r4, float px, float py, float pz) {
    float cx = r4[0], cy = r4[1], cz = r4[2], h = r4[3];
    uint64_t key = 0;
#pragma unroll
    for (int l = 0; l < kLevels; l++) {
        const uint32_t o = (px >= cx ? 1u : 0u) | (py >= cy ? 2u : 0u) | (pz >= cz ? 4u : 0u);
        key = (key << 3) | o;
        h = h * 0.5f;
        cx = cx + ((o & 1u) ? h : -h);
        cy = cy + ((o & 2u) ? h : -h);
        cz = cz + ((o & 4u) ? h : -h);
    }
    return key;
}

}  // namespace jc
