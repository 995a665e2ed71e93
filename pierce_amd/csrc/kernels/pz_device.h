// Device-side building blocks shared by the gfx950 kernels: node layout,
// ordered-float atomics, Morton codes, wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PZ_WAVE 64

namespace pzd {

// ------------------------------------------------------------------ layout
// BVH2 node: both children's boxes live in the parent so one 64-byte load
// decides both descents.  child >= 0: internal node index; child < 0: leaf
// ~child (position in the sorted primitive order).
struct __align__(16) Node {
    float4 a;  // c0.lo.x c0.lo.y c0.lo.z c0.hi.x
    float4 b;  // c0.hi.y c0.hi.z c1.lo.x c1.lo.y
    float4 c;  // c1.lo.z c1.hi.x c1.hi.y c1.hi.z
    int4 d;    // child0 child1 - -
};

// Node of a per-object BVH (the two-level join's index): its own box and its
// two child links (as in Node), 32 bytes.  A walk tests a node's box when it
// reaches the node, so a builder writes each record from the node's own box
// alone, with no leaf record read back.
struct __align__(16) BNode {
    float4 a;  // lo.x lo.y lo.z hi.x
    float4 b;  // hi.y hi.z child0 (bits) child1 (bits)
};
__device__ __forceinline__ int bnode_c0(const BNode& n) { return __float_as_int(n.b.z); }
__device__ __forceinline__ int bnode_c1(const BNode& n) { return __float_as_int(n.b.w); }

// Leaf triangle in sorted order: three vertices, w of v0 = original triangle
// index (bit pattern), w of v1 = object id (bit pattern).
struct __align__(16) LeafTri {
    float4 v0, v1, v2;
};

// Edge in per-object order: start.xyz + object id (bits), end.xyz + pad.
struct __align__(16) EdgeRec {
    float4 s, e;
};

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ bool boxes_meet(const float* lo, const float* hi, float lx, float ly, float lz, float hx,
                                           float hy, float hz) {
    return !(lx > hi[0] || hx < lo[0] || ly > hi[1] || hy < lo[1] || lz > hi[2] || hz < lo[2]);
}

// ----------------------------------------------------- ordered float atomics
// Monotone map float -> uint32 so that unsigned order == float order.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
__host__ __device__ inline uint32_t ord_pos_inf_min() { return 0xFFFFFFFFu; }  // identity for min
__host__ __device__ inline uint32_t ord_neg_inf_max() { return 0u; }          // identity for max

// ------------------------------------------------------------ Morton codes
__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
// 30-bit Morton code of p inside [lo, lo + ext] (ext may be 0 on an axis).
__device__ __forceinline__ uint32_t morton30(float px, float py, float pz, float lox, float loy, float loz,
                                             float inv_x, float inv_y, float inv_z) {
    float fx = fminf(fmaxf((px - lox) * inv_x, 0.0f), 1.0f);
    float fy = fminf(fmaxf((py - loy) * inv_y, 0.0f), 1.0f);
    float fz = fminf(fmaxf((pz - loz) * inv_z, 0.0f), 1.0f);
    uint32_t x = min((uint32_t)(fx * 1024.0f), 1023u);
    uint32_t y = min((uint32_t)(fy * 1024.0f), 1023u);
    uint32_t z = min((uint32_t)(fz * 1024.0f), 1023u);
    return (expand_bits10(x) << 2) | (expand_bits10(y) << 1) | expand_bits10(z);
}

// ------------------------------------------------------------- wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Exclusive prefix count of set predicates among lower lanes.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint64_t pack_pair(int32_t a, int32_t b) {
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

}  // namespace pzd
