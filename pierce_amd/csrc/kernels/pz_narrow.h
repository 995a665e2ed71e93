// Device helpers shared by the narrow-phase kernels (join.hip, sieve.hip):
// triangle corners, per-lane counters and the triangle-pair test built on the
// exact predicate chain of pz_predicates.h.
#pragma once

#include "pz_device.h"
#include "pz_kernels_join.h"
#include "pz_predicates.h"

namespace {

using namespace pzd;

__device__ __forceinline__ float3 vtx(const float* __restrict__ v, uint32_t i) {
    return make_float3(v[3ull * i], v[3ull * i + 1], v[3ull * i + 2]);
}

struct LaneStats {
    uint32_t tests, f64, exact, mis, traced;
    uint32_t probes = 0, contended = 0;  // hash-set slots probed / found holding another key
};

__device__ __forceinline__ void flush_stats(LaneStats& st, unsigned long long* stats) {
    uint32_t v[7] = {st.tests, st.f64, st.exact, st.mis, st.traced, st.probes, st.contended};
    for (int k = 0; k < 7; ++k) {
        uint32_t s = wave_sum(v[k]);
        if (__lane_id() == 0 && s) atomicAdd(&stats[k], (unsigned long long)s);
    }
}

__device__ __forceinline__ bool box_box(const float* alo, const float* ahi, const float* blo, const float* bhi) {
    return !(alo[0] > bhi[0] || ahi[0] < blo[0] || alo[1] > bhi[1] || ahi[1] < blo[1] || alo[2] > bhi[2] ||
             ahi[2] < blo[2]);
}

template <bool F32>
__device__ __forceinline__ bool tri_pair_test(const LeafTri& X, const LeafTri& Y, int dir_mask, bool allow_f32,
                                              LaneStats& st) {
    const float3 x0 = make_float3(X.v0.x, X.v0.y, X.v0.z), x1 = make_float3(X.v1.x, X.v1.y, X.v1.z),
                 x2 = make_float3(X.v2.x, X.v2.y, X.v2.z);
    const float3 y0 = make_float3(Y.v0.x, Y.v0.y, Y.v0.z), y1 = make_float3(Y.v1.x, Y.v1.y, Y.v1.z),
                 y2 = make_float3(Y.v2.x, Y.v2.y, Y.v2.z);
    Esc esc{0};
    bool hit = false;
    int tests = 0;
    if (dir_mask & 1) {
        ++tests;
        hit = seg_tri<F32>(x0, x1, y0, y1, y2, allow_f32, esc);
        if (!hit) { ++tests; hit = seg_tri<F32>(x1, x2, y0, y1, y2, allow_f32, esc); }
        if (!hit) { ++tests; hit = seg_tri<F32>(x2, x0, y0, y1, y2, allow_f32, esc); }
    }
    if (!hit && (dir_mask & 2)) {
        ++tests;
        hit = seg_tri<F32>(y0, y1, x0, x1, x2, allow_f32, esc);
        if (!hit) { ++tests; hit = seg_tri<F32>(y1, y2, x0, x1, x2, allow_f32, esc); }
        if (!hit) { ++tests; hit = seg_tri<F32>(y2, y0, x0, x1, x2, allow_f32, esc); }
    }
    st.tests += tests;
    if (esc.level >= 1) {
        if (esc.level == 1)
            ++st.f64;
        else
            ++st.exact;
        if (F32) {
            bool naive = false;
            if (dir_mask & 1)
                naive = seg_tri_naive(x0, x1, y0, y1, y2) || seg_tri_naive(x1, x2, y0, y1, y2) ||
                        seg_tri_naive(x2, x0, y0, y1, y2);
            if (!naive && (dir_mask & 2))
                naive = seg_tri_naive(y0, y1, x0, x1, x2) || seg_tri_naive(y1, y2, x0, x1, x2) ||
                        seg_tri_naive(y2, y0, x0, x1, x2);
            if (naive != hit) ++st.mis;
        }
    }
    return hit;
}

// tri_pair_test decided by the fp32 filter alone: 1 = hit, 0 = miss,
// -1 = undecided (some edge unproven and none proven to hit).
__device__ __forceinline__ int tri_pair_f32(const LeafTri& X, const LeafTri& Y, int dir_mask, LaneStats& st) {
    const float3 x0 = make_float3(X.v0.x, X.v0.y, X.v0.z), x1 = make_float3(X.v1.x, X.v1.y, X.v1.z),
                 x2 = make_float3(X.v2.x, X.v2.y, X.v2.z);
    const float3 y0 = make_float3(Y.v0.x, Y.v0.y, Y.v0.z), y1 = make_float3(Y.v1.x, Y.v1.y, Y.v1.z),
                 y2 = make_float3(Y.v2.x, Y.v2.y, Y.v2.z);
    bool open = false;
    int r;
    if (dir_mask & 1) {
        ++st.tests;
        if ((r = seg_tri_f32(x0, x1, y0, y1, y2)) > 0) return 1;
        open |= r < 0;
        ++st.tests;
        if ((r = seg_tri_f32(x1, x2, y0, y1, y2)) > 0) return 1;
        open |= r < 0;
        ++st.tests;
        if ((r = seg_tri_f32(x2, x0, y0, y1, y2)) > 0) return 1;
        open |= r < 0;
    }
    if (dir_mask & 2) {
        ++st.tests;
        if ((r = seg_tri_f32(y0, y1, x0, x1, x2)) > 0) return 1;
        open |= r < 0;
        ++st.tests;
        if ((r = seg_tri_f32(y1, y2, x0, x1, x2)) > 0) return 1;
        open |= r < 0;
        ++st.tests;
        if ((r = seg_tri_f32(y2, y0, x0, x1, x2)) > 0) return 1;
        open |= r < 0;
    }
    return open ? -1 : 0;
}

// found[] is written by other waves while a walk runs: read it with a relaxed
// device-scope atomic load, so every check fetches the current value.
__device__ __forceinline__ bool found_now(const uint32_t* found, uint32_t p) {
    return __hip_atomic_load(found + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

}  // namespace
