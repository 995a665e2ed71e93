// Linear BVH construction (T. Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012) over primitives
// sorted by 64-bit keys, built independently per segment: a segment is a run
// of sorted positions [begin, end) (one object's triangles, or the whole set);
// its internal nodes occupy slots [begin, end - 1) and its root is slot
// `begin` (or the single leaf when end - begin == 1).  Keys tie-break on the
// sorted position, so duplicate Morton codes still give a valid radix tree.
// Boxes are filled bottom-up with one counter per internal node: in LDS for
// nodes whose leaves all belong to one workgroup's block of 2,048 sorted
// positions, in HBM (agent-scope atomics) for the few that span blocks.
#include "pz_device.h"
#include "pz_kernels_bvh.h"

namespace {

using pzd::Node;

__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, int64_t i, int64_t j, int64_t b, int64_t e) {
    if (j < b || j >= e) return -1;
    uint64_t ki = keys[i], kj = keys[j];
    if (ki != kj) return __clzll((long long)(ki ^ kj));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__device__ __forceinline__ void seg_range(const uint32_t* __restrict__ seg_off, const uint64_t* __restrict__ keys,
                                          int seg_shift, uint64_t n, uint64_t i, int64_t& b, int64_t& e) {
    if (seg_off) {
        uint64_t s = keys[i] >> seg_shift;
        b = seg_off[s];
        e = seg_off[s + 1];
    } else {
        b = 0;
        e = (int64_t)n;
    }
}

__global__ void k_karras(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ seg_off,
                         int seg_shift, Node* __restrict__ nodes, uint32_t* __restrict__ parent_int,
                         uint32_t* __restrict__ parent_leaf, uint32_t* __restrict__ counters,
                         uint32_t* __restrict__ range_end) {
    uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n) return;
    int64_t b, e;
    seg_range(seg_off, keys, seg_shift, n, gi, b, e);
    const int64_t i = (int64_t)gi;
    if (i == e - 1) return;  // last position of a segment owns no internal node
    counters[i] = 0;
    const int d = (delta(keys, i, i + 1, b, e) - delta(keys, i, i - 1, b, e)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, i, i - d, b, e);
    int64_t lmax = 2;
    while (delta(keys, i, i + lmax * d, b, e) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, i, i + (l + t) * d, b, e) > dmin) l += t;
    const int64_t j = i + l * d;
    range_end[i] = (uint32_t)j;  // node i covers sorted positions [min(i, j), max(i, j)]
    const int dnode = delta(keys, i, j, b, e);
    int64_t s = 0;
    int64_t div = 2;
    while (true) {
        int64_t t = (l + div - 1) / div;
        if (delta(keys, i, i + (s + t) * d, b, e) > dnode) s += t;
        if (t <= 1) break;
        div <<= 1;
    }
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int c0, c1;
    if (lo == g) {
        c0 = ~(int)g;
        parent_leaf[g] = (uint32_t)i;
    } else {
        c0 = (int)g;
        parent_int[g] = (uint32_t)i;
    }
    if (hi == g + 1) {
        c1 = ~(int)(g + 1);
        parent_leaf[g + 1] = (uint32_t)i;
    } else {
        c1 = (int)(g + 1);
        parent_int[g + 1] = (uint32_t)i;
    }
    nodes[i].d = make_int4(c0, c1, 0, 0);
}

// Child boxes cross workgroups during the refit.  They are written and read
// with agent-scope relaxed atomics, which on gfx950 bypass the per-XCD L2
// (sc1), so no cache write-back fence is needed; the writer waits for its
// stores to complete (s_waitcnt) before incrementing the node counter, and
// the reader only loads after its increment has returned.
__device__ __forceinline__ void st_agent(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_agent(const float* p) {
    return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_child_box(Node* nd, int slot, const float* lo, const float* hi, int height) {
    float* f = reinterpret_cast<float*>(nd);
    const int o = slot == 0 ? 0 : 6;
    st_agent(f + o + 0, lo[0]);
    st_agent(f + o + 1, lo[1]);
    st_agent(f + o + 2, lo[2]);
    st_agent(f + o + 3, hi[0]);
    st_agent(f + o + 4, hi[1]);
    st_agent(f + o + 5, hi[2]);
    st_agent(f + 14 + slot, __int_as_float(height));
}
__device__ __forceinline__ void load_child_box(const Node* nd, int slot, float* lo, float* hi, int* height) {
    const float* f = reinterpret_cast<const float*>(nd);
    const int o = slot == 0 ? 0 : 6;
    lo[0] = ld_agent(f + o + 0);
    lo[1] = ld_agent(f + o + 1);
    lo[2] = ld_agent(f + o + 2);
    hi[0] = ld_agent(f + o + 3);
    hi[1] = ld_agent(f + o + 4);
    hi[2] = ld_agent(f + o + 5);
    *height = __float_as_int(ld_agent(f + 14 + slot));
}

// Child boxes of nodes whose whole leaf range lies in one workgroup's block
// never leave the CU: plain stores, a workgroup-scope fence and an LDS
// counter per node.  Only the few nodes that span blocks (the top levels of
// each segment) use the agent-scope protocol above.
// Leaves per workgroup: 2048 for large sets; 256 below 2M leaves so that a
// small set (an object index of 100k boxes) still spreads over every CU.
constexpr int kRefitBlock = 2048;
constexpr int kRefitBlockSmall = 256;
constexpr int kRefitThreads = 256;
inline int refit_block_for(uint64_t n) { return n >= (uint64_t)kRefitBlock * 1024 ? kRefitBlock : kRefitBlockSmall; }

__device__ __forceinline__ void store_child_box_local(Node* nd, int slot, const float* lo, const float* hi,
                                                      int height) {
    float* f = reinterpret_cast<float*>(nd);
    const int o = slot == 0 ? 0 : 6;
    for (int a = 0; a < 3; ++a) {
        f[o + a] = lo[a];
        f[o + 3 + a] = hi[a];
    }
    f[14 + slot] = __int_as_float(height);
}
__device__ __forceinline__ void load_child_box_local(const Node* nd, int slot, float* lo, float* hi, int* height) {
    const float* f = reinterpret_cast<const float*>(nd);
    const int o = slot == 0 ? 0 : 6;
    for (int a = 0; a < 3; ++a) {
        lo[a] = f[o + a];
        hi[a] = f[o + 3 + a];
    }
    *height = __float_as_int(f[14 + slot]);
}

template <int BLK, class BoxFn>
__device__ void refit_blocked(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ seg_off,
                              int seg_shift, Node* nodes, const uint32_t* __restrict__ parent_int,
                              const uint32_t* __restrict__ parent_leaf, uint32_t* counters,
                              const uint32_t* __restrict__ range_end, uint32_t* max_depth, BoxFn boxfn) {
    __shared__ uint32_t lcnt[BLK];
    const uint64_t B = (uint64_t)blockIdx.x * BLK;
    const uint64_t E = B + BLK < n ? B + BLK : n;
    for (int i = threadIdx.x; i < BLK; i += kRefitThreads) lcnt[i] = 0u;
    __syncthreads();
    for (uint64_t p = B + threadIdx.x; p < E; p += kRefitThreads) {
        int64_t b, e;
        seg_range(seg_off, keys, seg_shift, n, p, b, e);
        float lo[3], hi[3];
        boxfn(p, lo, hi);
        if (e - b == 1) continue;
        uint32_t node = parent_leaf[p];
        int child = ~(int)p;
        int height = 0;
        while (true) {
            Node* nd = nodes + node;
            const int slot = (nd->d.x == child) ? 0 : 1;
            const uint64_t j = range_end[node];
            const uint64_t rlo = node < j ? node : j, rhi = node < j ? j : node;
            float slo[3], shi[3];
            int sh;
            if (rlo >= B && rhi < E) {  // both subtrees were finished by this workgroup
                store_child_box_local(nd, slot, lo, hi, height);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                const uint32_t prev = atomicAdd(&lcnt[node - B], 1u);
                if (prev == 0) break;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                load_child_box_local(nd, 1 - slot, slo, shi, &sh);
            } else {
                store_child_box(nd, slot, lo, hi, height);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                const uint32_t prev =
                    __hip_atomic_fetch_add(&counters[node], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (prev == 0) break;
                load_child_box(nd, 1 - slot, slo, shi, &sh);
            }
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], slo[a]);
                hi[a] = fmaxf(hi[a], shi[a]);
            }
            height = 1 + (height > sh ? height : sh);
            if ((int64_t)node == b) {  // segment root
                atomicMax(max_depth, (uint32_t)height);
                break;
            }
            child = (int)node;
            node = parent_int[node];
        }
    }
}

struct LeafBoxFn {  // leaf records already in sorted order: coalesced reads
    const pzd::LeafTri* leaves;
    __device__ void operator()(uint64_t p, float* lo, float* hi) const {
        const pzd::LeafTri L = leaves[p];
        lo[0] = fminf(L.v0.x, fminf(L.v1.x, L.v2.x));
        lo[1] = fminf(L.v0.y, fminf(L.v1.y, L.v2.y));
        lo[2] = fminf(L.v0.z, fminf(L.v1.z, L.v2.z));
        hi[0] = fmaxf(L.v0.x, fmaxf(L.v1.x, L.v2.x));
        hi[1] = fmaxf(L.v0.y, fmaxf(L.v1.y, L.v2.y));
        hi[2] = fmaxf(L.v0.z, fmaxf(L.v1.z, L.v2.z));
    }
};
struct OrdBoxFn {  // boxes stored as ordered uints: 6 per primitive
    const uint32_t* obox;
    const uint32_t* order;
    __device__ void operator()(uint64_t p, float* lo, float* hi) const {
        const uint32_t* b = obox + 6ull * order[p];
        for (int a = 0; a < 3; ++a) {
            lo[a] = pzd::ord2f(b[a]);
            hi[a] = pzd::ord2f(b[3 + a]);
        }
    }
};

template <int BLK>
__global__ __launch_bounds__(kRefitThreads) void k_refit_leaves(const uint64_t* keys, uint64_t n,
                                                                 const uint32_t* seg_off, int seg_shift, Node* nodes,
                                                                 const uint32_t* parent_int,
                                                                 const uint32_t* parent_leaf, uint32_t* counters,
                                                                 const uint32_t* range_end, uint32_t* max_depth,
                                                                 LeafBoxFn fn) {
    refit_blocked<BLK>(keys, n, seg_off, seg_shift, nodes, parent_int, parent_leaf, counters, range_end, max_depth,
                       fn);
}
template <int BLK>
__global__ __launch_bounds__(kRefitThreads) void k_refit_boxes(const uint64_t* keys, uint64_t n,
                                                                const uint32_t* seg_off, int seg_shift, Node* nodes,
                                                                const uint32_t* parent_int,
                                                                const uint32_t* parent_leaf, uint32_t* counters,
                                                                const uint32_t* range_end, uint32_t* max_depth,
                                                                OrdBoxFn fn) {
    refit_blocked<BLK>(keys, n, seg_off, seg_shift, nodes, parent_int, parent_leaf, counters, range_end, max_depth,
                       fn);
}

__global__ void k_gather_leaf_tris(const uint32_t* __restrict__ order, uint64_t n, const float* __restrict__ verts,
                                   const uint32_t* __restrict__ idx, const int32_t* __restrict__ t2o,
                                   pzd::LeafTri* __restrict__ out) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t t = order[p];
    const float* v0 = verts + 3ull * idx[3ull * t];
    const float* v1 = verts + 3ull * idx[3ull * t + 1];
    const float* v2 = verts + 3ull * idx[3ull * t + 2];
    pzd::LeafTri L;
    L.v0 = make_float4(v0[0], v0[1], v0[2], __uint_as_float(t));
    L.v1 = make_float4(v1[0], v1[1], v1[2], __int_as_float(t2o[t]));
    L.v2 = make_float4(v2[0], v2[1], v2[2], 0.0f);
    out[p] = L;
}

inline unsigned grid1(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace pzk {

void lbvh_karras(const uint64_t* sorted_keys, uint64_t n, const uint32_t* seg_off, int seg_shift, pzd::Node* nodes,
                 uint32_t* parent_int, uint32_t* parent_leaf, uint32_t* counters, uint32_t* range_end,
                 hipStream_t s) {
    if (n == 0) return;
    KernelTimer kt_("karras", s);
    hipLaunchKernelGGL(k_karras, dim3(grid1(n)), dim3(256), 0, s, sorted_keys, n, seg_off, seg_shift, nodes, parent_int,
                       parent_leaf, counters, range_end);
}

void lbvh_refit_leaves(const uint64_t* sorted_keys, uint64_t n, const uint32_t* seg_off, int seg_shift,
                       const pzd::LeafTri* leaves, pzd::Node* nodes, const uint32_t* parent_int,
                       const uint32_t* parent_leaf, uint32_t* counters, const uint32_t* range_end,
                       uint32_t* max_depth, hipStream_t s) {
    if (n == 0) return;
    LeafBoxFn fn{leaves};
    KernelTimer kt_("refit", s);
    if (refit_block_for(n) == kRefitBlock)
        hipLaunchKernelGGL(k_refit_leaves<kRefitBlock>, dim3(grid1(n, kRefitBlock)), dim3(kRefitThreads), 0, s,
                           sorted_keys, n, seg_off, seg_shift, nodes, parent_int, parent_leaf, counters, range_end,
                           max_depth, fn);
    else
        hipLaunchKernelGGL(k_refit_leaves<kRefitBlockSmall>, dim3(grid1(n, kRefitBlockSmall)), dim3(kRefitThreads),
                           0, s, sorted_keys, n, seg_off, seg_shift, nodes, parent_int, parent_leaf, counters,
                           range_end, max_depth, fn);
}

void lbvh_refit_ord_boxes(const uint64_t* sorted_keys, const uint32_t* order, uint64_t n, const uint32_t* seg_off,
                          int seg_shift, const uint32_t* obox, pzd::Node* nodes, const uint32_t* parent_int,
                          const uint32_t* parent_leaf, uint32_t* counters, const uint32_t* range_end,
                          uint32_t* max_depth, hipStream_t s) {
    if (n == 0) return;
    OrdBoxFn fn{obox, order};
    if (refit_block_for(n) == kRefitBlock)
        hipLaunchKernelGGL(k_refit_boxes<kRefitBlock>, dim3(grid1(n, kRefitBlock)), dim3(kRefitThreads), 0, s,
                           sorted_keys, n, seg_off, seg_shift, nodes, parent_int, parent_leaf, counters, range_end,
                           max_depth, fn);
    else
        hipLaunchKernelGGL(k_refit_boxes<kRefitBlockSmall>, dim3(grid1(n, kRefitBlockSmall)), dim3(kRefitThreads), 0,
                           s, sorted_keys, n, seg_off, seg_shift, nodes, parent_int, parent_leaf, counters, range_end,
                           max_depth, fn);
}

void gather_leaf_tris(const uint32_t* order, uint64_t n, const float* verts, const uint32_t* idx, const int32_t* t2o,
                      pzd::LeafTri* out, hipStream_t s) {
    if (n == 0) return;
    KernelTimer kt_("gather_leaves", s);
    hipLaunchKernelGGL(k_gather_leaf_tris, dim3(grid1(n)), dim3(256), 0, s, order, n, verts, idx, t2o, out);
}

}  // namespace pzk
