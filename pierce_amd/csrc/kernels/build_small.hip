// Per-object LBVH built entirely in LDS, one workgroup per object.
//
// For an object of n <= 2048 triangles stored contiguously from first_tri:
//   1. 32-bit keys: a 21-bit Morton code of the centroid on a 127^3 grid over
//      the object's own box (never all ones) above the 11-bit local index, so
//      keys are distinct and the padding key 0xFFFFFFFF sorts last;
//   2. bitonic sort: each wavefront keeps a 128-key chunk in registers
//      (two keys per lane) and runs every stride <= 64 with lane shuffles;
//      only strides >= 128 go through LDS with a block barrier (10 of the 66
//      stages for 2048 keys);
//   3. agglomerative bottom-up build (Apetrei 2014, "Fast and simple
//      agglomerative LBVH construction"): every leaf climbs; a node covering
//      sorted leaves [l, r] joins the neighbouring split with the smaller
//      key XOR (split r, or l - 1); the first child to arrive at a split
//      stops, the second merges both boxes, writes the finished node record
//      straight to HBM and climbs on.  This builds the radix tree and refits
//      it in one pass, with no per-leaf box array in LDS.
//   4. the root must sit at the segment's first node (the traversals start
//      there): it is written at index 0, and internal node 0's record moves
//      to the root's split index with the one link into it patched.
// LDS is dynamic, sized by the largest object of the launch: 4 B per key
// (next power of two) + 33 B per internal node.
#include "pz_device.h"
#include "pz_kernels_join.h"

namespace {

using namespace pzd;

constexpr int kMax = (int)pzk::kSmallMax;

__device__ __forceinline__ float3 vtx(const float* __restrict__ v, uint32_t i) {
    return make_float3(v[3ull * i], v[3ull * i + 1], v[3ull * i + 2]);
}

// Size class of a needed object: 0..2 for the LDS builder, -1 for the global pipeline.
__device__ __forceinline__ int small_class(uint32_t n) {
    return n <= pzk::kClassMax[0] ? 0 : (n <= pzk::kClassMax[1] ? 1 : (n <= pzk::kClassMax[2] ? 2 : -1));
}

__global__ void k_need_split(const uint32_t* __restrict__ need, const uint32_t* __restrict__ tri_cnt,
                             const uint32_t* __restrict__ first_tri, const uint32_t* __restrict__ last_tri,
                             int32_t nobj, uint32_t* __restrict__ big, uint32_t* __restrict__ cls,
                             uint32_t* __restrict__ totals) {
    const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    uint32_t n = 0;
    if (o < nobj) {
        n = tri_cnt[o];
        const bool contiguous = n > 0 && last_tri[o] - first_tri[o] + 1 == n;
        c = (need[o] && contiguous) ? small_class(n) : -1;
        big[o] = (need[o] && c < 0) ? 1u : 0u;
        for (int k = 0; k < pzk::kSmallClasses; ++k) cls[(uint64_t)k * (nobj + 1) + o] = (c == k) ? 1u : 0u;
    }
    // largest object per class: wave maximum first, one atomic per wave and class
    for (int k = 0; k < pzk::kSmallClasses; ++k) {
        uint32_t m = c == k ? n : 0u;
        for (int sft = 32; sft > 0; sft >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, sft, 64));
        if (__lane_id() == 0 && m > 0) atomicMax(&totals[5 + k], m);
    }
}

// loc = big ? loc_big : nbig + (objects of lower classes) + loc within the class.
__global__ void k_merge_locs(const uint32_t* __restrict__ big, const uint32_t* __restrict__ loc_big,
                             const uint32_t* __restrict__ loc_cls, int32_t nobj, uint32_t* __restrict__ loc,
                             uint32_t* __restrict__ totals) {
    int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nbig = loc_big[nobj];
    uint32_t off[pzk::kSmallClasses + 1];
    off[0] = nbig;
    for (int k = 0; k < pzk::kSmallClasses; ++k) off[k + 1] = off[k] + loc_cls[(uint64_t)k * (nobj + 1) + nobj];
    if (o == 0) {
        totals[0] = nbig;
        for (int k = 0; k < pzk::kSmallClasses; ++k) totals[2 + k] = off[k + 1] - off[k];
    }
    if (o >= nobj) return;
    uint32_t v = loc_big[o];
    if (!big[o])
        for (int k = 0; k < pzk::kSmallClasses; ++k)
            if (loc_cls[(uint64_t)k * (nobj + 1) + o + 1] != loc_cls[(uint64_t)k * (nobj + 1) + o])
                v = off[k] + loc_cls[(uint64_t)k * (nobj + 1) + o];
    loc[o] = v;
    if (o == nobj - 1) loc[nobj] = off[pzk::kSmallClasses];
}

__global__ void k_big_prefix(const uint32_t* __restrict__ seg_off, uint32_t* __restrict__ totals) {
    if (threadIdx.x == 0 && blockIdx.x == 0) totals[1] = seg_off[totals[0]];
}

template <int T>
__global__ __launch_bounds__(T) void k_build_small(
    const float* __restrict__ verts, const uint32_t* __restrict__ idx, const int32_t* __restrict__ t2o,
    const int32_t* __restrict__ obj_of_loc, const uint32_t* __restrict__ first_tri,
    const uint32_t* __restrict__ tri_cnt, const uint32_t* __restrict__ obox, const uint32_t* __restrict__ seg_off,
    uint32_t l_begin, int cap, Node* __restrict__ nodes, LeafTri* __restrict__ leaves, uint32_t* __restrict__ height_out
) {
    extern __shared__ __align__(16) unsigned char lds[];
    int cpow = 1;
    while (cpow < cap) cpow <<= 1;
    uint32_t* key = reinterpret_cast<uint32_t*>(lds);                       // cpow
    float* nbox = reinterpret_cast<float*>(key + cpow);                    // (cap - 1) x 6
    int* bound = reinterpret_cast<int*>(nbox + 6 * (cap > 1 ? cap - 1 : 1)); // cap - 1
    int16_t* cid = reinterpret_cast<int16_t*>(bound + (cap > 1 ? cap - 1 : 1)); // (cap - 1) x 2
    uint8_t* hgt = reinterpret_cast<uint8_t*>(cid + 2 * (cap > 1 ? cap - 1 : 1));   // cap - 1
    __shared__ Node rec0;       // record of internal node 0 (moved to the root's index at the end)
    __shared__ int p0_info;     // parent of node 0 << 1 | side, or -1
    __shared__ int root_idx;
    const uint32_t l = l_begin + blockIdx.x;
    const int32_t o = obj_of_loc[l];
    const int n = (int)tri_cnt[o];
    const uint32_t t0 = first_tri[o];
    const uint32_t base = seg_off[l];
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    float lo[3], sc[3];
    for (int k = 0; k < 3; ++k) {
        const float bl = ord2f(obox[6ll * o + k]), bh = ord2f(obox[6ll * o + 3 + k]);
        lo[k] = bl;
        sc[k] = bh > bl ? 127.0f / (bh - bl) : 0.0f;
    }
    int npow = 1;
    while (npow < n) npow <<= 1;
    // 1. keys: Morton code of the centroid on a 127^3 grid (never all ones), local index below
    for (int p = tid; p < npow; p += T) {
        uint32_t kv = 0xFFFFFFFFu;
        if (p < n) {
            const uint32_t t = t0 + (uint32_t)p;
            const uint32_t i0 = idx[3ull * t], i1 = idx[3ull * t + 1], i2 = idx[3ull * t + 2];
            uint32_t q[3];
            for (int k = 0; k < 3; ++k) {
                const float c = (verts[3ull * i0 + k] + verts[3ull * i1 + k] + verts[3ull * i2 + k]) * (1.0f / 3.0f);
                const float f = fminf(fmaxf((c - lo[k]) * sc[k], 0.0f), 126.0f);
                q[k] = (uint32_t)f;
            }
            uint32_t m = 0;
            for (int b = 6; b >= 0; --b)
                m = (m << 3) | (((q[0] >> b) & 1u) << 2) | (((q[1] >> b) & 1u) << 1) | ((q[2] >> b) & 1u);
            kv = (m << 11) | (uint32_t)p;
        }
        key[p] = kv;
    }
    __syncthreads();
    // 2. bitonic sort, ascending.  A wave holds a 128-key chunk in registers
    // (lane L: positions c*128 + L and c*128 + 64 + L); strides <= 64 run with
    // lane shuffles, strides >= 128 in LDS with block barriers.
    auto reg_stages = [&](int size, int smax) {
        for (int c = wave; c * 128 < npow; c += T / 64) {
            const int p0 = c * 128 + lane, p1 = p0 + 64;
            uint32_t k0 = key[p0], k1 = key[p1];
            for (int sz = (smax ? size : 2); sz <= size; sz <<= 1) {
                for (int st = min(sz >> 1, 64); st > 0; st >>= 1) {
                    if (st == 64) {
                        const bool up = (p0 & sz) == 0;
                        const uint32_t a = min(k0, k1), b = max(k0, k1);
                        k0 = up ? a : b;
                        k1 = up ? b : a;
                    } else {
                        const bool lower = (lane & st) == 0;
                        const uint32_t o0 = (uint32_t)__shfl_xor((int)k0, st, 64);
                        const uint32_t o1 = (uint32_t)__shfl_xor((int)k1, st, 64);
                        const bool up0 = (p0 & sz) == 0, up1 = (p1 & sz) == 0;
                        k0 = (lower == up0) ? min(k0, o0) : max(k0, o0);
                        k1 = (lower == up1) ? min(k1, o1) : max(k1, o1);
                    }
                }
            }
            key[p0] = k0;
            key[p1] = k1;
        }
    };
    if (npow >= 128) {
        reg_stages(128, 0);  // sizes 2 .. 128 entirely in registers
        for (int size = 256; size <= npow; size <<= 1) {
            for (int st = size >> 1; st >= 128; st >>= 1) {
                __syncthreads();
                for (int i = tid; i < npow / 2; i += T) {
                    const int a = 2 * i - (i & (st - 1)), b = a + st;
                    const bool up = (a & size) == 0;
                    const uint32_t x = key[a], y = key[b];
                    if ((x > y) == up) {
                        key[a] = y;
                        key[b] = x;
                    }
                }
            }
            __syncthreads();
            reg_stages(size, 1);  // strides 64 .. 1 of this size
        }
    } else {
        for (int size = 2; size <= npow; size <<= 1)
            for (int st = size >> 1; st > 0; st >>= 1) {
                __syncthreads();
                for (int i = tid; i < npow / 2; i += T) {
                    const int a = 2 * i - (i & (st - 1)), b = a + st;
                    const bool up = (a & size) == 0;
                    const uint32_t x = key[a], y = key[b];
                    if ((x > y) == up) {
                        key[a] = y;
                        key[b] = x;
                    }
                }
            }
    }
    for (int i = tid; i < n - 1; i += T) bound[i] = -1;
    if (tid == 0) {
        p0_info = -1;
        root_idx = 0;
    }
    __syncthreads();
    // 3. agglomerative build: every leaf climbs; at each node the first child
    // to arrive stops, the second merges both boxes, writes the node and goes on.
    uint32_t my_h = 0;
    auto leaf_box = [&](int p, float* blo, float* bhi) {
        const uint32_t t = t0 + (key[p] & 0x7FFu);
        const uint32_t i0 = idx[3ull * t], i1 = idx[3ull * t + 1], i2 = idx[3ull * t + 2];
        for (int k = 0; k < 3; ++k) {
            const float a = verts[3ull * i0 + k], b = verts[3ull * i1 + k], c = verts[3ull * i2 + k];
            blo[k] = fminf(a, fminf(b, c));
            bhi[k] = fmaxf(a, fmaxf(b, c));
        }
    };
    auto delta = [&](int i) { return key[i] ^ key[i + 1]; };  // split i between sorted keys i and i+1
    for (int p = tid; p < n; p += T) {
        const uint32_t t = t0 + (key[p] & 0x7FFu);
        const uint32_t i0 = idx[3ull * t], i1 = idx[3ull * t + 1], i2 = idx[3ull * t + 2];
        const float3 a = vtx(verts, i0), b = vtx(verts, i1), c = vtx(verts, i2);
        LeafTri L;
        L.v0 = make_float4(a.x, a.y, a.z, __uint_as_float(t));
        L.v1 = make_float4(b.x, b.y, b.z, __int_as_float(t2o[t]));
        L.v2 = make_float4(c.x, c.y, c.z, 0.0f);
        leaves[base + p] = L;
        if (n < 2) continue;
        float blo[3] = {fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z))};
        float bhi[3] = {fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z))};
        int lft = p, rgt = p;
        int me = ~p;  // leaf: ~position; internal: split index
        int h = 0;
        for (int guard = 0; guard < 64; ++guard) {  // every climb ends at the root within 64 levels
            int parent, side;
            if (lft == 0 || (rgt != n - 1 && delta(rgt) < delta(lft - 1))) {
                parent = rgt;
                side = 0;
            } else {
                parent = lft - 1;
                side = 1;
            }
            cid[2 * parent + side] = (int16_t)me;
            if (me >= 0) hgt[me] = (uint8_t)h;
            // LDS-only ordering: the node records stored to HBM need no wait here
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            const int other = atomicExch(&bound[parent], side == 0 ? lft : rgt);
            if (other < 0) break;  // first to arrive: the sibling finishes this node
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            const int sib = cid[2 * parent + (1 - side)];
            float slo[3], shi[3];
            int sh = 0;
            if (sib < 0) {
                leaf_box(~sib, slo, shi);
            } else {
                for (int k = 0; k < 3; ++k) {
                    slo[k] = nbox[6 * sib + k];
                    shi[k] = nbox[6 * sib + 3 + k];
                }
                sh = hgt[sib];
            }
            if (side == 0) rgt = other; else lft = other;
            // node record: child 0 = left (side 0), child 1 = right
            const float* L0 = side == 0 ? blo : slo; const float* H0 = side == 0 ? bhi : shi;
            const float* L1 = side == 0 ? slo : blo; const float* H1 = side == 0 ? shi : bhi;
            const int c0 = side == 0 ? me : sib, c1 = side == 0 ? sib : me;
            Node nd;
            nd.a = make_float4(L0[0], L0[1], L0[2], H0[0]);
            nd.b = make_float4(H0[1], H0[2], L1[0], L1[1]);
            nd.c = make_float4(L1[2], H1[0], H1[1], H1[2]);
            nd.d = make_int4(c0, c1, 0, 0);  // local links, rebased below
            for (int k = 0; k < 3; ++k) {
                blo[k] = fminf(blo[k], slo[k]);
                bhi[k] = fmaxf(bhi[k], shi[k]);
                nbox[6 * parent + k] = blo[k];
                nbox[6 * parent + 3 + k] = bhi[k];
            }
            h = 1 + max(h, sh);
            // rebase links: leaves ~(base + pos), internal base + index (0 and the root swap places)
            int g0 = c0 < 0 ? ~(int)(base + (uint32_t)(~c0)) : (int)(base + (uint32_t)c0);
            int g1 = c1 < 0 ? ~(int)(base + (uint32_t)(~c1)) : (int)(base + (uint32_t)c1);
            nd.d = make_int4(g0, g1, 0, 0);
            if (c0 == 0) p0_info = (parent << 1) | 0;
            if (c1 == 0) p0_info = (parent << 1) | 1;
            me = parent;
            const bool root = (lft == 0 && rgt == n - 1);
            if (parent == 0) rec0 = nd;
            else if (root) nodes[base] = nd;  // the root lives at the segment's first node
            else nodes[base + parent] = nd;
            if (root) {
                root_idx = parent;
                my_h = (uint32_t)h;
                break;
            }
        }
    }
    __syncthreads();
    // 4. node 0's record goes to the root's index (unless node 0 is the root),
    // and the link to node 0 in its parent is redirected there.
    if (tid == 0 && n >= 2) {
        const int r = root_idx;
        if (r == 0) {
            nodes[base] = rec0;
        } else {
            nodes[base + r] = rec0;
            const int par = p0_info >> 1, sd = p0_info & 1;
            // links into internal node 0 were written as base + 0: point them at base + r
            int* link = reinterpret_cast<int*>(&nodes[par == r ? base : base + par].d) + sd;
            *link = (int)(base + (uint32_t)r);
        }
    }
    if (my_h) atomicMax(height_out, my_h);
}

inline unsigned grid1(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace pzk {

void need_split(const uint32_t* need, const uint32_t* tri_cnt, const uint32_t* first_tri, const uint32_t* last_tri,
                int32_t nobj, uint32_t* big, uint32_t* cls, uint32_t* totals, hipStream_t s) {
    if (nobj > 0)
        hipLaunchKernelGGL(k_need_split, dim3(grid1((uint64_t)nobj)), dim3(256), 0, s, need, tri_cnt, first_tri,
                           last_tri, nobj, big, cls, totals);
}

void merge_locs(const uint32_t* big, const uint32_t* loc_big, const uint32_t* loc_cls, int32_t nobj, uint32_t* loc,
                uint32_t* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_merge_locs, dim3(grid1((uint64_t)std::max(nobj, 1))), dim3(256), 0, s, big, loc_big,
                       loc_cls, nobj, loc, totals);
}

void big_prefix(const uint32_t* seg_off, uint32_t* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_big_prefix, dim3(1), dim3(64), 0, s, seg_off, totals);
}

size_t build_small_lds(uint32_t cap) {
    uint32_t cpow = 1;
    while (cpow < cap) cpow <<= 1;
    return 4ull * cpow + 33ull * std::max<uint32_t>(cap, 2) + 16;
}

bool build_small(const float* verts, const uint32_t* idx, const int32_t* t2o, const int32_t* obj_of_loc,
                 const uint32_t* first_tri, const uint32_t* tri_cnt, const uint32_t* obox, const uint32_t* seg_off,
                 uint32_t l_begin, uint32_t count, uint32_t cap, int threads, Node* nodes, LeafTri* leaves,
                 uint32_t* height, const char* timer, hipStream_t s) {
    if (count == 0) return true;
    if (cap == 0 || cap > pzk::kSmallMax) return false;
    const size_t lds = build_small_lds(cap);
    KernelTimer kt_(timer ? timer : "build_small", s);
    KernelTimer kt_all_("build_small", s);
#define PZ_LAUNCH_BS(T)                                                                                          \
    hipLaunchKernelGGL(k_build_small<T>, dim3(count), dim3(T), lds, s, verts, idx, t2o, obj_of_loc, first_tri, \
                       tri_cnt, obox, seg_off, l_begin, (int)cap, nodes, leaves, height)
    switch (threads) {
        case 128: PZ_LAUNCH_BS(128); break;
        case 256: PZ_LAUNCH_BS(256); break;
        case 512: PZ_LAUNCH_BS(512); break;
        case 1024: PZ_LAUNCH_BS(1024); break;
        default: return false;
    }
#undef PZ_LAUNCH_BS
    return true;
}

}  // namespace pzk
