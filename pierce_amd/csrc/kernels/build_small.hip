// Per-object LBVH built entirely in LDS, one workgroup per object.
//
// For an object of n <= 2048 triangles stored contiguously from first_tri:
//   1. one gather: each thread loads its local triangles q = tid + k * T
//      (the first 2048 / T of them stay in registers for the whole build) and
//      writes a 32-bit key: a 21-bit Morton code of the centroid on a 127^3
//      grid over the object's own box (never all ones) above the 11-bit local
//      index, so keys are distinct and the padding key 0xFFFFFFFF sorts last;
//   2. bitonic sort: a wavefront holds 256 keys in registers (four per lane);
//      strides 128 and 64 compare registers, strides 32 .. 4 swap lanes with
//      ds_bpermute and strides 2 and 1 with DPP quad permutes; only strides
//      >= 256 go through LDS with a block barrier (6 of the 66 stages for
//      2048 keys);
//   3. leaf records in sorted order, staged through LDS in two halves so the
//      48-byte records are written to HBM coalesced;
//   4. agglomerative bottom-up build (Apetrei 2014, "Fast and simple
//      agglomerative LBVH construction"), LDS only: each thread climbs from
//      the sorted positions of its own triangles.  A node covering sorted
//      leaves [l, r] joins the neighbouring split with the smaller key XOR
//      (split r, or l - 1).  The first child to arrive parks and stops; the
//      exchange on bound[split] hands its far range end, id and height to
//      the second, which merges both boxes and climbs on.  Boxes live in one
//      LDS array: leaf p's box at p until node p completes, node i's box at
//      i afterwards (every node containing leaf p completes after leaf p's
//      parent, which consumed it);
//   5. node records written in index order from LDS (coalesced); a leaf
//      child's box is re-read from its just-written leaf record.  The root
//      must sit at the segment's first node (the traversals start there): it
//      is written at index 0 and internal node 0's record at the root's
//      index, with the one link into node 0 redirected.
// LDS is dynamic, sized by the largest object of the launch: 4 B per key
// (next power of two, at least 2) + 34 B per triangle + 4 B.
#include <type_traits>

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

// Value of lane (lane ^ S) within the wavefront.
template <int S>
__device__ __forceinline__ uint32_t lane_swap(uint32_t v) {
    if constexpr (S == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (S == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((__lane_id() ^ S) << 2), (int)v);
    }
}

template <int T>
__global__ __launch_bounds__(T) void k_build_small(
    const float* __restrict__ verts, const uint32_t* __restrict__ idx, const int32_t* __restrict__ obj_of_loc,
    const uint32_t* __restrict__ first_tri, const uint32_t* __restrict__ tri_cnt, const uint32_t* __restrict__ obox, const uint32_t* __restrict__ seg_off,
    uint32_t l_begin, int cap, Node* __restrict__ nodes, LeafTri* __restrict__ leaves, uint32_t* __restrict__ height_out) {
    constexpr int PR = T >= 512 ? kMax / T : 4;  // local triangles per thread held in registers
    extern __shared__ __align__(16) unsigned char lds[];
    int cpow = 2;
    while (cpow < cap) cpow <<= 1;
    uint32_t* key = reinterpret_cast<uint32_t*>(lds);        // cpow
    float* box = reinterpret_cast<float*>(key + cpow);       // 6 x cap: leaf p's box at p, then node i's at i
    int* bound = reinterpret_cast<int*>(box + 6 * cap);      // cap
    int16_t* cid = reinterpret_cast<int16_t*>(bound + cap);  // 2 x cap + 2: children of node i
    uint16_t* inv = reinterpret_cast<uint16_t*>(cid + 2 * cap + 2);  // cap: local triangle -> sorted position
    float* stage = box;  // before the climb: 9 floats per triangle of one half (36 x ceil(cap / 2) <= 32 x cap + 4 B)
    __shared__ int root_idx;
    __shared__ uint32_t wg_h;
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
    auto morton = [&](const float3& a, const float3& b, const float3& c, int q) {
        const float cx[3] = {(a.x + b.x + c.x) * (1.0f / 3.0f), (a.y + b.y + c.y) * (1.0f / 3.0f),
                             (a.z + b.z + c.z) * (1.0f / 3.0f)};
        uint32_t qq[3];
        for (int k = 0; k < 3; ++k) qq[k] = (uint32_t)fminf(fmaxf((cx[k] - lo[k]) * sc[k], 0.0f), 126.0f);
        uint32_t m = 0;
        for (int b2 = 6; b2 >= 0; --b2)
            m = (m << 3) | (((qq[0] >> b2) & 1u) << 2) | (((qq[1] >> b2) & 1u) << 1) | ((qq[2] >> b2) & 1u);
        return (m << 11) | (uint32_t)q;
    };
    auto tri = [&](int q, float3& a, float3& b, float3& c) {
        const uint32_t t = t0 + (uint32_t)q;
        const uint32_t i0 = idx[3ull * t], i1 = idx[3ull * t + 1], i2 = idx[3ull * t + 2];
        a = vtx(verts, i0);
        b = vtx(verts, i1);
        c = vtx(verts, i2);
    };
    // 1. one gather; local triangle q = tid + k * T, the first PR in registers
    float3 ra[PR], rb[PR], rc[PR];
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int q = tid + k * T;
        if (q < n) tri(q, ra[k], rb[k], rc[k]);
    }
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int q = tid + k * T;
        if (q < npow) key[q] = q < n ? morton(ra[k], rb[k], rc[k], q) : 0xFFFFFFFFu;
    }
    for (int q = tid + PR * T; q < npow; q += T) {
        uint32_t kv = 0xFFFFFFFFu;
        if (q < n) {
            float3 a, b, c;
            tri(q, a, b, c);
            kv = morton(a, b, c, q);
        }
        key[q] = kv;
    }
    __syncthreads();
    // 2. bitonic sort, ascending
    auto cx = [](uint32_t& x, uint32_t& y, bool up) {
        const uint32_t a = min(x, y), b = max(x, y);
        x = up ? a : b;
        y = up ? b : a;
    };
    auto lds_stage = [&](int size, int st) {
        for (int i = tid; i < npow / 2; i += T) {
            const int a = 2 * i - (i & (st - 1)), b = a + st;
            const bool up = (a & size) == 0;
            const uint32_t x = key[a], y = key[b];
            if ((x > y) == up) {
                key[a] = y;
                key[b] = x;
            }
        }
    };
    if (npow >= 256) {
        // a wave holds 256 keys: lane L has positions c*256 + j*64 + L, j = 0..3;
        // strides 128 and 64 compare registers, strides <= 32 swap lanes
        auto reg4 = [&](int s_lo, int s_hi) {
            for (int c = wave; c * 256 < npow; c += T / 64) {
                const int pb = c * 256 + lane;
                uint32_t k0 = key[pb], k1 = key[pb + 64], k2 = key[pb + 128], k3 = key[pb + 192];
                for (int sz = s_lo; sz <= s_hi; sz <<= 1) {
                    if (sz >= 256) {
                        cx(k0, k2, (pb & sz) == 0);
                        cx(k1, k3, ((pb + 64) & sz) == 0);
                    }
                    if (sz >= 128) {
                        cx(k0, k1, (pb & sz) == 0);
                        cx(k2, k3, ((pb + 128) & sz) == 0);
                    }
                    const bool d0 = (pb & sz) == 0, d1 = ((pb + 64) & sz) == 0;
                    const bool d2 = ((pb + 128) & sz) == 0, d3 = ((pb + 192) & sz) == 0;
                    auto stage = [&](auto stc) {
                        constexpr int st = decltype(stc)::value;
                        if (2 * st > sz) return;
                        const bool lower = (lane & st) == 0;
                        const uint32_t o0 = lane_swap<st>(k0), o1 = lane_swap<st>(k1);
                        const uint32_t o2 = lane_swap<st>(k2), o3 = lane_swap<st>(k3);
                        k0 = (lower == d0) ? min(k0, o0) : max(k0, o0);
                        k1 = (lower == d1) ? min(k1, o1) : max(k1, o1);
                        k2 = (lower == d2) ? min(k2, o2) : max(k2, o2);
                        k3 = (lower == d3) ? min(k3, o3) : max(k3, o3);
                    };
                    stage(std::integral_constant<int, 32>{});
                    stage(std::integral_constant<int, 16>{});
                    stage(std::integral_constant<int, 8>{});
                    stage(std::integral_constant<int, 4>{});
                    stage(std::integral_constant<int, 2>{});
                    stage(std::integral_constant<int, 1>{});
                }
                key[pb] = k0;
                key[pb + 64] = k1;
                key[pb + 128] = k2;
                key[pb + 192] = k3;
            }
        };
        reg4(2, 256);
        for (int size = 512; size <= npow; size <<= 1) {
            for (int st = size >> 1; st >= 256; st >>= 1) {
                __syncthreads();
                lds_stage(size, st);
            }
            __syncthreads();
            reg4(size, size);
        }
    } else {
        for (int size = 2; size <= npow; size <<= 1)
            for (int st = size >> 1; st > 0; st >>= 1) {
                __syncthreads();
                lds_stage(size, st);
            }
    }
    __syncthreads();
    // 3. sorted position of every local triangle
    for (int p = tid; p < n; p += T) inv[key[p] & 0x7FFu] = (uint16_t)p;
    if (tid == 0) {
        root_idx = 0;
        wg_h = 0;
    }
    __syncthreads();
    // 3b. leaf records in sorted order, staged through LDS in two halves
    const int half = (n + 1) >> 1;
    for (int r = 0; r < 2; ++r) {
        const int h0 = r * half, h1 = min(n, h0 + half);
        auto put = [&](int q, const float3& a, const float3& b, const float3& c) {
            const int p = (int)inv[q];
            if (p < h0 || p >= h1) return;
            float* st = stage + 9 * (p - h0);
            st[0] = a.x; st[1] = a.y; st[2] = a.z;
            st[3] = b.x; st[4] = b.y; st[5] = b.z;
            st[6] = c.x; st[7] = c.y; st[8] = c.z;
        };
#pragma unroll
        for (int k = 0; k < PR; ++k) {
            const int q = tid + k * T;
            if (q < n) put(q, ra[k], rb[k], rc[k]);
        }
        for (int q = tid + PR * T; q < n; q += T) {
            float3 a, b, c;
            tri(q, a, b, c);
            put(q, a, b, c);
        }
        __syncthreads();
        for (int p = h0 + tid; p < h1; p += T) {
            const float* st = stage + 9 * (p - h0);
            LeafTri L;
            L.v0 = make_float4(st[0], st[1], st[2], __uint_as_float(t0 + (key[p] & 0x7FFu)));
            L.v1 = make_float4(st[3], st[4], st[5], __int_as_float(o));
            L.v2 = make_float4(st[6], st[7], st[8], 0.0f);
            leaves[base + p] = L;
        }
        __syncthreads();
    }
    for (int i = tid; i < n - 1; i += T) bound[i] = -1;
    __syncthreads();
    // 4. climb: a leaf's box sits at its position p; a finished node's box
    // moves to its own index i (leaf i's box is consumed by then, as every
    // node containing leaf i completes after leaf i's parent).
    auto climb = [&](int q, const float3& a, const float3& b, const float3& c) {
        if (n < 2) return;
        const int p = min((int)inv[q], n - 1);
        float blo[3] = {fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z))};
        float bhi[3] = {fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z))};
        {
            float* sb = box + 6 * p;
            for (int k = 0; k < 3; ++k) {
                sb[k] = blo[k];
                sb[3 + k] = bhi[k];
            }
        }
        int lft = p, rgt = p, me = ~p, h = 0;
        const uint32_t kp = key[p];
        uint32_t dL = p > 0 ? key[p - 1] ^ kp : 0u, dR = p < n - 1 ? kp ^ key[p + 1] : 0u;
        for (int guard = 0; guard < 64; ++guard) {
            const bool s0 = lft == 0 || (rgt != n - 1 && dR < dL);
            const int parent = s0 ? rgt : lft - 1;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            const int other = atomicExch(&bound[parent], (s0 ? lft : rgt) | ((me + 2048) << 11) | (h << 23));
            if (other < 0) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            const int oend = other & 0x7FF, sib = ((other >> 11) & 0xFFF) - 2048, sh = other >> 23;
            const float* so = box + 6 * (sib >= 0 ? sib : ~sib);
            const int ka = min(max(s0 ? oend : oend - 1, 0), n - 2);
            const uint32_t kx = key[ka], ky = key[ka + 1];
            for (int k = 0; k < 3; ++k) {
                blo[k] = fminf(blo[k], so[k]);
                bhi[k] = fmaxf(bhi[k], so[3 + k]);
            }
            if (s0) {
                rgt = oend;
                dR = kx ^ ky;
            } else {
                lft = oend;
                dL = kx ^ ky;
            }
            cid[2 * parent] = (int16_t)(s0 ? me : sib);
            cid[2 * parent + 1] = (int16_t)(s0 ? sib : me);
            float* sb = box + 6 * parent;
            for (int k = 0; k < 3; ++k) {
                sb[k] = blo[k];
                sb[3 + k] = bhi[k];
            }
            h = 1 + max(h, sh);
            me = parent;
            if (lft == 0 && rgt == n - 1) {
                root_idx = parent;
                wg_h = (uint32_t)h;
                break;
            }
        }
    };
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int q = tid + k * T;
        if (q < n) climb(q, ra[k], rb[k], rc[k]);
    }
    for (int q = tid + PR * T; q < n; q += T) {
        float3 a, b, c;
        tri(q, a, b, c);
        climb(q, a, b, c);
    }
    __syncthreads();
    // 5. node records in index order (root at 0, node 0 at the root's index)
    const int rt = root_idx;
    auto child_box = [&](int c, float* bx) {
        if (c >= 0) {
            for (int k = 0; k < 6; ++k) bx[k] = box[6 * c + k];
        } else {
            const LeafTri& L = leaves[base + (uint32_t)(~c)];
            const float4 v0 = L.v0, v1 = L.v1, v2 = L.v2;
            bx[0] = fminf(v0.x, fminf(v1.x, v2.x));
            bx[1] = fminf(v0.y, fminf(v1.y, v2.y));
            bx[2] = fminf(v0.z, fminf(v1.z, v2.z));
            bx[3] = fmaxf(v0.x, fmaxf(v1.x, v2.x));
            bx[4] = fmaxf(v0.y, fmaxf(v1.y, v2.y));
            bx[5] = fmaxf(v0.z, fmaxf(v1.z, v2.z));
        }
    };
    auto link = [&](int c) {
        if (c < 0) return ~(int)(base + (uint32_t)(~c));
        const int m = c == 0 ? rt : c;  // the root is no one's child
        return (int)(base + (uint32_t)m);
    };
    for (int i = tid; i < n - 1; i += T) {
        const int c0 = cid[2 * i], c1 = cid[2 * i + 1];
        float bl[6], br[6];
        child_box(c0, bl);
        child_box(c1, br);
        Node nd;
        nd.a = make_float4(bl[0], bl[1], bl[2], bl[3]);
        nd.b = make_float4(bl[4], bl[5], br[0], br[1]);
        nd.c = make_float4(br[2], br[3], br[4], br[5]);
        nd.d = make_int4(link(c0), link(c1), 0, 0);
        const int at = i == rt ? 0 : (i == 0 ? rt : i);
        nodes[base + at] = nd;
    }
    if (tid == 0 && wg_h) atomicMax(height_out, wg_h);
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
    uint32_t cpow = 2;
    while (cpow < cap) cpow <<= 1;
    return 4ull * cpow + 34ull * cap + 4;
}

bool build_small(const float* verts, const uint32_t* idx, const int32_t* obj_of_loc,
                 const uint32_t* first_tri, const uint32_t* tri_cnt, const uint32_t* obox, const uint32_t* seg_off,
                 uint32_t l_begin, uint32_t count, uint32_t cap, int threads, Node* nodes, LeafTri* leaves,
                 uint32_t* height, const char* timer, hipStream_t s) {
    if (count == 0) return true;
    if (cap == 0 || cap > pzk::kSmallMax) return false;
    const size_t lds = build_small_lds(cap);
    KernelTimer kt_(timer ? timer : "build_small", s);
    KernelTimer kt_all_("build_small", s);
#define PZ_LAUNCH_BS(T)                                                                                          \
    hipLaunchKernelGGL(k_build_small<T>, dim3(count), dim3(T), lds, s, verts, idx, obj_of_loc, first_tri, tri_cnt, \
                       obox, seg_off, l_begin, (int)cap, nodes, leaves, height)
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
