// Per-object LBVH built entirely in LDS, one workgroup per object.
//
// For an object of n <= kSmallMax triangles stored contiguously from
// first_tri: 30-bit Morton codes of the centroids inside the object's box,
// packed with the local index into distinct 64-bit keys; bitonic sort in LDS;
// Karras hierarchy over the sorted keys; bottom-up refit with LDS counters;
// then one coalesced write of the internal nodes and leaf triangles at the
// object's segment [seg_off[l], seg_off[l] + n).  Compared with the global
// pipeline (radix sort + Karras + atomic refit + gather), the triangles are
// read from HBM once and every intermediate stays on chip.
#include "pz_device.h"
#include "pz_kernels_join.h"

namespace {

using namespace pzd;

constexpr int kThreads = 1024;
constexpr int kMax = (int)pzk::kSmallMax;

__device__ __forceinline__ float3 vtx(const float* __restrict__ v, uint32_t i) {
    return make_float3(v[3ull * i], v[3ull * i + 1], v[3ull * i + 2]);
}

__device__ __forceinline__ void tri_box(const float* __restrict__ v, const uint32_t* __restrict__ idx, uint32_t t,
                                        float* lo, float* hi) {
    float3 a = vtx(v, idx[3ull * t]), b = vtx(v, idx[3ull * t + 1]), c = vtx(v, idx[3ull * t + 2]);
    lo[0] = fminf(a.x, fminf(b.x, c.x));
    lo[1] = fminf(a.y, fminf(b.y, c.y));
    lo[2] = fminf(a.z, fminf(b.z, c.z));
    hi[0] = fmaxf(a.x, fmaxf(b.x, c.x));
    hi[1] = fmaxf(a.y, fmaxf(b.y, c.y));
    hi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
}

__global__ void k_need_split(const uint32_t* __restrict__ need, const uint32_t* __restrict__ tri_cnt,
                             const uint32_t* __restrict__ first_tri, const uint32_t* __restrict__ last_tri,
                             int32_t nobj, uint32_t* __restrict__ big, uint32_t* __restrict__ small) {
    int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nobj) return;
    uint32_t n = tri_cnt[o];
    bool contiguous = n > 0 && last_tri[o] - first_tri[o] + 1 == n;
    bool sm = need[o] && contiguous && n <= (uint32_t)kMax;
    big[o] = (need[o] && !sm) ? 1u : 0u;
    small[o] = sm ? 1u : 0u;
}

__global__ void k_merge_locs(const uint32_t* __restrict__ big, const uint32_t* __restrict__ loc_big,
                             const uint32_t* __restrict__ loc_small, int32_t nobj, uint32_t* __restrict__ loc,
                             uint32_t* __restrict__ totals) {
    int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nbig = loc_big[nobj];
    if (o == 0) totals[0] = nbig;
    if (o >= nobj) return;
    loc[o] = big[o] ? loc_big[o] : nbig + loc_small[o];
    if (o == nobj - 1) loc[nobj] = nbig + loc_small[nobj];
}

__global__ void k_big_prefix(const uint32_t* __restrict__ seg_off, uint32_t* __restrict__ totals) {
    if (threadIdx.x == 0 && blockIdx.x == 0) totals[1] = seg_off[totals[0]];
}

__device__ __forceinline__ int delta64(const uint64_t* key, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    return __clzll((long long)(key[i] ^ key[j]));  // keys are distinct
}

__global__ __launch_bounds__(kThreads) void k_build_small(
    const float* __restrict__ verts, const uint32_t* __restrict__ idx, const int32_t* __restrict__ t2o,
    const int32_t* __restrict__ obj_of_loc, const uint32_t* __restrict__ first_tri,
    const uint32_t* __restrict__ tri_cnt, const uint32_t* __restrict__ obox, const uint32_t* __restrict__ seg_off,
    uint32_t l_begin, Node* __restrict__ nodes, LeafTri* __restrict__ leaves, uint32_t* __restrict__ height_out) {
    __shared__ uint64_t key[kMax];                 // 16 KB
    __shared__ float lbox[kMax][6];                // 48 KB: box of the leaf at sorted position p
    __shared__ float nbox[kMax][6];                // 48 KB: box of internal node i
    __shared__ int16_t ch[kMax][2];                // children: j >= 0 internal, < 0 leaf ~j
    __shared__ uint16_t par_leaf[kMax], par_int[kMax];
    __shared__ uint32_t cnt[kMax];
    __shared__ uint8_t hgt[kMax];

    const uint32_t l = l_begin + blockIdx.x;
    const int32_t o = obj_of_loc[l];
    const int n = (int)tri_cnt[o];
    const uint32_t t0 = first_tri[o];
    const uint32_t base = seg_off[l];
    const int tid = threadIdx.x;

    float lo[3], inv[3];
    for (int k = 0; k < 3; ++k) {
        float bl = ord2f(obox[6ll * o + k]), bh = ord2f(obox[6ll * o + 3 + k]);
        lo[k] = bl;
        inv[k] = bh > bl ? 1.0f / (bh - bl) : 0.0f;
    }
    int npow = 1;
    while (npow < n) npow <<= 1;

    // 1. keys
    for (int p = tid; p < npow; p += kThreads) {
        if (p < n) {
            uint32_t t = t0 + (uint32_t)p;
            float3 a = vtx(verts, idx[3ull * t]), b = vtx(verts, idx[3ull * t + 1]), c = vtx(verts, idx[3ull * t + 2]);
            float cx = (a.x + b.x + c.x) * (1.0f / 3.0f), cy = (a.y + b.y + c.y) * (1.0f / 3.0f),
                  cz = (a.z + b.z + c.z) * (1.0f / 3.0f);
            uint32_t m = morton30(cx, cy, cz, lo[0], lo[1], lo[2], inv[0], inv[1], inv[2]);
            key[p] = ((uint64_t)m << 32) | (uint32_t)p;
        } else {
            key[p] = ~0ull;
        }
    }
    __syncthreads();
    // 2. bitonic sort of npow keys (ascending).  Thread i owns compare-exchange
    // pair i, so for strides <= 64 every pair of a wave lies in that wave's own
    // 128-key segment and only stages with larger strides need a block barrier.
    static_assert(kThreads * 2 >= kMax, "one compare-exchange pair per thread");
    for (int size = 2; size <= npow; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const bool cross_wave = stride > PZ_WAVE;
            if (cross_wave) __syncthreads();
            const int i = tid;
            if (i < npow / 2) {
                int lo_i = 2 * i - (i & (stride - 1));
                int hi_i = lo_i + stride;
                bool up = (lo_i & size) == 0;
                uint64_t x = key[lo_i], y = key[hi_i];
                if ((x > y) == up) {
                    key[lo_i] = y;
                    key[hi_i] = x;
                }
            }
            if (cross_wave) __syncthreads();
            else __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    // leaf boxes in sorted order (the only other global read of the triangles)
    for (int p = tid; p < n; p += kThreads) {
        float blo[3], bhi[3];
        tri_box(verts, idx, t0 + (uint32_t)(key[p] & 0xFFFFFFFFu), blo, bhi);
        for (int k = 0; k < 3; ++k) {
            lbox[p][k] = blo[k];
            lbox[p][3 + k] = bhi[k];
        }
    }
    // 3. Karras hierarchy over sorted positions [0, n)
    for (int i = tid; i < n - 1; i += kThreads) {
        const int d = (delta64(key, i, i + 1, n) - delta64(key, i, i - 1, n)) >= 0 ? 1 : -1;
        const int dmin = delta64(key, i, i - d, n);
        int lmax = 2;
        while (delta64(key, i, i + lmax * d, n) > dmin) lmax <<= 1;
        int len = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta64(key, i, i + (len + t) * d, n) > dmin) len += t;
        const int j = i + len * d;
        const int dnode = delta64(key, i, j, n);
        int s = 0;
        for (int div = 2;; div <<= 1) {
            int t = (len + div - 1) / div;
            if (delta64(key, i, i + (s + t) * d, n) > dnode) s += t;
            if (t <= 1) break;
        }
        const int g = i + s * d + (d < 0 ? -1 : 0);
        const int lo_j = min(i, j), hi_j = max(i, j);
        if (lo_j == g) {
            ch[i][0] = (int16_t)~g;
            par_leaf[g] = (uint16_t)i;
        } else {
            ch[i][0] = (int16_t)g;
            par_int[g] = (uint16_t)i;
        }
        if (hi_j == g + 1) {
            ch[i][1] = (int16_t)~(g + 1);
            par_leaf[g + 1] = (uint16_t)i;
        } else {
            ch[i][1] = (int16_t)(g + 1);
            par_int[g + 1] = (uint16_t)i;
        }
        cnt[i] = 0;
    }
    __syncthreads();
    // 4. refit: climb from every leaf; the second arrival at a node computes it.
    uint32_t my_height = 0;
    if (n >= 2) {
        for (int p = tid; p < n; p += kThreads) {
            float blo[3], bhi[3];
            for (int k = 0; k < 3; ++k) {
                blo[k] = lbox[p][k];
                bhi[k] = lbox[p][3 + k];
            }
            int node = par_leaf[p];
            int child = ~p;
            int h = 0;
            while (true) {
                __threadfence_block();
                if (atomicAdd(&cnt[node], 1u) == 0) break;  // sibling still climbing
                __threadfence_block();
                const int sib = ch[node][0] == child ? ch[node][1] : ch[node][0];
                float slo[3], shi[3];
                int sh;
                if (sib < 0) {
                    for (int k = 0; k < 3; ++k) {
                        slo[k] = lbox[~sib][k];
                        shi[k] = lbox[~sib][3 + k];
                    }
                    sh = 0;
                } else {
                    for (int k = 0; k < 3; ++k) {
                        slo[k] = nbox[sib][k];
                        shi[k] = nbox[sib][3 + k];
                    }
                    sh = hgt[sib];
                }
                for (int k = 0; k < 3; ++k) {
                    blo[k] = fminf(blo[k], slo[k]);
                    bhi[k] = fmaxf(bhi[k], shi[k]);
                    nbox[node][k] = blo[k];
                    nbox[node][3 + k] = bhi[k];
                }
                h = 1 + max(h, sh);
                hgt[node] = (uint8_t)h;
                if (node == 0) {
                    my_height = (uint32_t)h;
                    break;
                }
                child = node;
                node = par_int[node];
            }
        }
    }
    __syncthreads();
    if (my_height) atomicMax(height_out, my_height);
    // 5. write internal nodes and leaves at the object's segment
    for (int i = tid; i < n - 1; i += kThreads) {
        float b[2][6];
        int gc[2];
        for (int sl = 0; sl < 2; ++sl) {
            const int c = ch[i][sl];
            if (c < 0) {
                for (int k = 0; k < 6; ++k) b[sl][k] = lbox[~c][k];
                gc[sl] = ~(int)(base + (uint32_t)(~c));
            } else {
                for (int k = 0; k < 6; ++k) b[sl][k] = nbox[c][k];
                gc[sl] = (int)(base + (uint32_t)c);
            }
        }
        Node nd;
        nd.a = make_float4(b[0][0], b[0][1], b[0][2], b[0][3]);
        nd.b = make_float4(b[0][4], b[0][5], b[1][0], b[1][1]);
        nd.c = make_float4(b[1][2], b[1][3], b[1][4], b[1][5]);
        nd.d = make_int4(gc[0], gc[1], 0, 0);
        nodes[base + i] = nd;
    }
    for (int p = tid; p < n; p += kThreads) {
        const uint32_t t = t0 + (uint32_t)(key[p] & 0xFFFFFFFFu);
        float3 a = vtx(verts, idx[3ull * t]), b = vtx(verts, idx[3ull * t + 1]), c = vtx(verts, idx[3ull * t + 2]);
        LeafTri L;
        L.v0 = make_float4(a.x, a.y, a.z, __uint_as_float(t));
        L.v1 = make_float4(b.x, b.y, b.z, __int_as_float(t2o[t]));
        L.v2 = make_float4(c.x, c.y, c.z, 0.0f);
        leaves[base + p] = L;
    }
}

inline unsigned grid1(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace pzk {

void need_split(const uint32_t* need, const uint32_t* tri_cnt, const uint32_t* first_tri, const uint32_t* last_tri,
                int32_t nobj, uint32_t* big, uint32_t* small, hipStream_t s) {
    if (nobj > 0)
        hipLaunchKernelGGL(k_need_split, dim3(grid1((uint64_t)nobj)), dim3(256), 0, s, need, tri_cnt, first_tri,
                           last_tri, nobj, big, small);
}

void merge_locs(const uint32_t* big, const uint32_t* loc_big, const uint32_t* loc_small, int32_t nobj, uint32_t* loc,
                uint32_t* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_merge_locs, dim3(grid1((uint64_t)std::max(nobj, 1))), dim3(256), 0, s, big, loc_big,
                       loc_small, nobj, loc, totals);
}

void big_prefix(const uint32_t* seg_off, uint32_t* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_big_prefix, dim3(1), dim3(64), 0, s, seg_off, totals);
}

void build_small(const float* verts, const uint32_t* idx, const int32_t* t2o, const int32_t* obj_of_loc,
                 const uint32_t* first_tri, const uint32_t* tri_cnt, const uint32_t* obox, const uint32_t* seg_off,
                 uint32_t l_begin, uint32_t count, Node* nodes, LeafTri* leaves, uint32_t* height, hipStream_t s) {
    if (count == 0) return;
    KernelTimer kt_("build_small", s);
    hipLaunchKernelGGL(k_build_small, dim3(count), dim3(kThreads), 0, s, verts, idx, t2o, obj_of_loc, first_tri,
                       tri_cnt, obox, seg_off, l_begin, nodes, leaves, height);
}

}  // namespace pzk
