// Per-object LBVH built entirely in LDS, one workgroup per object.
//
// For an object of n <= 2048 triangles stored contiguously from first_tri:
//   1. one gather: each thread loads its local triangles q = tid + k * T
//      (the first 2-4 of them stay in registers for the whole build) and
//      writes a 32-bit key: a 21-bit Morton code of the centroid on a 127^3
//      grid over the object's own box (never all ones) above the 11-bit local
//      index, so keys are distinct;
//   2. bitonic sort in which every comparator puts the smaller key first: a
//      size-s block starts by pairing position i with its mirror
//      i ^ (s - 1), then halves with strides s/4 .. 1.  The virtual keys at
//      positions >= n are +infinity and never move, so they take no LDS.  A
//      wavefront holds 256 keys in registers (four per lane): register pairs
//      for strides 128 and 64, ds_bpermute for lane strides 32 .. 4 and the
//      mirrors, DPP quad permutes for strides 2 and 1; only strides >= 256
//      go through LDS with a block barrier;
//   3. the sorted position of each local triangle, kept in registers; leaf
//      records in sorted order, staged through LDS in two halves so the
//      48-byte records are written to HBM coalesced;
//   4. agglomerative bottom-up build (Apetrei 2014, "Fast and simple
//      agglomerative LBVH construction"), LDS only: each thread climbs from
//      the sorted positions of its own triangles.  A node covering sorted
//      leaves [l, r] joins the neighbouring split with the smaller key XOR
//      (split r, or l - 1).  The first child to arrive stops; the exchange
//      on bound[split] hands its far range end, id and height to the second,
//      which merges both boxes, packs the two child ids into bound[split]
//      and climbs on.  Boxes live in one LDS array: leaf p's box at p until
//      node p completes, node i's box at i afterwards (every node containing
//      leaf p completes after leaf p's parent, which consumed it);
//   5. node records (own box + child links, 32 B) written in index order
//      from LDS (coalesced), two lanes per record; every internal node's box
//      is in LDS after the climb, so no leaf record is read back.  The root
//      must sit at the segment's first node (the traversals start there): it
//      is written at index 0 and internal node 0's record at the root's
//      index, with the one link into node 0 redirected.
// LDS is dynamic, 32 B per triangle of the launch's largest object: keys
// (4 B), boxes (24 B) and the exchange words (4 B), which hold the sorted
// positions until the climb starts.
//
// Chunk modes (objects over 2,048 triangles or not contiguous), one
// workgroup per run of up to kChunk triangles of one object, which writes its
// root for the per-object top tree (k_top_level) and leaves the top tree one
// node slot per chunk:
//  - own order (every big object contiguous): a run of the object's own
//    triangles; the Morton grid is over the run's centroid bounds (a block
//    reduction after the gather) and the run is sorted as in step 2;
//  - sorted: the big objects' triangles are Morton-sorted once over the
//    whole set, a run's keys are already ascending and step 2 is skipped.
#include <type_traits>

#include "pz_device.h"
#include "pz_kernels_join.h"

namespace {

using namespace pzd;

constexpr int kMax = (int)pzk::kSmallMax;
constexpr uint32_t kCh = pzk::kChunk;           // triangles per chunk of a big object
constexpr int kChThreads = pzk::kChunkThreads;  // workgroup size of the chunk builds

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
    // needed objects whose triangles are not contiguous (they need the sorted big-object build)
    const uint64_t nc = __ballot(o < nobj && need[o] && !(n > 0 && last_tri[o] - first_tri[o] + 1 == n));
    if (__lane_id() == 0 && nc) atomicAdd(&totals[8], (uint32_t)__popcll(nc));
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

// Chunk mode of the builder: the workgroup's triangles are a run of the big
// objects' Morton-sorted order instead of an object's contiguous range.
struct ChunkSrc {
    const uint64_t* keys;       // (segment << mbits) | mbits-bit Morton code, ascending
    const uint32_t* order;      // triangle id at each sorted position
    const uint32_t* chunk_off;  // first chunk of each big segment (nbig + 1 entries)
    uint32_t nbig;
    int mbits;
    pzk::ChunkRoot* ent;        // per chunk: root link and box
};

constexpr uint32_t kM30 = (1u << 30) - 1u;

// Largest l in [0, n) with off[l] <= g, for off[0] = 0 <= g < off[n].
__device__ __forceinline__ uint32_t seg_of_chunk(const uint32_t* __restrict__ off, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// MODE 0: one object per workgroup; 1: a chunk of the sorted big objects;
// 2: a chunk of a big object in its own triangle order.
template <int T, int MODE>
__global__ __launch_bounds__(T) void k_build_small(
    const float* __restrict__ verts, const uint32_t* __restrict__ idx, const int32_t* __restrict__ obj_of_loc,
    const uint32_t* __restrict__ first_tri, const uint32_t* __restrict__ tri_cnt, const uint32_t* __restrict__ obox, const uint32_t* __restrict__ seg_off,
    uint32_t l_begin, int cap, BNode* __restrict__ nodes, LeafTri* __restrict__ leaves, uint32_t* __restrict__ height_out,
    ChunkSrc ch) {
    constexpr bool CHUNK = MODE != 0;
    constexpr int PR = T == 1024 ? 2 : (T == 512 ? 3 : 4);  // local triangles per thread kept in registers
    static_assert(MODE != 2 || PR * T >= (int)kCh, "a natural-order chunk is held in registers whole");
    constexpr int NP = (kMax / T + 1) / 2;                 // sorted positions per thread, two per register
    extern __shared__ __align__(16) unsigned char lds[];
    const int capk = (cap + 3) & ~3;
    uint32_t* key = reinterpret_cast<uint32_t*>(lds);      // capk
    float* box = reinterpret_cast<float*>(key + capk);     // 6 x cap (staging before the climb)
    int* bound = reinterpret_cast<int*>(box + 6 * cap);    // cap: sorted positions, then exchange / child ids
    uint16_t* inv = reinterpret_cast<uint16_t*>(bound);
    float* stage = box;
    // a sort key's Morton part, top-aligned in 30 bits
    auto code30 = [&](uint64_t k) { return ((uint32_t)k & ((1u << ch.mbits) - 1u)) << (30 - ch.mbits); };
    // leaves at lbase + sorted position, nodes at nbase + node index (root at nbase)
    uint32_t l, t0, lbase, nbase, g = 0, cpl = 0, tn0 = 0;
    int32_t o;
    int n;
    if constexpr (CHUNK) {
        g = blockIdx.x;
        if (g >= ch.chunk_off[ch.nbig]) return;  // the grid is an upper bound on the chunk count
        l = seg_of_chunk(ch.chunk_off, ch.nbig, g);
        const uint32_t sb = seg_off[l], m = seg_off[l + 1] - sb;
        const uint32_t j = g - ch.chunk_off[l], nch = ch.chunk_off[l + 1] - ch.chunk_off[l];
        t0 = sb + j * kCh;  // first sorted position of the chunk
        n = (int)min(kCh, m - j * kCh);
        lbase = t0;
        nbase = nch == 1 ? t0 : t0 + 1;  // slot t0 belongs to the top tree
        o = obj_of_loc[l];
        if constexpr (MODE == 1) {
            // codes of a sorted run share their leading bits: keep the 21 below them
            const uint32_t x = code30(ch.keys[t0] ^ ch.keys[t0 + n - 1]);
            cpl = x ? (uint32_t)__clz((int)x) - 2u : 30u;
        } else {
            tn0 = first_tri[o] + j * kCh;  // the chunk's first triangle
        }
    } else {
        l = l_begin + blockIdx.x;
        o = obj_of_loc[l];
        n = (int)tri_cnt[o];
        t0 = first_tri[o];
        lbase = nbase = seg_off[l];
    }
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    (void)code30;
    float lo[3] = {0.0f, 0.0f, 0.0f}, sc[3] = {0.0f, 0.0f, 0.0f};
    if constexpr (!CHUNK)
        for (int k = 0; k < 3; ++k) {
            const float bl = ord2f(obox[6ll * o + k]), bh = ord2f(obox[6ll * o + 3 + k]);
            lo[k] = bl;
            sc[k] = bh > bl ? 127.0f / (bh - bl) : 0.0f;
        }
    int npow = 1;
    while (npow < n) npow <<= 1;
    auto morton = [&](const float3& a, const float3& b, const float3& c, int q) {
        if constexpr (MODE == 1) {  // ascending in q: the run is sorted already
            const uint32_t m = ((code30(ch.keys[t0 + q]) << cpl) & kM30) >> 9;
            return (m << 11) | (uint32_t)q;
        }
        const float cx[3] = {(a.x + b.x + c.x) * (1.0f / 3.0f), (a.y + b.y + c.y) * (1.0f / 3.0f),
                             (a.z + b.z + c.z) * (1.0f / 3.0f)};
        uint32_t qq[3];
        for (int k = 0; k < 3; ++k) qq[k] = (uint32_t)fminf(fmaxf((cx[k] - lo[k]) * sc[k], 0.0f), 126.0f);
        uint32_t m = 0;
        for (int b2 = 6; b2 >= 0; --b2)
            m = (m << 3) | (((qq[0] >> b2) & 1u) << 2) | (((qq[1] >> b2) & 1u) << 1) | ((qq[2] >> b2) & 1u);
        return (m << 11) | (uint32_t)q;
    };
    auto tri_id = [&](int q) {
        return MODE == 1 ? ch.order[t0 + (uint32_t)q] : (MODE == 2 ? tn0 : t0) + (uint32_t)q;
    };
    auto tri = [&](int q, float3& a, float3& b, float3& c) {
        const uint32_t t = tri_id(q);
        const uint32_t i0 = idx[3ull * t], i1 = idx[3ull * t + 1], i2 = idx[3ull * t + 2];
        a = vtx(verts, i0);
        b = vtx(verts, i1);
        c = vtx(verts, i2);
    };
    // 1. one gather
    float3 ra[PR > 0 ? PR : 1], rb[PR > 0 ? PR : 1], rc[PR > 0 ? PR : 1];
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int q = tid + k * T;
        if (q < n) tri(q, ra[k], rb[k], rc[k]);
    }
    if constexpr (MODE == 2) {  // the Morton grid over the chunk's centroid bounds
        __shared__ uint32_t cbox[6];
        if (tid < 6) cbox[tid] = tid < 3 ? 0xFFFFFFFFu : 0u;
        float cl[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int k = 0; k < PR; ++k)
            if (tid + k * T < n) {
                const float cx[3] = {(ra[k].x + rb[k].x + rc[k].x) * (1.0f / 3.0f),
                                     (ra[k].y + rb[k].y + rc[k].y) * (1.0f / 3.0f),
                                     (ra[k].z + rb[k].z + rc[k].z) * (1.0f / 3.0f)};
                for (int a = 0; a < 3; ++a) {
                    cl[a] = fminf(cl[a], cx[a]);
                    chi[a] = fmaxf(chi[a], cx[a]);
                }
            }
        for (int a = 0; a < 3; ++a)
            for (int sft = 32; sft > 0; sft >>= 1) {
                cl[a] = fminf(cl[a], __shfl_xor(cl[a], sft, 64));
                chi[a] = fmaxf(chi[a], __shfl_xor(chi[a], sft, 64));
            }
        __syncthreads();
        if (lane == 0 && tid < n)
            for (int a = 0; a < 3; ++a) {
                atomicMin(&cbox[a], f2ord(cl[a]));
                atomicMax(&cbox[3 + a], f2ord(chi[a]));
            }
        __syncthreads();
        for (int a = 0; a < 3; ++a) {
            const float bl = ord2f(cbox[a]), bh = ord2f(cbox[3 + a]);
            lo[a] = bl;
            sc[a] = bh > bl ? 127.0f / (bh - bl) : 0.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int q = tid + k * T;
        if (q < n) key[q] = morton(ra[k], rb[k], rc[k], q);
    }
    for (int q = tid + PR * T; q < n; q += T) {
        float3 a, b, c;
        tri(q, a, b, c);
        key[q] = morton(a, b, c, q);
    }
    __syncthreads();
    // 2. bitonic sort, every comparator ascending (a block's first step pairs
    // i with its mirror i ^ (size - 1)), so the virtual keys at positions >= n
    // (+inf) never move and need no storage.
    auto lds_mirror = [&](int s) {
        const int hs = s >> 1;
        for (int t = tid; t < npow / 2; t += T) {
            const int blk = t / hs, off = t - blk * hs;
            const int a = blk * s + off, b = blk * s + s - 1 - off;
            if (b < n) {
                const uint32_t x = key[a], y = key[b];
                if (x > y) {
                    key[a] = y;
                    key[b] = x;
                }
            }
        }
    };
    auto lds_half = [&](int st) {
        for (int t = tid; t < npow / 2; t += T) {
            const int a = 2 * t - (t & (st - 1)), b = a + st;
            if (b < n) {
                const uint32_t x = key[a], y = key[b];
                if (x > y) {
                    key[a] = y;
                    key[b] = x;
                }
            }
        }
    };
    // a wave holds 256 positions: lane L has c*256 + j*64 + L, j = 0..3
    auto lohi = [](uint32_t& x, uint32_t& y) {
        const uint32_t a = min(x, y), b = max(x, y);
        x = a;
        y = b;
    };
    auto lane_half = [&](auto stc, uint32_t& k0, uint32_t& k1, uint32_t& k2, uint32_t& k3) {
        constexpr int st = decltype(stc)::value;
        const bool lower = (lane & st) == 0;
        const uint32_t o0 = lane_swap<st>(k0), o1 = lane_swap<st>(k1), o2 = lane_swap<st>(k2), o3 = lane_swap<st>(k3);
        k0 = lower ? min(k0, o0) : max(k0, o0);
        k1 = lower ? min(k1, o1) : max(k1, o1);
        k2 = lower ? min(k2, o2) : max(k2, o2);
        k3 = lower ? min(k3, o3) : max(k3, o3);
    };
    auto bperm = [&](uint32_t v, int m) {  // value of lane (lane ^ m)
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane ^ m) << 2), (int)v);
    };
    auto reg_halves = [&](int s, uint32_t& k0, uint32_t& k1, uint32_t& k2, uint32_t& k3) {
        // half-cleaner strides min(s/4, 128) .. 1
        if (s >= 512) {
            lohi(k0, k2);
            lohi(k1, k3);
        }
        if (s >= 256) {
            lohi(k0, k1);
            lohi(k2, k3);
        }
        if (s >= 128) lane_half(std::integral_constant<int, 32>{}, k0, k1, k2, k3);
        if (s >= 64) lane_half(std::integral_constant<int, 16>{}, k0, k1, k2, k3);
        if (s >= 32) lane_half(std::integral_constant<int, 8>{}, k0, k1, k2, k3);
        if (s >= 16) lane_half(std::integral_constant<int, 4>{}, k0, k1, k2, k3);
        if (s >= 8) lane_half(std::integral_constant<int, 2>{}, k0, k1, k2, k3);
        if (s >= 4) lane_half(std::integral_constant<int, 1>{}, k0, k1, k2, k3);
    };
    auto reg_pass = [&](int s_lo, int s_hi) {  // s_lo == s_hi >= 512: half-cleaners only
        for (int c = wave; c * 256 < n; c += T / 64) {
            const int pb = c * 256 + lane;
            uint32_t k0 = pb < n ? key[pb] : ~0u, k1 = pb + 64 < n ? key[pb + 64] : ~0u;
            uint32_t k2 = pb + 128 < n ? key[pb + 128] : ~0u, k3 = pb + 192 < n ? key[pb + 192] : ~0u;
            if (s_lo >= 512) {
                reg_halves(s_lo, k0, k1, k2, k3);
            } else {
                for (int s = s_lo; s <= s_hi; s <<= 1) {
                    // mirror step: partner of position i is i ^ (s - 1)
                    if (s <= 64) {
                        const int m = s - 1;
                        const bool lower = (lane & (s >> 1)) == 0;
                        const uint32_t o0 = bperm(k0, m), o1 = bperm(k1, m), o2 = bperm(k2, m), o3 = bperm(k3, m);
                        k0 = lower ? min(k0, o0) : max(k0, o0);
                        k1 = lower ? min(k1, o1) : max(k1, o1);
                        k2 = lower ? min(k2, o2) : max(k2, o2);
                        k3 = lower ? min(k3, o3) : max(k3, o3);
                    } else if (s == 128) {  // j ^ 1, lane ^ 63
                        const uint32_t o0 = bperm(k1, 63), o1 = bperm(k0, 63), o2 = bperm(k3, 63), o3 = bperm(k2, 63);
                        k0 = min(k0, o0);
                        k1 = max(k1, o1);
                        k2 = min(k2, o2);
                        k3 = max(k3, o3);
                    } else {  // s == 256: j ^ 3, lane ^ 63
                        const uint32_t o0 = bperm(k3, 63), o1 = bperm(k2, 63), o2 = bperm(k1, 63), o3 = bperm(k0, 63);
                        k0 = min(k0, o0);
                        k1 = min(k1, o1);
                        k2 = max(k2, o2);
                        k3 = max(k3, o3);
                    }
                    reg_halves(s, k0, k1, k2, k3);
                }
            }
            if (pb < n) key[pb] = k0;
            if (pb + 64 < n) key[pb + 64] = k1;
            if (pb + 128 < n) key[pb + 128] = k2;
            if (pb + 192 < n) key[pb + 192] = k3;
        }
    };
    if (MODE != 1 && n > 1) {
        reg_pass(2, min(npow, 256));
        for (int s = 512; s <= npow; s <<= 1) {
            __syncthreads();
            lds_mirror(s);
            for (int st = s >> 2; st >= 256; st >>= 1) {
                __syncthreads();
                lds_half(st);
            }
            __syncthreads();
            reg_pass(s, s);
        }
    }
    __syncthreads();
    // 3. sorted positions (kept in registers for the climb), leaf records
    // staged through LDS in two halves
    for (int p = tid; p < n; p += T) inv[key[p] & 0x7FFu] = (uint16_t)p;
    __syncthreads();
    uint32_t ppos[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) ppos[k] = 0;
#pragma unroll
    for (int k = 0; k < 2 * NP; ++k) {
        const int q = tid + k * T;
        if (q < n) ppos[k >> 1] |= (uint32_t)inv[q] << (16 * (k & 1));
    }
    auto pos_of = [&](int k) { return (int)((ppos[k >> 1] >> (16 * (k & 1))) & 0xFFFFu); };
    // inv lives in bound, which is reset before the climb: every thread must
    // hold its positions first (the staging below also reuses the box area)
    __syncthreads();
    const int half = (n + 1) >> 1;
    for (int r = 0; r < 2; ++r) {
        const int h0 = r * half, h1 = min(n, h0 + half);
        auto put = [&](int k, const float3& a, const float3& b, const float3& c) {
            const int p = pos_of(k);
            if (p < h0 || p >= h1) return;
            float* st = stage + 9 * (p - h0);
            st[0] = a.x; st[1] = a.y; st[2] = a.z;
            st[3] = b.x; st[4] = b.y; st[5] = b.z;
            st[6] = c.x; st[7] = c.y; st[8] = c.z;
        };
#pragma unroll
        for (int k = 0; k < PR; ++k) {
            const int q = tid + k * T;
            if (q < n) put(k, ra[k], rb[k], rc[k]);
        }
        for (int k = PR; k < 2 * NP; ++k) {
            const int q = tid + k * T;
            if (q >= n) break;
            float3 a, b, c;
            tri(q, a, b, c);
            put(k, a, b, c);
        }
        __syncthreads();
        // four lanes per record, lane c < 3 writing corner c, so that each
        // store instruction covers 16 consecutive records
        {
            const int c = tid & 3;
            for (int p = h0 + (tid >> 2); p < h1; p += T / 4) {
                if (c < 3) {
                    const float* st = stage + 9 * (p - h0) + 3 * c;
                    const float w = c == 0 ? __uint_as_float(tri_id((int)(key[p] & 0x7FFu)))
                                           : (c == 1 ? __int_as_float(o) : 0.0f);
                    reinterpret_cast<float4*>(leaves + lbase + p)[c] = make_float4(st[0], st[1], st[2], w);
                }
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < n - 1; i += T) bound[i] = -1;
    __syncthreads();
    // 4. climb (as v8); the second arrival packs the children into bound[split]
    auto climb = [&](int p, const float3& a, const float3& b, const float3& c) {
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
            const int c0 = s0 ? me : sib, c1 = s0 ? sib : me;
            bound[parent] = (int)(((uint32_t)c0 & 0xFFFFu) | ((uint32_t)c1 << 16));
            float* sb = box + 6 * parent;
            for (int k = 0; k < 3; ++k) {
                sb[k] = blo[k];
                sb[3 + k] = bhi[k];
            }
            h = 1 + max(h, sh);
            me = parent;
            if (lft == 0 && rgt == n - 1) {
                box[6 * (n - 1)] = __int_as_float(parent);  // root index: leaf n-1's slot is free by now
                atomicMax(height_out, (uint32_t)h);
                break;
            }
        }
    };
    if (n >= 2) {
#pragma unroll
        for (int k = 0; k < PR; ++k) {
            const int q = tid + k * T;
            if (q < n) climb(pos_of(k), ra[k], rb[k], rc[k]);
        }
        for (int k = PR; k < 2 * NP; ++k) {
            const int q = tid + k * T;
            if (q >= n) break;
            float3 a, b, c;
            tri(q, a, b, c);
            climb(pos_of(k), a, b, c);
        }
    }
    __syncthreads();
    // 5. node records in index order (root at 0, node 0 at the root's index):
    // node i's own box is box[6 i ..] after the climb
    if (n >= 2) {
        const int rt = __float_as_int(box[6 * (n - 1)]);
        auto link = [&](int c) {
            if (c < 0) return ~(int)(lbase + (uint32_t)(~c));
            return (int)(nbase + (uint32_t)(c == 0 ? rt : c));
        };
        // two lanes per record, lane h writing its 16-byte half, so that each
        // store instruction covers 32 consecutive records
        const int h = tid & 1;
        for (int i = tid >> 1; i < n - 1; i += T / 2) {
            const float* bx = box + 6 * i;
            float4 q;
            if (h == 0) {
                q = make_float4(bx[0], bx[1], bx[2], bx[3]);
            } else {
                const uint32_t v = (uint32_t)bound[i];
                const int c0 = (int16_t)(v & 0xFFFFu), c1 = (int16_t)(v >> 16);
                q = make_float4(bx[4], bx[5], __int_as_float(link(c0)), __int_as_float(link(c1)));
            }
            reinterpret_cast<float4*>(nodes + nbase + (i == rt ? 0 : (i == 0 ? rt : i)))[h] = q;
        }
    }
    if constexpr (CHUNK) {  // the chunk's root for the top tree
        if (tid == 0) {
            pzk::ChunkRoot e;
            if (n >= 2) {
                const int rt = __float_as_int(box[6 * (n - 1)]);
                for (int k = 0; k < 3; ++k) {
                    e.lo[k] = box[6 * rt + k];
                    e.hi[k] = box[6 * rt + 3 + k];
                }
                e.link = (int32_t)nbase;
            } else {
                const float3 a = ra[0], b = rb[0], c = rc[0];
                e.lo[0] = fminf(a.x, fminf(b.x, c.x));
                e.lo[1] = fminf(a.y, fminf(b.y, c.y));
                e.lo[2] = fminf(a.z, fminf(b.z, c.z));
                e.hi[0] = fmaxf(a.x, fmaxf(b.x, c.x));
                e.hi[1] = fmaxf(a.y, fmaxf(b.y, c.y));
                e.hi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
                e.link = ~(int32_t)lbase;
            }
            e.pad = 0;
            ch.ent[g] = e;
        }
    }
}

// Number of chunks of each big segment.
__global__ void k_chunk_count(const uint32_t* __restrict__ seg_off, uint32_t nbig, uint32_t* __restrict__ cnt) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nbig) cnt[l] = (seg_off[l + 1] - seg_off[l] + kCh - 1u) / kCh;
}

// Whether each big object's own triangle order is spatially coherent enough
// for its chunks: thread per object, the chunk boxes' surface areas summed in
// chunk order against kCoherentArea times its box's.  A coherent order sums
// to a few box areas (tubes generated ring by ring 0.5-1.2, subdivided
// spheres 3.3-3.7 at 1,024-triangle chunks); chunks drawn from all over the
// object each cover about the whole box, K of them K areas.  Objects that
// fail are counted in *incoherent.
constexpr float kCoherentArea = 6.0f;
__device__ __forceinline__ float box_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
__global__ void k_chunk_check(const uint32_t* __restrict__ chunk_off, uint32_t nbig,
                              const int32_t* __restrict__ obj_of_loc, const uint32_t* __restrict__ obox,
                              const pzk::ChunkRoot* __restrict__ ent, uint32_t* __restrict__ incoherent) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (l < nbig && chunk_off[l + 1] - chunk_off[l] >= 2) {
        float sum = 0.0f;
        for (uint32_t g = chunk_off[l]; g < chunk_off[l + 1]; ++g) sum += box_area(ent[g].lo, ent[g].hi);
        const int32_t o = obj_of_loc[l];
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = ord2f(obox[6ll * o + k]);
            hi[k] = ord2f(obox[6ll * o + 3 + k]);
        }
        bad = !(sum <= kCoherentArea * box_area(lo, hi));
    }
    const uint64_t m = __ballot(bad);
    if (__lane_id() == 0 && m) atomicAdd(incoherent, (uint32_t)__popcll(m));
}

// One level of the top trees: an object with np >= 2 entries left (np =
// ceil(K / 2^(level-1))) pairs entries (2i, 2i+1) under a new node and moves
// an odd last entry up.  Nodes are numbered from the root down: the nodes
// made at this level follow those of every level above, so the root (the
// last level's single node) is top node 0, which is slot seg_off[l].
__global__ void k_top_level(const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ chunk_off,
                            uint32_t nbig, uint32_t level, const pzk::ChunkRoot* __restrict__ in,
                            pzk::ChunkRoot* __restrict__ out, BNode* __restrict__ nodes) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= chunk_off[nbig]) return;
    const uint32_t l = seg_of_chunk(chunk_off, nbig, g);
    const uint32_t cb = chunk_off[l], K = chunk_off[l + 1] - cb, i = g - cb;
    const uint32_t np = (uint32_t)(((uint64_t)K + (1ull << (level - 1)) - 1) >> (level - 1));
    if (np < 2) return;
    const uint32_t made = np / 2;
    if (i < made) {
        const pzk::ChunkRoot a = in[cb + 2 * i], b = in[cb + 2 * i + 1];
        uint32_t slot = i;
        for (uint32_t nl = (np + 1) / 2; nl >= 2; nl = (nl + 1) / 2) slot += nl / 2;
        const uint32_t node = seg_off[l] + slot * kCh;
        pzk::ChunkRoot e;
        for (int k = 0; k < 3; ++k) {
            e.lo[k] = fminf(a.lo[k], b.lo[k]);
            e.hi[k] = fmaxf(a.hi[k], b.hi[k]);
        }
        BNode nd;
        nd.a = make_float4(e.lo[0], e.lo[1], e.lo[2], e.hi[0]);
        nd.b = make_float4(e.hi[1], e.hi[2], __int_as_float(a.link), __int_as_float(b.link));
        nodes[node] = nd;
        e.link = (int32_t)node;
        e.pad = 0;
        out[cb + i] = e;
    } else if ((np & 1u) && i == made) {
        out[cb + i] = in[cb + np - 1];
    }
}

// The top levels add at most `levels` to the chunks' height.
__global__ void k_add_height(uint32_t* height, uint32_t levels) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *height += levels;
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

size_t build_small_lds(uint32_t cap) { return 32ull * std::max<uint32_t>(cap, 2); }

bool build_small(const float* verts, const uint32_t* idx, const int32_t* obj_of_loc,
                 const uint32_t* first_tri, const uint32_t* tri_cnt, const uint32_t* obox, const uint32_t* seg_off,
                 uint32_t l_begin, uint32_t count, uint32_t cap, int threads, BNode* nodes, LeafTri* leaves,
                 uint32_t* height, const char* timer, hipStream_t s) {
    if (count == 0) return true;
    if (cap == 0 || cap > pzk::kSmallMax) return false;
    const size_t lds = build_small_lds(cap);
    KernelTimer kt_(timer ? timer : "build_small", s);
    KernelTimer kt_all_("build_small", s);
#define PZ_LAUNCH_BS(T)                                                                                         \
    hipLaunchKernelGGL((k_build_small<T, 0>), dim3(count), dim3(T), lds, s, verts, idx, obj_of_loc, first_tri, \
                       tri_cnt, obox, seg_off, l_begin, (int)cap, nodes, leaves, height, ChunkSrc{})
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

uint64_t big_chunk_bound(uint64_t nbig, uint64_t nbig_tri) { return nbig_tri / kChunk + nbig; }

void build_big_chunked(const float* verts, const uint32_t* idx, const int32_t* obj_of_loc, const uint32_t* seg_off,
                       const uint32_t* first_tri, const uint32_t* obox, uint32_t nbig, uint64_t nbig_tri,
                       const uint64_t* sorted_keys, int mbits, const uint32_t* order, uint32_t* chunk_cnt,
                       uint32_t* chunk_off, void* scan_tmp, ChunkRoot* ent0, ChunkRoot* ent1, BNode* nodes,
                       LeafTri* leaves, uint32_t* height, uint32_t* incoherent, hipStream_t s) {
    if (nbig == 0 || nbig_tri == 0) return;
    const uint64_t bound = big_chunk_bound(nbig, nbig_tri);
    hipLaunchKernelGGL(k_chunk_count, dim3(grid1(nbig)), dim3(256), 0, s, seg_off, nbig, chunk_cnt);
    exclusive_scan_u32_u32(chunk_cnt, chunk_off, nbig, scan_tmp, s);
    {
        KernelTimer kt_("build_chunks", s);
        if (sorted_keys)
            hipLaunchKernelGGL((k_build_small<kChThreads, 1>), dim3((unsigned)bound), dim3(kChThreads),
                               build_small_lds(kChunk), s, verts, idx, obj_of_loc, nullptr, nullptr, nullptr, seg_off,
                               0u, (int)kChunk, nodes, leaves, height,
                               ChunkSrc{sorted_keys, order, chunk_off, nbig, mbits, ent0});
        else  // chunks in each object's own triangle order
            hipLaunchKernelGGL((k_build_small<kChThreads, 2>), dim3((unsigned)bound), dim3(kChThreads),
                               build_small_lds(kChunk), s, verts, idx, obj_of_loc, first_tri, nullptr, nullptr,
                               seg_off, 0u, (int)kChunk, nodes, leaves, height,
                               ChunkSrc{nullptr, nullptr, chunk_off, nbig, 0, ent0});
    }
    if (!sorted_keys)
        hipLaunchKernelGGL(k_chunk_check, dim3(grid1(nbig)), dim3(256), 0, s, chunk_off, nbig, obj_of_loc, obox, ent0,
                           incoherent);
    // an object has at most ceil(nbig_tri / kChunk) chunks
    const uint64_t kmax = (nbig_tri + kChunk - 1) / kChunk;
    uint32_t levels = 0;
    while ((1ull << levels) < kmax) ++levels;
    KernelTimer kt_("build_top", s);
    for (uint32_t lv = 1; lv <= levels; ++lv)
        hipLaunchKernelGGL(k_top_level, dim3(grid1(bound)), dim3(256), 0, s, seg_off, chunk_off, nbig, lv,
                           (lv & 1) ? ent0 : ent1, (lv & 1) ? ent1 : ent0, nodes);
    if (levels) hipLaunchKernelGGL(k_add_height, dim3(1), dim3(64), 0, s, height, levels);
}

}  // namespace pzk
