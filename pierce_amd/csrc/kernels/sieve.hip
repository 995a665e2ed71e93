// Pair sieve: the narrow phase of one candidate object pair without
// per-object trees, for gfx950.
//
// One workgroup per candidate pair (a, b) whose objects store their
// triangles contiguously, at most kSieveMaxTris each:
//  1. every triangle of both objects is read once (index triple and corners)
//     and kept if its box meets O = box(a) ∩ box(b): two triangles' boxes
//     can only meet inside O.
//  2. the side with fewer kept triangles (G) is binned into a uniform grid
//     over O in LDS, each triangle in every cell its box overlaps; cells are
//     about kCellScale times the mean triangle extent, at most 16 per axis
//     and kSieveCells in all.
//  3. every kept triangle of the other side looks up the cells its box
//     overlaps and tests the G triangles there whose boxes meet its own, each
//     pair once: in the cell of the larger of the two boxes' lower corners
//     (cells are a monotone function of the coordinate, so two boxes that
//     meet share that cell).  The triangle-pair test is the one the tree
//     walk runs (edges of each triangle against the other, by direction);
//     with the fp32 filter, a pair it leaves open goes to the escalation
//     list as (candidate pair, mesh1 triangle, mesh2 triangle) for
//     k_escalate_tris.  The workgroup stops at the first proven hit.
// A pair beyond a cap (kept triangles, grid references, object size or
// layout, escalation list) is flagged in defer[] and decided by the
// per-object tree path instead; the sieve then writes nothing else for it.
// Once the open triangle pairs are decided, k_sieve_parity casts the +Z
// parity rays of the sieved pairs that did not touch, through every triangle
// of the target object whose box meets the ray's box (the perturbed crossing
// of the tree path's parity kernel), so the exact predicate stages stay out
// of the sieve itself.
#include "pz_device.h"
#include "pz_kernels_join.h"
#include "pz_narrow.h"
#include "pz_predicates.h"

namespace {

using namespace pzd;

constexpr int kSieveThreads = 256;
constexpr uint32_t kSieveMaxTris = 8192;  // per object
constexpr int kSieveKeep = 2048;          // kept triangles per side
constexpr int kSieveG = 1280;             // grid-side triangles
constexpr int kSieveRefs = 4096;          // (cell, triangle) references
constexpr int kSieveCells = 512;
constexpr float kCellScale = 1.0f;
constexpr int kU = 2;                     // triangles per object a thread loads at a time
constexpr int kUS = 1;                    // the same in the grid lookups
// defer[] value of a pair left to the trees
constexpr uint32_t kDeferTrees = 1;

struct Tri {
    float3 a, b, c;
};

__device__ __forceinline__ Tri load_tri(const SieveSet& S, uint32_t t) {
    const uint32_t i0 = S.idx[3ull * t], i1 = S.idx[3ull * t + 1], i2 = S.idx[3ull * t + 2];
    return Tri{vtx(S.v, i0), vtx(S.v, i1), vtx(S.v, i2)};
}

// Corners of U triangles ids[0..U) with every index load issued before the
// corner loads, so a thread has all of them in flight at once.
template <int U>
__device__ __forceinline__ void load_tris(const SieveSet& S, const uint32_t* ids, Tri* t) {
    uint32_t ix[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        ix[u][0] = S.idx[3ull * ids[u]];
        ix[u][1] = S.idx[3ull * ids[u] + 1];
        ix[u][2] = S.idx[3ull * ids[u] + 2];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = Tri{vtx(S.v, ix[u][0]), vtx(S.v, ix[u][1]), vtx(S.v, ix[u][2])};
}

__device__ __forceinline__ void tri_box(const Tri& t, float* lo, float* hi) {
    lo[0] = fminf(t.a.x, fminf(t.b.x, t.c.x));
    lo[1] = fminf(t.a.y, fminf(t.b.y, t.c.y));
    lo[2] = fminf(t.a.z, fminf(t.b.z, t.c.z));
    hi[0] = fmaxf(t.a.x, fmaxf(t.b.x, t.c.x));
    hi[1] = fmaxf(t.a.y, fmaxf(t.b.y, t.c.y));
    hi[2] = fmaxf(t.a.z, fmaxf(t.b.z, t.c.z));
}

__device__ __forceinline__ LeafTri as_leaf(const Tri& t) {
    return LeafTri{make_float4(t.a.x, t.a.y, t.a.z, 0.f), make_float4(t.b.x, t.b.y, t.b.z, 0.f),
                   make_float4(t.c.x, t.c.y, t.c.z, 0.f)};
}

// Does the box [lo, hi] meet the ray box [p, (p.x, p.y, +inf)]?
__device__ __forceinline__ bool meets_ray(const float* lo, const float* hi, const float3& p) {
    return lo[0] <= p.x && hi[0] >= p.x && lo[1] <= p.y && hi[1] >= p.y && hi[2] >= p.z;
}

// The +Z parity rays a pair asks for, as the tree path's parity kernel casts
// them: dir 0 from the first vertex of the mesh2 object into the mesh1 object
// (containment; intersection's mesh2->mesh1 direction), dir 1 from the first
// vertex of the mesh1 object into the mesh2 object (intersection's
// mesh1->mesh2 direction).  cast[d]: the ray is wanted and starts inside the
// target's box footprint, below its top (a ray that does not crosses
// nothing).
struct PairRays {
    float3 o[2];
    bool want[2], cast[2];
    __device__ void init(const SieveSet& A, const SieveSet& B, int32_t oa, int32_t ob, uint32_t fa, uint32_t fb,
                         int meta) {
        const int pred = meta & 3, dm = meta >> 2;
        want[0] = pred == 2 || (pred == 1 && (dm & 2));
        want[1] = pred == 1 && (dm & 1);
        o[0] = want[0] ? vtx(B.v, B.idx[3ull * fb]) : make_float3(0.f, 0.f, 0.f);
        o[1] = want[1] ? vtx(A.v, A.idx[3ull * fa]) : make_float3(0.f, 0.f, 0.f);
        for (int d = 0; d < 2; ++d) {
            const SieveSet& T = d == 0 ? A : B;
            const int32_t ot = d == 0 ? oa : ob;
            const float x0 = ord2f(T.obox[6ll * ot]), y0 = ord2f(T.obox[6ll * ot + 1]);
            const float x1 = ord2f(T.obox[6ll * ot + 3]), y1 = ord2f(T.obox[6ll * ot + 4]);
            const float z1 = ord2f(T.obox[6ll * ot + 5]);
            cast[d] = want[d] && o[d].x >= x0 && o[d].x <= x1 && o[d].y >= y0 && o[d].y <= y1 && o[d].z <= z1;
        }
    }
};

// Ray candidates the sieve's first pass records: (pair, triangle | dir << 31)
// for every triangle of a ray's target whose box meets the ray's box.
struct RayList {
    uint2* list;
    unsigned long long* count;
    uint64_t cap;
    uint32_t* overflow;  // per pair: a candidate did not fit
};

// Appends lane's candidate (if any) with one atomic per wave.
__device__ __forceinline__ void append_ray(const RayList& R, bool m, uint32_t p, uint32_t t, int dir) {
    const uint64_t bal = __ballot(m);
    if (!bal) return;
    const int leader = __ffsll((long long)bal) - 1;
    unsigned long long base = 0;
    if (__lane_id() == leader) base = atomicAdd(R.count, (unsigned long long)__popcll(bal));
    base = __shfl((long long)base, leader, PZ_WAVE);
    if (m) {
        const unsigned long long slot = base + lanes_below(bal);
        if (slot < R.cap)
            R.list[slot] = make_uint2(p, t | ((uint32_t)dir << 31));
        else
            R.overflow[p] = 1u;
    }
}

struct Grid {
    float lo[3], inv[3];
    int R[3];
    __device__ int cell(float x, int k) const {
        float f = (x - lo[k]) * inv[k];
        f = fminf(fmaxf(f, 0.f), (float)(R[k] - 1));
        return (int)f;
    }
};

// Grid over O: cells about kCellScale times the mean extent ext_sum / n of the
// binned triangles' boxes per axis, at most 16 per axis and kSieveCells in all.
__device__ Grid make_grid(const float* olo, const float* ohi, const float* ext_sum, int n) {
    Grid g;
    int r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float e = ohi[k] - olo[k], m = ext_sum[k] / (float)n;
        const float q = (e > 0.f && m > 0.f) ? e / (kCellScale * m) : 1.f;
        r[k] = q >= 16.f ? 16 : (q >= 1.f ? (int)q : 1);
    }
    while (r[0] * r[1] * r[2] > kSieveCells) {  // halve the finest axis
        if (r[0] >= r[1] && r[0] >= r[2])
            r[0] = (r[0] + 1) / 2;
        else if (r[1] >= r[2])
            r[1] = (r[1] + 1) / 2;
        else
            r[2] = (r[2] + 1) / 2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float e = ohi[k] - olo[k];
        g.R[k] = r[k];
        g.lo[k] = olo[k];
        g.inv[k] = (r[k] > 1 && e > 0.f) ? (float)r[k] / e : 0.f;
    }
    return g;
}

// Exclusive scan of cnt[0..n) into start[0..n] (start[n] = total), n <= 2 T,
// by the whole workgroup.
template <int T>
__device__ void block_scan(const uint32_t* cnt, uint16_t* start, int n, uint32_t* wsum) {
    constexpr int W = T / PZ_WAVE;
    const int tid = threadIdx.x, lane = __lane_id(), w = tid / PZ_WAVE;
    const int i0 = 2 * tid;
    const uint32_t c0 = i0 < n ? cnt[i0] : 0u, c1 = i0 + 1 < n ? cnt[i0 + 1] : 0u;
    uint32_t x = c0 + c1, incl = x;
    for (int d = 1; d < PZ_WAVE; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, d, PZ_WAVE);
        if (lane >= d) incl += y;
    }
    if (lane == PZ_WAVE - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (int k = 0; k < w; ++k) base += wsum[k];
    const uint32_t excl = base + incl - x;
    if (i0 < n) start[i0] = (uint16_t)excl;
    if (i0 + 1 < n) start[i0 + 1] = (uint16_t)(excl + c0);
    if (tid == T - 1) {
        uint32_t tot = 0;
        for (int k = 0; k < W; ++k) tot += wsum[k];
        start[n] = (uint16_t)tot;
    }
    __syncthreads();
}

template <bool F32>
__global__ __launch_bounds__(kSieveThreads) void k_pair_sieve(
    const uint64_t* __restrict__ pairs, const uint32_t* __restrict__ work, uint32_t* __restrict__ ctr,
    SieveSet A, SieveSet B, const uint8_t* __restrict__ meta, int allow_f32, uint32_t* __restrict__ found,
    uint32_t* __restrict__ defer, RayList rays,
    uint4* __restrict__ esc, unsigned long long* __restrict__ esc_count, uint64_t esc_cap,
    unsigned long long* __restrict__ stats) {
    constexpr int T = kSieveThreads;
    __shared__ uint16_t keepA[kSieveKeep], keepB[kSieveKeep];
    __shared__ float gbox[6][kSieveG];
    __shared__ uint16_t refs[kSieveRefs];
    __shared__ uint16_t cstart[kSieveCells + 1];
    __shared__ uint32_t ccnt[kSieveCells];
    __shared__ uint32_t wsum[T / PZ_WAVE];
    __shared__ int s_nA, s_nB, s_nref, s_hit, s_defer;
    __shared__ float s_ext[3];
    __shared__ Grid s_grid;
    const int tid = threadIdx.x;
    LaneStats st{0, 0, 0, 0, 0};
    uint32_t n_sieved = 0;  // thread 0: pairs decided here
    unsigned long long n_tris = 0;  // and their objects' triangles
    __shared__ uint32_t s_item;
    while (true) {  // pairs of the work list, one at a time
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(&ctr[3], 1u);
        __syncthreads();
        if (s_item >= ctr[1]) break;  // workgroup-uniform
        // a pair handed on by the vertex-run sieve has its ray candidates already
        const uint32_t w = work[s_item], p = w & 0x7FFFFFFFu;
        const bool record_rays = (w >> 31) == 0;
        const uint64_t key = pairs[p];
        const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
        const uint32_t fa = A.first_tri[oa], ca = A.tri_cnt[oa], fb = B.first_tri[ob], cb = B.tri_cnt[ob];
        const bool fits = ca > 0 && cb > 0 && ca <= kSieveMaxTris && cb <= kSieveMaxTris &&
                          A.last_tri[oa] - fa + 1 == ca && B.last_tri[ob] - fb + 1 == cb;
        if (!fits) {  // workgroup-uniform
            if (tid == 0) defer[p] = kDeferTrees;
            continue;
        }
        const int dm = meta[p] >> 2;
        float olo[3], ohi[3];
        #pragma unroll
        for (int k = 0; k < 3; ++k) {
            olo[k] = fmaxf(ord2f(A.obox[6ll * oa + k]), ord2f(B.obox[6ll * ob + k]));
            ohi[k] = fminf(ord2f(A.obox[6ll * oa + 3 + k]), ord2f(B.obox[6ll * ob + 3 + k]));
        }
        PairRays pr;
        pr.init(A, B, oa, ob, fa, fb, meta[p]);
        const bool rays0 = record_rays && pr.cast[0], rays1 = record_rays && pr.cast[1];
        if (tid == 0) {
            s_nA = s_nB = s_nref = s_hit = s_defer = 0;
            s_ext[0] = s_ext[1] = s_ext[2] = 0.f;
        }
        __syncthreads();
        // 1. triangles whose boxes meet O, kU of each object per thread at a
        // time (positions past an object's end reload its first triangle)
        for (uint32_t k0 = tid; k0 < max(ca, cb); k0 += kU * T) {
            uint32_t ia[kU], ib[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t k = k0 + u * T;
                ia[u] = fa + (k < ca ? k : 0u);
                ib[u] = fb + (k < cb ? k : 0u);
            }
            Tri ta[kU], tb[kU];
            if (k0 < ca) load_tris<kU>(A, ia, ta);
            if (k0 < cb) load_tris<kU>(B, ib, tb);
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t k = k0 + u * T;
                float lo[3], hi[3];
                bool ra = false, rb = false;
                if (k < ca) {
                    tri_box(ta[u], lo, hi);
                    if (box_box(lo, hi, olo, ohi)) {
                        const int s = atomicAdd(&s_nA, 1);
                        if (s < kSieveKeep) keepA[s] = (uint16_t)k;
                    }
                    ra = rays0 && meets_ray(lo, hi, pr.o[0]);
                }
                if (k < cb) {
                    tri_box(tb[u], lo, hi);
                    if (box_box(lo, hi, olo, ohi)) {
                        const int s = atomicAdd(&s_nB, 1);
                        if (s < kSieveKeep) keepB[s] = (uint16_t)k;
                    }
                    rb = rays1 && meets_ray(lo, hi, pr.o[1]);
                }
                if (rays0) append_ray(rays, ra, p, fa + k, 0);
                if (rays1) append_ray(rays, rb, p, fb + k, 1);
            }
        }
        __syncthreads();
        const int nA = s_nA, nB = s_nB;
        if (nA > kSieveKeep || nB > kSieveKeep || (nA > 0 && nB > 0 && min(nA, nB) > kSieveG)) {
            if (tid == 0) defer[p] = kDeferTrees;
            __syncthreads();
            continue;
        }
        if (nA > 0 && nB > 0) {
            const bool gA = nA <= nB;  // the grid side
            const SieveSet G = gA ? A : B, S = gA ? B : A;
            const uint16_t* keepG = gA ? keepA : keepB;
            const uint16_t* keepS = gA ? keepB : keepA;
            const uint32_t fG = gA ? fa : fb, fS = gA ? fb : fa;
            const int nG = gA ? nA : nB, nS = gA ? nB : nA;
            // 2. grid-side boxes, mean extent
            float ex[3] = {0.f, 0.f, 0.f};
            for (int i0 = tid; i0 < nG; i0 += kU * T) {
                uint32_t ig[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) ig[u] = fG + keepG[i0 + u * T < nG ? i0 + u * T : 0];
                Tri tg[kU];
                load_tris<kU>(G, ig, tg);
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int i = i0 + u * T;
                    if (i >= nG) continue;
                    float lo[3], hi[3];
                    tri_box(tg[u], lo, hi);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        gbox[k][i] = lo[k];
                        gbox[3 + k][i] = hi[k];
                        ex[k] += hi[k] - lo[k];
                    }
                }
            }
            #pragma unroll
            for (int k = 0; k < 3; ++k) {
                float v = ex[k];
                for (int sft = 32; sft > 0; sft >>= 1) v += __shfl_xor(v, sft, PZ_WAVE);
                if (__lane_id() == 0) atomicAdd(&s_ext[k], v);
            }
            __syncthreads();
            if (tid == 0) s_grid = make_grid(olo, ohi, s_ext, nG);
            __syncthreads();
            const Grid g = s_grid;
            const int ncell = g.R[0] * g.R[1] * g.R[2];
            for (int c = tid; c < ncell; c += T) ccnt[c] = 0u;
            __syncthreads();
            // references per cell
            int nref = 0;
            for (int i = tid; i < nG; i += T) {
                int cl[3], ch[3];
                #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    cl[k] = g.cell(gbox[k][i], k);
                    ch[k] = g.cell(gbox[3 + k][i], k);
                }
                nref += (ch[0] - cl[0] + 1) * (ch[1] - cl[1] + 1) * (ch[2] - cl[2] + 1);
                if (nref > kSieveRefs) break;
                for (int z = cl[2]; z <= ch[2]; ++z)
                    for (int y = cl[1]; y <= ch[1]; ++y)
                        for (int x = cl[0]; x <= ch[0]; ++x) atomicAdd(&ccnt[(z * g.R[1] + y) * g.R[0] + x], 1u);
            }
            if (nref) atomicAdd(&s_nref, nref);
            __syncthreads();
            if (s_nref > kSieveRefs) {  // workgroup-uniform
                if (tid == 0) defer[p] = kDeferTrees;
                __syncthreads();
                continue;
            }
            block_scan<T>(ccnt, cstart, ncell, wsum);
            for (int c = tid; c < ncell; c += T) ccnt[c] = cstart[c];  // cursors
            __syncthreads();
            for (int i = tid; i < nG; i += T) {
                int cl[3], ch[3];
                #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    cl[k] = g.cell(gbox[k][i], k);
                    ch[k] = g.cell(gbox[3 + k][i], k);
                }
                for (int z = cl[2]; z <= ch[2]; ++z)
                    for (int y = cl[1]; y <= ch[1]; ++y)
                        for (int x = cl[0]; x <= ch[0]; ++x)
                            refs[atomicAdd(&ccnt[(z * g.R[1] + y) * g.R[0] + x], 1u)] = (uint16_t)i;
            }
            __syncthreads();
            // 3. the other side's kept triangles against the grid, kUS per
            // thread at a time
            for (int j0 = tid; j0 < nS; j0 += kUS * T) {
                if (*(volatile int*)&s_hit) break;
                uint32_t is[kUS];
#pragma unroll
                for (int u = 0; u < kUS; ++u) is[u] = fS + keepS[j0 + u * T < nS ? j0 + u * T : 0];
                Tri tsv[kUS];
                load_tris<kUS>(S, is, tsv);
                bool hit = false;
#pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    if (j0 + u * T >= nS || hit) continue;
                    const uint32_t ts = is[u];
                    const Tri t = tsv[u];
                    float lo[3], hi[3];
                    tri_box(t, lo, hi);
                    int cl[3], ch[3];
                    #pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        cl[k] = g.cell(lo[k], k);
                        ch[k] = g.cell(hi[k], k);
                    }
                    for (int z = cl[2]; z <= ch[2] && !hit; ++z)
                        for (int y = cl[1]; y <= ch[1] && !hit; ++y)
                            for (int x = cl[0]; x <= ch[0] && !hit; ++x) {
                                const int c = (z * g.R[1] + y) * g.R[0] + x;
                                for (int r = cstart[c]; r < cstart[c + 1]; ++r) {
                                    const int i = refs[r];
                                    const float glo[3] = {gbox[0][i], gbox[1][i], gbox[2][i]};
                                    const float ghi[3] = {gbox[3][i], gbox[4][i], gbox[5][i]};
                                    if (!box_box(lo, hi, glo, ghi)) continue;
                                    // once per pair: in the cell of the larger lower corner
                                    if (max(g.cell(glo[0], 0), cl[0]) != x || max(g.cell(glo[1], 1), cl[1]) != y ||
                                        max(g.cell(glo[2], 2), cl[2]) != z)
                                        continue;
                                    const uint32_t tg = fG + keepG[i];
                                    const Tri gt = load_tri(G, tg);
                                    const LeafTri X = as_leaf(gA ? gt : t), Y = as_leaf(gA ? t : gt);
                                    const uint32_t xa = gA ? tg : ts, yb = gA ? ts : tg;
                                    if constexpr (F32) {  // launched only with allow_f32
                                        const int rr = tri_pair_f32(X, Y, dm, st);
                                        hit = rr > 0;
                                        if (rr < 0) {
                                            const unsigned long long slot = atomicAdd(esc_count, 1ull);
                                            if (slot < esc_cap)
                                                esc[slot] = make_uint4((uint32_t)p, xa, yb, 0u);
                                            else
                                                s_defer = 1;
                                        }
                                    } else {
                                        hit = tri_pair_test<false>(X, Y, dm, false, st);
                                    }
                                    if (hit) break;
                                }
                            }
                }
                if (hit) s_hit = 1;
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (s_defer) {
                defer[p] = kDeferTrees;
            } else {
                ++n_sieved;
                n_tris += ca + cb;
                if (s_hit) found[p] = 1u;
            }
        }
        __syncthreads();
    }
    flush_stats(st, stats);
    if (tid == 0 && n_sieved) {
        atomicAdd(&stats[pzk::kStatSieved], (unsigned long long)n_sieved);
        atomicAdd(&stats[pzk::kStatSieveTris], n_tris);
    }
}

// ------------------------------------------------- sieve over vertex runs
// The same three steps for pairs of objects that own vertex runs of at most
// kRunV vertices.  k_sieve_classify writes each such pair a descriptor (its
// objects' triangle and vertex ranges, the box O, the parity rays it casts,
// and the grid side: the object expected to keep fewer triangles), so a
// workgroup reads one descriptor and then issues every global load of the
// pair at once: both vertex runs (copied to LDS, coalesced) and the first kUR
// triangles per thread of each object's index triples (kept in registers).
// The next work item is fetched while the current pair runs.
//  1. the grid side's triangles whose boxes meet O are kept as 16-bit
//     quantized boxes over O (GridQ) with their numbers in the object;
//  2. they are binned into up to kRunCells cells over O;
//  3. one pass over the other side keeps its triangles whose boxes meet O and
//     whose rows of cells hold references, again as quantized boxes, and the
//     walk tests those against the references with integer box tests.
// Triangle corners for the exact tests are read back through the index
// triples (a few box-meeting pairs per object pair).
constexpr int kRunV = 672;      // vertices per object
constexpr int kRunKeep = 1280;  // kept triangles per side
constexpr int kRunG = 1088;     // grid-side triangles
constexpr int kRunRefs = 6144;
constexpr int kUR = 3;  // measured against 2 and 4 on 1,280-triangle objects
constexpr int kWalkILP = 4;  // grid references a lane tests per step
constexpr int kRunThreads = 512;

// One vertex-run work item (5 x 16 B): x = pair, first triangles (A, B),
// meta byte | cast flags << 8 | grid side (1 = mesh2) << 10; y = first
// vertices (A, B), triangle count | vertex count << 16 (A, B); z, w, v = O's
// corners and the two ray origins.
struct RunDesc {
    uint4 x, y;
    float4 z, w, v;
};
constexpr int kRunDescWords = 5;

__device__ __forceinline__ RunDesc make_run_desc(const SieveSet& A, const SieveSet& B, uint32_t p, int32_t oa,
                                                 int32_t ob, uint8_t meta) {
    RunDesc d;
    const uint32_t fa = A.first_tri[oa], fb = B.first_tri[ob];
    PairRays pr;
    pr.init(A, B, oa, ob, fa, fb, meta);
    float olo[3], ohi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        olo[k] = fmaxf(ord2f(A.obox[6ll * oa + k]), ord2f(B.obox[6ll * ob + k]));
        ohi[k] = fminf(ord2f(A.obox[6ll * oa + 3 + k]), ord2f(B.obox[6ll * ob + 3 + k]));
    }
    const uint32_t va = A.vmin[oa], vb = B.vmin[ob];
    const uint32_t na = A.vmax[oa] - va + 1, nb = B.vmax[ob] - vb + 1;
    // the grid side: the object expected to keep fewer triangles (its triangle
    // count times the share of its box that O covers)
    float va_ = 1.f, vb_ = 1.f, vo = 1.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        va_ *= ord2f(A.obox[6ll * oa + 3 + k]) - ord2f(A.obox[6ll * oa + k]);
        vb_ *= ord2f(B.obox[6ll * ob + 3 + k]) - ord2f(B.obox[6ll * ob + k]);
        vo *= ohi[k] - olo[k];
    }
    const float ka = (float)A.tri_cnt[oa] * (va_ > 0.f ? fminf(1.f, vo / va_) : 1.f);
    const float kb = (float)B.tri_cnt[ob] * (vb_ > 0.f ? fminf(1.f, vo / vb_) : 1.f);
    d.x = make_uint4(p, fa, fb, (uint32_t)meta | (pr.cast[0] ? 0x100u : 0u) | (pr.cast[1] ? 0x200u : 0u) |
                                    (kb < ka ? 0x400u : 0u));
    d.y = make_uint4(va, vb, A.tri_cnt[oa] | (na << 16), B.tri_cnt[ob] | (nb << 16));
    d.z = make_float4(olo[0], olo[1], olo[2], ohi[0]);
    d.w = make_float4(ohi[1], ohi[2], pr.o[0].x, pr.o[0].y);
    d.v = make_float4(pr.o[0].z, pr.o[1].x, pr.o[1].y, pr.o[1].z);
    return d;
}

// Grid of the vertex-run sieve over O in 16-bit fixed point: q = (x - lo) *
// scale per axis, clamped to [0, 65535], lower corners rounded down and upper
// corners up (one step of margin each way), so a quantized box contains the
// part of the triangle's box inside O; cells are (q * R) >> 16.  Boxes and
// cells of both sides come from the same quantization, so two triangles whose
// quantized boxes meet share the cell of the larger lower corner.
constexpr int kRunCells = 2048;
struct GridQ {
    float lo[3], scale[3];
    uint32_t R[3];
    __device__ void quant(const float* blo, const float* bhi, uint32_t* qlo, uint32_t* qhi) const {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float a = (blo[k] - lo[k]) * scale[k], b = (bhi[k] - lo[k]) * scale[k];
            qlo[k] = (uint32_t)fminf(fmaxf(floorf(a) - 1.f, 0.f), 65535.f);
            qhi[k] = (uint32_t)fminf(fmaxf(ceilf(b) + 1.f, 0.f), 65535.f);
        }
    }
    __device__ uint32_t cell(uint32_t q, int k) const { return (q * R[k]) >> 16; }
};

__device__ GridQ make_grid_q(const float* olo, const float* ohi, const float* ext_sum, int n) {
    GridQ g;
    uint32_t r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float e = ohi[k] - olo[k], m = ext_sum[k] / (float)n;
        const float q = (e > 0.f && m > 0.f) ? e / (kCellScale * m) : 1.f;
        r[k] = q >= 16.f ? 16u : (q >= 1.f ? (uint32_t)q : 1u);
    }
    while (r[0] * r[1] * r[2] > (uint32_t)kRunCells) {  // halve the finest axis
        if (r[0] >= r[1] && r[0] >= r[2])
            r[0] = (r[0] + 1) / 2;
        else if (r[1] >= r[2])
            r[1] = (r[1] + 1) / 2;
        else
            r[2] = (r[2] + 1) / 2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float e = ohi[k] - olo[k];
        g.R[k] = r[k];
        g.lo[k] = olo[k];
        g.scale[k] = e > 0.f ? 65535.f / e : 0.f;
    }
    return g;
}

template <bool F32>
__global__ __launch_bounds__(kRunThreads, 4) void k_pair_sieve_runs(
    const RunDesc* __restrict__ rdesc, uint32_t* __restrict__ ctr, SieveSet A, SieveSet B,
    uint32_t* __restrict__ found, uint32_t* __restrict__ defer, uint32_t* __restrict__ gwork, RayList rays,
    uint4* __restrict__ esc, unsigned long long* __restrict__ esc_count, uint64_t esc_cap,
    unsigned long long* __restrict__ stats) {
    constexpr int T = kRunThreads;
    __shared__ float vx[2][kRunV], vy[2][kRunV], vz[2][kRunV];
    // grid-side kept triangles: quantized box (lo | hi << 16 per axis), then
    // lower cells (x | y << 5 | z << 10) | number in the object << 16
    __shared__ uint4 gq[kRunG];
    __shared__ uint16_t refs[kRunRefs];  // grid-side triangle | cell x << 11
    // other-side triangles with references in their cells: quantized box, number
    __shared__ uint4 cq[kRunKeep];
    __shared__ uint32_t wpairs[T / PZ_WAVE][2 * PZ_WAVE];  // per wave: box-meeting pairs to test
    __shared__ uint32_t ccnt[kRunCells + 1];  // counts, then inclusive ends, then starts (ccnt[n] = total)
    __shared__ int s_nG, s_nref, s_ncand, s_next, s_hit, s_defer;
    __shared__ float s_ext[3];
    __shared__ uint32_t s_item;
    const int tid = threadIdx.x, lane = __lane_id();
    LaneStats st{0, 0, 0, 0, 0};
    uint32_t n_sieved = 0;
    unsigned long long n_tris = 0;
    auto corner = [&](int side, uint32_t v) { return make_float3(vx[side][v], vy[side][v], vz[side][v]); };
    const uint32_t nrun = ctr[0];
    if (tid == 0) s_item = atomicAdd(&ctr[2], 1u);
    __syncthreads();
    uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
    while (true) {  // work items (classified by k_sieve_classify), one at a time
        if (item >= nrun) break;  // workgroup-uniform
        uint32_t nxt = 0;  // thread 0: the next item, fetched while this one runs
        if (tid == 0) nxt = atomicAdd(&ctr[2], 1u);
        const RunDesc D = rdesc[item];
        const uint32_t p = D.x.x, fa = D.x.y, fb = D.x.z;
        const int dm = (int)(D.x.w & 0xFFu) >> 2;
        const bool cast[2] = {(D.x.w & 0x100u) != 0, (D.x.w & 0x200u) != 0};
        const int gs = (D.x.w & 0x400u) ? 1 : 0, ss = 1 - gs;  // the grid side, the other side
        const uint32_t va = D.y.x, vb = D.y.y, ca = D.y.z & 0xFFFFu, cb = D.y.w & 0xFFFFu;
        const uint32_t na = D.y.z >> 16, nb = D.y.w >> 16;
        const float olo[3] = {D.z.x, D.z.y, D.z.z}, ohi[3] = {D.z.w, D.w.x, D.w.y};
        const float3 ro[2] = {make_float3(D.w.z, D.w.w, D.v.x), make_float3(D.v.y, D.v.z, D.v.w)};
        // O in 16-bit fixed point (the grid's resolution comes later)
        GridQ g;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float e = ohi[k] - olo[k];
            g.lo[k] = olo[k];
            g.scale[k] = e > 0.f ? 65535.f / e : 0.f;
            g.R[k] = 1;
        }
        // every global load of the pair in flight together: both vertex runs
        // and the first kUR triangles per thread of both objects' index triples
        float3 q[2][2];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const SieveSet& S = side == 0 ? A : B;
            const uint32_t v0 = side == 0 ? va : vb, nv = side == 0 ? na : nb;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t v = tid + h * T;
                q[side][h] = vtx(S.v, v0 + (v < nv ? v : 0u));
            }
        }
        uint32_t ix0[2][kUR][3];  // [0] grid side, [1] the other side
#pragma unroll
        for (int role = 0; role < 2; ++role) {
            const int side = role == 0 ? gs : ss;
            const uint32_t* idx = side == 0 ? A.idx : B.idx;
            const uint32_t f = side == 0 ? fa : fb, cnt = side == 0 ? ca : cb, v0 = side == 0 ? va : vb;
#pragma unroll
            for (int u = 0; u < kUR; ++u) {
                const uint64_t t = (uint64_t)f + (tid + u * T < cnt ? tid + u * T : 0u);
                ix0[role][u][0] = idx[3 * t] - v0;
                ix0[role][u][1] = idx[3 * t + 1] - v0;
                ix0[role][u][2] = idx[3 * t + 2] - v0;
            }
        }
        if (tid == 0) {
            s_nG = s_nref = s_ncand = s_next = s_hit = s_defer = 0;
            s_ext[0] = s_ext[1] = s_ext[2] = 0.f;
        }
        for (int c = tid; c <= kRunCells; c += T) ccnt[c] = 0u;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const uint32_t nv = side == 0 ? na : nb;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t v = tid + h * T;
                if (v < nv) {
                    vx[side][v] = q[side][h].x;
                    vy[side][v] = q[side][h].y;
                    vz[side][v] = q[side][h].z;
                }
            }
        }
        if (tid == 0) s_item = nxt;  // the vertex stores above waited for this atomic already
        __syncthreads();
        const uint32_t item_next = __builtin_amdgcn_readfirstlane(s_item);
        // ends this item on a path without a closing barrier of its own
        auto next_item = [&]() {
            item = item_next;
            __syncthreads();
        };
        // Visits the triangles of one side (index triples of the first chunk
        // from registers), casting the side's parity-ray candidates; fn(k, t,
        // lo, hi, m) gets each triangle with its box and whether it meets O.
        auto each_tri = [&](auto role_c, auto&& fn) {
            constexpr int role = decltype(role_c)::value;
            const int side = role == 0 ? gs : ss;
            const SieveSet& S = side == 0 ? A : B;
            const uint32_t f = side == 0 ? fa : fb, cnt = side == 0 ? ca : cb, v0 = side == 0 ? va : vb;
            for (uint32_t k0 = tid; k0 < cnt; k0 += kUR * T) {
                uint32_t ix[kUR][3];
                if (k0 == (uint32_t)tid) {
#pragma unroll
                    for (int u = 0; u < kUR; ++u) {
                        ix[u][0] = ix0[role][u][0];
                        ix[u][1] = ix0[role][u][1];
                        ix[u][2] = ix0[role][u][2];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < kUR; ++u) {
                        const uint64_t t = (uint64_t)f + (k0 + u * T < cnt ? k0 + u * T : 0u);
                        ix[u][0] = S.idx[3 * t] - v0;
                        ix[u][1] = S.idx[3 * t + 1] - v0;
                        ix[u][2] = S.idx[3 * t + 2] - v0;
                    }
                }
#pragma unroll
                for (int u = 0; u < kUR; ++u) {
                    const uint32_t k = k0 + u * T;
                    bool m = false, rm = false;
                    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
                    if (k < cnt) {
                        const Tri t{corner(side, ix[u][0]), corner(side, ix[u][1]), corner(side, ix[u][2])};
                        tri_box(t, lo, hi);
                        m = box_box(lo, hi, olo, ohi);
                        rm = cast[side] && meets_ray(lo, hi, ro[side]);
                    }
                    if (cast[side]) append_ray(rays, rm, p, f + k, side);
                    fn(k, lo, hi, m);
                }
            }
        };
        // wave-aggregated append of lanes with m to an LDS list (count in *n)
        auto append = [&](bool m, int* n) {
            const uint64_t bal = __ballot(m);
            int pos = -1;
            if (bal) {
                const int leader = __ffsll((long long)bal) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(n, __popcll(bal));
                base = __shfl(base, leader, PZ_WAVE);
                if (m) pos = base + (int)lanes_below(bal);
            }
            return pos;
        };
        // 1. grid-side triangles whose boxes meet O, kept as quantized boxes;
        // their box extents summed (for the grid's resolution)
        {
            float ex[3] = {0.f, 0.f, 0.f};
            each_tri(std::integral_constant<int, 0>{}, [&](uint32_t k, const float* lo, const float* hi, bool m) {
                const int pos = append(m, &s_nG);
                if (m) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) ex[a] += hi[a] - lo[a];
                    if (pos < kRunG) {
                        uint32_t qlo[3], qhi[3];
                        g.quant(lo, hi, qlo, qhi);
                        gq[pos] = make_uint4(qlo[0] | (qhi[0] << 16), qlo[1] | (qhi[1] << 16), qlo[2] | (qhi[2] << 16),
                                             k << 16);
                    }
                }
            });
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float v = ex[a];
                for (int sft = 32; sft > 0; sft >>= 1) v += __shfl_xor(v, sft, PZ_WAVE);
                if (lane == 0 && v != 0.f) atomicAdd(&s_ext[a], v);
            }
        }
        __syncthreads();
        const int nG = s_nG;
        if (nG > kRunG) {  // workgroup-uniform: to the gather sieve, with every ray candidate recorded
            if (cast[ss]) each_tri(std::integral_constant<int, 1>{}, [&](uint32_t, const float*, const float*, bool) {});
            if (tid == 0) gwork[atomicAdd(&ctr[1], 1u)] = p | 0x80000000u;
            next_item();
            continue;
        }
        if (nG == 0) {  // nothing of the grid side meets O: no triangle pair can meet
            if (cast[ss]) each_tri(std::integral_constant<int, 1>{}, [&](uint32_t, const float*, const float*, bool) {});  // its ray candidates
            if (tid == 0) {
                ++n_sieved;
                n_tris += ca + cb;
            }
            next_item();
            continue;
        }
        // 2. grid over O from the grid side's mean kept extent (every thread)
        {
            const GridQ gg = make_grid_q(olo, ohi, s_ext, nG);
            g.R[0] = gg.R[0];
            g.R[1] = gg.R[1];
            g.R[2] = gg.R[2];
        }
        const uint32_t ncell = g.R[0] * g.R[1] * g.R[2];
        auto box_cells = [&](const uint4& b, uint32_t* cl, uint32_t* ch) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t w = k == 0 ? b.x : (k == 1 ? b.y : b.z);
                cl[k] = g.cell(w & 0xFFFFu, k);
                ch[k] = g.cell(w >> 16, k);
            }
        };
        // references per cell; lower cells into the entry
        {
            int nref = 0;
            for (int i = tid; i < nG; i += T) {
                const uint4 b = gq[i];
                uint32_t cl[3], ch[3];
                box_cells(b, cl, ch);
                gq[i].w = (b.w & 0xFFFF0000u) | cl[0] | (cl[1] << 5) | (cl[2] << 10);
                nref += (ch[0] - cl[0] + 1) * (ch[1] - cl[1] + 1) * (ch[2] - cl[2] + 1);
                if (nref > kRunRefs) break;
                for (uint32_t z = cl[2]; z <= ch[2]; ++z)
                    for (uint32_t y = cl[1]; y <= ch[1]; ++y)
                        for (uint32_t x = cl[0]; x <= ch[0]; ++x) atomicAdd(&ccnt[(z * g.R[1] + y) * g.R[0] + x], 1u);
            }
            if (nref) atomicAdd(&s_nref, nref);
        }
        __syncthreads();
        if (s_nref > kRunRefs) {  // workgroup-uniform
            if (tid == 0) defer[p] = kDeferTrees;
            next_item();
            continue;
        }
        // inclusive scan of the counts by the first wave
        if (tid < PZ_WAVE) {
            const uint32_t per = (ncell + PZ_WAVE - 1) / PZ_WAVE;
            uint32_t run = 0;
            for (uint32_t j = 0; j < per; ++j) {
                const uint32_t c = tid * per + j;
                run += c < ncell ? ccnt[c] : 0u;
            }
            uint32_t incl = run;
            for (int d = 1; d < PZ_WAVE; d <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)incl, d, PZ_WAVE);
                if (lane >= d) incl += y;
            }
            uint32_t acc = incl - run;
            for (uint32_t j = 0; j < per; ++j) {
                const uint32_t c = tid * per + j;
                if (c < ncell) {
                    acc += ccnt[c];
                    ccnt[c] = acc;
                }
            }
            if (tid == PZ_WAVE - 1) ccnt[ncell] = incl;
        }
        __syncthreads();
        // scatter: each reference takes the slot below its cell's end, which
        // leaves ccnt[c] = start of cell c
        for (int i = tid; i < nG; i += T) {
            uint32_t cl[3], ch[3];
            box_cells(gq[i], cl, ch);
            for (uint32_t z = cl[2]; z <= ch[2]; ++z)
                for (uint32_t y = cl[1]; y <= ch[1]; ++y)
                    for (uint32_t x = cl[0]; x <= ch[0]; ++x)
                        refs[atomicSub(&ccnt[(z * g.R[1] + y) * g.R[0] + x], 1u) - 1u] = (uint16_t)(i | (x << 11));
        }
        __syncthreads();
        // 3a. the other side's triangles whose boxes meet O and whose rows of
        // cells hold references (most hold none), with their quantized boxes
        each_tri(std::integral_constant<int, 1>{}, [&](uint32_t k, const float* lo, const float* hi, bool m) {
            uint32_t qlo[3], qhi[3];
            bool any = false;
            if (m) {
                g.quant(lo, hi, qlo, qhi);
                uint32_t cl[3], ch[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    cl[a] = g.cell(qlo[a], a);
                    ch[a] = g.cell(qhi[a], a);
                }
                for (uint32_t z = cl[2]; z <= ch[2] && !any; ++z)
                    for (uint32_t y = cl[1]; y <= ch[1] && !any; ++y) {
                        const uint32_t c0 = (z * g.R[1] + y) * g.R[0];
                        any = ccnt[c0 + ch[0] + 1] > ccnt[c0 + cl[0]];  // the row's cells [cl0, ch0]
                    }
            }
            const int pos = append(any, &s_ncand);
            if (any && pos < kRunKeep)
                cq[pos] = make_uint4(qlo[0] | (qhi[0] << 16), qlo[1] | (qhi[1] << 16), qlo[2] | (qhi[2] << 16), k);
        });
        __syncthreads();
        const int ncand = s_ncand;
        if (ncand > kRunKeep) {  // workgroup-uniform
            if (tid == 0) gwork[atomicAdd(&ctr[1], 1u)] = p | 0x80000000u;  // to the gather sieve
            next_item();
            continue;
        }
        if (ncand > 0) {
            const uint32_t fG = gs == 0 ? fa : fb, fS = gs == 0 ? fb : fa;
            const uint32_t vG = gs == 0 ? va : vb, vS = gs == 0 ? vb : va;
            const SieveSet& SG = gs == 0 ? A : B;
            const SieveSet& SS = gs == 0 ? B : A;
            // 3b. lanes walk the rows of cells of those triangles, up to
            // kWalkILP references at a time, a lane taking the next triangle of
            // the list (block-wide counter) when its own is done; the
            // box-meeting pairs, each once, gather in a per-wave list that the
            // wave tests, a pair per lane, whenever 64 are waiting
            uint32_t* wq = wpairs[tid / PZ_WAVE];
            int wn = 0;  // pairs waiting (wave-uniform)
            bool live = false, exhausted = false;
            int jc = 0;
            uint32_t r = 0, rend = 0, cy = 0, cz = 0;
            uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0}, cl[3] = {0, 0, 0}, ch[3] = {0, 0, 0};
            auto test_pairs = [&](int n) {  // wave-uniform n <= 64
                if (lane < n) {
                    const uint32_t e = wq[lane];
                    const uint32_t ts = fS + cq[e >> 16].w, tg = fG + (gq[e & 0xFFFFu].w >> 16);
                    const Tri t{vtx(SS.v, SS.idx[3ull * ts]), vtx(SS.v, SS.idx[3ull * ts + 1]),
                                vtx(SS.v, SS.idx[3ull * ts + 2])};
                    const Tri gt{vtx(SG.v, SG.idx[3ull * tg]), vtx(SG.v, SG.idx[3ull * tg + 1]),
                                 vtx(SG.v, SG.idx[3ull * tg + 2])};
                    (void)vG;
                    (void)vS;
                    const LeafTri X = as_leaf(gs == 0 ? gt : t), Y = as_leaf(gs == 0 ? t : gt);
                    const uint32_t xa = gs == 0 ? tg : ts, yb = gs == 0 ? ts : tg;
                    bool hit;
                    if constexpr (F32) {
                        const int rr = tri_pair_f32(X, Y, dm, st);
                        hit = rr > 0;
                        if (rr < 0) {
                            const unsigned long long slot = atomicAdd(esc_count, 1ull);
                            if (slot < esc_cap)
                                esc[slot] = make_uint4((uint32_t)p, xa, yb, 0u);
                            else
                                s_defer = 1;
                        }
                    } else {
                        hit = tri_pair_test<false>(X, Y, dm, false, st);
                    }
                    if (hit) s_hit = 1;
                }
            };
            while (true) {
                if (__hip_atomic_load(&s_hit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;  // uniform
                const bool need = !live && !exhausted;
                const uint64_t nbal = __ballot(need);
                if (nbal) {
                    const int leader = __ffsll((long long)nbal) - 1;
                    int base = 0;
                    if (lane == leader) base = atomicAdd(&s_next, __popcll(nbal));
                    base = __shfl(base, leader, PZ_WAVE);
                    if (need) {
                        const int qq = base + (int)lanes_below(nbal);
                        if (qq < ncand) {
                            jc = qq;
                            const uint4 b = cq[qq];
                            qlo[0] = b.x & 0xFFFFu;
                            qhi[0] = b.x >> 16;
                            qlo[1] = b.y & 0xFFFFu;
                            qhi[1] = b.y >> 16;
                            qlo[2] = b.z & 0xFFFFu;
                            qhi[2] = b.z >> 16;
                            box_cells(b, cl, ch);
                            cy = cl[1];
                            cz = cl[2];
                            const uint32_t c0 = (cz * g.R[1] + cy) * g.R[0];
                            r = ccnt[c0 + cl[0]];
                            rend = ccnt[c0 + ch[0] + 1];
                            live = true;
                        } else {
                            exhausted = true;
                        }
                    }
                }
                // skip to the next reference: the next non-empty row of the range
                while (live && r >= rend) {
                    if (++cy > ch[1]) {
                        cy = cl[1];
                        if (++cz > ch[2]) {
                            live = false;
                            break;
                        }
                    }
                    const uint32_t c0 = (cz * g.R[1] + cy) * g.R[0];
                    r = ccnt[c0 + cl[0]];
                    rend = ccnt[c0 + ch[0] + 1];
                }
                // up to kWalkILP references of the current row per step, their
                // loads in flight together
                uint32_t gr[kWalkILP];
                bool match[kWalkILP];
                const uint32_t nr = live ? min((uint32_t)kWalkILP, rend - r) : 0u;
#pragma unroll
                for (int u = 0; u < kWalkILP; ++u) gr[u] = (uint32_t)u < nr ? refs[r + u] : 0u;
                r += nr;
#pragma unroll
                for (int u = 0; u < kWalkILP; ++u) {
                    const uint4 b = gq[gr[u] & 0x7FFu];
                    // boxes meet, and once per pair: in the cell of the larger lower corner
                    const bool meet = (qlo[0] <= (b.x >> 16)) & ((b.x & 0xFFFFu) <= qhi[0]) &
                                      (qlo[1] <= (b.y >> 16)) & ((b.y & 0xFFFFu) <= qhi[1]) &
                                      (qlo[2] <= (b.z >> 16)) & ((b.z & 0xFFFFu) <= qhi[2]);
                    const uint32_t gx = b.w & 31u, gy = (b.w >> 5) & 31u, gz = (b.w >> 10) & 31u;
                    match[u] = ((uint32_t)u < nr) & meet & (max(gx, cl[0]) == (gr[u] >> 11)) &
                               (max(gy, cl[1]) == cy) & (max(gz, cl[2]) == cz);
                }
#pragma unroll
                for (int u = 0; u < kWalkILP; ++u) {
                    const uint64_t mb = __ballot(match[u]);
                    if (match[u]) wq[wn + (int)lanes_below(mb)] = ((uint32_t)jc << 16) | (gr[u] & 0x7FFu);
                    wn += __popcll(mb);
                    if (wn >= PZ_WAVE) {
                        __builtin_amdgcn_wave_barrier();
                        test_pairs(PZ_WAVE);
                        const uint32_t rest = lane < wn - PZ_WAVE ? wq[PZ_WAVE + lane] : 0u;
                        __builtin_amdgcn_wave_barrier();
                        if (lane < wn - PZ_WAVE) wq[lane] = rest;
                        wn -= PZ_WAVE;
                    }
                }
                if (__ballot(live || !exhausted) == 0) break;
            }
            if (wn > 0 && !__hip_atomic_load(&s_hit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                __builtin_amdgcn_wave_barrier();
                test_pairs(wn);
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (s_defer) {
                defer[p] = kDeferTrees;
            } else {
                ++n_sieved;
                n_tris += ca + cb;
                if (s_hit) found[p] = 1u;
            }
        }
        if (ncand > 0)
            item = item_next;  // the walk closed with a barrier
        else
            next_item();
    }
    flush_stats(st, stats);
    if (tid == 0 && n_sieved) {
        atomicAdd(&stats[pzk::kStatSieved], (unsigned long long)n_sieved);
        atomicAdd(&stats[pzk::kStatSieveTris], n_tris);
    }
}

// The ray candidates the sieve recorded, thread per candidate: the perturbed
// crossing test of the tree path's parity kernel, crossings counted per
// (pair, direction) for pairs neither touching nor deferred.
__global__ __launch_bounds__(256) void k_sieve_rays(const uint2* __restrict__ list,
                                                   const unsigned long long* __restrict__ count, uint64_t cap,
                                                   const uint64_t* __restrict__ pairs, SieveSet A, SieveSet B,
                                                   const uint32_t* __restrict__ found,
                                                   const uint32_t* __restrict__ defer,
                                                   const uint32_t* __restrict__ overflow,
                                                   uint32_t* __restrict__ cross,
                                                   unsigned long long* __restrict__ stats) {
    const uint64_t n = *count < cap ? *count : cap;
    LaneStats st{0, 0, 0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint2 e = list[i];
        const uint32_t p = e.x, t = e.y & 0x7FFFFFFFu;
        const int dir = (int)(e.y >> 31);
        if (found[p] || defer[p] || overflow[p]) continue;
        const uint64_t key = pairs[p];
        const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
        const SieveSet Tg = dir == 0 ? A : B, Src = dir == 0 ? B : A;
        const float3 pt = vtx(Src.v, Src.idx[3ull * Src.first_tri[dir == 0 ? ob : oa]]);
        const Tri tr = load_tri(Tg, t);
        Esc esc{0};
        if (ray_cross(pt, tr.a, tr.b, tr.c, esc)) atomicAdd(&cross[2ull * p + dir], 1u);
        ++st.tests;
        if (esc.level == 2) ++st.exact;
    }
    flush_stats(st, stats);
    const uint32_t tests = wave_sum(st.tests);
    if (__lane_id() == 0 && tests) atomicAdd(&stats[pzk::kStatParityRays + 4], (unsigned long long)tests);
}

// Parity results of the sieved pairs, thread per (pair, direction): the rays
// the tree path would cast (wanted, pair not touching), their crossings, and
// inside[] bit dir for an odd count.  Pairs whose candidates overflowed are
// left to k_sieve_parity.
__global__ __launch_bounds__(256) void k_sieve_parity_finish(const uint64_t* __restrict__ pairs, uint64_t np,
                                                            const uint8_t* __restrict__ meta,
                                                            const uint32_t* __restrict__ found,
                                                            const uint32_t* __restrict__ defer,
                                                            const uint32_t* __restrict__ overflow,
                                                            const uint32_t* __restrict__ cross,
                                                            uint32_t* __restrict__ inside,
                                                            unsigned long long* __restrict__ stats) {
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t p = g >> 1;
    const int dir = (int)(g & 1);
    bool cast = false;
    uint32_t c = 0;
    if (p < np && !found[p] && !defer[p] && !overflow[p]) {
        const int pred = meta[p] & 3, dm = meta[p] >> 2;
        cast = (pred == 2 && dir == 0) || (pred == 1 && (dm & (dir == 0 ? 2 : 1)));
        if (cast) c = cross[2 * p + dir];
    }
    const bool odd = (c & 1u) != 0;
    if (odd) atomicOr(&inside[p], 1u << dir);
    // rays cast, triangles crossed, rays crossing something, points inside
    const uint32_t pv[4] = {cast ? 1u : 0u, c, c ? 1u : 0u, odd ? 1u : 0u};
    for (int k = 0; k < 4; ++k) {
        const uint32_t sum = wave_sum(pv[k]);
        if (__lane_id() == 0 && sum) atomicAdd(&stats[pzk::kStatParityRays + k], (unsigned long long)sum);
    }
}

// Parity rays of the sieved pairs, wave per (pair, direction), as the tree
// path's parity kernel casts them: direction 0 = first vertex of the mesh2
// object in the mesh1 object (containment; intersection's mesh2->mesh1
// direction), 1 = first vertex of the mesh1 object in the mesh2 object
// (intersection's mesh1->mesh2 direction); only for pairs not proven to
// touch; a ray tests every target triangle whose box meets the ray's box when
// it starts inside the target's box footprint and below its top.
__global__ __launch_bounds__(256) void k_sieve_parity(const uint64_t* __restrict__ pairs, uint64_t np,
                                                     SieveSet A, SieveSet B, const uint8_t* __restrict__ meta,
                                                     const uint32_t* __restrict__ found,
                                                     const uint32_t* __restrict__ defer,
                                                     const uint32_t* __restrict__ only,
                                                     uint32_t* __restrict__ inside,
                                                     unsigned long long* __restrict__ stats) {
    const uint64_t wid = ((uint64_t)blockIdx.x * 256 + threadIdx.x) / PZ_WAVE;
    const uint64_t p = wid >> 1;
    const int dir = (int)(wid & 1);
    const int lane = __lane_id();
    if (p >= np || found[p] || defer[p] || !only[p]) return;  // wave-uniform
    const int pred = meta[p] & 3, dm = meta[p] >> 2;
    if (!((pred == 2 && dir == 0) || (pred == 1 && (dm & (dir == 0 ? 2 : 1))))) return;
    const uint64_t key = pairs[p];
    const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
    const SieveSet Tg = dir == 0 ? A : B, Src = dir == 0 ? B : A;
    const int32_t ot = dir == 0 ? oa : ob, os = dir == 0 ? ob : oa;
    const float3 pt = vtx(Src.v, Src.idx[3ull * Src.first_tri[os]]);
    const float bx0 = ord2f(Tg.obox[6ll * ot]), by0 = ord2f(Tg.obox[6ll * ot + 1]);
    const float bx1 = ord2f(Tg.obox[6ll * ot + 3]), by1 = ord2f(Tg.obox[6ll * ot + 4]);
    const float bz1 = ord2f(Tg.obox[6ll * ot + 5]);
    // a ray starting outside the footprint or above the top crosses nothing
    // (it still counts as cast, as in the tree path)
    const bool inb = pt.x >= bx0 && pt.x <= bx1 && pt.y >= by0 && pt.y <= by1 && pt.z <= bz1;
    LaneStats st{0, 0, 0, 0, 0};
    uint32_t crossings = 0;
    const uint32_t f = Tg.first_tri[ot], n = inb ? Tg.tri_cnt[ot] : 0u;
    for (uint32_t k = lane; k < n; k += PZ_WAVE) {
        const Tri t = load_tri(Tg, f + k);
        float lo[3], hi[3];
        tri_box(t, lo, hi);
        if (!meets_ray(lo, hi, pt)) continue;
        Esc e{0};
        crossings += ray_cross(pt, t.a, t.b, t.c, e) ? 1u : 0u;
        ++st.tests;
        if (e.level == 2) ++st.exact;
    }
    const uint32_t cr = wave_sum(crossings), tests = wave_sum(st.tests);
    const bool odd = (cr & 1u) != 0;
    if (lane == 0 && odd) atomicOr(&inside[p], 1u << dir);
    flush_stats(st, stats);
    // rays cast, triangles crossed, rays crossing something, points inside, tests
    if (lane == 0) {
        const uint32_t pv[5] = {1u, cr, cr ? 1u : 0u, odd ? 1u : 0u, tests};
        for (int k = 0; k < 5; ++k)
            if (pv[k]) atomicAdd(&stats[pzk::kStatParityRays + k], (unsigned long long)pv[k]);
    }
}

// The leaf pairs the sieve's fp32 filter left open, decided with the fp64 and
// exact stages (as k_escalate, with corners read by triangle id).
__global__ __launch_bounds__(256) void k_escalate_tris(const uint4* __restrict__ list,
                                                      const unsigned long long* __restrict__ count, uint64_t cap,
                                                      SieveSet A, SieveSet B, const uint8_t* __restrict__ meta,
                                                      uint32_t* __restrict__ found,
                                                      unsigned long long* __restrict__ stats) {
    const uint64_t n = *count < cap ? *count : cap;
    LaneStats st{0, 0, 0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint4 e = list[i];
        if (found[e.x]) continue;
        const Tri X = load_tri(A, e.y), Y = load_tri(B, e.z);
        const int dm = meta[e.x] >> 2;
        Esc esc{0};
        bool hit = false;
        if (dm & 1)
            hit = seg_tri<false>(X.a, X.b, Y.a, Y.b, Y.c, false, esc) || seg_tri<false>(X.b, X.c, Y.a, Y.b, Y.c, false, esc) ||
                  seg_tri<false>(X.c, X.a, Y.a, Y.b, Y.c, false, esc);
        if (!hit && (dm & 2))
            hit = seg_tri<false>(Y.a, Y.b, X.a, X.b, X.c, false, esc) || seg_tri<false>(Y.b, Y.c, X.a, X.b, X.c, false, esc) ||
                  seg_tri<false>(Y.c, Y.a, X.a, X.b, X.c, false, esc);
        if (esc.level == 2)
            ++st.exact;
        else
            ++st.f64;
        bool naive = false;
        if (dm & 1)
            naive = seg_tri_naive(X.a, X.b, Y.a, Y.b, Y.c) || seg_tri_naive(X.b, X.c, Y.a, Y.b, Y.c) ||
                    seg_tri_naive(X.c, X.a, Y.a, Y.b, Y.c);
        if (!naive && (dm & 2))
            naive = seg_tri_naive(Y.a, Y.b, X.a, X.b, X.c) || seg_tri_naive(Y.b, Y.c, X.a, X.b, X.c) ||
                    seg_tri_naive(Y.c, Y.a, X.a, X.b, X.c);
        if (naive != hit) ++st.mis;
        if (hit) found[e.x] = 1u;
    }
    flush_stats(st, stats);
}

// Work lists of the sieve, thread per candidate pair: pairs of objects with
// contiguous triangles (at most kSieveMaxTris each) go to the vertex-run
// sieve when both objects own runs of at most kRunV vertices, else to the
// gather sieve; the others are deferred to the trees.  ctr = (vertex-run
// pairs, gather pairs, and each list's next-item counter).
__global__ __launch_bounds__(256) void k_sieve_classify(const uint64_t* __restrict__ pairs, uint64_t np, SieveSet A,
                                                       SieveSet B, const uint8_t* __restrict__ meta, int runs,
                                                       RunDesc* __restrict__ rdesc, uint32_t* __restrict__ gwork,
                                                       uint32_t* __restrict__ ctr, uint32_t* __restrict__ defer) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    int cls = -1;  // 0 vertex-run sieve, 1 gather sieve, 2 trees
    if (p < np) {
        const uint64_t key = pairs[p];
        const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
        const uint32_t fa = A.first_tri[oa], ca = A.tri_cnt[oa], fb = B.first_tri[ob], cb = B.tri_cnt[ob];
        const bool tris_fit = ca > 0 && cb > 0 && ca <= kSieveMaxTris && cb <= kSieveMaxTris &&
                              A.last_tri[oa] - fa + 1 == ca && B.last_tri[ob] - fb + 1 == cb;
        cls = tris_fit ? 1 : 2;
        if (tris_fit && runs && A.vmax[oa] - A.vmin[oa] < (uint32_t)kRunV && B.vmax[ob] - B.vmin[ob] < (uint32_t)kRunV)
            cls = 0;
        if (cls == 2) defer[p] = kDeferTrees;
    }
    for (int c = 0; c < 2; ++c) {
        const uint64_t bal = __ballot(cls == c);
        if (!bal) continue;
        const int leader = __ffsll((long long)bal) - 1;
        uint32_t base = 0;
        if (__lane_id() == leader) base = atomicAdd(&ctr[c], (uint32_t)__popcll(bal));
        base = (uint32_t)__shfl((int)base, leader, PZ_WAVE);
        if (cls == c) {
            const uint32_t slot = base + lanes_below(bal);
            if (c == 0) {
                const uint64_t key = pairs[p];
                rdesc[slot] = make_run_desc(A, B, (uint32_t)p, (int32_t)(key >> 32), (int32_t)(uint32_t)key, meta[p]);
            } else {
                gwork[slot] = (uint32_t)p;
            }
        }
    }
}


}  // namespace

namespace pzk {

void pair_sieve(const uint64_t* pairs, uint64_t np, const SieveSet& A, const SieveSet& B, const uint8_t* meta,
                bool f32, int allow_f32, uint32_t* found, uint32_t* defer, uint4* run_desc, uint32_t* gwork,
                uint32_t* ctr, uint2* rays, unsigned long long* nrays, uint64_t rays_cap, uint32_t* ray_overflow,
                uint4* esc, unsigned long long* esc_count, uint64_t esc_cap, unsigned blocks,
                unsigned long long* stats, hipStream_t s) {
    if (np == 0) return;
    const RayList rl{rays, nrays, rays_cap, ray_overflow};
    const bool fp32 = f32 && allow_f32;
    const bool runs = A.vmin && B.vmin;
    KernelTimer kt_("pair_sieve", s);
    static_assert(sizeof(RunDesc) == kRunDescWords * sizeof(uint4), "run descriptor layout");
    RunDesc* rdesc = reinterpret_cast<RunDesc*>(run_desc);
    hipLaunchKernelGGL(k_sieve_classify, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, pairs, np, A, B,
                       meta, runs ? 1 : 0, rdesc, gwork, ctr, defer);
    // persistent grids over the two work lists: pairs of objects with vertex
    // runs first; what they hand on, and pairs without runs, next
    const unsigned gr = (unsigned)std::min<uint64_t>(np, 2ull * blocks);
    if (runs) {
        if (fp32)
            hipLaunchKernelGGL(k_pair_sieve_runs<true>, dim3(gr), dim3(kRunThreads), 0, s, rdesc, ctr, A, B, found,
                               defer, gwork, rl, esc, esc_count, esc_cap, stats);
        else
            hipLaunchKernelGGL(k_pair_sieve_runs<false>, dim3(gr), dim3(kRunThreads), 0, s, rdesc, ctr, A, B, found,
                               defer, gwork, rl, esc, esc_count, esc_cap, stats);
    }
    const unsigned gg = (unsigned)std::min<uint64_t>(np, 4ull * blocks);
    if (fp32)
        hipLaunchKernelGGL(k_pair_sieve<true>, dim3(gg), dim3(kSieveThreads), 0, s, pairs, gwork, ctr, A, B, meta,
                           allow_f32, found, defer, rl, esc, esc_count, esc_cap, stats);
    else
        hipLaunchKernelGGL(k_pair_sieve<false>, dim3(gg), dim3(kSieveThreads), 0, s, pairs, gwork, ctr, A, B, meta,
                           allow_f32, found, defer, rl, esc, esc_count, esc_cap, stats);
}

void sieve_parity(const uint64_t* pairs, uint64_t np, const SieveSet& A, const SieveSet& B, const uint8_t* meta,
                  const uint32_t* found, const uint32_t* defer, const uint2* rays, const unsigned long long* nrays,
                  uint64_t rays_cap, const uint32_t* overflow, uint32_t* cross, uint32_t* inside,
                  unsigned long long* stats, hipStream_t s) {
    if (np == 0) return;
    KernelTimer kt_("sieve_parity", s);
    hipLaunchKernelGGL(k_sieve_rays, dim3(1024), dim3(256), 0, s, rays, nrays, rays_cap, pairs, A, B, found, defer,
                       overflow, cross, stats);
    hipLaunchKernelGGL(k_sieve_parity_finish, dim3((unsigned)((2 * np + 255) / 256)), dim3(256), 0, s, pairs, np,
                       meta, found, defer, overflow, cross, inside, stats);
    const uint64_t waves = 2 * np;  // pairs whose candidates did not fit: every target triangle
    hipLaunchKernelGGL(k_sieve_parity, dim3((unsigned)((waves * PZ_WAVE + 255) / 256)), dim3(256), 0, s, pairs, np, A,
                       B, meta, found, defer, overflow, inside, stats);
}

void escalate_tris(const uint4* list, const unsigned long long* count, uint64_t cap, const SieveSet& A,
                   const SieveSet& B, const uint8_t* meta, uint32_t* found, unsigned long long* stats, hipStream_t s) {
    KernelTimer kt_("escalate_tris", s);
    hipLaunchKernelGGL(k_escalate_tris, dim3(1024), dim3(256), 0, s, list, count, cap, A, B, meta, found, stats);
}

}  // namespace pzk
