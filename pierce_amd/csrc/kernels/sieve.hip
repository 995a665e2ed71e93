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
//     about twice the mean triangle extent, at most 16 per axis and
//     kSieveCells in all.
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
// kRunV vertices: both runs are first copied into LDS (coalesced), so the
// triangles' corners are LDS reads and the only other global reads are the
// objects' index triples, kUR per object per thread in flight at once.  A
// kept triangle is its three run-local vertex numbers and its number in the
// object (8 B); grid-side boxes are recomputed from LDS where needed.
constexpr int kRunV = 672;      // vertices per object
constexpr int kRunKeep = 1280;  // kept triangles per side
constexpr int kRunG = 1088;     // grid-side triangles
constexpr int kRunRefs = 3584;
constexpr int kUR = 4;
constexpr int kWalkILP = 4;  // grid references a lane tests per step
constexpr int kRunThreads = 512;

template <bool F32>
__global__ __launch_bounds__(kRunThreads, 4) void k_pair_sieve_runs(
    const uint64_t* __restrict__ pairs, const uint32_t* __restrict__ work, uint32_t* __restrict__ ctr,
    SieveSet A, SieveSet B, const uint8_t* __restrict__ meta, uint32_t* __restrict__ found,
    uint32_t* __restrict__ defer, uint32_t* __restrict__ gwork, RayList rays, uint4* __restrict__ esc,
    unsigned long long* __restrict__ esc_count, uint64_t esc_cap, unsigned long long* __restrict__ stats) {
    constexpr int T = kRunThreads;
    __shared__ float vx[2][kRunV], vy[2][kRunV], vz[2][kRunV];
    __shared__ uint2 keep[2][kRunKeep];  // (v0 | v1 << 16, v2 | k << 16)
    __shared__ float4 gb0[kRunG];  // grid-side boxes: lo.xyz, hi.x
    __shared__ float2 gb1[kRunG];  // hi.yz
    __shared__ uint16_t refs[kRunRefs];
    __shared__ uint16_t cand[kRunKeep];  // other-side triangles with references in their cells
    __shared__ uint32_t wpairs[T / PZ_WAVE][2 * PZ_WAVE];  // per wave: box-meeting pairs to test
    __shared__ uint16_t cstart[kSieveCells + 1];
    __shared__ uint32_t ccnt[kSieveCells];
    __shared__ uint32_t wsum[T / PZ_WAVE];
    __shared__ int s_n[2], s_nref, s_ncand, s_next, s_hit, s_defer;
    __shared__ float s_ext[3];
    __shared__ Grid s_grid;
    const int tid = threadIdx.x, lane = __lane_id();
    LaneStats st{0, 0, 0, 0, 0};
    uint32_t n_sieved = 0;
    unsigned long long n_tris = 0;
    auto corner = [&](int side, uint32_t v) { return make_float3(vx[side][v], vy[side][v], vz[side][v]); };
    auto kept_tri = [&](int side, uint2 r) {
        return Tri{corner(side, r.x & 0xFFFFu), corner(side, r.x >> 16), corner(side, r.y & 0xFFFFu)};
    };
    __shared__ uint32_t s_item;
    // a pair handed on to the gather sieve
    auto to_gather = [&](uint32_t pp) { gwork[atomicAdd(&ctr[1], 1u)] = pp | 0x80000000u; };
    while (true) {  // pairs of the work list (classified by k_sieve_classify), one at a time
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(&ctr[2], 1u);
        __syncthreads();
        if (s_item >= ctr[0]) break;  // workgroup-uniform
        const uint32_t p = work[s_item];
        const uint64_t key = pairs[p];
        const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
        const uint32_t fa = A.first_tri[oa], ca = A.tri_cnt[oa], fb = B.first_tri[ob], cb = B.tri_cnt[ob];
        const uint32_t va = A.vmin[oa], vb = B.vmin[ob];
        const uint32_t na = A.vmax[oa] - va + 1, nb = B.vmax[ob] - vb + 1;
        const int dm = meta[p] >> 2;
        float olo[3], ohi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            olo[k] = fmaxf(ord2f(A.obox[6ll * oa + k]), ord2f(B.obox[6ll * ob + k]));
            ohi[k] = fminf(ord2f(A.obox[6ll * oa + 3 + k]), ord2f(B.obox[6ll * ob + 3 + k]));
        }
        PairRays pr;
        pr.init(A, B, oa, ob, fa, fb, meta[p]);
        if (tid == 0) {
            s_n[0] = s_n[1] = s_nref = s_ncand = s_next = s_hit = s_defer = 0;
            s_ext[0] = s_ext[1] = s_ext[2] = 0.f;
        }
        // both vertex runs into LDS
        for (uint32_t v = tid; v < na; v += T) {
            const float3 q = vtx(A.v, va + v);
            vx[0][v] = q.x;
            vy[0][v] = q.y;
            vz[0][v] = q.z;
        }
        for (uint32_t v = tid; v < nb; v += T) {
            const float3 q = vtx(B.v, vb + v);
            vx[1][v] = q.x;
            vy[1][v] = q.y;
            vz[1][v] = q.z;
        }
        __syncthreads();
        // 1. triangles whose boxes meet O, appended a wave at a time
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const SieveSet S = side == 0 ? A : B;
            const uint32_t f = side == 0 ? fa : fb, cnt = side == 0 ? ca : cb, v0 = side == 0 ? va : vb;
            for (uint32_t k0 = tid; k0 < cnt; k0 += kUR * T) {
                uint32_t ix[kUR][3];
#pragma unroll
                for (int u = 0; u < kUR; ++u) {
                    const uint64_t t = (uint64_t)f + (k0 + u * T < cnt ? k0 + u * T : 0u);
                    ix[u][0] = S.idx[3 * t] - v0;
                    ix[u][1] = S.idx[3 * t + 1] - v0;
                    ix[u][2] = S.idx[3 * t + 2] - v0;
                }
#pragma unroll
                for (int u = 0; u < kUR; ++u) {
                    const uint32_t k = k0 + u * T;
                    bool m = false, rm = false;
                    if (k < cnt) {
                        const Tri t{corner(side, ix[u][0]), corner(side, ix[u][1]), corner(side, ix[u][2])};
                        float lo[3], hi[3];
                        tri_box(t, lo, hi);
                        m = box_box(lo, hi, olo, ohi);
                        rm = pr.cast[side] && meets_ray(lo, hi, pr.o[side]);
                    }
                    if (pr.cast[side]) append_ray(rays, rm, p, f + k, side);
                    const uint64_t bal = __ballot(m);
                    if (bal) {
                        const int leader = __ffsll((long long)bal) - 1;
                        int base = 0;
                        if (lane == leader) base = atomicAdd(&s_n[side], __popcll(bal));
                        base = __shfl(base, leader, PZ_WAVE);
                        const int pos = base + (int)lanes_below(bal);
                        if (m && pos < kRunKeep)
                            keep[side][pos] = make_uint2(ix[u][0] | (ix[u][1] << 16), ix[u][2] | (k << 16));
                    }
                }
            }
        }
        __syncthreads();
        const int nA = s_n[0], nB = s_n[1];
        if (nA > kRunKeep || nB > kRunKeep) {  // workgroup-uniform
            if (tid == 0) to_gather(p);
            continue;
        }
        if (nA > 0 && nB > 0 && min(nA, nB) > kRunG) {  // workgroup-uniform
            if (tid == 0) to_gather(p);
            continue;
        }
        if (nA > 0 && nB > 0) {
            const int gs = nA <= nB ? 0 : 1, ss = 1 - gs;  // the grid side, the other side
            const int nG = gs == 0 ? nA : nB, nS = gs == 0 ? nB : nA;
            const uint32_t fG = gs == 0 ? fa : fb, fS = gs == 0 ? fb : fa;
            // 2. grid-side boxes and mean extent, grid
            float ex[3] = {0.f, 0.f, 0.f};
            for (int i = tid; i < nG; i += T) {
                float lo[3], hi[3];
                tri_box(kept_tri(gs, keep[gs][i]), lo, hi);
#pragma unroll
                for (int k = 0; k < 3; ++k) ex[k] += hi[k] - lo[k];
                gb0[i] = make_float4(lo[0], lo[1], lo[2], hi[0]);
                gb1[i] = make_float2(hi[1], hi[2]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float v = ex[k];
                for (int sft = 32; sft > 0; sft >>= 1) v += __shfl_xor(v, sft, PZ_WAVE);
                if (lane == 0) atomicAdd(&s_ext[k], v);
            }
            __syncthreads();
            if (tid == 0) s_grid = make_grid(olo, ohi, s_ext, nG);
            __syncthreads();
            const Grid g = s_grid;
            const int ncell = g.R[0] * g.R[1] * g.R[2];
            for (int c = tid; c < ncell; c += T) ccnt[c] = 0u;
            __syncthreads();
            auto cells_of_g = [&](int i, int* cl, int* ch) {
                const float4 b0 = gb0[i];
                const float2 b1 = gb1[i];
                cl[0] = g.cell(b0.x, 0);
                cl[1] = g.cell(b0.y, 1);
                cl[2] = g.cell(b0.z, 2);
                ch[0] = g.cell(b0.w, 0);
                ch[1] = g.cell(b1.x, 1);
                ch[2] = g.cell(b1.y, 2);
            };
            int nref = 0;
            for (int i = tid; i < nG; i += T) {
                int cl[3], ch[3];
                cells_of_g(i, cl, ch);
                nref += (ch[0] - cl[0] + 1) * (ch[1] - cl[1] + 1) * (ch[2] - cl[2] + 1);
                if (nref > kRunRefs) break;
                for (int z = cl[2]; z <= ch[2]; ++z)
                    for (int y = cl[1]; y <= ch[1]; ++y)
                        for (int x = cl[0]; x <= ch[0]; ++x) atomicAdd(&ccnt[(z * g.R[1] + y) * g.R[0] + x], 1u);
            }
            if (nref) atomicAdd(&s_nref, nref);
            __syncthreads();
            if (s_nref > kRunRefs) {  // workgroup-uniform
                if (tid == 0) defer[p] = kDeferTrees;
                __syncthreads();
                continue;
            }
            block_scan<T>(ccnt, cstart, ncell, wsum);
            for (int c = tid; c < ncell; c += T) ccnt[c] = cstart[c];  // cursors
            __syncthreads();
            for (int i = tid; i < nG; i += T) {
                int cl[3], ch[3];
                cells_of_g(i, cl, ch);
                for (int z = cl[2]; z <= ch[2]; ++z)
                    for (int y = cl[1]; y <= ch[1]; ++y)
                        for (int x = cl[0]; x <= ch[0]; ++x)
                            refs[atomicAdd(&ccnt[(z * g.R[1] + y) * g.R[0] + x], 1u)] = (uint16_t)i;
            }
            __syncthreads();
            auto cells_of = [&](const float* lo, const float* hi, int* cl, int* ch) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    cl[k] = g.cell(lo[k], k);
                    ch[k] = g.cell(hi[k], k);
                }
            };
            // 3a. the other side's kept triangles whose cells hold references
            // (most hold none), appended a wave at a time
            for (int j0 = tid; j0 < nS; j0 += T) {
                float lo[3], hi[3];
                tri_box(kept_tri(ss, keep[ss][j0]), lo, hi);
                int cl[3], ch[3];
                cells_of(lo, hi, cl, ch);
                bool any = false;
                for (int z = cl[2]; z <= ch[2] && !any; ++z)
                    for (int y = cl[1]; y <= ch[1] && !any; ++y)
                        for (int x = cl[0]; x <= ch[0] && !any; ++x) {
                            const int c = (z * g.R[1] + y) * g.R[0] + x;
                            any = cstart[c + 1] > cstart[c];
                        }
                const uint64_t bal = __ballot(any);
                if (bal) {
                    const int leader = __ffsll((long long)bal) - 1;
                    int base = 0;
                    if (lane == leader) base = atomicAdd(&s_ncand, __popcll(bal));
                    base = __shfl(base, leader, PZ_WAVE);
                    if (any) cand[base + (int)lanes_below(bal)] = (uint16_t)j0;
                }
            }
            __syncthreads();
            // 3b. lanes walk the references in those triangles' cells one at a
            // time, a lane taking the next triangle of the list (block-wide
            // counter) when its own is done; the box-meeting pairs, each once,
            // gather in a per-wave list that the wave tests, a pair per lane,
            // whenever 64 are waiting
            const int ncand = s_ncand;
            uint32_t* wq = wpairs[tid / PZ_WAVE];
            int wn = 0;  // pairs waiting (wave-uniform)
            bool live = false, exhausted = false;
            int jc = 0, r = 0, rend = 0, cx = 0, cy = 0, cz = 0;
            float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
            int cl[3] = {0, 0, 0}, ch[3] = {0, 0, 0};
            auto test_pairs = [&](int n) {  // wave-uniform n <= 64
                if (lane < n) {
                    const uint32_t e = wq[lane];
                    const uint2 rs = keep[ss][e >> 16], rg = keep[gs][e & 0xFFFFu];
                    const Tri t = kept_tri(ss, rs), gt = kept_tri(gs, rg);
                    const LeafTri X = as_leaf(gs == 0 ? gt : t), Y = as_leaf(gs == 0 ? t : gt);
                    const uint32_t tg = fG + (rg.y >> 16), ts = fS + (rs.y >> 16);
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
                if (*(volatile int*)&s_hit) break;  // one LDS read for the wave: uniform
                const bool need = !live && !exhausted;
                const uint64_t nb = __ballot(need);
                if (nb) {
                    const int leader = __ffsll((long long)nb) - 1;
                    int base = 0;
                    if (lane == leader) base = atomicAdd(&s_next, __popcll(nb));
                    base = __shfl(base, leader, PZ_WAVE);
                    if (need) {
                        const int q = base + (int)lanes_below(nb);
                        if (q < ncand) {
                            jc = cand[q];
                            tri_box(kept_tri(ss, keep[ss][jc]), lo, hi);
                            cells_of(lo, hi, cl, ch);
                            cx = cl[0];
                            cy = cl[1];
                            cz = cl[2];
                            const int c = (cz * g.R[1] + cy) * g.R[0] + cx;
                            r = cstart[c];
                            rend = cstart[c + 1];
                            live = true;
                        } else {
                            exhausted = true;
                        }
                    }
                }
                // skip to the next reference: the next non-empty cell of the range
                while (live && r >= rend) {
                    if (++cx > ch[0]) {
                        cx = cl[0];
                        if (++cy > ch[1]) {
                            cy = cl[1];
                            if (++cz > ch[2]) {
                                live = false;
                                break;
                            }
                        }
                    }
                    const int c = (cz * g.R[1] + cy) * g.R[0] + cx;
                    r = cstart[c];
                    rend = cstart[c + 1];
                }
                // up to kWalkILP references of the current cell per step, their
                // loads in flight together
                int gi[kWalkILP];
                bool match[kWalkILP];
                const int nr = live ? min(kWalkILP, rend - r) : 0;
#pragma unroll
                for (int u = 0; u < kWalkILP; ++u) gi[u] = u < nr ? refs[r + u] : 0;
                r += nr;
#pragma unroll
                for (int u = 0; u < kWalkILP; ++u) {
                    const float4 b0 = gb0[gi[u]];
                    const float2 b1 = gb1[gi[u]];
                    const float glo[3] = {b0.x, b0.y, b0.z}, ghi[3] = {b0.w, b1.x, b1.y};
                    // once per pair: in the cell of the larger lower corner
                    match[u] = u < nr && box_box(lo, hi, glo, ghi) && max(g.cell(glo[0], 0), cl[0]) == cx &&
                               max(g.cell(glo[1], 1), cl[1]) == cy && max(g.cell(glo[2], 2), cl[2]) == cz;
                }
#pragma unroll
                for (int u = 0; u < kWalkILP; ++u) {
                    const uint64_t mb = __ballot(match[u]);
                    if (match[u]) wq[wn + (int)lanes_below(mb)] = ((uint32_t)jc << 16) | (uint32_t)gi[u];
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
            if (wn > 0 && !*(volatile int*)&s_hit) {
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
        __syncthreads();
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
                                                       SieveSet B, int runs, uint32_t* __restrict__ work,
                                                       uint32_t* __restrict__ gwork, uint32_t* __restrict__ ctr,
                                                       uint32_t* __restrict__ defer) {
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
        if (cls == c) (c == 0 ? work : gwork)[base + lanes_below(bal)] = (uint32_t)p;
    }
}

__global__ void k_fill_ctr(uint32_t* ctr) {
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0u;
}

}  // namespace

namespace pzk {

void pair_sieve(const uint64_t* pairs, uint64_t np, const SieveSet& A, const SieveSet& B, const uint8_t* meta,
                bool f32, int allow_f32, uint32_t* found, uint32_t* defer, uint32_t* work, uint32_t* gwork,
                uint32_t* ctr, uint2* rays, unsigned long long* nrays, uint64_t rays_cap, uint32_t* ray_overflow,
                uint4* esc, unsigned long long* esc_count, uint64_t esc_cap, unsigned blocks,
                unsigned long long* stats, hipStream_t s) {
    if (np == 0) return;
    const RayList rl{rays, nrays, rays_cap, ray_overflow};
    const bool fp32 = f32 && allow_f32;
    const bool runs = A.vmin && B.vmin;
    KernelTimer kt_("pair_sieve", s);
    hipLaunchKernelGGL(k_fill_ctr, dim3(1), dim3(64), 0, s, ctr);
    hipLaunchKernelGGL(k_sieve_classify, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, pairs, np, A, B,
                       runs ? 1 : 0, work, gwork, ctr, defer);
    // persistent grids over the two work lists: pairs of objects with vertex
    // runs first; what they hand on, and pairs without runs, next
    const unsigned gr = (unsigned)std::min<uint64_t>(np, 2ull * blocks);
    if (runs) {
        if (fp32)
            hipLaunchKernelGGL(k_pair_sieve_runs<true>, dim3(gr), dim3(kRunThreads), 0, s, pairs, work, ctr, A, B,
                               meta, found, defer, gwork, rl, esc, esc_count, esc_cap, stats);
        else
            hipLaunchKernelGGL(k_pair_sieve_runs<false>, dim3(gr), dim3(kRunThreads), 0, s, pairs, work, ctr, A, B,
                               meta, found, defer, gwork, rl, esc, esc_count, esc_cap, stats);
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
