// Join kernels for gfx950.
//
// Two strategies share these kernels:
//  * two-level (default): object boxes -> object-level LBVH of mesh2 ->
//    broad phase (object pairs whose boxes meet) -> per-object LBVHs of the
//    objects that occur in a pair -> narrow phase: each wave takes a task
//    (pair, direction, chunk of source edges) and stops as soon as any lane
//    proves an intersection for that pair -> +Z parity for containment.
//  * global (the reference's shape, pierce/query/src/cuda/mesh_overlap_edges.cu:105-208):
//    every edge of one set is traced through one BVH over all triangles of
//    the other set and every hit is kept, then deduplicated by sort+unique
//    (two-pass sizing) or by an open-addressing hash set.
// Traversal keeps a per-lane stack in LDS (64 lanes x STACK ints per wave).
#include "pz_device.h"
#include "pz_kernels_join.h"
#include "pz_predicates.h"
#include "pz_narrow.h"

namespace {

using namespace pzd;

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / PZ_WAVE;

inline unsigned grid1(uint64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------ object stats
// Per object: box (ordered-uint atomics), triangle count, first and last
// triangle.  Each wavefront walks a contiguous run of kStatsRun triangles and
// keeps lane-local running sums for the object it is on; when the 64
// triangles of a step all belong to that object they just accumulate, and the
// sums are flushed (wave reduction + one atomic per field) only when the
// object changes.  Objects' triangles are usually contiguous, so a run of
// 1,024 triangles costs a couple of flushes instead of one set of atomics
// per 64 triangles.  A step that crosses a boundary also costs one flush: the
// ending object's lanes join the sums before it, the next object's after it.
constexpr int kStatsRun = 1024;

__global__ __launch_bounds__(kBlock) void k_object_stats(const float* __restrict__ verts, uint64_t nv,
                                                         const uint32_t* __restrict__ idx,
                                                         const int32_t* __restrict__ t2o, uint64_t nt, int32_t nobj,
                                                         uint32_t* __restrict__ obox, uint32_t* __restrict__ tri_cnt,
                                                         uint32_t* __restrict__ first_tri,
                                                         uint32_t* __restrict__ last_tri) {
    const int lane = __lane_id();
    const uint64_t begin = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / PZ_WAVE * kStatsRun;
    const uint64_t end = begin + kStatsRun < nt ? begin + kStatsRun : nt;
    int32_t run_o = -1;  // wave-uniform object of the running sums
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t cnt = 0, tmin = 0xFFFFFFFFu, tmax = 0u;
    auto flush = [&]() {  // whole wave
        if (run_o < 0) return;
        float l[3], h[3];
        for (int a = 0; a < 3; ++a) {
            l[a] = lo[a];
            h[a] = hi[a];
            for (int sft = 32; sft > 0; sft >>= 1) {
                l[a] = fminf(l[a], __shfl_xor(l[a], sft, 64));
                h[a] = fmaxf(h[a], __shfl_xor(h[a], sft, 64));
            }
            lo[a] = INFINITY;
            hi[a] = -INFINITY;
        }
        uint32_t c = cnt, mn = tmin, mx = tmax;
        for (int sft = 32; sft > 0; sft >>= 1) {
            c += (uint32_t)__shfl_xor((int)c, sft, 64);
            mn = min(mn, (uint32_t)__shfl_xor((int)mn, sft, 64));
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, sft, 64));
        }
        if (lane == 0) {
            for (int a = 0; a < 3; ++a) {
                atomicMin(&obox[6ull * run_o + a], f2ord(l[a]));
                atomicMax(&obox[6ull * run_o + 3 + a], f2ord(h[a]));
            }
            atomicAdd(&tri_cnt[run_o], c);
            atomicMin(&first_tri[run_o], mn);
            atomicMax(&last_tri[run_o], mx);
        }
        cnt = 0;
        tmin = 0xFFFFFFFFu;
        tmax = 0u;
        run_o = -1;
    };
    auto add = [&](bool mine, uint64_t t, const float* tl, const float* th) {  // into the running sums
        if (!mine) return;
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], tl[a]);
            hi[a] = fmaxf(hi[a], th[a]);
        }
        ++cnt;
        tmin = min(tmin, (uint32_t)t);
        tmax = max(tmax, (uint32_t)t);
    };
    for (uint64_t step = begin; step < end; step += PZ_WAVE) {
        const uint64_t t = step + lane;
        bool valid = t < end;
        int32_t o = -1;
        float tl[3] = {INFINITY, INFINITY, INFINITY}, th[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (valid) {
            o = t2o[t];
            const uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
            valid = o >= 0 && o < nobj && i0 < nv && i1 < nv && i2 < nv;
            if (valid) {
                const float3 a = vtx(verts, i0), b = vtx(verts, i1), c = vtx(verts, i2);
                tl[0] = fminf(a.x, fminf(b.x, c.x));
                tl[1] = fminf(a.y, fminf(b.y, c.y));
                tl[2] = fminf(a.z, fminf(b.z, c.z));
                th[0] = fmaxf(a.x, fmaxf(b.x, c.x));
                th[1] = fmaxf(a.y, fmaxf(b.y, c.y));
                th[2] = fmaxf(a.z, fmaxf(b.z, c.z));
            }
        }
        const uint64_t vm = __ballot(valid);
        if (vm == 0) continue;
        // objects of the first and the last valid lane (wave-uniform)
        const int32_t o0 = __shfl(o, __ffsll((long long)vm) - 1, 64);
        const int32_t o1 = __shfl(o, 63 - __clzll((long long)vm), 64);
        if (o0 != run_o) {
            flush();
            run_o = o0;
        }
        add(valid && o == o0, t, tl, th);
        if (__ballot(valid && o != o0) == 0) continue;  // one object
        // An object boundary: the first object is complete once its lanes are
        // in the running sums, the last one starts them anew, and lanes of
        // objects wholly inside these 64 triangles use per-lane atomics.
        flush();
        const bool second = o1 != o0;
        if (second) {
            run_o = o1;
            add(valid && o == o1, t, tl, th);
        }
        if (valid && o != o0 && (!second || o != o1)) {
            for (int a = 0; a < 3; ++a) {
                atomicMin(&obox[6ull * o + a], f2ord(tl[a]));
                atomicMax(&obox[6ull * o + 3 + a], f2ord(th[a]));
            }
            atomicAdd(&tri_cnt[o], 1u);
            atomicMin(&first_tri[o], (uint32_t)t);
            atomicMax(&last_tri[o], (uint32_t)t);
        }
    }
    flush();
}

// ------------------------------------------- object bounds over vertex runs
// A dataset whose objects own disjoint, gap-free vertex runs (every vertex of
// object o's run [vbeg, vend] is referenced by o's triangles and by no other
// object's) has per-object boxes that are the boxes of those runs, so the
// per-join bounds read each vertex once (12 B) instead of every triangle's
// index triple, object id and three gathered corners.  The runs, and the
// per-object triangle ranges, are topology: they are found once, at upload,
// by the two kernels below and checked on the host.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int sft = 32; sft > 0; sft >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, sft, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int sft = 32; sft > 0; sft >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, sft, 64));
    return v;
}

// Per object: first and last triangle, triangle count, least and greatest
// vertex index; per vertex: the object referencing it (a second object sets
// the clash flag).  Thread per triangle; a wave whose valid triangles share
// one object reduces before its atomics.
__global__ __launch_bounds__(kBlock) void k_topo_tris(const uint32_t* __restrict__ idx, const int32_t* __restrict__ t2o,
                                                     uint64_t nt, int32_t nobj, uint32_t* __restrict__ first,
                                                     uint32_t* __restrict__ last, uint32_t* __restrict__ cnt,
                                                     uint32_t* __restrict__ vmin, uint32_t* __restrict__ vmax,
                                                     int32_t* __restrict__ owner, uint32_t* __restrict__ clash) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const int32_t o = t < nt ? t2o[t] : -1;
    const bool valid = o >= 0 && o < nobj;
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    bool cl = false;
    if (valid) {
        for (int k = 0; k < 3; ++k) {
            const uint32_t i = idx[3 * t + k];
            mn = min(mn, i);
            mx = max(mx, i);
            const int32_t old = atomicCAS(&owner[i], -1, o);
            cl |= old != -1 && old != o;
        }
    }
    if (cl) atomicOr(clash, 1u);
    const uint64_t vm = __ballot(valid);
    if (vm == 0) return;
    const int32_t o0 = __shfl(o, __ffsll((long long)vm) - 1, 64);
    if (__ballot(valid && o != o0) == 0) {
        const uint32_t tmn = wave_min_u32(valid ? (uint32_t)t : 0xFFFFFFFFu);
        const uint32_t tmx = wave_max_u32(valid ? (uint32_t)t : 0u);
        const uint32_t vmn = wave_min_u32(mn), vmx = wave_max_u32(mx);
        if (__lane_id() == 0) {
            atomicMin(&first[o0], tmn);
            atomicMax(&last[o0], tmx);
            atomicAdd(&cnt[o0], (uint32_t)__popcll(vm));
            atomicMin(&vmin[o0], vmn);
            atomicMax(&vmax[o0], vmx);
        }
    } else if (valid) {
        atomicMin(&first[o], (uint32_t)t);
        atomicMax(&last[o], (uint32_t)t);
        atomicAdd(&cnt[o], 1u);
        atomicMin(&vmin[o], mn);
        atomicMax(&vmax[o], mx);
    }
}

// Vertices some object references.
__global__ __launch_bounds__(kBlock) void k_topo_verts(const int32_t* __restrict__ owner, uint64_t nv,
                                                      unsigned long long* __restrict__ referenced) {
    const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = wave_sum((uint32_t)(v < nv && owner[v] >= 0));
    if (__lane_id() == 0 && n) atomicAdd(referenced, (unsigned long long)n);
}

// Wave per chunk (object, first vertex) of at most kRunChunk vertices of one
// run: ordered-uint minima and maxima (a total order, so the result does not
// depend on the order of the reduction), one set of atomics per chunk.
constexpr uint32_t kRunChunk = 4096;
constexpr int kBoundsBatch = 12;  // measured against 8 and 16 on 642-vertex objects

__global__ __launch_bounds__(kBlock) void k_object_bounds_runs(const float* __restrict__ verts,
                                                              const uint2* __restrict__ chunks, uint64_t nch,
                                                              const uint32_t* __restrict__ vmax,
                                                              uint32_t* __restrict__ obox) {
    const uint64_t c = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / PZ_WAVE;
    if (c >= nch) return;  // wave-uniform
    const uint2 ch = chunks[c];
    const uint32_t o = ch.x, v0 = ch.y;
    const uint32_t v1 = min(v0 + kRunChunk - 1u, vmax[o]);  // last vertex of the chunk
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    // kBoundsBatch vertex loads per lane in flight at once (a chunk of 768
    // vertices in one round); past the chunk's end a lane repeats its last vertex
    for (uint32_t vb = v0 + __lane_id(); vb <= v1; vb += kBoundsBatch * PZ_WAVE) {
        float3 p[kBoundsBatch];
#pragma unroll
        for (int j = 0; j < kBoundsBatch; ++j) p[j] = vtx(verts, min(vb + j * PZ_WAVE, v1));
#pragma unroll
        for (int j = 0; j < kBoundsBatch; ++j) {
            const uint32_t q[3] = {f2ord(p[j].x), f2ord(p[j].y), f2ord(p[j].z)};
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], q[a]);
                hi[a] = max(hi[a], q[a]);
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min_u32(lo[a]);
        hi[a] = wave_max_u32(hi[a]);
    }
    if (__lane_id() == 0)
        for (int a = 0; a < 3; ++a) {
            atomicMin(&obox[6ull * o + a], lo[a]);
            atomicMax(&obox[6ull * o + 3 + a], hi[a]);
        }
}

__global__ __launch_bounds__(kBlock) void k_edge_count(const int32_t* __restrict__ eo, uint64_t ne, int32_t nobj,
                                                       uint32_t* __restrict__ edge_cnt) {
    uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    int32_t o = e < ne ? eo[e] : -1;
    bool valid = o >= 0 && o < nobj;
    uint64_t vm = __ballot(valid);
    if (vm == 0) return;
    int leader = __ffsll((long long)vm) - 1;
    int32_t o0 = __shfl(o, leader, 64);
    if (__ballot(valid && o == o0) == vm) {
        if (__lane_id() == leader) atomicAdd(&edge_cnt[o0], (uint32_t)__popcll(vm));
    } else if (valid) {
        atomicAdd(&edge_cnt[o], 1u);
    }
}

// Union of object boxes (ordered uints) into scene[6].
// Scene box of the object boxes: a grid-stride loop over objects, then wave
// and block reductions, so each block issues one atomic per bound.
constexpr unsigned kSceneBlocks = 64;
__global__ __launch_bounds__(kBlock) void k_scene_box(const uint32_t* __restrict__ obox, int32_t nobj,
                                                      uint32_t* __restrict__ scene) {
    __shared__ uint32_t part[kBlock / PZ_WAVE][6];
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (int32_t o = blockIdx.x * kBlock + threadIdx.x; o < nobj; o += gridDim.x * kBlock)
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], obox[6ll * o + a]);
            v[3 + a] = max(v[3 + a], obox[6ll * o + 3 + a]);
        }
    for (int a = 0; a < 3; ++a)
        for (int sft = 32; sft > 0; sft >>= 1) {
            v[a] = min(v[a], (uint32_t)__shfl_xor((int)v[a], sft, 64));
            v[3 + a] = max(v[3 + a], (uint32_t)__shfl_xor((int)v[3 + a], sft, 64));
        }
    const int w = threadIdx.x / PZ_WAVE;
    if (__lane_id() == 0)
        for (int a = 0; a < 6; ++a) part[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / PZ_WAVE; ++k)
            for (int a = 0; a < 3; ++a) {
                v[a] = min(v[a], part[k][a]);
                v[3 + a] = max(v[3 + a], part[k][3 + a]);
            }
        for (int a = 0; a < 3; ++a) {
            atomicMin(&scene[a], v[a]);
            atomicMax(&scene[3 + a], v[3 + a]);
        }
    }
}

__device__ __forceinline__ void load_scene(const uint32_t* scene, float* lo, float* inv) {
    for (int a = 0; a < 3; ++a) {
        float l = ord2f(scene[a]), h = ord2f(scene[3 + a]);
        if (!(l <= h)) {  // empty scene
            l = 0.f;
            h = 0.f;
        }
        lo[a] = l;
        float ext = h - l;
        inv[a] = ext > 0.f ? 1.0f / ext : 0.0f;
    }
}

// Object keys: (part << 30) | Morton code of the box centre inside the scene box.
__global__ __launch_bounds__(kBlock) void k_object_keys(const uint32_t* __restrict__ obox, int32_t nobj,
                                                        const uint32_t* __restrict__ scene,
                                                        const uint32_t* __restrict__ part_of,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    int32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= nobj) return;
    float lo[3], inv[3];
    load_scene(scene, lo, inv);
    float c[3];
    bool empty = false;
    for (int a = 0; a < 3; ++a) {
        float l = ord2f(obox[6ll * o + a]), h = ord2f(obox[6ll * o + 3 + a]);
        empty |= !(l <= h);
        c[a] = 0.5f * l + 0.5f * h;
    }
    const uint64_t part = part_of ? (uint64_t)part_of[o] : 0ull;
    keys[o] = (part << 30) | (empty ? 0ull : morton30(c[0], c[1], c[2], lo[0], lo[1], lo[2], inv[0], inv[1], inv[2]));
    vals[o] = (uint32_t)o;
}

// --------------------------------------------------------------- broad phase
// A wave per mesh1 object walks the mesh2 object BVH of its part with its box:
// lane L takes the subtree reached by the bits of L from the part's root
// (kSplitDepth levels; a leaf met on the way belongs to the lane whose
// remaining bits are zero), so an object whose box meets thousands of boxes
// (a long branching tube) is walked by 64 lanes, not one.  The lanes step in
// lockstep, one node each, and place each step's hits by ballot prefix, so
// the output order is fixed by the tree.
// Mode 0: count pairs per mesh1 object into cnt[a]; 1: write them at off[a],
// those below cap (the append form is a count, a scan and this write).
constexpr int kSplitDepth = 6;  // 2^6 = PZ_WAVE subtrees
template <int Mode>
__global__ __launch_bounds__(kBlock) void k_broad(const uint32_t* __restrict__ oboxA, int32_t nA,
                                                  const uint32_t* __restrict__ oboxB,
                                                  const Node* __restrict__ nodesB, const uint32_t* __restrict__ orderB,
                                                  int32_t nB, const uint32_t* __restrict__ part_of_a,
                                                  const uint32_t* __restrict__ boff, uint32_t* __restrict__ cnt,
                                                  const uint64_t* __restrict__ off, uint64_t* __restrict__ out,
                                                  uint64_t cap) {
    static_assert(Mode == 0 || Mode == 1, "count or write");
    static_assert((1 << kSplitDepth) == PZ_WAVE, "one subtree per lane");
    const int32_t a = (int32_t)(blockIdx.x * kWavesPerBlock + threadIdx.x / PZ_WAVE);
    if (a >= nA) return;  // wave-uniform
    const int lane = __lane_id();
    // mesh2 objects of this object's part occupy sorted positions [b0, b0 + nb)
    const uint32_t part = part_of_a ? part_of_a[a] : 0u;
    const int32_t b0 = boff ? (int32_t)boff[part] : 0;
    const int32_t nb = boff ? (int32_t)(boff[part + 1] - boff[part]) : nB;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = ord2f(oboxA[6ll * a + k]);
        hi[k] = ord2f(oboxA[6ll * a + 3 + k]);
    }
    if (!(nb > 0 && lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2])) {  // wave-uniform
        if (Mode == 0 && lane == 0) cnt[a] = 0u;
        return;
    }
    // Visits the part's tree, handing each step's hits of every lane to step():
    // (v0, e0) and (v1, e1) are a lane's first and second hit of the step.
    auto walk = [&](auto&& step) {
        if (nb == 1) {  // single object: no internal node; test its box directly
            bool v = false;
            int32_t b = 0;
            if (lane == 0) {
                b = (int32_t)orderB[b0];
                const uint32_t* bb = oboxB + 6ll * b;
                v = boxes_meet(lo, hi, ord2f(bb[0]), ord2f(bb[1]), ord2f(bb[2]), ord2f(bb[3]), ord2f(bb[4]),
                               ord2f(bb[5]));
            }
            step(v, b, false, 0);
            return;
        }
        // descent to this lane's subtree
        bool alive = true;
        int cur = b0;
        for (int d = 0; d < kSplitDepth; ++d) {
            bool v = false;
            int32_t b = 0;
            if (alive) {
                const Node nd = nodesB[cur];
                const int bit = (lane >> (kSplitDepth - 1 - d)) & 1;
                const bool h = bit ? boxes_meet(lo, hi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w)
                                   : boxes_meet(lo, hi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
                const int ch = bit ? nd.d.y : nd.d.x;
                if (!h) {
                    alive = false;
                } else if (ch < 0) {  // a leaf above the split depth: one lane reports it
                    alive = false;
                    if ((lane & ((1 << (kSplitDepth - 1 - d)) - 1)) == 0) {
                        v = true;
                        b = (int32_t)orderB[~ch];
                    }
                } else {
                    cur = ch;
                }
            }
            step(v, b, false, 0);
        }
        // the subtree below cur (its own box meets), depth first
        int stack[64];
        int sp = 0;
        while (__ballot(alive) != 0) {
            bool v0 = false, v1 = false;
            int32_t e0 = 0, e1 = 0;
            if (alive) {
                const Node nd = nodesB[cur];
                bool h0 = boxes_meet(lo, hi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
                bool h1 = boxes_meet(lo, hi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
                const int c0 = nd.d.x, c1 = nd.d.y;
                if (h0 && c0 < 0) {
                    v0 = true;
                    e0 = (int32_t)orderB[~c0];
                    h0 = false;
                }
                if (h1 && c1 < 0) {
                    v1 = true;
                    e1 = (int32_t)orderB[~c1];
                    h1 = false;
                }
                if (h0 && h1) {
                    stack[sp++] = c1;
                    cur = c0;
                } else if (h0) {
                    cur = c0;
                } else if (h1) {
                    cur = c1;
                } else if (sp == 0) {
                    alive = false;
                } else {
                    cur = stack[--sp];
                }
            }
            step(v0, e0, v1, e1);
        }
    };
    // the wave's hits: counted, or written from base in step order and lane order
    auto count = [&]() {
        uint32_t n = 0;
        walk([&](bool v0, int32_t, bool v1, int32_t) { n += (uint32_t)v0 + (uint32_t)v1; });
        for (int sft = 32; sft > 0; sft >>= 1) n += (uint32_t)__shfl_xor((int)n, sft, PZ_WAVE);
        return n;
    };
    auto write = [&](uint64_t base, uint64_t limit) {
        walk([&](bool v0, int32_t e0, bool v1, int32_t e1) {
            const uint64_t m0 = __ballot(v0), m1 = __ballot(v1);
            const uint32_t t0 = (uint32_t)__popcll(m0);
            const uint64_t p0 = base + lanes_below(m0), p1 = base + t0 + lanes_below(m1);
            if (v0 && p0 < limit) out[p0] = pack_pair(a, e0);
            if (v1 && p1 < limit) out[p1] = pack_pair(a, e1);
            base += t0 + (uint32_t)__popcll(m1);
        });
    };
    if constexpr (Mode == 0) {
        const uint32_t n = count();
        if (lane == 0) cnt[a] = n;
    } else {
        const uint64_t base = off[a];
        if (base < cap) write(base, cap);
    }
}

// ------------------------------------------------ broad phase over a hashed grid
// Cells are kGridCellScale times the mesh2 objects' mean box extent per axis, on an unbounded
// integer lattice hashed into nb (a power of two) buckets, so the grid needs
// no scene box and spreads dense clusters as well as sparse sets.  A mesh2
// object is referenced in every cell its box overlaps (bucket of the cell,
// entry = object and cell), or, when that is more than kGridSpan cells, kept
// in a list every query tests.  A mesh1 object visits the cells of its box and
// accepts a referenced object once: in the cell of the larger of the two
// boxes' lower corners (cells are a monotone function of the coordinate, so
// two meeting boxes share that cell); entries of other cells that share the
// bucket are skipped by their cell.
constexpr uint32_t kGridSpan = 27;    // cells a referenced mesh2 box may overlap
// Cell edge over the mean box extent: each reference costs an atomic and a
// scattered write, while larger cells only add box tests where objects are
// dense (measured 1, 1.5, 2, 3 and 4 on 100k vs 100k sparse ellipsoids).
constexpr float kGridCellScale = 3.f;
constexpr uint64_t kQuerySpan = 4096; // cells beyond which a mesh1 box tests every mesh2 box
constexpr int32_t kCellClamp = 1 << 20;

__device__ __forceinline__ uint32_t cell_hash(int32_t x, int32_t y, int32_t z, uint32_t nb) {
    return (((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u)) & (nb - 1u);
}

struct GridCells {
    float inv[3];
    __device__ int32_t cell(float x, int k) const {
        const float f = floorf(x * inv[k]);
        return (int32_t)fminf(fmaxf(f, (float)-kCellClamp), (float)kCellClamp);
    }
};

// stats[0..2]: per-axis sums of the extents of the non-empty mesh2 boxes, stats[3]: their number.
__device__ __forceinline__ GridCells grid_cells(const float* __restrict__ stats) {
    GridCells g;
    const float n = fmaxf(stats[3], 1.f);
    float m[3], mx = 0.f;
    for (int k = 0; k < 3; ++k) {
        m[k] = stats[k] / n;
        mx = fmaxf(mx, m[k]);
    }
    for (int k = 0; k < 3; ++k) {
        // a flat axis takes the largest mean extent (or 1 when every box is a point)
        const float h = m[k] > 0.f ? m[k] : (mx > 0.f ? mx : 1.f);
        g.inv[k] = 1.f / (kGridCellScale * h);
    }
    return g;
}

__device__ __forceinline__ bool load_box(const uint32_t* __restrict__ obox, int32_t o, float* lo, float* hi) {
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
        lo[k] = ord2f(obox[6ll * o + k]);
        hi[k] = ord2f(obox[6ll * o + 3 + k]);
        ok &= lo[k] <= hi[k];
    }
    return ok;
}

// Sums of the non-empty mesh2 boxes' extents (a grid-stride loop, wave and
// block reductions, four atomics per block).
__global__ __launch_bounds__(kBlock) void k_grid_stats(const uint32_t* __restrict__ obox, int32_t n,
                                                       float* __restrict__ stats) {
    __shared__ float part[kBlock / PZ_WAVE][4];
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    for (int32_t o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) {
        float lo[3], hi[3];
        if (load_box(obox, o, lo, hi)) {
            for (int k = 0; k < 3; ++k) v[k] += hi[k] - lo[k];
            v[3] += 1.f;
        }
    }
    const int w = threadIdx.x / PZ_WAVE;
    for (int k = 0; k < 4; ++k) {
        const float t = wave_sum(v[k]);
        if (__lane_id() == 0) part[w][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        float t = 0.f;
        for (int i = 0; i < kBlock / PZ_WAVE; ++i) t += part[i][threadIdx.x];
        atomicAdd(&stats[threadIdx.x], t);
    }
}

constexpr int kGridLanes = 8;  // lanes per object of the grid's reference and query kernels

// kGridLanes lanes per mesh2 object.  Mode 0: count references per bucket
// (and list the oversized boxes); mode 1: place each reference in its
// bucket's range [start, start + count), counting the bucket down (the counts
// end at zero).
template <int Mode>
__global__ __launch_bounds__(kBlock) void k_grid_refs(const uint32_t* __restrict__ obox, int32_t n,
                                                      const float* __restrict__ stats, uint32_t nb,
                                                      uint32_t* __restrict__ counts, const uint32_t* __restrict__ start,
                                                      uint4* __restrict__ entries, float4* __restrict__ ebox,
                                                      uint32_t* __restrict__ big, uint32_t* __restrict__ nbig) {
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const int32_t o = (int32_t)(gid / kGridLanes);
    const uint32_t sub = (uint32_t)(gid % kGridLanes);
    if (o >= n) return;
    float lo[3], hi[3];
    if (!load_box(obox, o, lo, hi)) return;
    const GridCells g = grid_cells(stats);
    int32_t cl[3], ch[3];
    uint64_t span = 1;
    for (int k = 0; k < 3; ++k) {
        cl[k] = g.cell(lo[k], k);
        ch[k] = g.cell(hi[k], k);
        span *= (uint64_t)(ch[k] - cl[k] + 1);
    }
    if (span > kGridSpan) {
        if (Mode == 0 && sub == 0) big[atomicAdd(nbig, 1u)] = (uint32_t)o;
        return;
    }
    const uint32_t rx = (uint32_t)(ch[0] - cl[0] + 1), ry = (uint32_t)(ch[1] - cl[1] + 1);
    for (uint32_t i = sub; i < (uint32_t)span; i += kGridLanes) {
        const int32_t x = cl[0] + (int32_t)(i % rx), y = cl[1] + (int32_t)((i / rx) % ry), z = cl[2] + (int32_t)(i / (rx * ry));
        const uint32_t b = cell_hash(x, y, z, nb);
        if (Mode == 0) {
            atomicAdd(&counts[b], 1u);
        } else {
            const uint32_t e = start[b] + atomicSub(&counts[b], 1u) - 1u;
            entries[e] = make_uint4((uint32_t)o, (uint32_t)x, (uint32_t)y, (uint32_t)z);
            ebox[2ull * e] = make_float4(lo[0], lo[1], lo[2], hi[0]);
            ebox[2ull * e + 1] = make_float4(hi[1], hi[2], 0.f, 0.f);
        }
    }
}

// The grid's zeroes: bucket counts (nb + 1), extent sums, oversized count.
__global__ void k_grid_clear(uint32_t* __restrict__ counts, uint64_t n, float* __restrict__ stats,
                             uint32_t* __restrict__ nbig) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) counts[i] = 0u;
    if (i < 4) stats[i] = 0.f;
    if (i == 0) *nbig = 0u;
}

// kGridLanes lanes per mesh1 object: the mesh2 objects whose boxes meet its
// box, each once (modes as k_broad: count, write at off[a], append with
// capacity).  The lanes share the object's cells and the oversized list; each
// tests the entries of its cells' buckets (their boxes stored beside them).
template <int Mode>
__global__ __launch_bounds__(kBlock) void k_grid_broad(const uint32_t* __restrict__ oboxA, int32_t nA,
                                                       const uint32_t* __restrict__ oboxB, int32_t nB,
                                                       const float* __restrict__ stats, uint32_t nb,
                                                       const uint32_t* __restrict__ start,
                                                       const uint4* __restrict__ entries,
                                                       const float4* __restrict__ ebox,
                                                       const uint32_t* __restrict__ big,
                                                       const uint32_t* __restrict__ nbig, uint32_t* __restrict__ cnt,
                                                       const uint64_t* __restrict__ off, uint64_t* __restrict__ out,
                                                       unsigned long long* __restrict__ cursor, uint64_t cap) {
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const int32_t a = (int32_t)(gid / kGridLanes);
    const int sub = (int)(gid % kGridLanes), lane = __lane_id();
    if (Mode != 2 && (int64_t)(gid - sub) >= (int64_t)nA * kGridLanes) return;  // whole groups only
    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {-1.f, -1.f, -1.f};
    const bool active = a < nA && load_box(oboxA, a, lo, hi);
    const GridCells g = grid_cells(stats);
    int32_t cl[3] = {0, 0, 0}, ch[3] = {-1, -1, -1};
    uint64_t span = 0;
    if (active) {
        span = 1;
        for (int k = 0; k < 3; ++k) {
            cl[k] = g.cell(lo[k], k);
            ch[k] = g.cell(hi[k], k);
            span *= (uint64_t)(ch[k] - cl[k] + 1);
        }
    }
    const int32_t rx = ch[0] - cl[0] + 1, ry = ch[1] - cl[1] + 1;
    auto box_meets = [&](const float* blo, const float* bhi) {
        return blo[0] <= hi[0] && bhi[0] >= lo[0] && blo[1] <= hi[1] && bhi[1] >= lo[1] && blo[2] <= hi[2] &&
               bhi[2] >= lo[2];
    };
    auto visit = [&](auto&& emit) {
        if (!active) return;
        float blo[3], bhi[3];
        if (span > kQuerySpan) {  // a box over very many cells: test every referenced mesh2 box
            for (int32_t b = sub; b < nB; b += kGridLanes) {
                if (!load_box(oboxB, b, blo, bhi) || !box_meets(blo, bhi)) continue;
                uint64_t sb = 1;
                for (int k = 0; k < 3; ++k) sb *= (uint64_t)(g.cell(bhi[k], k) - g.cell(blo[k], k) + 1);
                if (sb <= kGridSpan) emit(b);  // the oversized ones follow below
            }
        } else {
            for (uint32_t i = (uint32_t)sub; i < (uint32_t)span; i += kGridLanes) {
                const int32_t x = cl[0] + (int32_t)(i % (uint32_t)rx);
                const int32_t y = cl[1] + (int32_t)((i / (uint32_t)rx) % (uint32_t)ry);
                const int32_t z = cl[2] + (int32_t)(i / ((uint32_t)rx * (uint32_t)ry));
                const uint32_t bk = cell_hash(x, y, z, nb);
                for (uint32_t e = start[bk], e1 = start[bk + 1]; e < e1; ++e) {
                    const uint4 en = entries[e];
                    const float4 b0 = ebox[2ull * e], b1 = ebox[2ull * e + 1];
                    if ((int32_t)en.y != x || (int32_t)en.z != y || (int32_t)en.w != z) continue;
                    blo[0] = b0.x;
                    blo[1] = b0.y;
                    blo[2] = b0.z;
                    bhi[0] = b0.w;
                    bhi[1] = b1.x;
                    bhi[2] = b1.y;
                    if (!box_meets(blo, bhi)) continue;
                    // once per pair: in the cell of the larger lower corner
                    if (max(g.cell(blo[0], 0), cl[0]) != x || max(g.cell(blo[1], 1), cl[1]) != y ||
                        max(g.cell(blo[2], 2), cl[2]) != z)
                        continue;
                    emit((int32_t)en.x);
                }
            }
        }
        const uint32_t nbg = *nbig;
        for (uint32_t i = (uint32_t)sub; i < nbg; i += kGridLanes) {
            const int32_t b = (int32_t)big[i];
            if (load_box(oboxB, b, blo, bhi) && box_meets(blo, bhi)) emit(b);
        }
    };
    // per lane count, then the group's (count) or the wave's (append) prefix
    uint32_t n = 0;
    visit([&](int32_t) { ++n; });
    uint32_t incl = n;  // inclusive prefix over the wave
    for (int sft = 1; sft < PZ_WAVE; sft <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)incl, sft, PZ_WAVE);
        if (lane >= sft) incl += u;
    }
    const int first = lane - sub;  // the group's first lane
    const uint32_t prev = (uint32_t)__shfl((int)incl, first > 0 ? first - 1 : 0, PZ_WAVE);  // every lane takes part
    const uint32_t before = first > 0 ? prev : 0u;
    if (Mode == 0) {
        const uint32_t group = (uint32_t)__shfl((int)incl, first + kGridLanes - 1, PZ_WAVE) - before;
        if (sub == 0 && a < nA) cnt[a] = group;
    } else if (Mode == 1) {
        if (active && n > 0) {
            const uint64_t base = off[a] + (incl - n - before);
            uint32_t i = 0;
            visit([&](int32_t b) { out[base + i++] = pack_pair(a, b); });
        }
    } else {
        const uint32_t total = (uint32_t)__shfl((int)incl, PZ_WAVE - 1, PZ_WAVE);
        unsigned long long wbase = 0ull;
        if (lane == 0 && total > 0) wbase = atomicAdd(cursor, (unsigned long long)total);
        wbase = __shfl(wbase, 0, PZ_WAVE);
        const uint64_t base = wbase + incl - n;
        if (active && n > 0 && base < cap) {
            uint32_t i = 0;
            visit([&](int32_t b) {
                if (base + i < cap) out[base + i] = pack_pair(a, b);
                ++i;
            });
        }
    }
}

// Broad phase without an object index, for small mesh2 sets: each mesh1
// object tests the box of every mesh2 object of its part (the part's objects
// have ids [boff[part], boff[part + 1])).  Same modes and output as k_broad,
// except that one object's pairs are grouped by lane, not by id.
// - S lanes (a power of two dividing the wave) share one mesh1 object; lane j
//   of the group tests tile entries j, j + S, ...
// - The block stages the mesh2 boxes its objects need (a contiguous id range:
//   parts are contiguous in both sets) through LDS, 256 at a time; lanes read
//   them as broadcasts.  All lanes stay for the barriers.
// - Count: the group's sum.  Write: lane j writes at off[a] plus the counts
//   of lanes < j of its group.  Append: one reservation per wave.
template <int Mode, bool Parts, int S>
__global__ __launch_bounds__(kBlock) void k_broad_brute(const uint32_t* __restrict__ oboxA, int32_t nA,
                                                        const uint32_t* __restrict__ oboxB, int32_t nB,
                                                        const uint32_t* __restrict__ part_of_a,
                                                        const uint32_t* __restrict__ boff, uint32_t* __restrict__ cnt,
                                                        const uint64_t* __restrict__ off, uint64_t* __restrict__ out,
                                                        unsigned long long* __restrict__ cursor, uint64_t cap) {
    constexpr int kPer = kBlock / S;  // mesh1 objects per block
    __shared__ float tile[kBlock][6];
    const int lane = __lane_id();
    const int j = (int)threadIdx.x % S;
    const int32_t first = blockIdx.x * kPer;
    const int32_t a = first + (int32_t)threadIdx.x / S;
    const bool in = a < nA;
    int32_t b0 = 0, b1 = nB, bb0 = 0, bb1 = nB;  // this object's and the block's mesh2 ids
    if (Parts) {
        const int32_t last = min(nA, first + kPer) - 1;
        bb0 = (int32_t)boff[part_of_a[first]];
        bb1 = (int32_t)boff[part_of_a[last] + 1];
        const uint32_t part = part_of_a[in ? a : first];
        b0 = (int32_t)boff[part];
        b1 = (int32_t)boff[part + 1];
    }
    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {-1.f, -1.f, -1.f};
    if (in) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = ord2f(oboxA[6ll * a + k]);
            hi[k] = ord2f(oboxA[6ll * a + 3 + k]);
        }
    }
    const bool active = in && b1 > b0 && lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2];
    auto visit = [&](auto&& emit) {
        for (int32_t t0 = bb0; t0 < bb1; t0 += kBlock) {
            __syncthreads();
            const int32_t bl = t0 + (int32_t)threadIdx.x;
            if (bl < bb1)
                for (int k = 0; k < 6; ++k) tile[threadIdx.x][k] = ord2f(oboxB[6ll * bl + k]);
            __syncthreads();
            if (active) {
                const int32_t i0 = max(b0 - t0, 0), i1 = min(b1 - t0, min(kBlock, bb1 - t0));
                for (int32_t i = i0 + j; i < i1; i += S)
                    if (boxes_meet(lo, hi, tile[i][0], tile[i][1], tile[i][2], tile[i][3], tile[i][4], tile[i][5]))
                        emit(t0 + i);
            }
        }
    };
    uint32_t n = 0;
    visit([&](int32_t) { ++n; });
    if (Mode == 0) {
        uint32_t sum = n;
        for (int sft = S / 2; sft > 0; sft >>= 1) sum += (uint32_t)__shfl_xor((int)sum, sft, PZ_WAVE);
        if (in && j == 0) cnt[a] = sum;
        return;
    }
    uint64_t base;
    if (Mode == 1) {
        uint32_t incl = n;  // inclusive prefix within the group
        for (int sft = 1; sft < S; sft <<= 1) {
            const uint32_t u = (uint32_t)__shfl_up((int)incl, sft, PZ_WAVE);
            if (j >= sft) incl += u;
        }
        base = (in ? off[a] : 0ull) + incl - n;
    } else {
        uint32_t incl = n;  // inclusive prefix over the wave
        for (int sft = 1; sft < PZ_WAVE; sft <<= 1) {
            const uint32_t u = (uint32_t)__shfl_up((int)incl, sft, PZ_WAVE);
            if (lane >= sft) incl += u;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, PZ_WAVE - 1, PZ_WAVE);
        unsigned long long wbase = 0ull;
        if (lane == 0 && total > 0) wbase = atomicAdd(cursor, (unsigned long long)total);
        wbase = __shfl(wbase, 0, PZ_WAVE);
        base = wbase + incl - n;
    }
    uint32_t i = 0;
    visit([&](int32_t b) {
        if (Mode == 1 || base + i < cap) out[base + i] = pack_pair(a, b);
        ++i;
    });
}

// Per-pair predicate and direction mask: meta = predicate | dir_mask << 2
// (predicate 0 overlap, 1 intersection, 2 containment; dir bit 0: mesh1 edges
// and points, bit 1: mesh2's).  Batches take them from the pair's part.
__global__ void k_pair_meta(const uint64_t* __restrict__ pairs, uint64_t np, const uint32_t* __restrict__ part_of_a,
                            const uint8_t* __restrict__ meta_of_part, uint8_t uniform, uint8_t* __restrict__ meta) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    meta[i] = part_of_a ? meta_of_part[part_of_a[(uint32_t)(pairs[i] >> 32)]] : uniform;
}

// Per candidate pair: meta (as k_pair_meta) and the zeroes of the per-pair
// flag arrays; the first thread also zeroes the join's single counters.
__global__ void k_pair_prep(const uint64_t* __restrict__ pairs, uint64_t np, const uint32_t* __restrict__ part_of_a,
                            const uint8_t* __restrict__ meta_of_part, uint8_t uniform, uint8_t* __restrict__ meta,
                            pzk::PairPrepZero z) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i <= np)
        for (int k = 0; k < pzk::PairPrepZero::kFlags; ++k)
            if (z.flags[k]) z.flags[k][i] = 0u;
    if (i < np) meta[i] = part_of_a ? meta_of_part[part_of_a[(uint32_t)(pairs[i] >> 32)]] : uniform;
    if (i == 0) {
        for (int k = 0; k < pzk::PairPrepZero::kCounters; ++k)
            if (z.counters[k]) *z.counters[k] = 0ull;
        for (uint32_t w = 0; w < z.nwords; ++w) z.words[w] = 0u;
    }
}

// ------------------------------------------------- needed objects / layout
__global__ void k_mark_needed(const uint64_t* __restrict__ pairs, uint64_t np, uint32_t* __restrict__ needA,
                              uint32_t* __restrict__ needB) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    uint64_t k = pairs[i];
    needA[(uint32_t)(k >> 32)] = 1u;
    needB[(uint32_t)k] = 1u;
}

// loc = exclusive scan of need; fills per-needed-object counts.
__global__ void k_needed_layout(const uint32_t* __restrict__ need, const uint32_t* __restrict__ loc, int32_t nobj,
                                const uint32_t* __restrict__ tri_cnt, const uint32_t* __restrict__ edge_cnt,
                                int32_t* __restrict__ obj_of_loc, uint32_t* __restrict__ seg_tri_cnt,
                                uint32_t* __restrict__ seg_edge_cnt) {
    int32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= nobj || !need[o]) return;
    uint32_t l = loc[o];
    obj_of_loc[l] = o;
    seg_tri_cnt[l] = tri_cnt[o];
    seg_edge_cnt[l] = edge_cnt[o];
}

// Wave-aggregated slot reservation inside object segments.
__device__ __forceinline__ uint32_t reserve_slot(bool valid, uint32_t l, uint32_t* cursor) {
    uint64_t vm = __ballot(valid);
    if (vm == 0) return 0;
    int leader = __ffsll((long long)vm) - 1;
    uint32_t l0 = (uint32_t)__shfl((int)l, leader, 64);
    uint32_t slot = 0;
    if (__ballot(valid && l == l0) == vm) {
        uint32_t base = 0;
        if (__lane_id() == leader) base = atomicAdd(&cursor[l0], (uint32_t)__popcll(vm));
        base = (uint32_t)__shfl((int)base, leader, 64);
        slot = base + lanes_below(vm);
    } else if (valid) {
        slot = atomicAdd(&cursor[l], 1u);
    }
    return slot;
}

// Triangles of needed objects -> (loc << 30 | Morton in object box, triangle).
// Objects whose triangles are contiguous place triangle t at t - first_tri;
// the others reserve slots through per-object cursors.
__global__ __launch_bounds__(kBlock) void k_tri_keys_two_level(
    const float* __restrict__ verts, const uint32_t* __restrict__ idx, const int32_t* __restrict__ t2o, uint64_t nt,
    int32_t nobj, const uint32_t* __restrict__ need, const uint32_t* __restrict__ loc,
    const uint32_t* __restrict__ obox, const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ first_tri,
    const uint32_t* __restrict__ last_tri, const uint32_t* __restrict__ tri_cnt, uint32_t* __restrict__ cursor,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, uint64_t nv, int mbits) {
    uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    int32_t o = t < nt ? t2o[t] : -1;
    bool valid = o >= 0 && o < nobj && need[o];
    uint32_t i0 = 0, i1 = 0, i2 = 0;
    if (valid) {
        i0 = idx[3 * t];
        i1 = idx[3 * t + 1];
        i2 = idx[3 * t + 2];
        valid = i0 < nv && i1 < nv && i2 < nv;
    }
    const bool direct = valid && last_tri[o] - first_tri[o] + 1 == tri_cnt[o];
    uint32_t l = valid ? loc[o] : 0u;
    uint32_t slot = reserve_slot(valid && !direct, l, cursor);
    if (!valid) return;
    if (direct) slot = (uint32_t)t - first_tri[o];
    float3 a = vtx(verts, i0), b = vtx(verts, i1), c = vtx(verts, i2);
    float lo[3], inv[3];
    for (int k = 0; k < 3; ++k) {
        float bl = ord2f(obox[6ll * o + k]), bh = ord2f(obox[6ll * o + 3 + k]);
        lo[k] = bl;
        inv[k] = bh > bl ? 1.0f / (bh - bl) : 0.0f;
    }
    float cx = (a.x + b.x + c.x) * (1.0f / 3.0f), cy = (a.y + b.y + c.y) * (1.0f / 3.0f),
          cz = (a.z + b.z + c.z) * (1.0f / 3.0f);
    uint64_t pos = (uint64_t)seg_off[l] + slot;
    keys[pos] = ((uint64_t)l << mbits) |
                (morton30(cx, cy, cz, lo[0], lo[1], lo[2], inv[0], inv[1], inv[2]) >> (30 - mbits));
    vals[pos] = (uint32_t)t;
}

// Triangles of a whole set -> (Morton in scene box, triangle); invalid ones get
// key 2^30 - 1 and an empty box at refit time is avoided by the caller
// counting only valid triangles first.
__global__ __launch_bounds__(kBlock) void k_tri_keys_global(const float* __restrict__ verts,
                                                            const uint32_t* __restrict__ idx, uint64_t nt,
                                                            const uint32_t* __restrict__ scene,
                                                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nt) return;
    float lo[3], inv[3];
    load_scene(scene, lo, inv);
    float3 a = vtx(verts, idx[3 * t]), b = vtx(verts, idx[3 * t + 1]), c = vtx(verts, idx[3 * t + 2]);
    float cx = (a.x + b.x + c.x) * (1.0f / 3.0f), cy = (a.y + b.y + c.y) * (1.0f / 3.0f),
          cz = (a.z + b.z + c.z) * (1.0f / 3.0f);
    keys[t] = morton30(cx, cy, cz, lo[0], lo[1], lo[2], inv[0], inv[1], inv[2]);
    vals[t] = (uint32_t)t;
}

// Edges of needed objects -> EdgeRec in per-object segments.
__global__ __launch_bounds__(kBlock) void k_edge_scatter(const float* __restrict__ es, const float* __restrict__ ee,
                                                         const int32_t* __restrict__ eo, uint64_t ne, int32_t nobj,
                                                         const uint32_t* __restrict__ need,
                                                         const uint32_t* __restrict__ loc,
                                                         const uint32_t* __restrict__ eseg_off,
                                                         uint32_t* __restrict__ cursor, EdgeRec* __restrict__ out) {
    uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    int32_t o = e < ne ? eo[e] : -1;
    bool valid = o >= 0 && o < nobj && need[o];
    uint32_t l = valid ? loc[o] : 0u;
    uint32_t slot = reserve_slot(valid, l, cursor);
    if (!valid) return;
    EdgeRec r;
    r.s = make_float4(es[3 * e], es[3 * e + 1], es[3 * e + 2], __int_as_float(o));
    r.e = make_float4(ee[3 * e], ee[3 * e + 1], ee[3 * e + 2], 0.0f);
    out[(uint64_t)eseg_off[l] + slot] = r;
}

// ------------------------------------------------------------- traversal
template <bool F32>
__device__ __forceinline__ bool leaf_test(const LeafTri& L, const float3& p, const float3& q, bool allow_f32,
                                          LaneStats& st) {
    float3 a = make_float3(L.v0.x, L.v0.y, L.v0.z), b = make_float3(L.v1.x, L.v1.y, L.v1.z),
           c = make_float3(L.v2.x, L.v2.y, L.v2.z);
    Esc esc{0};
    bool hit = seg_tri<F32>(p, q, a, b, c, allow_f32, esc);
    ++st.tests;
    if (esc.level >= 1) {
        if (esc.level == 1)
            ++st.f64;
        else
            ++st.exact;
        if (F32 && seg_tri_naive(p, q, a, b, c) != hit) ++st.mis;
    }
    return hit;
}

// Two-level narrow phase.  Task = (pair, direction, chunk of kChunk source
// edges); one wave per task; early exit per pair.
constexpr int kChunk = 256;

template <int STACK, bool F32>
__global__ __launch_bounds__(kBlock, 4) void k_narrow(const uint64_t* __restrict__ pairs, uint64_t np,
                                                   const uint64_t* __restrict__ task_off, SetView A, SetView B,
                                                   uint32_t* __restrict__ found, const uint8_t* __restrict__ meta,
                                                   int allow_f32, unsigned long long* __restrict__ stats) {
    __shared__ int stk[kWavesPerBlock * STACK * PZ_WAVE];
    const int lane = __lane_id(), w = threadIdx.x / PZ_WAVE;
    int* my = stk + (w * STACK) * PZ_WAVE + lane;
    const uint64_t task = (uint64_t)blockIdx.x * kWavesPerBlock + w;
    const uint64_t ntasks = task_off[np];
    if (task >= ntasks) return;
    // pair = upper_bound(task_off, task) - 1
    uint64_t lo = 0, hi = np;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (task_off[mid] <= task)
            lo = mid;
        else
            hi = mid;
    }
    const uint64_t p = lo;
    if (found[p]) return;
    const int dir_mask = meta[p] >> 2;
    const uint64_t key = pairs[p];
    const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
    const uint32_t la = A.loc[oa], lb = B.loc[ob];
    const uint32_t ea0 = A.eseg_off[la], ea1 = A.eseg_off[la + 1];
    const uint32_t eb0 = B.eseg_off[lb], eb1 = B.eseg_off[lb + 1];
    const uint64_t chunks_a = (dir_mask & 1) ? (ea1 - ea0 + kChunk - 1) / kChunk : 0;
    uint64_t c = task - task_off[p];
    // direction 0: edges of A against triangles of B; direction 1: the reverse
    const bool rev = c >= chunks_a;
    if (rev) c -= chunks_a;
    const SetView& S = rev ? B : A;
    const SetView& T = rev ? A : B;
    const uint32_t e_begin = (rev ? eb0 : ea0) + (uint32_t)c * kChunk;
    const uint32_t e_end = min(e_begin + kChunk, rev ? eb1 : ea1);
    const int32_t ot = rev ? oa : ob;
    const uint32_t lt = rev ? la : lb;
    const uint32_t t0 = T.seg_off[lt], t1 = T.seg_off[lt + 1];
    const int root = (t1 - t0 >= 2) ? (int)t0 : ~(int)t0;
    float tlo[3], thi[3];
    for (int k = 0; k < 3; ++k) {
        tlo[k] = ord2f(T.obox[6ll * ot + k]);
        thi[k] = ord2f(T.obox[6ll * ot + 3 + k]);
    }
    LaneStats st{0, 0, 0, 0, 0};
    bool any_hit = false;
    for (uint32_t base = e_begin; base < e_end && !any_hit; base += PZ_WAVE) {
        const uint32_t e = base + lane;
        bool act = e < e_end;
        float3 pp = make_float3(0, 0, 0), qq = pp;
        float elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
        if (act) {
            EdgeRec r = S.edges[e];
            pp = make_float3(r.s.x, r.s.y, r.s.z);
            qq = make_float3(r.e.x, r.e.y, r.e.z);
            elo[0] = fminf(pp.x, qq.x);
            elo[1] = fminf(pp.y, qq.y);
            elo[2] = fminf(pp.z, qq.z);
            ehi[0] = fmaxf(pp.x, qq.x);
            ehi[1] = fmaxf(pp.y, qq.y);
            ehi[2] = fmaxf(pp.z, qq.z);
            act = boxes_meet(elo, ehi, tlo[0], tlo[1], tlo[2], thi[0], thi[1], thi[2]);
        }
        if (act) ++st.traced;
        int cur = root;
        int sp = 0;
        bool hit = false;
        while (__ballot(act) != 0) {
            if (act) {
                if (cur >= 0) {  // own-box node: descend if the edge's box meets it
                    const BNode nd = T.nodes[cur];
                    if (boxes_meet(elo, ehi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y)) {
                        my[sp * PZ_WAVE] = bnode_c1(nd);
                        ++sp;
                        cur = bnode_c0(nd);
                    } else if (sp > 0) {
                        --sp;
                        cur = my[sp * PZ_WAVE];
                    } else {
                        act = false;
                    }
                } else {
                    const LeafTri L = T.leaves[~cur];
                    if (leaf_test<F32>(L, pp, qq, allow_f32 != 0, st)) {
                        hit = true;
                        act = false;
                    } else if (sp > 0) {
                        --sp;
                        cur = my[sp * PZ_WAVE];
                    } else {
                        act = false;
                    }
                }
            }
            if (__ballot(hit) != 0) break;
        }
        any_hit = __ballot(hit) != 0;
    }
    if (any_hit && lane == 0) found[p] = 1u;
    flush_stats(st, stats);
}

// Dual-tree narrow phase: one wave per candidate pair walks pairs of
// (mesh1 node, mesh2 node) whose boxes meet, from the two object roots, with
// a LIFO queue in LDS (64 entries popped per step, children pairs pushed
// after a wave prefix sum).  At a pair of leaves the edges of each triangle
// are tested against the other triangle (direction bit 0: mesh1 edges,
// bit 1: mesh2 edges), which is the edge-against-triangle predicate on the
// triangles' own edges.  The wave stops at the first proven intersection.
// A queue overflow marks the pair for the per-edge kernel instead of
// dropping work.  The queue length Q (node pairs per wave) is a launch
// parameter: a shorter queue leaves LDS for more resident waves, a longer one
// sends fewer pairs to the per-edge kernel.
constexpr int kDualWaves = 4;

__device__ __forceinline__ void leaf_box(const LeafTri& L, float* lo, float* hi) {
    lo[0] = fminf(L.v0.x, fminf(L.v1.x, L.v2.x));
    lo[1] = fminf(L.v0.y, fminf(L.v1.y, L.v2.y));
    lo[2] = fminf(L.v0.z, fminf(L.v1.z, L.v2.z));
    hi[0] = fmaxf(L.v0.x, fmaxf(L.v1.x, L.v2.x));
    hi[1] = fmaxf(L.v0.y, fmaxf(L.v1.y, L.v2.y));
    hi[2] = fmaxf(L.v0.z, fmaxf(L.v1.z, L.v2.z));
}

// With F32 the walk runs the fp32 filter only: a leaf pair it cannot decide
// goes to esc_list (pair, mesh1 leaf, mesh2 leaf) for k_escalate, so the
// fp64 and exact stages (and their registers) stay out of this kernel; if
// esc_list is full the pair is sent to the per-edge kernel instead.
//
// Queue entries are (children of an x node, children of a y node): four
// links, kNone marking the second of a side that is a single leaf (a leaf's
// "children" are itself), or a leaf pair to test, marked by kTest in the
// second slot.  A step loads the up-to-four records of a node entry in one
// batch (own-box nodes, or leaf triangles), tests the up-to-four box pairs
// and pushes one entry per meeting pair: the pair's children links, taken
// from the records just loaded, or the leaf pair.  A lane thus runs at most
// one triangle-pair test per step.
constexpr int kNone = 0x7FFFFFFF;
constexpr int kTest = 0x7FFFFFFE;

// Box and child links of one record of a side (a leaf offers itself).
struct Elem {
    float lo[3], hi[3];
    int c0, c1;
    bool leaf;
};

template <class SetV>
__device__ __forceinline__ void load_elem(const SetV& S, int link, Elem& e) {
    if (link >= 0) {
        const BNode nd = S.nodes[link];
        e.lo[0] = nd.a.x; e.lo[1] = nd.a.y; e.lo[2] = nd.a.z;
        e.hi[0] = nd.a.w; e.hi[1] = nd.b.x; e.hi[2] = nd.b.y;
        e.c0 = bnode_c0(nd);
        e.c1 = bnode_c1(nd);
        e.leaf = false;
    } else {
        leaf_box(S.leaves[~link], e.lo, e.hi);
        e.c0 = link;
        e.c1 = kNone;
        e.leaf = true;
    }
}

// The entry a walk starts from: the children of both roots (the roots' own
// boxes are the objects' boxes, which the broad phase found to meet).
template <class SetV>
__device__ __forceinline__ void root_children(const SetV& S, uint32_t s0, uint32_t s1, int& c0, int& c1) {
    if (s1 - s0 >= 2) {
        const BNode nd = S.nodes[s0];
        c0 = bnode_c0(nd);
        c1 = bnode_c1(nd);
    } else {
        c0 = ~(int)s0;
        c1 = kNone;
    }
}

template <bool F32, int QLEN>
__global__ __launch_bounds__(kDualWaves * PZ_WAVE, 2) void k_narrow_dual(
    const uint64_t* __restrict__ pairs, SetView A, SetView B, uint32_t* __restrict__ found,
    uint32_t* __restrict__ fallback, const uint8_t* __restrict__ meta, int allow_f32,
    const DualTask* __restrict__ tasks_in, uint64_t ntasks, unsigned long long* __restrict__ next_task, int chunk,
    DualTask* __restrict__ tasks_out, unsigned long long* __restrict__ nout, uint64_t out_cap,
    uint4* __restrict__ esc_list, unsigned long long* __restrict__ esc_count, uint64_t esc_cap,
    unsigned long long* __restrict__ stats) {
    __shared__ int4 q[kDualWaves][QLEN];
    const int lane = __lane_id(), w = threadIdx.x / PZ_WAVE;
    int4* Q = q[w];  // ring buffer: entries [bot, top) modulo QLEN, LIFO at the top
    LaneStats st{0, 0, 0, 0, 0};
    uint32_t walks = 0;  // tasks this wave walked
    while (true) {
        unsigned long long t0 = 0;
        if (lane == 0) t0 = atomicAdd(next_task, (unsigned long long)chunk);
        t0 = __shfl(t0, 0, PZ_WAVE);
        if (t0 >= ntasks) break;
        const uint64_t t1 = t0 + (uint64_t)chunk < ntasks ? t0 + (uint64_t)chunk : ntasks;
        for (uint64_t t = t0; t < t1; ++t) {
            uint32_t p;
            int4 root;
            if (tasks_in) {
                const DualTask tk = tasks_in[t];
                p = tk.pair;
                root = make_int4(tk.x0, tk.x1, tk.y0, tk.y1);
            } else {
                p = (uint32_t)t;
                const uint64_t key = pairs[p];
                const int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
                const uint32_t la = A.loc[oa], lb = B.loc[ob];
                const uint32_t a0 = A.seg_off[la], a1 = A.seg_off[la + 1], b0 = B.seg_off[lb], b1 = B.seg_off[lb + 1];
                root_children(A, a0, a1, root.x, root.y);
                root_children(B, b0, b1, root.z, root.w);
            }
            if (found_now(found, p)) continue;  // another task of this pair already proved a hit
            ++walks;
            const int dir_mask = meta[p] >> 2;
            if (lane == 0) Q[0] = root;
            __builtin_amdgcn_wave_barrier();
            int bot = 0, top = 1;
            bool hit = false, ovf = false, esc_full = false;
            int iter = 0;
            while (top > bot) {
                if ((++iter & 15) == 0 && found_now(found, p)) break;
                const int qn = top - bot;
                const int take = qn < PZ_WAVE ? qn : PZ_WAVE;
                const bool act = lane < take;
                const int4 e = act ? Q[(top - 1 - lane) % QLEN] : make_int4(kNone, kNone, kNone, kNone);
                top -= take;
                int4 o0 = make_int4(0, 0, 0, 0), o1 = o0, o2 = o0, o3 = o0;
                int nout_l = 0;
                if (act && e.y == kTest) {
                    const LeafTri LA = A.leaves[~e.x], LB = B.leaves[~e.z];
                    if (F32) {
                        const int r = tri_pair_f32(LA, LB, dir_mask, st);
                        hit = r > 0;
                        if (r < 0) {
                            const unsigned long long slot = atomicAdd(esc_count, 1ull);
                            if (slot < esc_cap)
                                esc_list[slot] = make_uint4(p, (uint32_t)~e.x, (uint32_t)~e.z, 0u);
                            else
                                esc_full = true;
                        }
                    } else {
                        hit = tri_pair_test<false>(LA, LB, dir_mask, false, st);
                    }
                } else if (act) {
                    // all records of the entry in one batch of loads
                    Elem x[2], y[2];
                    const int nx = e.y == kNone ? 1 : 2, ny = e.w == kNone ? 1 : 2;
                    load_elem(A, e.x, x[0]);
                    if (nx > 1) load_elem(A, e.y, x[1]);
                    load_elem(B, e.z, y[0]);
                    if (ny > 1) load_elem(B, e.w, y[1]);
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            if (i >= nx || j >= ny || !box_box(x[i].lo, x[i].hi, y[j].lo, y[j].hi)) continue;
                            const bool lp = x[i].leaf && y[j].leaf;
                            const int4 v = lp ? make_int4(x[i].c0, kTest, y[j].c0, 0)
                                              : make_int4(x[i].c0, x[i].c1, y[j].c0, y[j].c1);
                            if (nout_l == 0) o0 = v;
                            else if (nout_l == 1) o1 = v;
                            else if (nout_l == 2) o2 = v;
                            else o3 = v;
                            ++nout_l;
                        }
                }
                if (__ballot(hit) != 0) break;
                // append the children entries: nout_l is 0..4, so its wave-wide
                // prefix is the bit-weighted count of lower lanes in three
                // ballots (no cross-lane data movement)
                const uint64_t b0 = __ballot(nout_l & 1), b1 = __ballot(nout_l & 2), b2 = __ballot(nout_l & 4);
                const uint64_t below = (1ull << lane) - 1ull;
                const int excl = __popcll(b0 & below) + 2 * __popcll(b1 & below) + 4 * __popcll(b2 & below);
                const int incl = excl + nout_l;
                const int total = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
                if (top - bot + total > QLEN) {
                    // hand the oldest half (the shallowest subtrees) to the next round
                    const int k = (top - bot) / 2;
                    unsigned long long base = 0;
                    if (lane == 0 && k > 0) base = atomicAdd(nout, (unsigned long long)k);
                    base = __shfl(base, 0, PZ_WAVE);
                    const int kk = base < out_cap ? (int)((out_cap - base) < (uint64_t)k ? out_cap - base : k) : 0;
                    for (int i = lane; i < kk; i += PZ_WAVE) {
                        const int4 v = Q[(bot + i) % QLEN];
                        DualTask tk;
                        tk.pair = p;
                        tk.x0 = v.x;
                        tk.x1 = v.y;
                        tk.y0 = v.z;
                        tk.y1 = v.w;
                        tk.pad[0] = tk.pad[1] = tk.pad[2] = 0;
                        tasks_out[base + i] = tk;
                    }
                    bot += kk;
                    if (top - bot + total > QLEN) {
                        ovf = true;
                        break;
                    }
                }
                const int pos = top + incl - nout_l;
                if (nout_l > 0) Q[pos % QLEN] = o0;
                if (nout_l > 1) Q[(pos + 1) % QLEN] = o1;
                if (nout_l > 2) Q[(pos + 2) % QLEN] = o2;
                if (nout_l > 3) Q[(pos + 3) % QLEN] = o3;
                top += total;
                __builtin_amdgcn_wave_barrier();
            }
            const bool any_hit = __ballot(hit) != 0;
            const bool any_full = __ballot(esc_full) != 0;
            if (lane == 0) {
                if (any_hit) found[p] = 1u;
                else if (ovf || any_full) fallback[p] = 1u;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    flush_stats(st, stats);
    if (lane == 0 && walks) atomicAdd(&stats[pzk::kStatWalks], (unsigned long long)walks);
}

__global__ void k_mark_task_pairs(const DualTask* __restrict__ tasks, uint64_t n, uint32_t* __restrict__ fallback) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) fallback[tasks[i].pair] = 1u;
}

// The leaf pairs the fp32 filter left open, decided with the fp64 and exact
// stages (grid-stride over min(*count, cap) entries, so no host read-back).
// Counters as in tri_pair_test: one fp64 or exact escalation per entry, and
// a misclassification when unguarded fp32 would have answered otherwise.
__global__ __launch_bounds__(kBlock) void k_escalate(const uint4* __restrict__ list,
                                                     const unsigned long long* __restrict__ count, uint64_t cap,
                                                     SetView A, SetView B, const uint8_t* __restrict__ meta,
                                                     uint32_t* __restrict__ found, unsigned long long* __restrict__ stats) {
    const uint64_t n = *count < cap ? *count : cap;
    LaneStats st{0, 0, 0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const uint4 e = list[i];
        if (found[e.x]) continue;
        const LeafTri X = A.leaves[e.y], Y = B.leaves[e.z];
        const int dm = meta[e.x] >> 2;
        const float3 x0 = make_float3(X.v0.x, X.v0.y, X.v0.z), x1 = make_float3(X.v1.x, X.v1.y, X.v1.z),
                     x2 = make_float3(X.v2.x, X.v2.y, X.v2.z);
        const float3 y0 = make_float3(Y.v0.x, Y.v0.y, Y.v0.z), y1 = make_float3(Y.v1.x, Y.v1.y, Y.v1.z),
                     y2 = make_float3(Y.v2.x, Y.v2.y, Y.v2.z);
        Esc esc{0};
        bool hit = false;
        if (dm & 1)
            hit = seg_tri<false>(x0, x1, y0, y1, y2, false, esc) || seg_tri<false>(x1, x2, y0, y1, y2, false, esc) ||
                  seg_tri<false>(x2, x0, y0, y1, y2, false, esc);
        if (!hit && (dm & 2))
            hit = seg_tri<false>(y0, y1, x0, x1, x2, false, esc) || seg_tri<false>(y1, y2, x0, x1, x2, false, esc) ||
                  seg_tri<false>(y2, y0, x0, x1, x2, false, esc);
        if (esc.level == 2) ++st.exact;
        else ++st.f64;
        bool naive = false;
        if (dm & 1)
            naive = seg_tri_naive(x0, x1, y0, y1, y2) || seg_tri_naive(x1, x2, y0, y1, y2) ||
                    seg_tri_naive(x2, x0, y0, y1, y2);
        if (!naive && (dm & 2))
            naive = seg_tri_naive(y0, y1, x0, x1, x2) || seg_tri_naive(y1, y2, x0, x1, x2) ||
                    seg_tri_naive(y2, y0, x0, x1, x2);
        if (naive != hit) ++st.mis;
        if (hit) found[e.x] = 1u;
    }
    flush_stats(st, stats);
}

// Parity: thread per (pair, direction).  Counts +Z crossings of the
// perturbed ray from the first vertex of one object against the other's
// triangles.  kind 0: containment (point of mesh2 object in mesh1 object);
// kind 1: intersection (both directions, sets inside[] bit 0 / bit 1).
template <int STACK>
__global__ __launch_bounds__(kBlock, 4) void k_parity(const uint64_t* __restrict__ pairs, uint64_t np,
                                                   const uint32_t* __restrict__ found, SetView A, SetView B,
                                                   const float* __restrict__ vA, const uint32_t* __restrict__ iA,
                                                   const uint32_t* __restrict__ ftA, const float* __restrict__ vB,
                                                   const uint32_t* __restrict__ iB, const uint32_t* __restrict__ ftB,
                                                   const uint8_t* __restrict__ meta, uint32_t* __restrict__ inside,
                                                   unsigned long long* __restrict__ stats) {
    __shared__ int stk[kWavesPerBlock * STACK * PZ_WAVE];
    const int lane = __lane_id(), w = threadIdx.x / PZ_WAVE;
    int* my = stk + (w * STACK) * PZ_WAVE + lane;
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t p = gid >> 1;
    const int dir = (int)(gid & 1);  // 0: point of B in A; 1: point of A in B
    // containment: dir 0 only; intersection: dir 0 belongs to the mesh2->mesh1
    // direction (bit 1), dir 1 to mesh1->mesh2 (bit 0); overlap: none
    bool act = false;
    if (p < np && !found[p]) {
        const int pred = meta[p] & 3, dm = meta[p] >> 2;
        act = (pred == 2 && dir == 0) || (pred == 1 && (dm & (dir == 0 ? 2 : 1)));
    }
    LaneStats st{0, 0, 0, 0, 0};
    float3 pt = make_float3(0, 0, 0);
    int cur = 0, sp = 0;
    bool odd = false;
    const bool cast = act;
    uint32_t crossings = 0;
    const SetView* T = dir == 0 ? &A : &B;
    int32_t ot = 0;
    if (act) {
        uint64_t key = pairs[p];
        int32_t oa = (int32_t)(key >> 32), ob = (int32_t)(uint32_t)key;
        if (dir == 0) {
            pt = vtx(vB, iB[3ull * ftB[ob]]);
            ot = oa;
        } else {
            pt = vtx(vA, iA[3ull * ftA[oa]]);
            ot = ob;
        }
        uint32_t lt = T->loc[ot];
        uint32_t t0 = T->seg_off[lt], t1 = T->seg_off[lt + 1];
        cur = (t1 - t0 >= 2) ? (int)t0 : ~(int)t0;
        // the ray must start inside the target's box footprint and below its top
        float bx0 = ord2f(T->obox[6ll * ot]), by0 = ord2f(T->obox[6ll * ot + 1]);
        float bx1 = ord2f(T->obox[6ll * ot + 3]), by1 = ord2f(T->obox[6ll * ot + 4]);
        float bz1 = ord2f(T->obox[6ll * ot + 5]);
        act = pt.x >= bx0 && pt.x <= bx1 && pt.y >= by0 && pt.y <= by1 && pt.z <= bz1;
    }
    const float rlo[3] = {pt.x, pt.y, pt.z}, rhi[3] = {pt.x, pt.y, INFINITY};
    while (act) {
        if (cur >= 0) {  // own-box node: descend if the ray's box meets it
            const BNode nd = T->nodes[cur];
            if (boxes_meet(rlo, rhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y)) {
                my[sp * PZ_WAVE] = bnode_c1(nd);
                ++sp;
                cur = bnode_c0(nd);
                continue;
            }
        } else {
            const LeafTri L = T->leaves[~cur];
            Esc esc{0};
            float3 a = make_float3(L.v0.x, L.v0.y, L.v0.z), b = make_float3(L.v1.x, L.v1.y, L.v1.z),
                   c = make_float3(L.v2.x, L.v2.y, L.v2.z);
            if (ray_cross(pt, a, b, c, esc)) {
                odd = !odd;
                ++crossings;
            }
            ++st.tests;
            if (esc.level == 2) ++st.exact;
        }
        if (sp == 0) break;
        --sp;
        cur = my[sp * PZ_WAVE];
    }
    if (p < np && odd) atomicOr(&inside[p], 1u << dir);
    flush_stats(st, stats);
    // rays cast, triangles crossed, rays crossing something, points inside, tests
    const uint32_t pv[5] = {cast ? 1u : 0u, crossings, crossings ? 1u : 0u, odd ? 1u : 0u, st.tests};
    for (int k = 0; k < 5; ++k) {
        const uint32_t sum = wave_sum(pv[k]);
        if (lane == 0 && sum) atomicAdd(&stats[pzk::kStatParityRays + k], (unsigned long long)sum);
    }
}

// Containment-ray tracking (pierce_intersection --enable-tracking,
// pierce/query/src/cuda/mesh_intersection_anyhit.cu:46-240): thread per
// source object.  The +Z ray from the first vertex of the object's first
// triangle walks the set-wide BVH of the other set; every triangle it crosses
// (the perturbed crossing of k_parity) adds a hit to that triangle's object
// in the source's list.  The list keeps objects in the order the walk meets
// them, at most max_targets; hits on objects beyond that count as overflow.
// Sources without triangles cast no ray (iterations 0).
template <int STACK>
__global__ __launch_bounds__(kBlock, 4) void k_track_rays(const float* __restrict__ vS, const uint32_t* __restrict__ iS,
                                                       const uint32_t* __restrict__ ftS,
                                                       const uint32_t* __restrict__ tcS, int32_t nsrc,
                                                       const Node* __restrict__ nodes,
                                                       const LeafTri* __restrict__ leaves, uint64_t ntri,
                                                       int max_targets, int32_t* __restrict__ ids,
                                                       uint32_t* __restrict__ hits, uint32_t* __restrict__ count,
                                                       uint32_t* __restrict__ overflow, uint32_t* __restrict__ iters,
                                                       unsigned long long* __restrict__ stats) {
    __shared__ int stk[kWavesPerBlock * STACK * PZ_WAVE];
    const int lane = __lane_id(), w = threadIdx.x / PZ_WAVE;
    int* my = stk + (w * STACK) * PZ_WAVE + lane;
    const int64_t src = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    LaneStats st{0, 0, 0, 0, 0};
    bool act = src < nsrc && tcS[src] > 0;
    float3 pt = make_float3(0, 0, 0);
    if (act) pt = vtx(vS, iS[3ull * ftS[src]]);
    int32_t* my_ids = ids + src * (int64_t)max_targets;
    uint32_t* my_hits = hits + src * (int64_t)max_targets;
    uint32_t n = 0, ovf = 0;
    const float rlo[3] = {pt.x, pt.y, pt.z}, rhi[3] = {pt.x, pt.y, INFINITY};
    int cur = ntri >= 2 ? 0 : ~0, sp = 0;
    bool walk = act && ntri > 0;
    while (walk) {
        if (cur >= 0) {
            const Node nd = nodes[cur];
            const bool h0 = boxes_meet(rlo, rhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
            const bool h1 = boxes_meet(rlo, rhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
            if (h0 && h1) {
                my[sp * PZ_WAVE] = nd.d.y;
                ++sp;
                cur = nd.d.x;
                continue;
            } else if (h0) {
                cur = nd.d.x;
                continue;
            } else if (h1) {
                cur = nd.d.y;
                continue;
            }
        } else {
            const LeafTri L = leaves[~cur];
            Esc esc{0};
            const float3 a = make_float3(L.v0.x, L.v0.y, L.v0.z), b = make_float3(L.v1.x, L.v1.y, L.v1.z),
                         c = make_float3(L.v2.x, L.v2.y, L.v2.z);
            ++st.tests;
            if (ray_cross(pt, a, b, c, esc)) {
                const int32_t obj = (int32_t)__float_as_uint(L.v1.w);
                uint32_t i = 0;
                while (i < n && my_ids[i] != obj) ++i;
                if (i < n) {
                    ++my_hits[i];
                } else if (n < (uint32_t)max_targets) {
                    my_ids[n] = obj;
                    my_hits[n] = 1u;
                    ++n;
                } else {
                    ++ovf;
                }
            }
            if (esc.level == 2) ++st.exact;
        }
        if (sp == 0) break;
        --sp;
        cur = my[sp * PZ_WAVE];
    }
    if (src < nsrc) {
        count[src] = n;
        overflow[src] = ovf;
        iters[src] = act ? 1u : 0u;
    }
    flush_stats(st, stats);
    uint32_t crossings = 0, inside = 0;
    for (uint32_t i = 0; i < n; ++i) {
        crossings += my_hits[i];
        inside += my_hits[i] & 1u;
    }
    // rays cast, triangles crossed, rays crossing something, targets containing the point, tests
    const uint32_t pv[5] = {act ? 1u : 0u, crossings + ovf, (crossings + ovf) ? 1u : 0u, inside, st.tests};
    for (int k = 0; k < 5; ++k) {
        const uint32_t sum = wave_sum(pv[k]);
        if (lane == 0 && sum) atomicAdd(&stats[pzk::kStatParityRays + k], (unsigned long long)sum);
    }
}

// Pair keys of a set-wide containment-ray pass: source s's list entries
// [0, count[s]) go to keys[off[s] + i] as (mesh1 << 32 | mesh2), with the
// source on the mesh2 side when src_is_mesh2; entries with an even hit count
// (the point is outside that target) write ~0, which sorts last.
__global__ __launch_bounds__(kBlock) void k_parity_keys(const int32_t* __restrict__ ids,
                                                        const uint32_t* __restrict__ hits,
                                                        const uint32_t* __restrict__ count,
                                                        const uint64_t* __restrict__ off, int32_t nsrc,
                                                        int max_targets, int src_is_mesh2,
                                                        uint64_t* __restrict__ keys) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= nsrc) return;
    const uint32_t n = count[s];
    for (uint32_t i = 0; i < n; ++i) {
        const int64_t e = s * (int64_t)max_targets + i;
        const int32_t t = ids[e];
        keys[off[s] + i] = (hits[e] & 1u) ? (src_is_mesh2 ? pack_pair(t, (int32_t)s) : pack_pair((int32_t)s, t))
                                          : ~0ull;
    }
}

// Keys of `in` (not ~0) absent from the sorted unique array `ref`, appended
// to out with type `type`; one wave reservation per wavefront.
__global__ __launch_bounds__(kBlock) void k_key_difference(const uint64_t* __restrict__ in, uint64_t n,
                                                           const uint64_t* __restrict__ ref, uint64_t nref,
                                                           uint64_t* __restrict__ out, uint32_t* __restrict__ out_type,
                                                           uint32_t type, unsigned long long* __restrict__ count) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool keep = false;
    uint64_t k = ~0ull;
    if (i < n) {
        k = in[i];
        if (k != ~0ull) {
            uint64_t lo = 0, hi = nref;  // first ref >= k
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (ref[mid] < k) lo = mid + 1;
                else hi = mid;
            }
            keep = !(lo < nref && ref[lo] == k);
        }
    }
    const uint64_t m = __ballot(keep);
    if (m == 0) return;
    unsigned long long base = 0;
    const int leader = __ffsll((long long)m) - 1;
    if (__lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl((long long)base, leader, 64);
    if (keep) {
        const uint64_t pos = base + lanes_below(m);
        out[pos] = k;
        out_type[pos] = type;
    }
}

// Ballot + prefix-sum compaction of selected candidate pairs.
// sel: 0 drop; else keep with type value (sel - 1) in vals.
__global__ __launch_bounds__(kBlock) void k_select_pairs(const uint64_t* __restrict__ pairs, uint64_t np,
                                                         const uint32_t* __restrict__ found,
                                                         const uint32_t* __restrict__ inside,
                                                         const uint8_t* __restrict__ meta,
                                                         int include_overlap, uint64_t* __restrict__ out_keys,
                                                         uint32_t* __restrict__ out_type,
                                                         unsigned long long* __restrict__ count,
                                                         unsigned long long* __restrict__ touching) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t type = 0;
    bool keep = false;
    const uint64_t mt = __ballot(i < np && found[i] != 0);  // pairs proven touching
    if (mt && __lane_id() == 0) atomicAdd(touching, (unsigned long long)__popcll(mt));
    if (i < np) {
        const int predicate = meta[i] & 3;
        bool ov = found[i] != 0;
        uint32_t in = inside ? inside[i] : 0u;
        if (predicate == 0) {
            keep = ov;
        } else if (predicate == 1) {
            keep = ov || in != 0;
        } else {
            if (ov) {
                keep = include_overlap != 0;
                type = 1;
            } else {
                keep = (in & 1u) != 0;
            }
        }
    }
    uint64_t m = __ballot(keep);
    if (m == 0) return;
    unsigned long long base = 0;
    const int leader = __ffsll((long long)m) - 1;
    if (__lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl((long long)base, leader, 64);
    if (keep) {
        uint64_t pos = base + lanes_below(m);
        out_keys[pos] = pairs[i];
        out_type[pos] = type;
    }
}

// Tasks per pair: chunks of source edges in the enabled directions.
__global__ void k_task_counts(const uint64_t* __restrict__ pairs, uint64_t np, const uint32_t* __restrict__ locA,
                              const uint32_t* __restrict__ esegA, const uint32_t* __restrict__ locB,
                              const uint32_t* __restrict__ esegB, const uint8_t* __restrict__ meta,
                              uint32_t* __restrict__ counts) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    const int dir_mask = meta[i] >> 2;
    uint64_t key = pairs[i];
    uint32_t la = locA[(uint32_t)(key >> 32)], lb = locB[(uint32_t)key];
    uint32_t ea = esegA[la + 1] - esegA[la], eb = esegB[lb + 1] - esegB[lb];
    uint32_t c = 0;
    if (dir_mask & 1) c += (ea + kChunk - 1) / kChunk;
    if (dir_mask & 2) c += (eb + kChunk - 1) / kChunk;
    counts[i] = c;
}

// ------------------------------------------------- global (reference-shaped)
__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Insert into a power-of-two open-addressing set; false when the probe
// budget runs out (the caller flags overflow and re-runs exactly).
__device__ __forceinline__ bool hash_insert(unsigned long long* table, uint64_t mask, uint64_t key, LaneStats& st) {
    uint64_t h = fmix64(key) & mask;
    for (uint32_t probe = 0; probe < 4096; ++probe) {
        ++st.probes;
        unsigned long long cur = table[h];
        if (cur == (unsigned long long)key) return true;
        if (cur == ~0ull) {
            unsigned long long prev = atomicCAS(&table[h], ~0ull, (unsigned long long)key);
            if (prev == ~0ull || prev == (unsigned long long)key) return true;
        }
        ++st.contended;
        h = (h + 1) & mask;
    }
    return false;
}

// Mode 0: count hits per edge; 1: write keys at off[e]; 2: hash insert;
// 3: count leaves whose boxes meet the edge box (candidate tests, no predicate).
// One lane per source edge; every hit on a distinct target object kept
// (consecutive hits on one object collapse).
template <int STACK, bool F32, int Mode>
__global__ __launch_bounds__(kBlock, 4) void k_trace_global(
    const float* __restrict__ es, const float* __restrict__ ee, const int32_t* __restrict__ eo, uint64_t ne,
    const Node* __restrict__ nodes, const LeafTri* __restrict__ leaves, uint64_t ntri, int swap, int allow_f32,
    uint32_t* __restrict__ counts, const uint64_t* __restrict__ off, uint64_t* __restrict__ out,
    unsigned long long* __restrict__ table, uint64_t mask, uint32_t* __restrict__ overflow,
    unsigned long long* __restrict__ stats) {
    __shared__ int stk[kWavesPerBlock * STACK * PZ_WAVE];
    const int lane = __lane_id(), w = threadIdx.x / PZ_WAVE;
    int* my = stk + (w * STACK) * PZ_WAVE + lane;
    const uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool act = e < ne && ntri > 0 && eo[e] >= 0;
    LaneStats st{0, 0, 0, 0, 0};
    float3 pp = make_float3(0, 0, 0), qq = pp;
    float elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
    int32_t src = 0;
    if (act) {
        pp = make_float3(es[3 * e], es[3 * e + 1], es[3 * e + 2]);
        qq = make_float3(ee[3 * e], ee[3 * e + 1], ee[3 * e + 2]);
        src = eo[e];
        elo[0] = fminf(pp.x, qq.x);
        elo[1] = fminf(pp.y, qq.y);
        elo[2] = fminf(pp.z, qq.z);
        ehi[0] = fmaxf(pp.x, qq.x);
        ehi[1] = fmaxf(pp.y, qq.y);
        ehi[2] = fmaxf(pp.z, qq.z);
        ++st.traced;
    }
    int cur = ntri >= 2 ? 0 : ~0;
    if (ntri == 1 && act) {
        // single-triangle set: the root is the leaf; its box is tested here
        const LeafTri L = leaves[0];
        float bl[3] = {fminf(L.v0.x, fminf(L.v1.x, L.v2.x)), fminf(L.v0.y, fminf(L.v1.y, L.v2.y)),
                       fminf(L.v0.z, fminf(L.v1.z, L.v2.z))};
        float bh[3] = {fmaxf(L.v0.x, fmaxf(L.v1.x, L.v2.x)), fmaxf(L.v0.y, fmaxf(L.v1.y, L.v2.y)),
                       fmaxf(L.v0.z, fmaxf(L.v1.z, L.v2.z))};
        act = boxes_meet(elo, ehi, bl[0], bl[1], bl[2], bh[0], bh[1], bh[2]);
    }
    int sp = 0;
    uint32_t n = 0;
    int32_t last = -1;
    uint64_t base = (Mode == 1 && e < ne) ? off[e] : 0;
    while (act) {
        if (cur >= 0) {
            const Node nd = nodes[cur];
            bool h0 = boxes_meet(elo, ehi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
            bool h1 = boxes_meet(elo, ehi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
            if (h0 && h1) {
                my[sp * PZ_WAVE] = nd.d.y;
                ++sp;
                cur = nd.d.x;
                continue;
            } else if (h0) {
                cur = nd.d.x;
                continue;
            } else if (h1) {
                cur = nd.d.y;
                continue;
            }
        } else {
            const LeafTri L = leaves[~cur];
            const int32_t tgt = __float_as_int(L.v1.w);
            if (Mode == 3) {
                ++n;  // candidate count only: boxes meet
            } else if (tgt >= 0 && tgt != last && leaf_test<F32>(L, pp, qq, allow_f32 != 0, st)) {
                last = tgt;
                uint64_t key = swap ? pack_pair(tgt, src) : pack_pair(src, tgt);
                if (Mode == 1) out[base + n] = key;
                if (Mode == 2 && !hash_insert(table, mask, key, st)) *overflow = 1u;
                ++n;
            }
        }
        if (sp == 0) break;
        --sp;
        cur = my[sp * PZ_WAVE];
    }
    if ((Mode == 0 || Mode == 3) && e < ne) counts[e] = n;
    flush_stats(st, stats);
}

__global__ __launch_bounds__(kBlock) void k_compact_table(const unsigned long long* __restrict__ table, uint64_t size,
                                                          uint64_t* __restrict__ out, uint64_t cap,
                                                          unsigned long long* __restrict__ count) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool keep = i < size && table[i] != ~0ull;
    uint64_t m = __ballot(keep);
    if (m == 0) return;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (__lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl((long long)base, leader, 64);
    if (keep) {
        uint64_t pos = base + lanes_below(m);
        if (pos < cap) out[pos] = table[i];
    }
}

__global__ __launch_bounds__(kBlock) void k_compact_flagged(const uint64_t* __restrict__ pairs, uint64_t np,
                                                           const uint32_t* __restrict__ flag, uint64_t* __restrict__ out,
                                                           uint32_t* __restrict__ out_index,
                                                           unsigned long long* __restrict__ count) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool keep = i < np && flag[i] != 0;
    uint64_t m = __ballot(keep);
    if (m == 0) return;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (__lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl((long long)base, leader, 64);
    if (keep) {
        uint64_t pos = base + lanes_below(m);
        out[pos] = pairs[i];
        out_index[pos] = (uint32_t)i;
    }
}

__global__ void k_scatter_found(const uint32_t* __restrict__ sub, const uint32_t* __restrict__ index, uint64_t n,
                                uint32_t* __restrict__ found) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n && sub[i]) found[index[i]] = sub[i];
}

template <class T>
__global__ void k_fill_any(T* p, T v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_init_boxes(uint32_t* obox, int64_t n) {
    int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= n) return;
    for (int a = 0; a < 3; ++a) {
        obox[6 * o + a] = 0xFF800000u;      // ordered +inf
        obox[6 * o + 3 + a] = 0x007FFFFFu;  // ordered -inf
    }
}

template <int STACK, bool F32>
void launch_narrow(unsigned g, hipStream_t s, const uint64_t* pairs, uint64_t np, const uint64_t* task_off, SetView A,
                   SetView B, uint32_t* found, const uint8_t* meta, int allow_f32, unsigned long long* stats) {
    {
        pzk::KernelTimer kt_("narrow", s);
        hipLaunchKernelGGL((k_narrow<STACK, F32>), dim3(g), dim3(kBlock), 0, s, pairs, np, task_off, A, B, found, meta,
                           allow_f32, stats);
    }
}

template <int STACK, bool F32>
void launch_trace(int mode, unsigned g, hipStream_t s, const float* es, const float* ee, const int32_t* eo,
                  uint64_t ne, const Node* nodes, const LeafTri* leaves, uint64_t ntri, int swap, int allow_f32,
                  uint32_t* counts, const uint64_t* off, uint64_t* out, unsigned long long* table, uint64_t mask,
                  uint32_t* overflow, unsigned long long* stats) {
    if (mode == 0)
        hipLaunchKernelGGL((k_trace_global<STACK, F32, 0>), dim3(g), dim3(kBlock), 0, s, es, ee, eo, ne, nodes, leaves,
                           ntri, swap, allow_f32, counts, off, out, table, mask, overflow, stats);
    else if (mode == 1)
        hipLaunchKernelGGL((k_trace_global<STACK, F32, 1>), dim3(g), dim3(kBlock), 0, s, es, ee, eo, ne, nodes, leaves,
                           ntri, swap, allow_f32, counts, off, out, table, mask, overflow, stats);
    else if (mode == 2)
        hipLaunchKernelGGL((k_trace_global<STACK, F32, 2>), dim3(g), dim3(kBlock), 0, s, es, ee, eo, ne, nodes, leaves,
                           ntri, swap, allow_f32, counts, off, out, table, mask, overflow, stats);
    else
        hipLaunchKernelGGL((k_trace_global<STACK, F32, 3>), dim3(g), dim3(kBlock), 0, s, es, ee, eo, ne, nodes, leaves,
                           ntri, swap, allow_f32, counts, off, out, table, mask, overflow, stats);
}

}  // namespace

namespace pzk {

int stack_for_height(uint32_t h) { return h < 32 ? 32 : (h < 64 ? 64 : 0); }

void object_stats(const float* verts, uint64_t nv, const uint32_t* idx, const int32_t* t2o, uint64_t nt, int32_t nobj,
                  uint32_t* obox, uint32_t* tri_cnt, uint32_t* first_tri, uint32_t* last_tri, hipStream_t s) {
    if (nobj <= 0) return;
    hipLaunchKernelGGL(k_init_boxes, dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, obox, (int64_t)nobj);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, tri_cnt, 0u,
                       (uint64_t)nobj);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, first_tri,
                       0xFFFFFFFFu, (uint64_t)nobj);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, last_tri, 0u,
                       (uint64_t)nobj);
    KernelTimer kt_("object_stats", s);
    const uint64_t waves = (nt + kStatsRun - 1) / kStatsRun;
    if (nt)
        hipLaunchKernelGGL(k_object_stats, dim3(grid1(waves * PZ_WAVE)), dim3(kBlock), 0, s, verts, nv, idx, t2o, nt,
                           nobj, obox,
                           tri_cnt, first_tri, last_tri);
}

void topology_scan(const uint32_t* idx, const int32_t* t2o, uint64_t nt, uint64_t nv, int32_t nobj, uint32_t* first,
                   uint32_t* last, uint32_t* cnt, uint32_t* vmin, uint32_t* vmax, int32_t* owner, uint32_t* clash,
                   unsigned long long* referenced, hipStream_t s) {
    const uint64_t n = (uint64_t)std::max(nobj, 1);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1(n)), dim3(kBlock), 0, s, first, 0xFFFFFFFFu, n);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1(n)), dim3(kBlock), 0, s, last, 0u, n);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1(n)), dim3(kBlock), 0, s, cnt, 0u, n);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1(n)), dim3(kBlock), 0, s, vmin, 0xFFFFFFFFu, n);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1(n)), dim3(kBlock), 0, s, vmax, 0u, n);
    if (nv)
        hipLaunchKernelGGL((k_fill_any<int32_t>), dim3(grid1(nv)), dim3(kBlock), 0, s, owner, (int32_t)-1, nv);
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(1), dim3(kBlock), 0, s, clash, 0u, (uint64_t)1);
    hipLaunchKernelGGL((k_fill_any<unsigned long long>), dim3(1), dim3(kBlock), 0, s, referenced, 0ull, (uint64_t)1);
    if (nt)
        hipLaunchKernelGGL(k_topo_tris, dim3(grid1(nt)), dim3(kBlock), 0, s, idx, t2o, nt, nobj, first, last, cnt,
                           vmin, vmax, owner, clash);
    if (nv) hipLaunchKernelGGL(k_topo_verts, dim3(grid1(nv)), dim3(kBlock), 0, s, owner, nv, referenced);
}

uint32_t run_chunk_vertices() { return kRunChunk; }

void object_bounds_runs(const float* verts, const uint2* chunks, uint64_t nch, const uint32_t* vmax, int32_t nobj,
                        uint32_t* obox, hipStream_t s) {
    if (nobj <= 0) return;
    hipLaunchKernelGGL(k_init_boxes, dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, obox, (int64_t)nobj);
    KernelTimer kt_("object_bounds_runs", s);
    if (nch)
        hipLaunchKernelGGL(k_object_bounds_runs, dim3(grid1(nch * PZ_WAVE)), dim3(kBlock), 0, s, verts, chunks, nch,
                           vmax, obox);
}

void edge_count(const int32_t* eo, uint64_t ne, int32_t nobj, uint32_t* edge_cnt, hipStream_t s) {
    if (nobj <= 0) return;
    hipLaunchKernelGGL((k_fill_any<uint32_t>), dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, edge_cnt, 0u,
                       (uint64_t)nobj);
    KernelTimer kt_("edge_count", s);
    if (ne) hipLaunchKernelGGL(k_edge_count, dim3(grid1(ne)), dim3(kBlock), 0, s, eo, ne, nobj, edge_cnt);
}

void scene_box(const uint32_t* obox, int32_t nobj, uint32_t* scene, hipStream_t s) {
    hipLaunchKernelGGL(k_init_boxes, dim3(1), dim3(kBlock), 0, s, scene, (int64_t)1);
    if (nobj > 0)
        hipLaunchKernelGGL(k_scene_box, dim3(std::min(grid1((uint64_t)nobj), kSceneBlocks)), dim3(kBlock), 0, s, obox,
                           nobj, scene);
}

void object_keys(const uint32_t* obox, int32_t nobj, const uint32_t* scene, const uint32_t* part_of, uint64_t* keys,
                 uint32_t* vals, hipStream_t s) {
    if (nobj > 0)
        hipLaunchKernelGGL(k_object_keys, dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, obox, nobj, scene, part_of,
                           keys, vals);
}

// Lanes per mesh1 object of the box scan: enough lanes to fill the GPU.
inline int brute_lanes(int32_t nA) { return nA <= 4096 ? 64 : (nA <= 16384 ? 16 : (nA <= 65536 ? 4 : 1)); }
#define PZ_BRUTE_S(M, P, Sn, ...)                                                                          \
    hipLaunchKernelGGL((k_broad_brute<M, P, Sn>), dim3((unsigned)(((uint64_t)nA * Sn + kBlock - 1) / kBlock)), \
                       dim3(kBlock), 0, s, oboxA, nA, oboxB, nB, part_of_a, boff, __VA_ARGS__)
#define PZ_BRUTE_P(M, P, ...)                                   \
    switch (brute_lanes(nA)) {                                  \
        case 64: PZ_BRUTE_S(M, P, 64, __VA_ARGS__); break;      \
        case 16: PZ_BRUTE_S(M, P, 16, __VA_ARGS__); break;      \
        case 4: PZ_BRUTE_S(M, P, 4, __VA_ARGS__); break;        \
        default: PZ_BRUTE_S(M, P, 1, __VA_ARGS__); break;       \
    }
#define PZ_BRUTE(M, ...)                          \
    do {                                          \
        if (boff) {                               \
            PZ_BRUTE_P(M, true, __VA_ARGS__)      \
        } else {                                  \
            PZ_BRUTE_P(M, false, __VA_ARGS__)     \
        }                                         \
    } while (0)

void broad_count(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, const Node* nodesB, const uint32_t* orderB,
                 int32_t nB, const uint32_t* part_of_a, const uint32_t* boff, uint32_t* cnt, hipStream_t s) {
    if (nA > 0 && !nodesB)
        PZ_BRUTE(0, cnt, (const uint64_t*)nullptr, (uint64_t*)nullptr, (unsigned long long*)nullptr, (uint64_t)0);
    else if (nA > 0)
        hipLaunchKernelGGL((k_broad<0>), dim3(grid1((uint64_t)nA * PZ_WAVE)), dim3(kBlock), 0, s, oboxA, nA, oboxB, nodesB, orderB, nB,
                           part_of_a, boff, cnt, (const uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t)0);
}
void broad_write(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, const Node* nodesB, const uint32_t* orderB,
                 int32_t nB, const uint32_t* part_of_a, const uint32_t* boff, const uint64_t* off, uint64_t* out,
                 hipStream_t s) {
    if (nA > 0 && !nodesB)
        PZ_BRUTE(1, (uint32_t*)nullptr, off, out, (unsigned long long*)nullptr, (uint64_t)0);
    else if (nA > 0)
        hipLaunchKernelGGL((k_broad<1>), dim3(grid1((uint64_t)nA * PZ_WAVE)), dim3(kBlock), 0, s, oboxA, nA, oboxB, nodesB, orderB, nB,
                           part_of_a, boff, (uint32_t*)nullptr, off, out, ~0ull);
}
void broad_append(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, const Node* nodesB,
                  const uint32_t* orderB, int32_t nB, const uint32_t* part_of_a, const uint32_t* boff, uint64_t* out,
                  unsigned long long* cursor, uint64_t cap, uint32_t* cnt, uint64_t* off, void* scan_tmp,
                  hipStream_t s) {
    if (nA > 0 && !nodesB) {
        PZ_BRUTE(2, (uint32_t*)nullptr, (const uint64_t*)nullptr, out, cursor, cap);
    } else if (nA > 0) {  // count, scan, write below cap: no atomics; the total is the scan's last entry
        broad_count(oboxA, nA, oboxB, nodesB, orderB, nB, part_of_a, boff, cnt, s);
        exclusive_scan_u32_u64(cnt, off, (uint64_t)nA, scan_tmp, s);
        hipLaunchKernelGGL((k_broad<1>), dim3(grid1((uint64_t)nA * PZ_WAVE)), dim3(kBlock), 0, s, oboxA, nA, oboxB,
                           nodesB, orderB, nB, part_of_a, boff, (uint32_t*)nullptr, off, out, cap);
        hipMemcpyAsync(cursor, off + nA, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
    }
}

#undef PZ_BRUTE
#undef PZ_BRUTE_P
#undef PZ_BRUTE_S

uint64_t grid_entry_bound(int32_t nB) { return (uint64_t)std::max(nB, 1) * kGridSpan; }

void grid_build(const uint32_t* oboxB, int32_t nB, const ObjGrid& g, void* scan_tmp, hipStream_t s) {
    KernelTimer kt_("grid_build", s);
    hipLaunchKernelGGL(k_grid_clear, dim3(grid1((uint64_t)g.nb + 1)), dim3(kBlock), 0, s, g.counts, (uint64_t)g.nb + 1,
                       g.stats, g.nbig);
    if (nB <= 0) return;
    hipLaunchKernelGGL(k_grid_stats, dim3(std::min(grid1((uint64_t)nB), kSceneBlocks)), dim3(kBlock), 0, s, oboxB, nB,
                       g.stats);
    hipLaunchKernelGGL((k_grid_refs<0>), dim3(grid1((uint64_t)nB * kGridLanes)), dim3(kBlock), 0, s, oboxB, nB,
                       (const float*)g.stats, g.nb, g.counts, (const uint32_t*)nullptr, (uint4*)nullptr,
                       (float4*)nullptr, g.big, g.nbig);
    exclusive_scan_u32_u32(g.counts, g.start, (uint64_t)g.nb + 1, scan_tmp, s);  // start[nb] = every reference
    hipLaunchKernelGGL((k_grid_refs<1>), dim3(grid1((uint64_t)nB * kGridLanes)), dim3(kBlock), 0, s, oboxB, nB,
                       (const float*)g.stats, g.nb, g.counts, (const uint32_t*)g.start, g.entries, g.ebox, g.big,
                       g.nbig);
}
void grid_count(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, int32_t nB, const ObjGrid& g, uint32_t* cnt,
                hipStream_t s) {
    if (nA > 0)
        hipLaunchKernelGGL((k_grid_broad<0>), dim3(grid1((uint64_t)nA * kGridLanes)), dim3(kBlock), 0, s, oboxA, nA,
                           oboxB, nB, (const float*)g.stats, g.nb, (const uint32_t*)g.start, (const uint4*)g.entries,
                           (const float4*)g.ebox,
                           (const uint32_t*)g.big, (const uint32_t*)g.nbig, cnt, (const uint64_t*)nullptr,
                           (uint64_t*)nullptr, (unsigned long long*)nullptr, (uint64_t)0);
}
void grid_write(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, int32_t nB, const ObjGrid& g,
                const uint64_t* off, uint64_t* out, hipStream_t s) {
    if (nA > 0)
        hipLaunchKernelGGL((k_grid_broad<1>), dim3(grid1((uint64_t)nA * kGridLanes)), dim3(kBlock), 0, s, oboxA, nA,
                           oboxB, nB, (const float*)g.stats, g.nb, (const uint32_t*)g.start, (const uint4*)g.entries,
                           (const float4*)g.ebox,
                           (const uint32_t*)g.big, (const uint32_t*)g.nbig, (uint32_t*)nullptr, off, out,
                           (unsigned long long*)nullptr, (uint64_t)0);
}
void grid_append(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, int32_t nB, const ObjGrid& g, uint64_t* out,
                 unsigned long long* cursor, uint64_t cap, hipStream_t s) {
    if (nA > 0)
        hipLaunchKernelGGL((k_grid_broad<2>), dim3(grid1((uint64_t)nA * kGridLanes)), dim3(kBlock), 0, s, oboxA, nA,
                           oboxB, nB, (const float*)g.stats, g.nb, (const uint32_t*)g.start, (const uint4*)g.entries,
                           (const float4*)g.ebox,
                           (const uint32_t*)g.big, (const uint32_t*)g.nbig, (uint32_t*)nullptr,
                           (const uint64_t*)nullptr, out, cursor, cap);
}

void mark_needed(const uint64_t* pairs, uint64_t np, uint32_t* needA, uint32_t* needB, hipStream_t s) {
    if (np) hipLaunchKernelGGL(k_mark_needed, dim3(grid1(np)), dim3(kBlock), 0, s, pairs, np, needA, needB);
}

void needed_layout(const uint32_t* need, const uint32_t* loc, int32_t nobj, const uint32_t* tri_cnt,
                   const uint32_t* edge_cnt, int32_t* obj_of_loc, uint32_t* seg_tri_cnt, uint32_t* seg_edge_cnt,
                   hipStream_t s) {
    if (nobj > 0)
        hipLaunchKernelGGL(k_needed_layout, dim3(grid1((uint64_t)nobj)), dim3(kBlock), 0, s, need, loc, nobj, tri_cnt,
                           edge_cnt, obj_of_loc, seg_tri_cnt, seg_edge_cnt);
}

void tri_keys_two_level(const float* verts, uint64_t nv, const uint32_t* idx, const int32_t* t2o, uint64_t nt,
                        int32_t nobj, const uint32_t* need, const uint32_t* loc, const uint32_t* obox,
                        const uint32_t* seg_off, const uint32_t* first_tri, const uint32_t* last_tri,
                        const uint32_t* tri_cnt, uint32_t* cursor, uint64_t* keys, uint32_t* vals, int mbits,
                        hipStream_t s) {
    KernelTimer kt_("tri_keys", s);
    if (nt)
        hipLaunchKernelGGL(k_tri_keys_two_level, dim3(grid1(nt)), dim3(kBlock), 0, s, verts, idx, t2o, nt, nobj, need,
                           loc, obox, seg_off, first_tri, last_tri, tri_cnt, cursor, keys, vals, nv, mbits);
}

void tri_keys_global(const float* verts, const uint32_t* idx, uint64_t nt, const uint32_t* scene, uint64_t* keys,
                     uint32_t* vals, hipStream_t s) {
    if (nt) hipLaunchKernelGGL(k_tri_keys_global, dim3(grid1(nt)), dim3(kBlock), 0, s, verts, idx, nt, scene, keys, vals);
}

void edge_scatter(const float* es, const float* ee, const int32_t* eo, uint64_t ne, int32_t nobj,
                  const uint32_t* need, const uint32_t* loc, const uint32_t* eseg_off, uint32_t* cursor,
                  EdgeRec* out, hipStream_t s) {
    KernelTimer kt_("edge_scatter", s);
    if (ne)
        hipLaunchKernelGGL(k_edge_scatter, dim3(grid1(ne)), dim3(kBlock), 0, s, es, ee, eo, ne, nobj, need, loc,
                           eseg_off, cursor, out);
}

void task_counts(const uint64_t* pairs, uint64_t np, const uint32_t* locA, const uint32_t* esegA,
                 const uint32_t* locB, const uint32_t* esegB, const uint8_t* meta, uint32_t* counts, hipStream_t s) {
    if (np)
        hipLaunchKernelGGL(k_task_counts, dim3(grid1(np)), dim3(kBlock), 0, s, pairs, np, locA, esegA, locB, esegB,
                           meta, counts);
}

bool narrow(const uint64_t* pairs, uint64_t np, const uint64_t* task_off, uint64_t ntasks, const SetView& A,
            const SetView& B, uint32_t* found, const uint8_t* meta, bool f32, int allow_f32, uint32_t height,
            unsigned long long* stats, hipStream_t s) {
    if (ntasks == 0) return true;
    int stack = stack_for_height(height);
    if (stack == 0) return false;
    unsigned g = (unsigned)((ntasks + kWavesPerBlock - 1) / kWavesPerBlock);
    if (stack == 32) {
        if (f32)
            launch_narrow<32, true>(g, s, pairs, np, task_off, A, B, found, meta, allow_f32, stats);
        else
            launch_narrow<32, false>(g, s, pairs, np, task_off, A, B, found, meta, allow_f32, stats);
    } else {
        if (f32)
            launch_narrow<64, true>(g, s, pairs, np, task_off, A, B, found, meta, allow_f32, stats);
        else
            launch_narrow<64, false>(g, s, pairs, np, task_off, A, B, found, meta, allow_f32, stats);
    }
    return true;
}

bool narrow_dual(const uint64_t* pairs, const SetView& A, const SetView& B, uint32_t* found, uint32_t* fallback,
                 const uint8_t* meta, bool f32, int allow_f32, int queue, const DualTask* tasks_in, uint64_t ntasks,
                 unsigned long long* next_task, int chunk, DualTask* tasks_out, unsigned long long* nout,
                 uint64_t out_cap, uint4* esc_list, unsigned long long* esc_count, uint64_t esc_cap, unsigned blocks,
                 unsigned long long* stats, hipStream_t s) {
    f32 = f32 && allow_f32 != 0;  // the fp32-only walk needs the filter's input range
    if (ntasks == 0) return true;
    if (chunk < 1) chunk = 1;
    const uint64_t need = (ntasks + (uint64_t)chunk * kDualWaves - 1) / ((uint64_t)chunk * kDualWaves);
    const unsigned g = (unsigned)(need < blocks ? need : blocks);
    KernelTimer kt_("narrow_dual", s);
#define PZ_LAUNCH_DUAL(F, Qn)                                                                                  \
    hipLaunchKernelGGL((k_narrow_dual<F, Qn>), dim3(g), dim3(kDualWaves * PZ_WAVE), 0, s, pairs, A, B, found,   \
                       fallback, meta, allow_f32, tasks_in, ntasks, next_task, chunk, tasks_out, nout, out_cap, \
                       esc_list, esc_count, esc_cap, stats)
#define PZ_DUAL_Q(Qn)                 \
    case Qn:                          \
        if (f32)                      \
            PZ_LAUNCH_DUAL(true, Qn);  \
        else                          \
            PZ_LAUNCH_DUAL(false, Qn); \
        break;
    switch (queue) {
        PZ_DUAL_Q(256)
        PZ_DUAL_Q(512)
        PZ_DUAL_Q(1024)
        PZ_DUAL_Q(1536)
        default:
            return false;
    }
#undef PZ_DUAL_Q
#undef PZ_LAUNCH_DUAL
    return true;
}

void mark_task_pairs(const DualTask* tasks, uint64_t n, uint32_t* fallback, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_mark_task_pairs, dim3(grid1(n)), dim3(kBlock), 0, s, tasks, n, fallback);
}

void escalate(const uint4* list, const unsigned long long* count, uint64_t cap, const SetView& A, const SetView& B,
              const uint8_t* meta, uint32_t* found, unsigned long long* stats, hipStream_t s) {
    if (cap == 0) return;
    const uint64_t g = std::min<uint64_t>(grid1(cap), 1024);
    hipLaunchKernelGGL(k_escalate, dim3((unsigned)g), dim3(kBlock), 0, s, list, count, cap, A, B, meta, found, stats);
}

int narrow_dual_waves() { return kDualWaves; }

bool parity(const uint64_t* pairs, uint64_t np, const uint32_t* found, const SetView& A, const SetView& B,
            const float* vA, const uint32_t* iA, const uint32_t* ftA, const float* vB, const uint32_t* iB,
            const uint32_t* ftB, const uint8_t* meta, uint32_t* inside, uint32_t height,
            unsigned long long* stats, hipStream_t s) {
    if (np == 0) return true;
    int stack = stack_for_height(height);
    if (stack == 0) return false;
    uint64_t n = np * 2;
    KernelTimer kt_("parity", s);
    if (stack == 32)
        hipLaunchKernelGGL((k_parity<32>), dim3(grid1(n)), dim3(kBlock), 0, s, pairs, np, found, A, B, vA, iA, ftA, vB,
                           iB, ftB, meta, inside, stats);
    else
        hipLaunchKernelGGL((k_parity<64>), dim3(grid1(n)), dim3(kBlock), 0, s, pairs, np, found, A, B, vA, iA, ftA, vB,
                           iB, ftB, meta, inside, stats);
    return true;
}

bool track_rays(const float* vS, const uint32_t* iS, const uint32_t* ftS, const uint32_t* tcS, int32_t nsrc,
                const Node* nodes, const LeafTri* leaves, uint64_t ntri, uint32_t height, int max_targets,
                int32_t* ids, uint32_t* hits, uint32_t* count, uint32_t* overflow, uint32_t* iters,
                unsigned long long* stats, hipStream_t s) {
    if (nsrc <= 0) return true;
    const int stack = stack_for_height(height);
    if (stack == 0) return false;
    KernelTimer kt_("track_rays", s);
    if (stack == 32)
        hipLaunchKernelGGL((k_track_rays<32>), dim3(grid1((uint64_t)nsrc)), dim3(kBlock), 0, s, vS, iS, ftS, tcS, nsrc,
                           nodes, leaves, ntri, max_targets, ids, hits, count, overflow, iters, stats);
    else
        hipLaunchKernelGGL((k_track_rays<64>), dim3(grid1((uint64_t)nsrc)), dim3(kBlock), 0, s, vS, iS, ftS, tcS, nsrc,
                           nodes, leaves, ntri, max_targets, ids, hits, count, overflow, iters, stats);
    return true;
}

void parity_keys(const int32_t* ids, const uint32_t* hits, const uint32_t* count, const uint64_t* off, int32_t nsrc,
                 int max_targets, bool src_is_mesh2, uint64_t* keys, hipStream_t s) {
    if (nsrc > 0)
        hipLaunchKernelGGL(k_parity_keys, dim3(grid1((uint64_t)nsrc)), dim3(kBlock), 0, s, ids, hits, count, off, nsrc,
                           max_targets, src_is_mesh2 ? 1 : 0, keys);
}

void key_difference(const uint64_t* in, uint64_t n, const uint64_t* ref, uint64_t nref, uint64_t* out,
                    uint32_t* out_type, uint32_t type, unsigned long long* count, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_key_difference, dim3(grid1(n)), dim3(kBlock), 0, s, in, n, ref, nref, out, out_type, type,
                           count);
}

void select_pairs(const uint64_t* pairs, uint64_t np, const uint32_t* found, const uint32_t* inside,
                  const uint8_t* meta, int include_overlap, uint64_t* out_keys, uint32_t* out_type,
                  unsigned long long* count, unsigned long long* touching, hipStream_t s) {
    if (np)
        hipLaunchKernelGGL(k_select_pairs, dim3(grid1(np)), dim3(kBlock), 0, s, pairs, np, found, inside, meta,
                           include_overlap, out_keys, out_type, count, touching);
}

bool trace_global(int mode, const float* es, const float* ee, const int32_t* eo, uint64_t ne, const Node* nodes,
                  const LeafTri* leaves, uint64_t ntri, uint32_t height, int swap, bool f32, int allow_f32,
                  uint32_t* counts, const uint64_t* off, uint64_t* out, unsigned long long* table, uint64_t mask,
                  uint32_t* overflow, unsigned long long* stats, hipStream_t s) {
    if (ne == 0) return true;
    int stack = stack_for_height(height);
    if (stack == 0) return false;
    unsigned g = grid1(ne);
    KernelTimer kt_("trace_global", s);
    if (stack == 32) {
        if (f32)
            launch_trace<32, true>(mode, g, s, es, ee, eo, ne, nodes, leaves, ntri, swap, allow_f32, counts, off, out,
                                   table, mask, overflow, stats);
        else
            launch_trace<32, false>(mode, g, s, es, ee, eo, ne, nodes, leaves, ntri, swap, allow_f32, counts, off, out,
                                    table, mask, overflow, stats);
    } else {
        if (f32)
            launch_trace<64, true>(mode, g, s, es, ee, eo, ne, nodes, leaves, ntri, swap, allow_f32, counts, off, out,
                                   table, mask, overflow, stats);
        else
            launch_trace<64, false>(mode, g, s, es, ee, eo, ne, nodes, leaves, ntri, swap, allow_f32, counts, off, out,
                                    table, mask, overflow, stats);
    }
    return true;
}

void compact_table(const unsigned long long* table, uint64_t size, uint64_t* out, uint64_t cap,
                   unsigned long long* count, hipStream_t s) {
    if (size) hipLaunchKernelGGL(k_compact_table, dim3(grid1(size)), dim3(kBlock), 0, s, table, size, out, cap, count);
}

void pair_meta(const uint64_t* pairs, uint64_t np, const uint32_t* part_of_a, const uint8_t* meta_of_part,
               uint8_t uniform, uint8_t* meta, hipStream_t s) {
    if (np)
        hipLaunchKernelGGL(k_pair_meta, dim3(grid1(np)), dim3(kBlock), 0, s, pairs, np, part_of_a, meta_of_part,
                           uniform, meta);
}

void pair_prep(const uint64_t* pairs, uint64_t np, const uint32_t* part_of_a, const uint8_t* meta_of_part,
               uint8_t uniform, uint8_t* meta, const PairPrepZero& z, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_prep, dim3(grid1(np + 1)), dim3(kBlock), 0, s, pairs, np, part_of_a, meta_of_part, uniform,
                       meta, z);
}

void compact_flagged(const uint64_t* pairs, uint64_t np, const uint32_t* flag, uint64_t* out, uint32_t* out_index,
                     unsigned long long* count, hipStream_t s) {
    if (np) hipLaunchKernelGGL(k_compact_flagged, dim3(grid1(np)), dim3(kBlock), 0, s, pairs, np, flag, out, out_index, count);
}

void scatter_found(const uint32_t* sub_found, const uint32_t* index, uint64_t n, uint32_t* found, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_scatter_found, dim3(grid1(n)), dim3(kBlock), 0, s, sub_found, index, n, found);
}

void fill_ull(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL((k_fill_any<unsigned long long>), dim3(grid1(n)), dim3(kBlock), 0, s, p, v, n);
}

}  // namespace pzk
