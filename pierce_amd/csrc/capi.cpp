// C ABI implementation: context, datasets in HBM, and the join pipelines.
//
// Pipelines (all device work on the context stream; the host waits only where
// a size must be known to launch the next step):
//   two-level : object stats -> mesh2 object LBVH -> broad phase (two-pass or
//               estimate-sized append with exact recount on overflow) ->
//               per-object LBVHs + grouped edges of the objects in some pair ->
//               narrow phase with early exit -> parity -> select -> sort.
//   global    : set-wide triangle LBVH per set -> every edge of one set traced
//               through the other's BVH (reference shape,
//               pierce/query/src/cuda/mesh_overlap_edges.cu:105-208) -> two-pass
//               count/scan/write/sort/unique
//               (pierce/query/src/applications/pierce_overlap_two_pass_main.cpp:39-156)
//               or hash-set dedup (pierce/query/src/applications/pierce_overlap_main.cpp:74-187).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "host/host_mesh.hpp"
#include "kernels/pz_kernels_est.h"
#include "kernels/pz_kernels_join.h"
#include "pierce_hip.h"
#include "pz_internal.h"

using namespace pz;

namespace pz {
// predicate | direction bits << 2 of one join description
uint8_t pair_meta_of(const pz_join_desc& d) {
    const int dir = d.predicate == PZ_PRED_CONTAINMENT
                        ? 3
                        : (d.direction == PZ_DIR_1TO2 ? 1 : (d.direction == PZ_DIR_2TO1 ? 2 : 3));
    return (uint8_t)(d.predicate | (dir << 2));
}

// Finds whether the objects of a just-uploaded dataset own disjoint, gap-free
// vertex runs and, if so, keeps the per-object triangle ranges and the
// chunks of the runs.  The runs are ordered by object id (every non-empty
// object's least vertex beyond the previous one's greatest), no vertex is
// referenced by two objects, and the runs' lengths add up to the number of
// referenced vertices, so every vertex of a run is referenced by its object.
void analyze_topology(pz_ctx* c, pz_dataset* d) {
    const uint64_t n = (uint64_t)std::max(d->nobj, 1);
    if (d->nobj <= 0 || d->nt == 0 || d->nv >= (1ull << 32) - pzk::run_chunk_vertices()) return;
    uint32_t* topo = nullptr;  // first, last, cnt, vmax, vmin (n each)
    int32_t* owner = nullptr;
    uint32_t* flags = nullptr;  // clash flag, then the referenced-vertex count (8 B)
    auto cleanup = [&]() {
        for (void* p : {(void*)owner, (void*)flags})
            if (p) (void)hipFree(p);
    };
    try {
        HIP_OK(hipMalloc((void**)&topo, 20 * n));
        uint32_t* vmin = topo + 4 * n;
        HIP_OK(hipMalloc((void**)&owner, 4 * std::max<uint64_t>(d->nv, 1)));
        HIP_OK(hipMalloc((void**)&flags, 16));
        pzk::topology_scan(d->idx, d->t2o, d->nt, d->nv, d->nobj, topo, topo + n, topo + 2 * n, vmin, topo + 3 * n,
                           owner, flags, reinterpret_cast<unsigned long long*>(flags + 2), c->s);
        HIP_OK(hipGetLastError());
        std::vector<uint32_t> cnt(n), lo(n), hi(n), fl(4);
        HIP_OK(hipMemcpyAsync(cnt.data(), topo + 2 * n, 4 * n, hipMemcpyDeviceToHost, c->s));
        HIP_OK(hipMemcpyAsync(lo.data(), vmin, 4 * n, hipMemcpyDeviceToHost, c->s));
        HIP_OK(hipMemcpyAsync(hi.data(), topo + 3 * n, 4 * n, hipMemcpyDeviceToHost, c->s));
        HIP_OK(hipMemcpyAsync(fl.data(), flags, 16, hipMemcpyDeviceToHost, c->s));
        HIP_OK(hipStreamSynchronize(c->s));
        uint64_t referenced = (uint64_t)fl[2] | ((uint64_t)fl[3] << 32), run_total = 0;
        bool ok = fl[0] == 0;
        int64_t prev = -1;
        std::vector<uint2> chunks;
        const uint32_t K = pzk::run_chunk_vertices();
        for (uint64_t o = 0; o < n && ok; ++o) {
            if (cnt[o] == 0) continue;
            ok = (int64_t)lo[o] > prev && hi[o] >= lo[o];
            prev = hi[o];
            run_total += (uint64_t)hi[o] - lo[o] + 1;
            for (uint64_t v = lo[o]; v <= hi[o]; v += K) chunks.push_back(make_uint2((uint32_t)o, (uint32_t)v));
        }
        ok = ok && run_total == referenced;
        cleanup();
        owner = nullptr;
        flags = nullptr;
        if (!ok) {
            (void)hipFree(topo);
            return;
        }
        d->topo = topo;
        topo = nullptr;
        HIP_OK(hipMalloc((void**)&d->topo_chunks, 8 * std::max<size_t>(chunks.size(), 1)));
        if (!chunks.empty())
            HIP_OK(hipMemcpy(d->topo_chunks, chunks.data(), 8 * chunks.size(), hipMemcpyHostToDevice));
        d->topo_nchunks = chunks.size();
        d->vruns = true;
    } catch (...) {
        cleanup();
        if (topo) (void)hipFree(topo);
        throw;
    }
}

uint64_t next_dataset_uid() {
    static std::atomic<uint64_t> next{1};
    return next++;
}

SetPrep prep_set(pz_ctx* c, const pz_dataset* d, const std::string& tag) {
    SetPrep p;
    size_t n = std::max<int32_t>(d->nobj, 1);
    p.obox = buf<uint32_t>(c, tag + ".obox", 6 * n);
    p.edge_cnt = buf<uint32_t>(c, tag + ".edge_cnt", n);
    if (d->vruns) {  // triangle ranges from the upload's topology, boxes from the vertex runs
        p.first_tri = d->topo;
        p.last_tri = d->topo + n;
        p.tri_cnt = d->topo + 2 * n;
        pzk::object_bounds_runs(d->v, d->topo_chunks, d->topo_nchunks, d->topo + 3 * n, d->nobj, p.obox, c->s);
        return p;
    }
    p.tri_cnt = buf<uint32_t>(c, tag + ".tri_cnt", n);
    p.first_tri = buf<uint32_t>(c, tag + ".first_tri", n);
    p.last_tri = buf<uint32_t>(c, tag + ".last_tri", n);
    pzk::object_stats(d->v, d->nv, d->idx, d->t2o, d->nt, d->nobj, p.obox, p.tri_cnt, p.first_tri, p.last_tri,
                      c->s);
    return p;
}
}  // namespace pz

namespace {

// Phase timer: events recorded on the stream at phase boundaries.  The first
// event is recorded when the timer is made, so the first phase also holds the
// host's set-up before its first launch and the phases cover the whole call
// up to the final synchronisation.
struct Phases {
    pz_ctx* c;
    pz_join_result* r;
    int n = 0;
    explicit Phases(pz_ctx* ctx, pz_join_result* res) : c(ctx), r(res) {
        HIP_OK(hipEventRecord(c->ev[0], c->s));
    }
    void begin(const char* name) {
        if (n >= PZ_MAX_PHASES) return;
        std::strncpy(r->phases[n].name, name, sizeof(r->phases[n].name) - 1);
        if (n > 0) HIP_OK(hipEventRecord(c->ev[n], c->s));
        ++n;
    }
    void finish() {
        if (n == 0) return;
        HIP_OK(hipEventRecord(c->ev[n], c->s));
        HIP_OK(hipEventSynchronize(c->ev[n]));
        for (int i = 0; i < n; ++i) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
            r->phases[i].ms = ms;
        }
        r->num_phases = n;
    }
};

// ------------------------------------------------------------- set prep

ObjBvh build_object_bvh(pz_ctx* c, const pz_dataset* d, const SetPrep& p, const std::string& tag) {
    // Batched sets: one subtree per part (keys carry the part above the Morton bits).
    const bool parts = !d->part_obj_off.empty();
    const uint64_t nparts = parts ? d->part_obj_off.size() - 1 : 1;
    const uint64_t n = std::max<int32_t>(d->nobj, 0);
    ObjBvh b;
    uint32_t* scene = buf<uint32_t>(c, tag + ".scene", 8);
    uint64_t* keys = buf<uint64_t>(c, tag + ".okeys", n + 1);
    uint32_t* vals = buf<uint32_t>(c, tag + ".ovals", n + 1);
    uint64_t* keys2 = buf<uint64_t>(c, tag + ".okeys2", n + 1);
    uint32_t* vals2 = buf<uint32_t>(c, tag + ".ovals2", n + 1);
    b.nodes = buf<pzd::Node>(c, tag + ".onodes", n + 1);
    uint32_t* pint = buf<uint32_t>(c, tag + ".opint", n + 1);
    uint32_t* pleaf = buf<uint32_t>(c, tag + ".opleaf", n + 1);
    uint32_t* cnt = buf<uint32_t>(c, tag + ".ocnt", n + 1);
    b.height = buf<uint32_t>(c, tag + ".oheight", 1);
    void* tmp = buf<char>(c, "sort.tmp.obj", pzk::radix_sort_temp_bytes(n));
    pzk::scene_box(p.obox, d->nobj, scene, c->s);
    pzk::object_keys(p.obox, d->nobj, scene, parts ? d->part_of : nullptr, keys, vals, c->s);
    bool alt = pzk::radix_sort_pairs(keys, vals, keys2, vals2, n, 0, 30 + (parts ? bits_for(nparts) : 0), tmp, c->s);
    uint64_t* sk = alt ? keys2 : keys;
    b.order = alt ? vals2 : vals;
    pzk::fill_u32(b.height, 0u, 1, c->s);
    const uint32_t* seg = parts ? d->part_off : nullptr;
    uint32_t* rng = buf<uint32_t>(c, tag + ".orange", n + 1);
    pzk::lbvh_karras(sk, n, seg, 30, b.nodes, pint, pleaf, cnt, rng, c->s);
    pzk::lbvh_refit_ord_boxes(sk, b.order, n, seg, 30, p.obox, b.nodes, pint, pleaf, cnt, rng, b.height, c->s);
    return b;
}

// Hashed grid of one set's object boxes for the broad phase of a single join:
// twice as many buckets as objects (a power of two).
pzk::ObjGrid build_object_grid(pz_ctx* c, const pz_dataset* d, const SetPrep& p) {
    const int32_t n = std::max<int32_t>(d->nobj, 0);
    pzk::ObjGrid g;
    uint32_t nb = 1024;
    while (nb < 2ull * (uint64_t)n && nb < (1u << 30)) nb <<= 1;
    g.nb = nb;
    g.counts = buf<uint32_t>(c, "grid.counts", (size_t)nb + 1);
    g.start = buf<uint32_t>(c, "grid.start", (size_t)nb + 1);
    g.entries = buf<uint4>(c, "grid.entries", pzk::grid_entry_bound(n));
    g.ebox = buf<float4>(c, "grid.ebox", 2 * pzk::grid_entry_bound(n));
    g.big = buf<uint32_t>(c, "grid.big", (size_t)std::max(n, 1));
    g.nbig = buf<uint32_t>(c, "grid.nbig", 1);
    g.stats = buf<float>(c, "grid.stats", 4);
    void* tmp = buf<char>(c, "scan.tmp.grid", pzk::scan_temp_bytes((uint64_t)nb + 1));
    pzk::grid_build(p.obox, n, g, tmp, c->s);
    return g;
}

// Per-object LBVHs and grouped edges of the needed objects of one set.
struct Needed {
    uint32_t* need;
    uint32_t* loc;
    uint32_t* seg_off;
    uint32_t* eseg_off;
    int32_t* obj_of_loc;
};

Needed needed_layout(pz_ctx* c, const pz_dataset* d, const SetPrep& p, const std::string& tag) {
    Needed nd;
    size_t n = std::max<int32_t>(d->nobj, 1);
    nd.need = buf<uint32_t>(c, tag + ".need", n);
    nd.loc = buf<uint32_t>(c, tag + ".loc", n + 1);
    nd.seg_off = buf<uint32_t>(c, tag + ".seg_off", n + 1);
    nd.eseg_off = buf<uint32_t>(c, tag + ".eseg_off", n + 1);
    nd.obj_of_loc = buf<int32_t>(c, tag + ".obj_of_loc", n);
    return nd;
}

struct SetAccel {
    pzd::BNode* nodes;  // per-object BVHs, own-box nodes
    pzd::LeafTri* leaves;
    pzd::EdgeRec* edges;
    uint32_t* height;
};

void init_result(pz_join_result* r) {
    std::memset(r, 0, sizeof(*r));
}

// Result counters from the device counter block (pzk::kStatSlots values).
void set_counters(pz_join_result* r, const std::vector<unsigned long long>& st) {
    r->tests_total = st[0];
    r->tests_fp64 = st[1];
    r->tests_exact = st[2];
    r->tests_fp32_misclassified = st[3];
    r->edges_traced = st[4];
    r->hash_accesses = st[5];
    r->hash_contentions = st[6];
    r->parity_rays = st[pzk::kStatParityRays];
    r->parity_crossings = st[pzk::kStatParityRays + 1];
    r->parity_rays_crossing = st[pzk::kStatParityRays + 2];
    r->parity_inside = st[pzk::kStatParityRays + 3];
    r->parity_tests = st[pzk::kStatParityRays + 4];
    r->walks_started = st[pzk::kStatWalks] + st[4];
    r->pairs_touching = st[pzk::kStatTouching];
    r->sieve_triangles = st[pzk::kStatSieveTris];
}

// Final stage shared by both strategies: sort unique keys (+types), copy to host.
// The result's host arrays from n pairs (and, with types, n u32 pair types)
// that have landed in pinned memory.
void land_pairs(pz_join_result* r, const char* h_pairs, const uint32_t* h_types, uint64_t n, bool want_types) {
    r->num_pairs = n;
    r->pairs = static_cast<pz_pair*>(std::malloc(std::max<uint64_t>(n, 1) * sizeof(pz_pair)));
    if (!r->pairs) throw PzError(PZ_E_NOMEM, "host result allocation failed");
    if (n) std::memcpy(r->pairs, h_pairs, n * sizeof(pz_pair));
    if (want_types) {
        r->pair_types = static_cast<uint8_t*>(std::malloc(std::max<uint64_t>(n, 1)));
        if (!r->pair_types) throw PzError(PZ_E_NOMEM, "host result allocation failed");
        for (uint64_t i = 0; i < n; ++i) r->pair_types[i] = static_cast<uint8_t>(h_types[i]);
    }
}

void finish_pairs(pz_ctx* c, uint64_t* keys, uint32_t* types, uint64_t n, int32_t nA, bool want_types,
                  pz_join_result* r) {
    uint64_t* keys2 = buf<uint64_t>(c, "fin.keys2", n + 1);
    uint32_t* types2 = buf<uint32_t>(c, "fin.types2", n + 1);
    bool alt;
    if (n > pzk::kSortTileKeys && n <= pzk::kSortTilesMax) {  // sorted tiles merged by rank: two launches
        alt = pzk::sort_tiles(keys, types, keys2, types2, n, c->s);
    } else {
        void* tmp = buf<char>(c, "sort.tmp.fin", pzk::radix_sort_temp_bytes(n));
        const int bits = 32 + bits_for((uint64_t)std::max<int32_t>(nA, 1));
        alt = types ? pzk::radix_sort_pairs(keys, types, keys2, types2, n, 0, bits, tmp, c->s)
                    : pzk::radix_sort_keys(keys, keys2, n, 0, bits, tmp, c->s);
    }
    uint64_t* sk = alt ? keys2 : keys;
    uint32_t* st = alt ? types2 : types;
    // the pairs (and types) land in pinned memory straight from the kernel, then host copies
    const size_t pb = n * sizeof(pz_pair), tb = want_types ? n * 4 : 0;
    const size_t at = (pb + tb + 7) & ~size_t(7);
    char* h = static_cast<char*>(pinned(c, at + 8));
    unsigned long long* hn = reinterpret_cast<unsigned long long*>(h + at);  // the count, for the counted kernel
    *hn = n;
    pzk::keys_to_pairs_counted(sk, want_types ? st : nullptr, hn, n, reinterpret_cast<int32_t*>(h),
                               reinterpret_cast<uint32_t*>(h + pb), c->s);
    HIP_OK(hipStreamSynchronize(c->s));
    land_pairs(r, h, reinterpret_cast<const uint32_t*>(h + pb), n, want_types);
}

// ------------------------------------------------------------- estimates
// Whole-set grid estimate of pierce/query/src/applications/pierce_overlap_main.cpp:376-471
// (the same steps as pierce_intersection_main.cpp:213-311 and
// pierce_containment_main.cpp:65-135): cells of b are matched to a's by index
// (a's last entry wins for a repeated index), the per-cell formula of
// pierce/query/src/cuda/estimated_overlap.cu:8-41 is summed over the matches, and
// the sum is divided by the replication factor alpha.
void estimate_grid(const pz_grid_cell* ga, uint64_t na, const pz_grid_cell* gb, uint64_t nb, float cell_size,
                   float eps, float gamma, bool alpha_corr, pz_grid_estimate* o) {
    *o = pz_grid_estimate{};
    o->alpha = 1.0f;
    o->avg_vol_ratio[0] = o->avg_vol_ratio[1] = 1.0f;
    if (!na || !nb || !(cell_size > 0)) return;
    o->has_grid = 1;
    struct K {
        int32_t x, y, z;
        bool operator<(const K& k) const { return std::tie(x, y, z) < std::tie(k.x, k.y, k.z); }
    };
    std::map<K, const pz_grid_cell*> ma;
    for (uint64_t i = 0; i < na; ++i) ma[K{ga[i].ix, ga[i].iy, ga[i].iz}] = &ga[i];
    std::vector<pz_grid_cell> A, B;
    for (uint64_t i = 0; i < nb; ++i) {
        auto it = ma.find(K{gb[i].ix, gb[i].iy, gb[i].iz});
        if (it != ma.end()) {
            A.push_back(*it->second);
            B.push_back(gb[i]);
        }
    }
    float cell_vol = cell_size * cell_size * cell_size;
    o->matched_cells = (int32_t)A.size();
    float raw = A.empty() ? 0.0f
                          : pz_estimate_overlap_selectivity(A.data(), B.data(), (int32_t)A.size(), cell_vol, eps, gamma);
    o->raw_pairs = raw;
    // touch-weighted means over every cell of one side (an empty side: size 0, ratio 1)
    auto avg = [](const pz_grid_cell* g, uint64_t n, bool size) {
        double t = 0;
        long long w = 0;
        for (uint64_t i = 0; i < n; ++i)
            if (g[i].touch_count > 0) {
                t += (double)(size ? g[i].avg_size_mean : g[i].vol_ratio) * g[i].touch_count;
                w += g[i].touch_count;
            }
        return w ? (float)(t / w) : (size ? 0.0f : 1.0f);
    };
    o->avg_size[0] = avg(ga, na, true);
    o->avg_size[1] = avg(gb, nb, true);
    o->avg_vol_ratio[0] = avg(ga, na, false);
    o->avg_vol_ratio[1] = avg(gb, nb, false);
    for (int s = 0; s < 2; ++s) o->effective_size[s] = o->avg_size[s] * std::cbrt(o->avg_vol_ratio[s]);
    float comb = o->effective_size[0] + o->effective_size[1];
    float mink = comb * comb * comb;
    if (cell_vol < 1e-9f) cell_vol = 1e-9f;
    if (alpha_corr) o->alpha = std::max(1.0f, mink / cell_vol);
    o->estimated_pairs = (int64_t)(raw / o->alpha);
}

double grid_estimate(const pz_dataset* a, const pz_dataset* b, float gamma, float eps, bool alpha_corr) {
    pz_grid_estimate e;
    estimate_grid(a->grid.data(), a->grid.size(), b->grid.data(), b->grid.size(), a->cell_size, eps, gamma,
                  alpha_corr, &e);
    return e.has_grid ? (double)e.estimated_pairs : -1.0;
}

// ------------------------------------------------------------- two-level
// meta_of_part: device predicate | dir << 2 per part for batched sets (null: from d).
// ------------------------------------------------------------ HIP graphs
// A fixed-shape segment of device work (its launches depend only on the
// datasets) is run eagerly the first time its key is seen, captured as a HIP
// graph the second time, and replayed after that.  The pool pointers the
// segment uses are stored with the graph; any buffer move (pool epoch change)
// drops the graph and starts over.  Only kernels are captured (no memset or
// copy nodes), and nothing is captured while per-kernel profiling is on.
}  // namespace
struct GraphSeg {
    uint64_t k1 = 0, k2 = 0;
    int tag = 0;
    uint64_t epoch = 0;
    hipGraphExec_t exec = nullptr;
    SegOut out;
};
namespace {

void drop_graph(GraphSeg* g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    g->exec = nullptr;
}

template <class F>
SegOut run_segment(pz_ctx* c, uint64_t k1, uint64_t k2, int tag, pz_join_result* r, F&& enqueue) {
    GraphSeg* g = nullptr;
    for (GraphSeg* e : c->segs)
        if (e->k1 == k1 && e->k2 == k2 && e->tag == tag) g = e;
    if (!c->graphs || pzk::prof_enabled()) return enqueue();  // plain launches, nothing recorded
    if (g && g->epoch == c->pool_epoch) {
        if (!g->exec) {  // second use with unchanged buffers: capture (no allocation can happen)
            hipGraph_t graph = nullptr;
            HIP_OK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
            SegOut o;
            try {
                o = enqueue();
            } catch (...) {
                (void)hipStreamEndCapture(c->s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            HIP_OK(hipStreamEndCapture(c->s, &graph));
            const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess || g->epoch != c->pool_epoch) {
                drop_graph(g);
                throw PzError(PZ_E_HIP, "graph capture of a fixed-shape segment failed");
            }
            g->out = o;
        }
        HIP_OK(hipGraphLaunch(g->exec, c->s));
        if (r) ++r->graph_launches;
        return g->out;
    }
    SegOut o = enqueue();
    if (!g) {
        if (c->segs.size() >= 16) {  // keep the most recent keys
            drop_graph(c->segs.front());
            delete c->segs.front();
            c->segs.erase(c->segs.begin());
        }
        g = new GraphSeg();
        g->k1 = k1;
        g->k2 = k2;
        g->tag = tag;
        c->segs.push_back(g);
    }
    drop_graph(g);
    g->epoch = c->pool_epoch;
    g->out = o;
    return o;
}

enum { kSegPrep = 1, kSegObjBvh = 2, kSegLayout = 3, kSegObjGrid = 4 };
// Largest mesh2 object count whose broad phase tests every box instead of
// walking an object LBVH (a walk costs tens of microseconds of dependent
// loads whatever the set size; a full scan of this many boxes costs less).
constexpr int32_t kBruteMaxObjects = 4096;

// preds_out (optional) receives the learned per-part predictions.
void join_two_level(pz_ctx* c, const pz_dataset* A, const pz_dataset* B, const pz_join_desc& d, pz_join_result* r,
                    Phases& ph, const uint8_t* meta_of_part = nullptr,
                    std::vector<SizePrediction>* preds_out = nullptr,
                    std::vector<uint64_t>* part_cand_out = nullptr) {
    const bool batched = meta_of_part != nullptr;
    hipStream_t s = c->s;
    unsigned long long* stats = buf<unsigned long long>(c, "stats", pzk::kStatSlots);
    pzk::fill_ull(stats, 0ull, pzk::kStatSlots, s);
    const int32_t nA = A->nobj, nB = B->nobj;

    ph.begin("Object Bounds");
    const SegOut prep = run_segment(c, A->uid, B->uid, kSegPrep, r, [&]() {
        SegOut o;
        o.pa = prep_set(c, A, "A");
        o.pb = prep_set(c, B, "B");
        return o;
    });
    const SetPrep pa = prep.pa, pb = prep.pb;
    // LEARNED: features and inference run on the side stream while the main
    // stream builds the object index (both only read the object statistics).
    uint32_t npred = 0;
    if (d.sizing == PZ_SIZING_LEARNED) npred = mlp_predict_async(c, d.model, A, B, pa, pb, meta_of_part, pair_meta_of(d));

    // Few mesh2 objects: the broad phase tests every box (k_broad_brute) and
    // no object index is built; otherwise a single join hashes the mesh2
    // boxes into a grid of their mean extent, and a batched join (whose parts
    // keep to themselves) walks an object LBVH with one subtree per part.
    const bool brute = d.broad_mode == 2 || (d.broad_mode == 0 && nB <= kBruteMaxObjects);
    const bool grid = !brute && !batched && (d.broad_mode == 0 || d.broad_mode == 3);
    ph.begin("Build Object Index");
    ObjBvh ob;
    pzk::ObjGrid og{};
    if (grid)
        og = run_segment(c, B->uid, 0, kSegObjGrid, r, [&]() {
                 SegOut o;
                 o.og = build_object_grid(c, B, pb);
                 return o;
             }).og;
    else if (!brute)
        ob = run_segment(c, B->uid, 0, kSegObjBvh, r, [&]() {
                 SegOut o;
                 o.ob = build_object_bvh(c, B, pb, "B");
                 return o;
             }).ob;

    ph.begin("Broad Phase");
    // The broad-phase walk keeps a 64-entry stack: the object LBVH must be shallower.
    const uint32_t oheight = (brute || grid) ? 0u : readback<uint32_t>(c, {ob.height})[0];
    double learned_upper = 0;  // LEARNED: calibrated upper bound of the candidate-pair count
    if (d.sizing == PZ_SIZING_LEARNED) {
        std::vector<SizePrediction> pr = mlp_predict_wait(c, d.model, npred);
        double cand = 0, pairs = 0;
        for (const SizePrediction& p : pr) {
            cand += p.cand;
            pairs += p.pairs;
            learned_upper += p.cand_upper;
        }
        r->predicted = (uint64_t)std::llround(cand);
        r->predicted_pairs = (uint64_t)std::llround(pairs);
        if (preds_out) *preds_out = std::move(pr);
    }
    if (oheight >= 64)
        throw PzError(PZ_E_LIMIT, "object BVH height " + std::to_string(oheight) + " exceeds the broad-phase stack (63)");
    uint64_t np = 0;
    uint64_t* pairs = nullptr;
    auto two_pass = [&]() {
        uint32_t* cnt = buf<uint32_t>(c, "bp.cnt", (size_t)std::max(nA, 1));
        uint64_t* off = buf<uint64_t>(c, "bp.off", (size_t)std::max(nA, 1) + 1);
        void* tmp = buf<char>(c, "scan.tmp", pzk::scan_temp_bytes((uint64_t)std::max(nA, 1)));
        if (grid)
            pzk::grid_count(pa.obox, nA, pb.obox, nB, og, cnt, s);
        else
            pzk::broad_count(pa.obox, nA, pb.obox, ob.nodes, ob.order, nB, batched ? A->part_of : nullptr,
                             batched ? B->part_off : nullptr, cnt, s);
        pzk::exclusive_scan_u32_u64(cnt, off, (uint64_t)std::max(nA, 0), tmp, s);
        np = readback<uint64_t>(c, {off + std::max(nA, 0)})[0];
        pairs = buf<uint64_t>(c, "bp.pairs", np + 1);
        if (grid)
            pzk::grid_write(pa.obox, nA, pb.obox, nB, og, off, pairs, s);
        else
            pzk::broad_write(pa.obox, nA, pb.obox, ob.nodes, ob.order, nB, batched ? A->part_of : nullptr,
                             batched ? B->part_off : nullptr, off, pairs, s);
    };
    if (d.sizing == PZ_SIZING_TWO_PASS) {
        two_pass();
        r->capacity = np;
    } else {
        uint64_t cap = d.capacity;
        if (d.sizing == PZ_SIZING_ESTIMATE) {
            // The buffer holds candidate object pairs, not result pairs: size it from
            // the expected number of box pairs that meet (a 16^3 grid over both sets'
            // object boxes, the estimator's feature 12), with room for twice the margin.
            const double pred = candidate_estimate(c, A, B, pa, pb);
            r->predicted = (uint64_t)std::llround(pred);
            const float margin = d.margin > 0 ? d.margin : 0.25f;
            cap = (uint64_t)std::ceil(std::min(pred, 0x1p40) * (1.0 + 2.0 * margin)) + 1024;
        } else if (d.sizing == PZ_SIZING_LEARNED) {
            cap = (uint64_t)std::ceil(std::min(learned_upper, 0x1p40)) + 64;
        }
        // never more slots than object pairs (or 2^32): a larger count overflows into the exact recount
        cap = std::min<uint64_t>(cap, std::min<uint64_t>((uint64_t)std::max(nA, 1) * (uint64_t)std::max(nB, 1) + 64,
                                                         1ull << 32));
        cap = std::max<uint64_t>(cap, 1);
        r->capacity = cap;
        pairs = buf<uint64_t>(c, "bp.pairs", cap + 1);
        unsigned long long* cursor = buf<unsigned long long>(c, "bp.cursor", 1);
        pzk::fill_ull(cursor, 0ull, 1, s);
        if (grid)
            pzk::grid_append(pa.obox, nA, pb.obox, nB, og, pairs, cursor, cap, s);
        else
            pzk::broad_append(pa.obox, nA, pb.obox, ob.nodes, ob.order, nB, batched ? A->part_of : nullptr,
                              batched ? B->part_off : nullptr, pairs, cursor, cap,
                              buf<uint32_t>(c, "bp.cnt", (size_t)std::max(nA, 1)),
                              buf<uint64_t>(c, "bp.off", (size_t)std::max(nA, 1) + 1),
                              buf<char>(c, "scan.tmp", pzk::scan_temp_bytes((uint64_t)std::max(nA, 1))), s);
        np = readback<unsigned long long>(c, {cursor})[0];
        if (np > cap) {
            r->overflowed = 1;
            two_pass();
        }
    }
    r->num_candidates = np;
    if (batched && part_cand_out) {  // candidate object pairs of each query
        const uint32_t nparts = (uint32_t)(A->part_obj_off.size() - 1);
        unsigned long long* pc = buf<unsigned long long>(c, "bp.part_cnt", nparts);
        pzk::part_pair_counts(pairs, np, A->part_of, nparts, pc, s);
        part_cand_out->resize(nparts);
        HIP_OK(hipMemcpyAsync(part_cand_out->data(), pc, sizeof(uint64_t) * nparts, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    if (np == 0) {  // no boxes meet: nothing to index or test
        ph.begin("Download Results");
        uint64_t* none = buf<uint64_t>(c, "sel.keys", 1);
        finish_pairs(c, none, nullptr, 0, nA,
                     batched || (d.predicate == PZ_PRED_CONTAINMENT && d.include_overlap_pairs), r);
        return;
    }

    // Per-pair meta = predicate | direction bits << 2 (bit 0: mesh1 edges vs mesh2
    // triangles and mesh1 points in mesh2 objects; bit 1: the reverse).
    uint8_t* meta = buf<uint8_t>(c, "np.meta", np + 1);
    const bool f32 = d.precision == PZ_PREC_FP32_FILTERED;
    const int allow_f32 = (std::max(A->max_abs, B->max_abs) <= 0x1p39f) ? 1 : 0;
    uint32_t* found = buf<uint32_t>(c, "np.found", np + 1);
    uint32_t* inside = nullptr;  // parity bits, for predicates and batches that need them
    if (d.predicate != PZ_PRED_OVERLAP || batched) inside = buf<uint32_t>(c, "np.inside", np + 1);
    unsigned long long* ocnt = buf<unsigned long long>(c, "sel.cnt", 1);
    // the pair sieve's flags and counters (narrow_mode 0)
    const bool sieve = d.narrow_mode == 0;
    uint32_t* defer = sieve ? buf<uint32_t>(c, "sv.defer", np + 1) : nullptr;
    unsigned long long* sctr = sieve ? buf<unsigned long long>(c, "sv.ctr", 2) : nullptr;  // open triangle pairs, deferred pairs
    unsigned long long* nrays = sieve ? buf<unsigned long long>(c, "sv.nrays", 1) : nullptr;
    uint32_t* rovf = sieve ? buf<uint32_t>(c, "sv.ray_overflow", np + 1) : nullptr;
    uint32_t* wctr = sieve ? buf<uint32_t>(c, "sv.wctr", 8) : nullptr;
    {
        pzk::PairPrepZero z{};
        z.flags[0] = found;
        z.flags[1] = inside;
        z.flags[2] = defer;
        z.flags[3] = (sieve && inside) ? rovf : nullptr;
        z.counters[0] = ocnt;
        z.counters[1] = sctr;
        z.counters[2] = sieve ? sctr + 1 : nullptr;
        z.counters[3] = nrays;
        z.words = wctr;
        z.nwords = sieve ? 8u : 0u;
        pzk::pair_prep(pairs, np, batched ? A->part_of : nullptr, meta_of_part, pz::pair_meta_of(d), meta, z, s);
    }

    // The tree path: per-object BVHs of the objects of the given pairs, the
    // dual-tree walk (or the per-edge kernel) and the parity rays through the
    // trees; writes found[] and inside[] of those pairs.
    auto tree_path = [&](const uint64_t* pairs, uint64_t np, uint8_t* meta, uint32_t* found, uint32_t* inside) {
    ph.begin("Build Mesh Index");
        Needed na = needed_layout(c, A, pa, "A"), nb = needed_layout(c, B, pb, "B");
        const size_t sA = std::max(nA, 1), sB = std::max(nB, 1);
        pzk::fill_u32(na.need, 0u, sA, s);
        pzk::fill_u32(nb.need, 0u, sB, s);
        pzk::mark_needed(pairs, np, na.need, nb.need, s);
        struct Tot {
            uint64_t nneed, ntri, nedge, nbig, nbig_tri;
            uint32_t ncls[pzk::kSmallClasses], cls_max[pzk::kSmallClasses];  // LDS-built objects per size class
            uint32_t nscattered;  // needed objects whose triangles are not contiguous
        };
        // Needed objects of one set: objects built by the global pipeline ("big":
        // more than kSmallMax triangles or not stored contiguously) take the first
        // segment indices, LDS-built ones follow.
        auto layout = [&](const pz_dataset* D, const SetPrep& P, Needed& N, const std::string& tag) -> uint32_t* {
            size_t n = std::max(D->nobj, 1);
            const uint64_t no = (uint64_t)std::max(D->nobj, 0);
            uint32_t* stc = buf<uint32_t>(c, tag + ".seg_tri_cnt", n);
            uint32_t* sec = buf<uint32_t>(c, tag + ".seg_edge_cnt", n);
            uint32_t* big = buf<uint32_t>(c, tag + ".big", n);
            uint32_t* cls = buf<uint32_t>(c, tag + ".cls", pzk::kSmallClasses * (n + 1));
            uint32_t* loc_big = buf<uint32_t>(c, tag + ".loc_big", n + 1);
            uint32_t* loc_cls = buf<uint32_t>(c, tag + ".loc_cls", pzk::kSmallClasses * (n + 1));
            uint32_t* totals = buf<uint32_t>(c, tag + ".totals", 9);
            pzk::fill_u32(stc, 0u, n, s);
            pzk::fill_u32(sec, 0u, n, s);
            pzk::fill_u32(totals, 0u, 9, s);
            void* tmp = buf<char>(c, "scan.tmp", pzk::scan_temp_bytes(n + 1));
            pzk::need_split(N.need, P.tri_cnt, P.first_tri, P.last_tri, D->nobj, big, cls, totals, s);
            pzk::exclusive_scan_u32_u32(big, loc_big, no, tmp, s);
            for (int k = 0; k < pzk::kSmallClasses; ++k)
                pzk::exclusive_scan_u32_u32(cls + (size_t)k * (no + 1), loc_cls + (size_t)k * (no + 1), no, tmp, s);
            pzk::merge_locs(big, loc_big, loc_cls, D->nobj, N.loc, totals, s);
            pzk::needed_layout(N.need, N.loc, D->nobj, P.tri_cnt, P.edge_cnt, N.obj_of_loc, stc, sec, s);
            pzk::exclusive_scan_u32_u32(stc, N.seg_off, no, tmp, s);
            pzk::big_prefix(N.seg_off, totals, s);
            return big;
        };
        // about 40 small launches whose shapes depend only on both sets' object
        // counts: one graph from the second join of the same datasets on
        const SegOut lay = run_segment(c, A->uid, B->uid, kSegLayout, r, [&]() {
            SegOut o;
            o.big[0] = layout(A, pa, na, "A");
            o.big[1] = layout(B, pb, nb, "B");
            return o;
        });
        uint32_t* bigA = lay.big[0];
        uint32_t* bigB = lay.big[1];
        uint32_t* totA = buf<uint32_t>(c, "A.totals", 9);
        uint32_t* totB = buf<uint32_t>(c, "B.totals", 9);
        // needed objects and their triangles, then totals[0..8] of each set, in one synchronisation
        uint32_t* hb = static_cast<uint32_t*>(pinned(c, 22 * sizeof(uint32_t)));
        pzk::HostGather g{{na.loc + std::max(nA, 0), na.seg_off + std::max(nA, 0), totA, nb.loc + std::max(nB, 0),
                           nb.seg_off + std::max(nB, 0), totB},
                          {1, 1, 9, 1, 1, 9},
                          6};
        pzk::gather_to_host(g, hb, s);
        HIP_OK(hipStreamSynchronize(s));
        auto tot_of = [&](const uint32_t* h) {
            Tot t{h[0], h[1], 0, h[2], h[3], {}, {}, h[10]};
            for (int k = 0; k < pzk::kSmallClasses; ++k) {
                t.ncls[k] = h[4 + k];
                t.cls_max[k] = h[7 + k];
            }
            return t;
        };
        const Tot ta = tot_of(hb), tb = tot_of(hb + 11);

        // Big objects are chunked in their own triangle order when every one is
        // contiguous; `incoherent` then counts those whose order is not spatially
        // coherent, and the set is built again with sorted = true.
        auto build = [&](const pz_dataset* D, const SetPrep& P, const Needed& N, const Tot& T, uint32_t* big,
                         const std::string& tag, uint32_t* incoherent, bool sorted) -> SetAccel {
            SetAccel acc;
            const uint64_t n = T.ntri, nb_tri = T.nbig_tri;
            acc.nodes = buf<pzd::BNode>(c, tag + ".nodes", n + 1);
            acc.leaves = buf<pzd::LeafTri>(c, tag + ".leaves", n + 1);
            acc.height = buf<uint32_t>(c, tag + ".height", 1);
            pzk::fill_u32(acc.height, 0u, 1, s);
            uint32_t* cursor = buf<uint32_t>(c, tag + ".cursor", (size_t)std::max(D->nobj, 1));
            if (nb_tri > 0) {  // global pipeline for the big objects (segment prefix)
                const uint64_t* sk = nullptr;
                const uint32_t* order = nullptr;
                int mbits = 0;
                const bool natural = !sorted && T.nscattered == 0;
                if (!natural) {
                    // sort buffers are shared by both sets' builds (one after the other on the stream)
                    uint64_t* keys = buf<uint64_t>(c, "big.tkeys", nb_tri + 1);
                    uint32_t* vals = buf<uint32_t>(c, "big.tvals", nb_tri + 1);
                    uint64_t* keys2 = buf<uint64_t>(c, "big.tkeys2", nb_tri + 1);
                    uint32_t* vals2 = buf<uint32_t>(c, "big.tvals2", nb_tri + 1);
                    pzk::fill_u32(cursor, 0u, (size_t)std::max(D->nobj, 1), s);
                    // Morton bits of the keys: of 30, 27, 24 and 21, the most that need the
                    // fewest 8-bit radix passes beside the segment bits (chunks only need
                    // a coherent order; each chunk's tree uses the bits below its common prefix)
                    const int sb = bits_for(std::max<uint64_t>(T.nbig, 1));
                    mbits = 30;
                    for (int m : {27, 24, 21})
                        if ((m + sb + 7) / 8 < (mbits + sb + 7) / 8) mbits = m;
                    pzk::tri_keys_two_level(D->v, D->nv, D->idx, D->t2o, D->nt, D->nobj, big, N.loc, P.obox,
                                            N.seg_off, P.first_tri, P.last_tri, P.tri_cnt, cursor, keys, vals, mbits, s);
                    void* tmp = buf<char>(c, "sort.tmp." + tag, pzk::radix_sort_temp_bytes(nb_tri));
                    bool alt = pzk::radix_sort_pairs(keys, vals, keys2, vals2, nb_tri, 0, mbits + sb, tmp, s);
                    sk = alt ? keys2 : keys;
                    order = alt ? vals2 : vals;
                }
                // runs of the sorted (or the objects' own) order through the LDS builder,
                // then a top tree per object
                const uint64_t nch = pzk::big_chunk_bound(T.nbig, nb_tri);
                uint32_t* ccnt = buf<uint32_t>(c, tag + ".chunk_cnt", T.nbig + 1);
                uint32_t* coff = buf<uint32_t>(c, tag + ".chunk_off", T.nbig + 1);
                auto* e0 = buf<pzk::ChunkRoot>(c, tag + ".chunk_ent0", nch);
                auto* e1 = buf<pzk::ChunkRoot>(c, tag + ".chunk_ent1", nch);
                void* stmp = buf<char>(c, "scan.tmp", pzk::scan_temp_bytes(T.nbig + 1));
                pzk::build_big_chunked(D->v, D->idx, N.obj_of_loc, N.seg_off, P.first_tri, P.obox, (uint32_t)T.nbig,
                                       nb_tri, sk, mbits, order, ccnt, coff, stmp, e0, e1, acc.nodes, acc.leaves,
                                       acc.height, incoherent, s);
            }
            static const char* const kTimer[pzk::kSmallClasses] = {"build_small.c0", "build_small.c1",
                                                                    "build_small.c2"};
            uint32_t l0 = (uint32_t)T.nbig;
            for (int k = 0; k < pzk::kSmallClasses; ++k) {  // one launch per size class
                if (!pzk::build_small(D->v, D->idx, N.obj_of_loc, P.first_tri, P.tri_cnt, P.obox, N.seg_off, l0,
                                      T.ncls[k], T.cls_max[k], c->tune.build_threads[k], acc.nodes, acc.leaves,
                                      acc.height, kTimer[k], s))
                    throw PzError(PZ_E_INVALID, "invalid LDS-builder configuration");
                l0 += T.ncls[k];
            }
            acc.edges = nullptr;
            return acc;
        };
        uint32_t* incoherent = buf<uint32_t>(c, "big.incoherent", 2);
        const bool natA = ta.nbig_tri > 0 && ta.nscattered == 0, natB = tb.nbig_tri > 0 && tb.nscattered == 0;
        if (natA || natB) pzk::fill_u32(incoherent, 0u, 2, s);
        SetAccel xa = build(A, pa, na, ta, bigA, "A", incoherent, false);
        SetAccel xb = build(B, pb, nb, tb, bigB, "B", incoherent + 1, false);
        r->chunk_sorted_sets = (ta.nbig_tri > 0 && !natA) + (tb.nbig_tri > 0 && !natB);
        if (natA || natB) {  // a set with an incoherent big object is built again from the sorted order
            const std::vector<uint32_t> bad = readback<uint32_t>(c, {incoherent, incoherent + 1});
            if (bad[0]) xa = build(A, pa, na, ta, bigA, "A", nullptr, true);
            if (bad[1]) xb = build(B, pb, nb, tb, bigB, "B", nullptr, true);
            r->chunk_sorted_sets += (bad[0] != 0) + (bad[1] != 0);
        }
        r->index_triangles = ta.ntri + tb.ntri;

        ph.begin("Narrow Phase");
        uint32_t* fallback = buf<uint32_t>(c, "np.fallback", np + 1);
        pzk::fill_u32(fallback, 0u, np + 1, s);
        pzd::SetView va{xa.nodes, xa.leaves, na.seg_off, nullptr, na.eseg_off, na.loc, pa.obox};
        pzd::SetView vb{xb.nodes, xb.leaves, nb.seg_off, nullptr, nb.eseg_off, nb.loc, pb.obox};
        uint64_t nfb = np;
        const uint64_t* fb_pairs = pairs;
        uint32_t* fb_index = nullptr;
        uint32_t height = 0;  // of the per-object BVHs of both sets
        if (d.narrow_mode != 1) {  // the dual-tree walk
            const int qclass = (ta.nbig + tb.nbig > 0) ? 1 : 0;  // pairs with objects beyond the LDS builder
            c->last_queue_class = qclass;
            // Work queue of (pair, subtree pair) tasks, run in rounds: round 0 takes
            // every candidate pair from its roots; a wave whose LDS queue fills up
            // hands its shallowest subtree pairs to the next round.
            // spill capacity per round (16 B a task): tasks beyond it send their pair to
            // the per-edge kernel, which is slow on big objects, so keep it generous
            const uint64_t cap_tasks = std::min<uint64_t>(std::max<uint64_t>(64 * np, 1ull << 20), 1ull << 24);
            pzd::DualTask* tin = buf<pzd::DualTask>(c, "np.tasks0", cap_tasks);
            pzd::DualTask* tout = buf<pzd::DualTask>(c, "np.tasks1", cap_tasks);
            unsigned long long* ctr = buf<unsigned long long>(c, "np.task_ctr", 3);  // next task, spilled, open leaf pairs
            // leaf pairs the fp32 walk leaves open, decided after the rounds
            const uint64_t cap_esc = std::min<uint64_t>(std::max<uint64_t>(np, 1 << 16), 1ull << 22);
            uint4* esc = buf<uint4>(c, "np.esc", cap_esc);
            pzk::fill_ull(ctr + 2, 0ull, 1, s);
            if (c->persistent_blocks == 0) c->persistent_blocks = (unsigned)c->cus * 8u;
            const pzd::DualTask* cur = nullptr;
            uint64_t ntasks = np, spilled_total = 0;
            int rounds = 0;
            uint64_t* fbp = buf<uint64_t>(c, "np.fb_pairs", np + 1);
            fb_index = buf<uint32_t>(c, "np.fb_index", np + 1);
            unsigned long long* fbc = buf<unsigned long long>(c, "np.fb_count", 1);
            // after the last round: decide the open leaf pairs, list the fallback pairs
            // The escalation list then starts empty again, so each open leaf pair is
            // decided (and counted) once, whatever the number of rounds.
            auto tail = [&]() {
                pzk::escalate(esc, ctr + 2, cap_esc, va, vb, meta, found, stats, s);
                pzk::fill_ull(ctr + 2, 0ull, 1, s);
                pzk::fill_ull(fbc, 0ull, 1, s);
                pzk::compact_flagged(pairs, np, fallback, fbp, fb_index, fbc, s);
            };
            while (ntasks > 0) {
                pzk::fill_ull(ctr, 0ull, 2, s);
                if (!pzk::narrow_dual(pairs, va, vb, found, fallback, meta, f32, allow_f32, c->tune.narrow_queue[qclass],
                                      cur, ntasks, ctr, c->tune.narrow_chunk[qclass], tout, ctr + 1, cap_tasks, esc,
                                      ctr + 2, cap_esc, c->persistent_blocks, stats, s))
                    throw PzError(PZ_E_INVALID, "invalid narrow-phase queue length");
                ++rounds;
                // The tail runs as if this were the last round (it usually is), so one
                // read-back brings the spill count, the fallback count and both heights.
                tail();
                uint32_t* hb = static_cast<uint32_t*>(pinned(c, 6 * sizeof(uint32_t)));
                pzk::HostGather g{{reinterpret_cast<const uint32_t*>(ctr + 1), reinterpret_cast<const uint32_t*>(fbc),
                                   xa.height, xb.height},
                                  {2, 2, 1, 1},
                                  4};
                pzk::gather_to_host(g, hb, s);
                HIP_OK(hipStreamSynchronize(s));
                const uint64_t spilled = std::min<uint64_t>(hb[0] | ((uint64_t)hb[1] << 32), cap_tasks);
                nfb = hb[2] | ((uint64_t)hb[3] << 32);
                height = std::max(hb[4], hb[5]);
                spilled_total += spilled;
                if (spilled == 0) break;
                if (rounds >= 32) {  // leave what is left to the per-edge kernel
                    pzk::mark_task_pairs(tout, spilled, fallback, s);
                    tail();
                    nfb = readback<unsigned long long>(c, {fbc})[0];
                    break;
                }
                std::swap(tin, tout);
                cur = tin;
                ntasks = spilled;
            }
            r->narrow_rounds = rounds;
            r->tasks_spilled = spilled_total;
            fb_pairs = fbp;
        } else {
            auto hh = readback<uint32_t>(c, {xa.height, xb.height});
            height = std::max(hh[0], hh[1]);
        }
        r->index_height = height;
        if (pzk::stack_for_height(height) == 0)
            throw PzError(PZ_E_LIMIT, "per-object BVH height " + std::to_string(height) + " exceeds the traversal stack (63)");
        r->narrow_fallback_pairs = d.narrow_mode != 1 ? nfb : 0;  // pairs re-run by the per-edge kernel
        if (nfb > 0) {
            // Per-edge narrow phase: every edge of the .pre payload regrouped by object.
            auto edges_of = [&](const pz_dataset* D, const SetPrep& P, Needed& N, const std::string& tag) {
                const size_t n = std::max(D->nobj, 1);
                const uint64_t no = (uint64_t)std::max(D->nobj, 0);
                uint32_t* stc = buf<uint32_t>(c, tag + ".seg_tri_cnt", n);
                uint32_t* sec = buf<uint32_t>(c, tag + ".seg_edge_cnt", n);
                pzk::edge_count(D->eo, D->ne, D->nobj, P.edge_cnt, s);
                pzk::fill_u32(sec, 0u, n, s);
                pzk::needed_layout(N.need, N.loc, D->nobj, P.tri_cnt, P.edge_cnt, N.obj_of_loc, stc, sec, s);
                void* tmp = buf<char>(c, "scan.tmp", pzk::scan_temp_bytes(n + 1));
                pzk::exclusive_scan_u32_u32(sec, N.eseg_off, no, tmp, s);
                uint64_t ne = readback<uint32_t>(c, {N.eseg_off + no})[0];
                pzd::EdgeRec* edges = buf<pzd::EdgeRec>(c, tag + ".edges", ne + 1);
                uint32_t* cursor = buf<uint32_t>(c, tag + ".cursor", n);
                pzk::fill_u32(cursor, 0u, n, s);
                pzk::edge_scatter(D->es, D->ee, D->eo, D->ne, D->nobj, N.need, N.loc, N.eseg_off, cursor, edges, s);
                r->index_edges += ne;
                return edges;
            };
            va.edges = edges_of(A, pa, na, "A");
            vb.edges = edges_of(B, pb, nb, "B");
            uint32_t* tcnt = buf<uint32_t>(c, "np.tcnt", nfb + 1);
            uint64_t* toff = buf<uint64_t>(c, "np.toff", nfb + 2);
            void* tmp = buf<char>(c, "scan.tmp.np", pzk::scan_temp_bytes(nfb + 1));
            uint8_t* fb_meta = meta;
            if (fb_index) {  // meta of the compacted fallback pairs
                fb_meta = buf<uint8_t>(c, "np.fb_meta", nfb + 1);
                pzk::pair_meta(fb_pairs, nfb, batched ? A->part_of : nullptr, meta_of_part, pz::pair_meta_of(d), fb_meta,
                               s);
            }
            pzk::task_counts(fb_pairs, nfb, na.loc, na.eseg_off, nb.loc, nb.eseg_off, fb_meta, tcnt, s);
            pzk::exclusive_scan_u32_u64(tcnt, toff, nfb, tmp, s);
            const uint64_t ntasks = readback<uint64_t>(c, {toff + nfb})[0];
            uint32_t* sub_found = fb_index ? buf<uint32_t>(c, "np.sub_found", nfb + 1) : found;
            if (fb_index) pzk::fill_u32(sub_found, 0u, nfb + 1, s);
            pzk::narrow(fb_pairs, nfb, toff, ntasks, va, vb, sub_found, fb_meta, f32, allow_f32, height, stats, s);
            if (fb_index) pzk::scatter_found(sub_found, fb_index, nfb, found, s);
        }

        if (inside) {
            ph.begin("Containment Parity");
            pzk::parity(pairs, np, found, va, vb, A->v, A->idx, pa.first_tri, B->v, B->idx, pb.first_tri, meta, inside,
                        height, stats, s);
        }
    };

    // Result pairs: the candidates that pass their predicate, as keys (and pair
    // types), then sorted and copied to the host.
    uint64_t* okeys = buf<uint64_t>(c, "sel.keys", np + 1);
    uint32_t* otype = buf<uint32_t>(c, "sel.type", np + 1);
    const bool want_types = batched || (d.predicate == PZ_PRED_CONTAINMENT && d.include_overlap_pairs);
    bool selected = false;  // the first selection finds the count zeroed by the pair prep
    auto select = [&]() {
        if (selected) pzk::fill_ull(ocnt, 0ull, 1, s);
        selected = true;
        // batches keep overlap pairs of containment queries (typed) and filter per query on the host
        pzk::select_pairs(pairs, np, found, inside, meta, batched ? 1 : d.include_overlap_pairs, okeys, otype, ocnt,
                          stats + pzk::kStatTouching, s);
    };
    constexpr uint64_t kSpec = pzk::kSortTilesMax;           // result pairs of the speculative download
    constexpr size_t kHead = (2 + pzk::kStatSlots) * 8;      // counts and counters at the front of the landing area
    // One read-back into h: [0] *deferred (0 without), [1] result pairs, [2..] the counters.
    auto gather_counts = [&](const unsigned long long* deferred, void* h) {
        pzk::HostGather g{{reinterpret_cast<const uint32_t*>(deferred ? deferred : ocnt),
                           reinterpret_cast<const uint32_t*>(ocnt), reinterpret_cast<const uint32_t*>(stats)},
                          {2, 2, 2 * pzk::kStatSlots},
                          3};
        pzk::gather_to_host(g, static_cast<uint32_t*>(h), s);
        HIP_OK(hipStreamSynchronize(s));
        unsigned long long* hs = static_cast<unsigned long long*>(h);
        if (!deferred) hs[0] = 0;
        return static_cast<const unsigned long long*>(hs);
    };

    // Pairs the tree path decides: all of them, or the ones the pair sieve defers.
    const uint64_t* tpairs = pairs;
    uint64_t ntp = np;
    uint8_t* tmeta = meta;
    uint32_t* tindex = nullptr;  // sieve: position of each deferred pair in pairs[]
    if (sieve) {
        ph.begin("Pair Sieve");
        const uint64_t cap_esc = std::min<uint64_t>(std::max<uint64_t>(4 * np, 1 << 16), 1ull << 22);
        uint4* esc = buf<uint4>(c, "sv.esc", cap_esc);
        auto sieve_set = [](const pz_dataset* D, const SetPrep& P) {
            const uint64_t n = (uint64_t)std::max(D->nobj, 1);
            return pzd::SieveSet{D->v, D->idx, P.first_tri, P.tri_cnt, P.last_tri, P.obox,
                                 D->vruns ? D->topo + 4 * n : nullptr, D->vruns ? D->topo + 3 * n : nullptr};
        };
        const pzd::SieveSet sa = sieve_set(A, pa), sb = sieve_set(B, pb);
        uint4* rdesc = buf<uint4>(c, "sv.rdesc", 5 * (np + 1));
        uint32_t* gwork = buf<uint32_t>(c, "sv.gwork", np + 1);
        // parity-ray candidates: about a dozen target triangles per ray
        const uint64_t cap_rays = inside ? std::min<uint64_t>(std::max<uint64_t>(64 * np, 1 << 16), 1ull << 26) : 1;
        uint2* rays = buf<uint2>(c, "sv.rays", cap_rays);
        pzk::pair_sieve(pairs, np, sa, sb, meta, f32, allow_f32, found, defer, rdesc, gwork, wctr,
                        inside ? rays : nullptr, nrays, inside ? cap_rays : 0, rovf, esc, sctr, cap_esc,
                        (unsigned)c->cus, stats, s);
        if (f32 && allow_f32) pzk::escalate_tris(esc, sctr, cap_esc, sa, sb, meta, found, stats, s);
        if (inside) {
            ph.begin("Containment Parity");
            uint32_t* cross = buf<uint32_t>(c, "sv.cross", 2 * np + 2);
            pzk::fill_u32(cross, 0u, 2 * np + 2, s);
            pzk::sieve_parity(pairs, np, sa, sb, meta, found, defer, rays, nrays, cap_rays, rovf, cross, inside,
                              stats, s);
        }
        uint64_t* dp = buf<uint64_t>(c, "sv.pairs", np + 1);
        tindex = buf<uint32_t>(c, "sv.index", np + 1);
        pzk::compact_flagged(pairs, np, defer, dp, tindex, sctr + 1, s);
        tpairs = dp;
        // Download as if the sieve deferred nothing (the usual case), reading
        // back the deferred count with the results: up to one sort tile of
        // result pairs needs no further round trip.
        ph.begin("Download Results");
        select();
        // sorted tiles merged by rank, with the count read on the device; the
        // pairs (and types) land in pinned memory straight from the kernel
        const uint64_t spec = std::min<uint64_t>(kSpec, np);
        uint64_t* okeys2 = buf<uint64_t>(c, "fin.keys2", spec + 1);
        uint32_t* otype2 = buf<uint32_t>(c, "fin.types2", spec + 1);
        pzk::sort_tiles_counted(okeys, want_types ? otype : nullptr, okeys2, want_types ? otype2 : nullptr, ocnt, spec,
                                s);
        char* hp = static_cast<char*>(pinned(c, kHead + spec * (sizeof(pz_pair) + 4)));
        pzk::keys_to_pairs_counted(okeys2, want_types ? otype2 : nullptr, ocnt, spec,
                                   reinterpret_cast<int32_t*>(hp + kHead),
                                   reinterpret_cast<uint32_t*>(hp + kHead + spec * sizeof(pz_pair)), s);
        const unsigned long long* hs = gather_counts(sctr + 1, hp);
        ntp = hs[0];
        r->sieve_pairs = np - ntp;
        if (ntp == 0) {
            const uint64_t nout = hs[1];
            const std::vector<unsigned long long> st(hs + 2, hs + 2 + pzk::kStatSlots);
            if (nout <= spec)
                land_pairs(r, hp + kHead, reinterpret_cast<const uint32_t*>(hp + kHead + spec * sizeof(pz_pair)),
                           nout, want_types);
            else  // more result pairs than the tiles hold: sort them with the count now known
                finish_pairs(c, okeys, want_types ? otype : nullptr, nout, nA, want_types, r);
            set_counters(r, st);
            return;
        }
        // the sieve deferred pairs to the trees: they are decided, then the results selected again
        pzk::fill_ull(stats + pzk::kStatTouching, 0ull, 1, s);
        tmeta = buf<uint8_t>(c, "sv.meta", ntp + 1);
        pzk::pair_meta(dp, ntp, batched ? A->part_of : nullptr, meta_of_part, pz::pair_meta_of(d), tmeta, s);
    }
    if (ntp > 0) {
        uint32_t* tfound = found;
        uint32_t* tinside = inside;
        if (tindex) {
            tfound = buf<uint32_t>(c, "sv.tfound", ntp + 1);
            pzk::fill_u32(tfound, 0u, ntp + 1, s);
            if (inside) {
                tinside = buf<uint32_t>(c, "sv.tinside", ntp + 1);
                pzk::fill_u32(tinside, 0u, ntp + 1, s);
            }
        }
        tree_path(tpairs, ntp, tmeta, tfound, tinside);
        if (tindex) {
            pzk::scatter_found(tfound, tindex, ntp, found, s);
            if (inside) pzk::scatter_found(tinside, tindex, ntp, inside, s);
        }
    }

    ph.begin("Download Results");
    select();
    const unsigned long long* hs = gather_counts(nullptr, pinned(c, kHead));
    const std::vector<unsigned long long> st(hs + 2, hs + 2 + pzk::kStatSlots);
    finish_pairs(c, okeys, want_types ? otype : nullptr, hs[1], nA, want_types, r);
    set_counters(r, st);
}

// ------------------------------------------------------------- global
struct GlobalBvh {
    pzd::Node* nodes;
    pzd::LeafTri* leaves;
    uint32_t height;
};

GlobalBvh build_global(pz_ctx* c, const pz_dataset* D, const SetPrep& P, const std::string& tag) {
    hipStream_t s = c->s;
    const uint64_t n = D->nt;
    GlobalBvh g;
    uint32_t* scene = buf<uint32_t>(c, tag + ".gscene", 8);
    pzk::scene_box(P.obox, D->nobj, scene, s);
    // build temporaries are shared by both sets' builds; only nodes and leaves stay per set
    uint64_t* keys = buf<uint64_t>(c, "gbuild.keys", n + 1);
    uint32_t* vals = buf<uint32_t>(c, "gbuild.vals", n + 1);
    uint64_t* keys2 = buf<uint64_t>(c, "gbuild.keys2", n + 1);
    uint32_t* vals2 = buf<uint32_t>(c, "gbuild.vals2", n + 1);
    pzk::tri_keys_global(D->v, D->idx, n, scene, keys, vals, s);
    void* tmp = buf<char>(c, "sort.tmp." + tag, pzk::radix_sort_temp_bytes(n));
    bool alt = pzk::radix_sort_pairs(keys, vals, keys2, vals2, n, 0, 30, tmp, s);
    uint64_t* sk = alt ? keys2 : keys;
    uint32_t* order = alt ? vals2 : vals;
    g.nodes = buf<pzd::Node>(c, tag + ".gnodes", n + 1);
    uint32_t* pint = buf<uint32_t>(c, "gbuild.pint", n + 1);
    uint32_t* pleaf = buf<uint32_t>(c, "gbuild.pleaf", n + 1);
    uint32_t* cnt = buf<uint32_t>(c, "gbuild.cnt", n + 1);
    uint32_t* h = buf<uint32_t>(c, tag + ".gheight", 1);
    pzk::fill_u32(h, 0u, 1, s);
    uint32_t* rng = buf<uint32_t>(c, "gbuild.range", n + 1);
    pzk::lbvh_karras(sk, n, nullptr, 0, g.nodes, pint, pleaf, cnt, rng, s);
    // leaf records first: the refit then reads its boxes from them in sorted
    // order instead of gathering every triangle a second time
    g.leaves = buf<pzd::LeafTri>(c, tag + ".gleaves", n + 1);
    pzk::gather_leaf_tris(order, n, D->v, D->idx, D->t2o, g.leaves, s);
    pzk::lbvh_refit_leaves(sk, n, nullptr, 0, g.leaves, g.nodes, pint, pleaf, cnt, rng, h, s);
    g.height = readback<uint32_t>(c, {h})[0];
    if (pzk::stack_for_height(g.height) == 0)
        throw PzError(PZ_E_LIMIT, "set-wide BVH height " + std::to_string(g.height) + " exceeds the traversal stack (63)");
    return g;
}

// Points in objects the reference's way (mesh_intersection_anyhit.cu:46-240,
// mesh_containment.cu:224-293): one +Z ray per source object from the first
// vertex of its first triangle through the set-wide BVH of the other set,
// counting crossed triangles per target object.  Returns the device keys of
// the (source, target) list entries, ~0 where the count is even, and their
// number.  A ray that meets more than 256 objects is re-cast with room for
// 4,096; beyond that the join fails (PZ_E_LIMIT) instead of dropping targets.
std::pair<uint64_t*, uint64_t> containment_rays(pz_ctx* c, const pz_dataset* S, const SetPrep& PS, const GlobalBvh& gT,
                                                uint64_t ntT, bool src_is_mesh2, const std::string& tag,
                                                unsigned long long* stats) {
    hipStream_t s = c->s;
    const int32_t ns = std::max<int32_t>(S->nobj, 0);
    if (ns == 0 || ntT == 0) return {nullptr, 0};
    uint32_t* cnt = buf<uint32_t>(c, tag + ".ray_cnt", 3 * (size_t)ns);
    uint64_t* off = buf<uint64_t>(c, tag + ".ray_off", (size_t)ns + 1);
    uint64_t* ovo = buf<uint64_t>(c, tag + ".ray_ovf_off", (size_t)ns + 1);
    void* tmp = buf<char>(c, "scan.tmp.ray", pzk::scan_temp_bytes((uint64_t)ns));
    for (int cap : {256, 4096}) {
        int32_t* ids = buf<int32_t>(c, "ray.ids", (size_t)ns * cap);
        uint32_t* hits = buf<uint32_t>(c, "ray.hits", (size_t)ns * cap);
        if (!pzk::track_rays(S->v, S->idx, PS.first_tri, PS.tri_cnt, ns, gT.nodes, gT.leaves, ntT, gT.height, cap, ids,
                             hits, cnt, cnt + ns, cnt + 2 * ns, stats, s))
            throw PzError(PZ_E_LIMIT, "set-wide BVH too tall for the containment ray's stack");
        pzk::exclusive_scan_u32_u64(cnt, off, (uint64_t)ns, tmp, s);
        pzk::exclusive_scan_u32_u64(cnt + ns, ovo, (uint64_t)ns, tmp, s);
        const auto tot = readback<uint64_t>(c, {off + ns, ovo + ns});
        if (tot[1] != 0) {
            if (cap == 4096)
                throw PzError(PZ_E_LIMIT, "a containment ray met more than 4096 objects");
            continue;
        }
        uint64_t* keys = buf<uint64_t>(c, tag + ".ray_keys", tot[0] + 1);
        pzk::parity_keys(ids, hits, cnt, off, ns, cap, src_is_mesh2, keys, s);
        return {keys, tot[0]};
    }
    return {nullptr, 0};
}

void join_global(pz_ctx* c, const pz_dataset* A, const pz_dataset* B, const pz_join_desc& d, pz_join_result* r,
                 Phases& ph) {
    hipStream_t s = c->s;
    unsigned long long* stats = buf<unsigned long long>(c, "stats", pzk::kStatSlots);
    pzk::fill_ull(stats, 0ull, pzk::kStatSlots, s);
    ph.begin("Build Mesh1 Index");
    SetPrep pa = prep_set(c, A, "A"), pb = prep_set(c, B, "B");
    // LEARNED: features and inference overlap the set-wide index builds on the side stream
    uint32_t npred = 0;
    if (d.sizing == PZ_SIZING_LEARNED) npred = mlp_predict_async(c, d.model, A, B, pa, pb, nullptr, pair_meta_of(d));
    // direction bits (containment always runs both edge directions)
    const int dir_mask = pair_meta_of(d) >> 2;
    const bool do12 = (dir_mask & 1) != 0, do21 = (dir_mask & 2) != 0;
    GlobalBvh ga{}, gb{};
    if (do21) ga = build_global(c, A, pa, "A");
    ph.begin("Build Mesh2 Index");
    if (do12) gb = build_global(c, B, pb, "B");
    r->index_triangles = (do21 ? A->nt : 0) + (do12 ? B->nt : 0);
    r->index_height = std::max(ga.height, gb.height);
    const bool f32 = d.precision == PZ_PREC_FP32_FILTERED;
    const int allow_f32 = (std::max(A->max_abs, B->max_abs) <= 0x1p39f) ? 1 : 0;
    uint32_t* overflow = buf<uint32_t>(c, "g.overflow", 1);

    auto run_two_pass = [&](uint64_t* out_n) -> uint64_t* {
        uint64_t n1 = do12 ? A->ne : 0, n2 = do21 ? B->ne : 0;
        uint32_t* cnt = buf<uint32_t>(c, "g.cnt", n1 + n2 + 1);
        uint64_t* off = buf<uint64_t>(c, "g.off", n1 + n2 + 2);
        void* tmp = buf<char>(c, "scan.tmp.g", pzk::scan_temp_bytes(n1 + n2 + 1));
        ph.begin("Raytrace_Mesh1ToMesh2_Pass1");
        if (do12)
            pzk::trace_global(0, A->es, A->ee, A->eo, n1, gb.nodes, gb.leaves, B->nt, gb.height, 0, f32, allow_f32,
                              cnt, nullptr, nullptr, nullptr, 0, overflow, stats, s);
        ph.begin("Raytrace_Mesh2ToMesh1_Pass1");
        if (do21)
            pzk::trace_global(0, B->es, B->ee, B->eo, n2, ga.nodes, ga.leaves, A->nt, ga.height, 1, f32, allow_f32,
                              cnt + n1, nullptr, nullptr, nullptr, 0, overflow, stats, s);
        pzk::exclusive_scan_u32_u64(cnt, off, n1 + n2, tmp, s);
        uint64_t total = readback<uint64_t>(c, {off + n1 + n2})[0];
        uint64_t* keys = buf<uint64_t>(c, "g.keys", total + 1);
        ph.begin("Raytrace_Mesh1ToMesh2_Pass2");
        if (do12)
            pzk::trace_global(1, A->es, A->ee, A->eo, n1, gb.nodes, gb.leaves, B->nt, gb.height, 0, f32, allow_f32,
                              nullptr, off, keys, nullptr, 0, overflow, stats, s);
        ph.begin("Raytrace_Mesh2ToMesh1_Pass2");
        if (do21)
            pzk::trace_global(1, B->es, B->ee, B->eo, n2, ga.nodes, ga.leaves, A->nt, ga.height, 1, f32, allow_f32,
                              nullptr, off + n1, keys, nullptr, 0, overflow, stats, s);
        *out_n = total;
        return keys;
    };

    uint64_t raw = 0;
    uint64_t* keys = nullptr;
    uint64_t nuniq = 0;
    uint64_t* uniq = nullptr;
    bool done = false;
    double learned_upper = 0;  // LEARNED: calibrated upper bound of the unique-pair count
    if (d.sizing == PZ_SIZING_LEARNED) {
        std::vector<SizePrediction> pr = mlp_predict_wait(c, d.model, npred);
        r->predicted = (uint64_t)std::llround(pr[0].pairs);
        r->predicted_pairs = r->predicted;
        learned_upper = pr[0].pairs_upper;
    }
    if (d.dedup == PZ_DEDUP_HASH || d.sizing != PZ_SIZING_TWO_PASS) {
        uint64_t expected = d.sizing == PZ_SIZING_FIXED ? d.capacity : std::max<uint64_t>(r->predicted, 1);
        if (d.sizing == PZ_SIZING_LEARNED) expected = (uint64_t)std::ceil(std::min(std::max(learned_upper, 1.0), 0x1p40));
        float lf = d.load_factor > 0 ? d.load_factor : 0.5f;
        uint64_t slots = 1024;
        if (d.sizing == PZ_SIZING_FIXED)
            slots = std::max<uint64_t>(d.capacity, 1);
        else
            slots = (uint64_t)std::ceil((double)expected / lf);
        // power-of-two table; estimate-sized tables keep a 1024-slot floor
        uint64_t cap2 = d.sizing == PZ_SIZING_FIXED ? 16 : 1024;
        while (cap2 < slots && cap2 < (1ull << 31)) cap2 <<= 1;  // at most 2^31 slots (16 GiB)
        r->capacity = cap2;
        unsigned long long* table = buf<unsigned long long>(c, "g.table", cap2);
        pzk::fill_ull(table, ~0ull, cap2, s);
        pzk::fill_u32(overflow, 0u, 1, s);
        ph.begin("Raytrace_Hash_Mesh1ToMesh2");
        if (do12)
            pzk::trace_global(2, A->es, A->ee, A->eo, A->ne, gb.nodes, gb.leaves, B->nt, gb.height, 0, f32,
                              allow_f32, nullptr, nullptr, nullptr, table, cap2 - 1, overflow, stats, s);
        ph.begin("Raytrace_Hash_Mesh2ToMesh1");
        if (do21)
            pzk::trace_global(2, B->es, B->ee, B->eo, B->ne, ga.nodes, ga.leaves, A->nt, ga.height, 1, f32,
                              allow_f32, nullptr, nullptr, nullptr, table, cap2 - 1, overflow, stats, s);
        ph.begin("compact_hash_table_pairs");
        unsigned long long* cnt = buf<unsigned long long>(c, "g.ccnt", 1);
        pzk::fill_ull(cnt, 0ull, 1, s);
        uniq = buf<uint64_t>(c, "g.uniq", cap2 + 1);
        pzk::compact_table(table, cap2, uniq, cap2, cnt, s);
        HIP_OK(hipMemcpyAsync(pinned(c, 16), overflow, 4, hipMemcpyDeviceToHost, s));
        uint32_t ovf = 0;
        unsigned long long cc = 0;
        HIP_OK(hipStreamSynchronize(s));
        std::memcpy(&ovf, c->pinned, 4);
        cc = readback<unsigned long long>(c, {cnt})[0];
        if (ovf) {
            r->overflowed = 1;  // fall back to the exact two-pass count
        } else {
            nuniq = cc;
            done = true;
            r->num_candidates = cc;
        }
    }
    if (!done) {
        keys = run_two_pass(&raw);
        r->num_candidates = raw;
        ph.begin("gpu deduplication");
        uint64_t* keys2 = buf<uint64_t>(c, "g.keys2", raw + 1);
        void* tmp = buf<char>(c, "sort.tmp.g", pzk::radix_sort_temp_bytes(raw));
        bool alt = pzk::radix_sort_keys(keys, keys2, raw, 0, 32 + bits_for((uint64_t)std::max(A->nobj, 1)), tmp, s);
        uint64_t* sk = alt ? keys2 : keys;
        uniq = buf<uint64_t>(c, "g.uniq", raw + 1);
        unsigned long long* ucnt = buf<unsigned long long>(c, "g.ucnt", 1);
        void* utmp = buf<char>(c, "uniq.tmp", pzk::unique_temp_bytes(raw));
        pzk::unique_u64(sk, raw, reinterpret_cast<uint64_t*>(uniq), reinterpret_cast<uint64_t*>(ucnt), utmp, s);
        nuniq = readback<unsigned long long>(c, {ucnt})[0];
        if (d.sizing == PZ_SIZING_TWO_PASS) r->capacity = raw;
    }
    r->pairs_touching = nuniq;
    uint32_t* types = nullptr;
    bool want_types = false;
    if (d.predicate != PZ_PRED_OVERLAP) {
        // the overlap keys sorted (the hash set's compaction leaves them in slot order)
        if (done) {
            uint64_t* u2 = buf<uint64_t>(c, "g.uniq2", nuniq + 1);
            void* tmp = buf<char>(c, "sort.tmp.g", pzk::radix_sort_temp_bytes(nuniq));
            if (pzk::radix_sort_keys(uniq, u2, nuniq, 0, 32 + bits_for((uint64_t)std::max(A->nobj, 1)), tmp, s))
                uniq = u2;
        }
        ph.begin("Containment Parity Rays");
        std::pair<uint64_t*, uint64_t> p12{nullptr, 0}, p21{nullptr, 0};
        if (d.predicate == PZ_PRED_CONTAINMENT) {
            p21 = containment_rays(c, B, pb, ga, A->nt, true, "B", stats);  // points of B in objects of A
        } else {
            if (do12) p12 = containment_rays(c, A, pa, gb, B->nt, false, "A", stats);
            if (do21) p21 = containment_rays(c, B, pb, ga, A->nt, true, "B", stats);
        }
        const uint64_t np = nuniq + p12.second + p21.second;
        uint64_t* all = buf<uint64_t>(c, "g.pred_keys", np + 1);
        uint32_t* ty = buf<uint32_t>(c, "g.pred_types", np + 1);
        if (d.predicate == PZ_PRED_INTERSECTION) {  // overlap pairs and every odd ray entry, sorted unique
            if (nuniq) HIP_OK(hipMemcpyAsync(all, uniq, 8 * nuniq, hipMemcpyDeviceToDevice, s));
            if (p12.second) HIP_OK(hipMemcpyAsync(all + nuniq, p12.first, 8 * p12.second, hipMemcpyDeviceToDevice, s));
            if (p21.second)
                HIP_OK(hipMemcpyAsync(all + nuniq + p12.second, p21.first, 8 * p21.second, hipMemcpyDeviceToDevice, s));
            uint64_t* all2 = buf<uint64_t>(c, "g.pred_keys2", np + 1);
            void* tmp = buf<char>(c, "sort.tmp.g", pzk::radix_sort_temp_bytes(np));
            uint64_t* sk = pzk::radix_sort_keys(all, all2, np, 0, 64, tmp, s) ? all2 : all;
            uint64_t* u = buf<uint64_t>(c, "g.pred_uniq", np + 1);
            unsigned long long* ucnt = buf<unsigned long long>(c, "g.ucnt", 1);
            void* utmp = buf<char>(c, "uniq.tmp", pzk::unique_temp_bytes(np));
            pzk::unique_u64(sk, np, u, reinterpret_cast<uint64_t*>(ucnt), utmp, s);
            uint64_t nu = readback<unsigned long long>(c, {ucnt})[0];
            if (nu && readback<uint64_t>(c, {u + nu - 1})[0] == ~0ull) --nu;  // the even entries' key
            uniq = u;
            nuniq = nu;
        } else {  // containment: odd entries without surface contact; overlap pairs typed 1 on request
            unsigned long long* oc = buf<unsigned long long>(c, "g.pred_cnt", 1);
            pzk::fill_ull(oc, 0ull, 1, s);
            pzk::key_difference(p21.first, p21.second, uniq, nuniq, all, ty, 0u, oc, s);
            if (d.include_overlap_pairs) pzk::key_difference(uniq, nuniq, nullptr, 0, all, ty, 1u, oc, s);
            nuniq = readback<unsigned long long>(c, {oc})[0];
            uniq = all;
            types = ty;
            want_types = d.include_overlap_pairs != 0;
        }
    }
    ph.begin("Download Results");
    finish_pairs(c, uniq, want_types ? types : nullptr, nuniq, A->nobj, want_types, r);
    std::vector<unsigned long long> st(pzk::kStatSlots);
    HIP_OK(hipMemcpyAsync(st.data(), stats, 8 * pzk::kStatSlots, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const uint64_t touching = r->pairs_touching;
    set_counters(r, st);
    r->pairs_touching = touching;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

int32_t pz_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

pz_status pz_ctx_create(int32_t device, pz_ctx** out) {
    if (!out) return fail(PZ_E_INVALID, "pz_ctx_create: null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(PZ_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(PZ_E_INVALID, "device index out of range");
    return guarded([&]() -> pz_status {
        std::unique_ptr<pz_ctx> c(new pz_ctx());
        c->device = device;
        HIP_OK(hipSetDevice(device));
        HIP_OK(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device));
        c->cus = std::max(c->cus, 1);
        HIP_OK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        for (auto& e : c->ev) HIP_OK(hipEventCreate(&e));
        *out = c.release();
        return PZ_OK;
    });
}

pz_status pz_ctx_mem_info(pz_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!c || !free_bytes || !total_bytes) return fail(PZ_E_INVALID, "pz_ctx_mem_info: null argument");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        size_t f = 0, t = 0;
        HIP_OK(hipMemGetInfo(&f, &t));
        *free_bytes = f;
        *total_bytes = t;
        return PZ_OK;
    });
}

pz_status pz_ctx_trim(pz_ctx* c) {
    if (!c) return fail(PZ_E_INVALID, "pz_ctx_trim: null context");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        HIP_OK(hipStreamSynchronize(c->s));
        HIP_OK(hipStreamSynchronize(c->s2));
        for (GraphSeg* g : c->segs) {
            drop_graph(g);
            delete g;
        }
        c->segs.clear();
        for (auto& kv : c->bufs)
            if (kv.second.p) HIP_OK(hipFree(kv.second.p));
        c->bufs.clear();
        ++c->pool_epoch;
        for (uint64_t& g : c->slot_gen) ++g;  // datasets instanced into slots lost their arrays
        return PZ_OK;
    });
}

void pz_ctx_destroy(pz_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->s);
    (void)hipStreamSynchronize(c->s2);
    for (GraphSeg* g : c->segs) {
        drop_graph(g);
        delete g;
    }
    for (auto& kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    c->bufs.clear();
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->pinned_est) (void)hipHostFree(c->pinned_est);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(c->ev_fork);
    (void)hipEventDestroy(c->ev_join);
    (void)hipStreamDestroy(c->s2);
    (void)hipStreamDestroy(c->s);
    delete c;
}

pz_status pz_dataset_upload_parts(pz_ctx* c, const pz_mesh_view* v, const std::vector<uint64_t>* part_obj_off,
                                  pz_dataset** out);

pz_status pz_dataset_upload(pz_ctx* c, const pz_mesh_view* v, pz_dataset** out) {
    return pz_dataset_upload_parts(c, v, nullptr, out);
}

pz_status pz_dataset_upload_parts(pz_ctx* c, const pz_mesh_view* v, const std::vector<uint64_t>* part_obj_off,
                                  pz_dataset** out) {
    if (!c || !v || !out) return fail(PZ_E_INVALID, "pz_dataset_upload: null argument");
    if (v->num_edges == 0 || !v->edge_starts || !v->edge_ends || !v->edge_objects)
        return fail(PZ_E_NO_EDGES,
                    "dataset has no precomputed edges; re-run preprocessing with edge extraction enabled");
    if (v->num_triangles >= (1ull << 31) || v->num_edges >= (1ull << 31))
        return fail(PZ_E_LIMIT, "a set is limited to 2^31 - 1 triangles and edges");
    if (v->num_triangles && (!v->indices || !v->triangle_to_object || !v->vertices))
        return fail(PZ_E_INVALID, "null geometry array");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        std::unique_ptr<pz_dataset> d(new pz_dataset());
        d->ctx = c;
        d->uid = pz::next_dataset_uid();
        d->nv = v->num_vertices;
        d->nt = v->num_triangles;
        d->ne = v->num_edges;
        // largest object id, index bounds and largest |coordinate|, over 256 slices in parallel
        constexpr uint64_t kSlices = 256;
        std::vector<int32_t> smx(kSlices, -1);
        std::vector<float> sma(kSlices, 0.f);
        std::vector<uint8_t> sbad(kSlices, 0);
        pz::parallel_for(kSlices, [&](uint64_t k) {
            auto range = [&](uint64_t n) { return std::make_pair(n * k / kSlices, n * (k + 1) / kSlices); };
            int32_t mx = -1;
            float ma = 0.f;
            bool bad = false;
            auto [t0, t1] = range(d->nt);
            for (uint64_t t = t0; t < t1; ++t) mx = std::max(mx, v->triangle_to_object[t]);
            for (uint64_t i = 3 * t0; i < 3 * t1; ++i) bad |= v->indices[i] >= d->nv;
            auto [e0, e1] = range(d->ne);
            for (uint64_t e = e0; e < e1; ++e) mx = std::max(mx, v->edge_objects[e]);
            for (uint64_t i = 3 * e0; i < 3 * e1; ++i)
                ma = std::max(ma, std::max(std::fabs(v->edge_starts[i]), std::fabs(v->edge_ends[i])));
            auto [v0, v1] = range(d->nv);
            for (uint64_t i = 3 * v0; i < 3 * v1; ++i) ma = std::max(ma, std::fabs(v->vertices[i]));
            smx[k] = mx;
            sma[k] = ma;
            sbad[k] = bad;
        });
        int32_t mx = -1;
        float ma = 0.f;
        for (uint64_t k = 0; k < kSlices; ++k) {
            if (sbad[k]) throw PzError(PZ_E_INVALID, "triangle index beyond the vertex array");
            mx = std::max(mx, smx[k]);
            ma = std::max(ma, sma[k]);
        }
        d->nobj = mx + 1;
        d->max_abs = std::isfinite(ma) ? ma : INFINITY;
        auto up = [&](void** dst, const void* src, size_t bytes) {
            HIP_OK(hipMalloc(dst, std::max<size_t>(bytes, 16)));
            if (bytes) HIP_OK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        };
        try {
            up((void**)&d->v, v->vertices, 12 * d->nv);
            up((void**)&d->idx, v->indices, 12 * d->nt);
            up((void**)&d->t2o, v->triangle_to_object, 4 * d->nt);
            up((void**)&d->es, v->edge_starts, 12 * d->ne);
            up((void**)&d->ee, v->edge_ends, 12 * d->ne);
            up((void**)&d->eo, v->edge_objects, 4 * d->ne);
        } catch (...) {
            pz_dataset_release(d.release());
            throw;
        }
        if (v->grid_cells && v->num_grid_cells) d->grid.assign(v->grid_cells, v->grid_cells + v->num_grid_cells);
        d->cell_size = v->grid_cell_size;
        if (part_obj_off) {
            d->part_obj_off = *part_obj_off;
            if (d->part_obj_off.back() != (uint64_t)std::max(d->nobj, 0))
                d->nobj = (int32_t)d->part_obj_off.back();  // trailing empty parts own no objects
            std::vector<uint32_t> part_of((size_t)std::max(d->nobj, 1), 0u), off(d->part_obj_off.size());
            for (size_t p = 0; p + 1 < d->part_obj_off.size(); ++p)
                for (uint64_t o = d->part_obj_off[p]; o < d->part_obj_off[p + 1]; ++o) part_of[o] = (uint32_t)p;
            for (size_t p = 0; p < off.size(); ++p) off[p] = (uint32_t)d->part_obj_off[p];
            try {
                HIP_OK(hipMalloc((void**)&d->part_of, part_of.size() * 4));
                HIP_OK(hipMemcpy(d->part_of, part_of.data(), part_of.size() * 4, hipMemcpyHostToDevice));
                HIP_OK(hipMalloc((void**)&d->part_off, off.size() * 4));
                HIP_OK(hipMemcpy(d->part_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
            } catch (...) {
                pz_dataset_release(d.release());
                throw;
            }
        }
        try {
            pz::analyze_topology(c, d.get());
        } catch (...) {
            pz_dataset_release(d.release());
            throw;
        }
        *out = d.release();
        return PZ_OK;
    });
}

pz_status pz_dataset_load_pre(pz_ctx* c, const char* path, pz_dataset** out) {
    pz_host_mesh* m = nullptr;
    pz_status st = pz_pre_read(path, &m);
    if (st != PZ_OK) return st;
    pz_mesh_view v = m->view();
    st = pz_dataset_upload(c, &v, out);
    pz_host_mesh_free(m);
    return st;
}

void pz_dataset_release(pz_dataset* d) {
    if (!d) return;
    if (d->slot >= 0) {  // the arrays are the context's (instancing slot)
        delete d;
        return;
    }
    for (void* p : {(void*)d->v, (void*)d->idx, (void*)d->t2o, (void*)d->es, (void*)d->ee, (void*)d->eo,
                    (void*)d->part_of, (void*)d->part_off, (void*)d->topo, (void*)d->topo_chunks})
        if (p) (void)hipFree(p);
    delete d;
}

pz_status pz_dataset_info(const pz_dataset* d, uint64_t out[5]) {
    if (!d || !out) return fail(PZ_E_INVALID, "pz_dataset_info: null argument");
    out[0] = d->nv;
    out[1] = d->nt;
    out[2] = d->ne;
    out[3] = (uint64_t)std::max(d->nobj, 0);
    out[4] = d->grid.size();
    return PZ_OK;
}

pz_status pz_dataset_flags(const pz_dataset* d, uint32_t* flags) {
    if (!d || !flags) return fail(PZ_E_INVALID, "pz_dataset_flags: null argument");
    *flags = d->vruns ? PZ_DATASET_VERTEX_RUNS : 0u;
    return PZ_OK;
}

void pz_join_desc_defaults(pz_join_desc* d) {
    std::memset(d, 0, sizeof(*d));
    d->predicate = PZ_PRED_OVERLAP;
    d->direction = PZ_DIR_BOTH;
    d->strategy = PZ_STRATEGY_TWO_LEVEL;
    d->precision = PZ_PREC_FP32_FILTERED;
    d->sizing = PZ_SIZING_TWO_PASS;
    d->dedup = PZ_DEDUP_SORT;
    d->load_factor = 0.5f;
    d->margin = 0.25f;
    d->gamma = 0.8f;
    d->epsilon = 0.001f;
}

pz_status pz_join(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, const pz_join_desc* desc, pz_join_result* r) {
    if (!c || !a || !b || !desc || !r) return fail(PZ_E_INVALID, "pz_join: null argument");
    if (a->ctx != c || b->ctx != c) return fail(PZ_E_INVALID, "datasets belong to another context");
    if (!dataset_live(a) || !dataset_live(b))
        return fail(PZ_E_INVALID, "a dataset's instancing slot has been reused by a later instantiation");
    pz_join_desc d = *desc;
    if (d.predicate < 0 || d.predicate > 2) return fail(PZ_E_INVALID, "unknown predicate");
    if (d.direction < 0 || d.direction > 2) return fail(PZ_E_INVALID, "unknown query direction");
    if (d.sizing < 0 || d.sizing > 3) return fail(PZ_E_INVALID, "unknown sizing mode");
    if (d.load_factor < 0 || d.load_factor > 1) return fail(PZ_E_INVALID, "load factor must be in (0, 1]");
    if (d.sizing == PZ_SIZING_LEARNED && !d.model) return fail(PZ_E_INVALID, "learned sizing needs a model");
    if (d.keep_device_result) return fail(PZ_E_INVALID, "keep_device_result is reserved");
    if (d.broad_mode < 0 || d.broad_mode > 3) return fail(PZ_E_INVALID, "unknown broad-phase mode");
    init_result(r);
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        double t0 = now_ms();
        Phases ph(c, r);
        if (d.sizing == PZ_SIZING_ESTIMATE) {
            // the reference's grid estimate of result pairs; the global strategy
            // sizes its hash set with it, the two-level one sizes its candidate
            // buffer from a candidate estimate of its own
            ph.begin("Selectivity Estimation");
            double est = grid_estimate(a, b, d.gamma > 0 ? d.gamma : 0.8f, d.epsilon > 0 ? d.epsilon : 0.001f,
                                       !d.no_alpha_correction);
            r->predicted_pairs = est > 0 ? (uint64_t)est : 0;
            r->predicted = r->predicted_pairs;
        }
        if (d.strategy == PZ_STRATEGY_GLOBAL)
            join_global(c, a, b, d, r, ph);
        else
            join_two_level(c, a, b, d, r, ph);
        ph.finish();
        r->ms_total = now_ms() - t0;
        return PZ_OK;
    });
}

void pz_join_result_release(pz_join_result* r) {
    if (!r) return;
    std::free(r->pairs);
    std::free(r->pair_types);
    r->pairs = nullptr;
    r->pair_types = nullptr;
    r->num_pairs = 0;
}

pz_status pz_count_candidate_tests(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, uint64_t* out) {
    if (!c || !a || !b || !out) return fail(PZ_E_INVALID, "pz_count_candidate_tests: null argument");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        hipStream_t s = c->s;
        unsigned long long* stats = buf<unsigned long long>(c, "stats", pzk::kStatSlots);
        SetPrep pa = prep_set(c, a, "A"), pb = prep_set(c, b, "B");
        GlobalBvh ga = build_global(c, a, pa, "A");
        GlobalBvh gb = build_global(c, b, pb, "B");
        uint64_t n1 = a->ne, n2 = b->ne;
        uint32_t* cnt = buf<uint32_t>(c, "g.cnt", n1 + n2 + 1);
        uint64_t* off = buf<uint64_t>(c, "g.off", n1 + n2 + 2);
        void* tmp = buf<char>(c, "scan.tmp.g", pzk::scan_temp_bytes(n1 + n2 + 1));
        uint32_t* overflow = buf<uint32_t>(c, "g.overflow", 1);
        pzk::trace_global(3, a->es, a->ee, a->eo, n1, gb.nodes, gb.leaves, b->nt, gb.height, 0, false, 0, cnt,
                          nullptr, nullptr, nullptr, 0, overflow, stats, s);
        pzk::trace_global(3, b->es, b->ee, b->eo, n2, ga.nodes, ga.leaves, a->nt, ga.height, 1, false, 0, cnt + n1,
                          nullptr, nullptr, nullptr, 0, overflow, stats, s);
        pzk::exclusive_scan_u32_u64(cnt, off, n1 + n2, tmp, s);
        *out = readback<uint64_t>(c, {off + n1 + n2})[0];
        return PZ_OK;
    });
}

pz_status pz_track_containment_rays(pz_ctx* c, const pz_dataset* src, const pz_dataset* tgt, int32_t max_targets,
                                    int32_t* target_ids, uint32_t* target_hits, uint32_t* candidate_count,
                                    uint32_t* candidate_overflow, uint32_t* iterations) {
    if (!c || !src || !tgt) return fail(PZ_E_INVALID, "pz_track_containment_rays: null argument");
    if (max_targets < 1 || max_targets > 4096)
        return fail(PZ_E_INVALID, "pz_track_containment_rays: max_targets must be in [1, 4096]");
    const uint64_t ns = (uint64_t)std::max<int32_t>(src->nobj, 0);
    if (ns && (!target_ids || !target_hits || !candidate_count || !candidate_overflow || !iterations))
        return fail(PZ_E_INVALID, "pz_track_containment_rays: null output array");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        hipStream_t s = c->s;
        const uint64_t slots = ns * (uint64_t)max_targets;
        std::fill(target_ids, target_ids + slots, -1);
        std::fill(target_hits, target_hits + slots, 0u);
        std::fill(candidate_count, candidate_count + ns, 0u);
        std::fill(candidate_overflow, candidate_overflow + ns, 0u);
        std::fill(iterations, iterations + ns, 0u);
        if (ns == 0) return PZ_OK;
        unsigned long long* stats = buf<unsigned long long>(c, "stats", pzk::kStatSlots);
        pzk::fill_ull(stats, 0ull, pzk::kStatSlots, s);
        SetPrep ps = prep_set(c, src, "A");
        SetPrep pt = prep_set(c, tgt, "B");
        GlobalBvh gt{};
        if (tgt->nt) gt = build_global(c, tgt, pt, "B");
        int32_t* ids = buf<int32_t>(c, "trk.ids", slots);
        uint32_t* hits = buf<uint32_t>(c, "trk.hits", slots);
        uint32_t* cnt = buf<uint32_t>(c, "trk.cnt", 3 * ns);
        if (!pzk::track_rays(src->v, src->idx, ps.first_tri, ps.tri_cnt, src->nobj, gt.nodes, gt.leaves, tgt->nt,
                             gt.height, max_targets, ids, hits, cnt, cnt + ns, cnt + 2 * ns, stats, s))
            throw PzError(PZ_E_LIMIT, "set-wide BVH too tall for the tracking ray's stack");
        std::vector<int32_t> h_ids(slots);
        std::vector<uint32_t> h_hits(slots), h_cnt(3 * ns);
        HIP_OK(hipMemcpyAsync(h_ids.data(), ids, slots * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(h_hits.data(), hits, slots * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(h_cnt.data(), cnt, 3 * ns * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        std::vector<std::pair<int32_t, uint32_t>> odd;
        for (uint64_t o = 0; o < ns; ++o) {
            const uint64_t base = o * (uint64_t)max_targets;
            const uint32_t n = h_cnt[o];
            odd.clear();
            for (uint32_t i = 0; i < n; ++i)
                if (h_hits[base + i] & 1u) odd.emplace_back(h_ids[base + i], h_hits[base + i]);
            std::sort(odd.begin(), odd.end());
            for (size_t i = 0; i < odd.size(); ++i) {
                target_ids[base + i] = odd[i].first;
                target_hits[base + i] = odd[i].second;
            }
            candidate_count[o] = n;
            candidate_overflow[o] = h_cnt[ns + o];
            iterations[o] = h_cnt[2 * ns + o];
        }
        return PZ_OK;
    });
}

pz_status pz_dataset_upload_batch(pz_ctx* c, uint32_t nparts, const pz_mesh_view* views, pz_dataset** out) {
    if (!c || !views || !out || nparts == 0) return fail(PZ_E_INVALID, "pz_dataset_upload_batch: bad argument");
    return guarded([&]() -> pz_status {
        // Concatenate: vertex indices, object ids and edge object ids are shifted
        // by the running vertex / object totals; part p owns objects [off[p], off[p+1]).
        pz_host_mesh m;
        std::vector<uint64_t> part_off(nparts + 1, 0);
        uint64_t nv = 0, nt = 0, ne = 0;
        for (uint32_t p = 0; p < nparts; ++p) {
            nv += views[p].num_vertices;
            nt += views[p].num_triangles;
            ne += views[p].num_edges;
        }
        if (nv >= (1ull << 32)) throw PzError(PZ_E_LIMIT, "more than 2^32 vertices in one batch side");
        m.vertices.reserve(3 * nv);
        m.indices.reserve(3 * nt);
        m.tri_to_object.reserve(nt);
        m.edge_starts.reserve(3 * ne);
        m.edge_ends.reserve(3 * ne);
        m.edge_objects.reserve(ne);
        uint64_t vbase = 0, obase = 0;
        for (uint32_t p = 0; p < nparts; ++p) {
            const pz_mesh_view& v = views[p];
            if (v.num_edges == 0) throw PzError(PZ_E_NO_EDGES, "batch part " + std::to_string(p) + " has no edges");
            int32_t mx = -1;
            for (uint64_t t = 0; t < v.num_triangles; ++t) mx = std::max(mx, v.triangle_to_object[t]);
            for (uint64_t e = 0; e < v.num_edges; ++e) mx = std::max(mx, v.edge_objects[e]);
            m.vertices.insert(m.vertices.end(), v.vertices, v.vertices + 3 * v.num_vertices);
            for (uint64_t i = 0; i < 3 * v.num_triangles; ++i) {
                if (v.indices[i] >= v.num_vertices) throw PzError(PZ_E_INVALID, "triangle index beyond the vertex array");
                m.indices.push_back((uint32_t)(v.indices[i] + vbase));
            }
            for (uint64_t t = 0; t < v.num_triangles; ++t)
                m.tri_to_object.push_back(v.triangle_to_object[t] < 0 ? -1 : (int32_t)(v.triangle_to_object[t] + obase));
            m.edge_starts.insert(m.edge_starts.end(), v.edge_starts, v.edge_starts + 3 * v.num_edges);
            m.edge_ends.insert(m.edge_ends.end(), v.edge_ends, v.edge_ends + 3 * v.num_edges);
            for (uint64_t e = 0; e < v.num_edges; ++e)
                m.edge_objects.push_back(v.edge_objects[e] < 0 ? -1 : (int32_t)(v.edge_objects[e] + obase));
            vbase += v.num_vertices;
            obase += (uint64_t)(mx + 1);
            part_off[p + 1] = obase;
        }
        if (obase >= (1ull << 31)) throw PzError(PZ_E_LIMIT, "more than 2^31 objects in one batch side");
        pz_mesh_view cv = m.view();
        pz_status st = pz_dataset_upload_parts(c, &cv, &part_off, out);
        if (st != PZ_OK) throw PzError(st, pz_last_error());
        return PZ_OK;
    });
}

pz_status pz_join_batch(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, uint32_t nq, const pz_join_desc* descs,
                        pz_join_result* results, pz_join_result* summary) {
    if (!c || !a || !b || !descs || !results || nq == 0) return fail(PZ_E_INVALID, "pz_join_batch: bad argument");
    if (a->part_obj_off.size() != (size_t)nq + 1 || b->part_obj_off.size() != (size_t)nq + 1)
        return fail(PZ_E_INVALID, "pz_join_batch: both datasets must come from pz_dataset_upload_batch with nq parts");
    if (a->ctx != c || b->ctx != c) return fail(PZ_E_INVALID, "datasets belong to another context");
    if (!dataset_live(a) || !dataset_live(b))
        return fail(PZ_E_INVALID, "a dataset's instancing slot has been reused by a later instantiation");
    for (uint32_t q = 0; q < nq; ++q) {
        if (descs[q].predicate < 0 || descs[q].predicate > 2) return fail(PZ_E_INVALID, "unknown predicate");
        init_result(&results[q]);
    }
    if (descs[0].strategy != PZ_STRATEGY_TWO_LEVEL)
        return fail(PZ_E_INVALID, "batched joins use the two-level strategy");
    if (descs[0].broad_mode < 0 || descs[0].broad_mode > 2) return fail(PZ_E_INVALID, "unknown broad-phase mode");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        pz_join_result all;
        init_result(&all);
        double t0 = now_ms();
        Phases ph(c, &all);
        pz_join_desc d = descs[0];
        std::vector<uint8_t> hmeta(nq);
        for (uint32_t q = 0; q < nq; ++q) hmeta[q] = pz::pair_meta_of(descs[q]);
        uint8_t* dmeta = buf<uint8_t>(c, "batch.meta", nq);
        HIP_OK(hipMemcpyAsync(dmeta, hmeta.data(), nq, hipMemcpyHostToDevice, c->s));
        if (d.sizing == PZ_SIZING_LEARNED && !d.model)
            throw PzError(PZ_E_INVALID, "learned sizing needs a model");
        if (d.sizing == PZ_SIZING_ESTIMATE) d.sizing = PZ_SIZING_TWO_PASS;  // the grid estimate is per dataset
        std::vector<SizePrediction> preds;
        std::vector<uint64_t> all_part_cand;
        join_two_level(c, a, b, d, &all, ph, dmeta, &preds, &all_part_cand);
        ph.finish();
        all.ms_total = now_ms() - t0;
        // split by part: pairs are sorted by (mesh1, mesh2) global ids, so parts come in order
        std::vector<std::vector<pz_pair>> per(nq);
        std::vector<std::vector<uint8_t>> pert(nq);
        for (uint64_t i = 0; i < all.num_pairs; ++i) {
            const pz_pair g = all.pairs[i];
            const uint64_t ga = (uint64_t)g.object_id_mesh1, gb = (uint64_t)g.object_id_mesh2;
            const uint32_t q = (uint32_t)(std::upper_bound(a->part_obj_off.begin(), a->part_obj_off.end(), ga) -
                                          a->part_obj_off.begin() - 1);
            const uint8_t type = all.pair_types ? all.pair_types[i] : 0;
            const bool want_overlap = descs[q].predicate == PZ_PRED_CONTAINMENT && descs[q].include_overlap_pairs;
            if (type == 1 && !want_overlap) continue;
            per[q].push_back({(int32_t)(ga - a->part_obj_off[q]), (int32_t)(gb - b->part_obj_off[q])});
            pert[q].push_back(type);
        }
        for (uint32_t q = 0; q < nq; ++q) {
            pz_join_result& r = results[q];
            r.num_pairs = per[q].size();
            r.pairs = static_cast<pz_pair*>(std::malloc(std::max<size_t>(per[q].size(), 1) * sizeof(pz_pair)));
            if (!r.pairs) throw PzError(PZ_E_NOMEM, "host result allocation failed");
            std::copy(per[q].begin(), per[q].end(), r.pairs);
            if (descs[q].predicate == PZ_PRED_CONTAINMENT && descs[q].include_overlap_pairs) {
                r.pair_types = static_cast<uint8_t*>(std::malloc(std::max<size_t>(pert[q].size(), 1)));
                if (!r.pair_types) throw PzError(PZ_E_NOMEM, "host result allocation failed");
                std::copy(pert[q].begin(), pert[q].end(), r.pair_types);
            }
            r.ms_total = all.ms_total;
            if (q < preds.size()) {
                r.predicted = (uint64_t)std::llround(preds[q].cand);
                r.predicted_pairs = (uint64_t)std::llround(preds[q].pairs);
                r.capacity = (uint64_t)std::ceil(preds[q].cand_upper);
            }
            if (q < all_part_cand.size()) r.num_candidates = all_part_cand[q];
        }
        if (summary) {
            *summary = all;
            summary->pairs = nullptr;
            summary->pair_types = nullptr;
        }
        std::free(all.pairs);
        std::free(all.pair_types);
        return PZ_OK;
    });
}

pz_status pz_tune_get(const pz_ctx* c, pz_tune* out) {
    if (!c || !out) return fail(PZ_E_INVALID, "pz_tune_get: null argument");
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < pzk::kSmallClasses; ++k) out->build_threads[k] = c->tune.build_threads[k];
    for (int k = 0; k < 2; ++k) {
        out->narrow_queue[k] = c->tune.narrow_queue[k];
        out->narrow_chunk[k] = c->tune.narrow_chunk[k];
    }
    return PZ_OK;
}

static bool valid_queue(int q) { return q == 256 || q == 512 || q == 1024 || q == 1536; }

static bool valid_threads(int t) { return t == 128 || t == 256 || t == 512 || t == 1024; }

pz_status pz_tune_set(pz_ctx* c, const pz_tune* in) {
    if (!c || !in) return fail(PZ_E_INVALID, "pz_tune_set: null argument");
    for (int k = 0; k < pzk::kSmallClasses; ++k)
        if (!valid_threads(in->build_threads[k])) return fail(PZ_E_INVALID, "build_threads must be 128, 256, 512 or 1024");
    if (!valid_queue(in->narrow_queue[0]) || !valid_queue(in->narrow_queue[1]))
        return fail(PZ_E_INVALID, "narrow_queue must be 256, 512, 1024 or 1536");
    for (int k = 0; k < 2; ++k)
        if (in->narrow_chunk[k] < 1 || in->narrow_chunk[k] > 64)
            return fail(PZ_E_INVALID, "narrow_chunk must be in [1, 64]");
    for (int k = 0; k < pzk::kSmallClasses; ++k) c->tune.build_threads[k] = in->build_threads[k];
    for (int k = 0; k < 2; ++k) {
        c->tune.narrow_queue[k] = in->narrow_queue[k];
        c->tune.narrow_chunk[k] = in->narrow_chunk[k];
    }
    return PZ_OK;
}

pz_status pz_autotune(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, const pz_join_desc* desc,
                      pz_tune_report* report) {
    if (!c || !a || !b || !desc || !report) return fail(PZ_E_INVALID, "pz_autotune: null argument");
    if (desc->strategy != PZ_STRATEGY_TWO_LEVEL) return fail(PZ_E_INVALID, "pz_autotune tunes the two-level strategy");
    std::memset(report, 0, sizeof(*report));
    static const int kCand[4] = {128, 256, 512, 1024};
    static const char* const kTimer[pzk::kSmallClasses] = {"build_small.c0", "build_small.c1", "build_small.c2"};
    const TuneTable saved = c->tune;
    const bool was_on = pzk::prof_enabled();
    pz_join_result r;
    // The tuned parameters are the per-object trees' (builder workgroups, the
    // dual-tree walk's queue and chunk): the pair sieve would leave most
    // pairs out of them, so the tuning joins send every pair to the trees.
    pz_join_desc td = *desc;
    if (td.narrow_mode == 0) td.narrow_mode = 2;
    auto run = [&]() -> pz_status {
        pz_status st = pz_join(c, a, b, &td, &r);
        if (st == PZ_OK) pz_join_result_release(&r);
        ++report->joins_run;
        return st;
    };
    pz_status st = run();  // warm-up: buffers grown, code loaded
    for (int i = 0; i < 4 && st == PZ_OK; ++i) {  // LDS builder workgroup size, all classes at once
        for (int k = 0; k < pzk::kSmallClasses; ++k) c->tune.build_threads[k] = kCand[i];
        pzk::prof_reset();
        pzk::prof_enable(true);
        st = run();
        pzk::prof_enable(false);
        for (int k = 0; k < pzk::kSmallClasses; ++k) {
            double ms = 0;
            uint64_t n = 0;
            pzk::prof_query(kTimer[k], &ms, &n);
            report->build_ms[k][i] = n ? ms : -1.0;
        }
    }
    c->tune = saved;
    if (st == PZ_OK)
        for (int k = 0; k < pzk::kSmallClasses; ++k) {
            int best = -1;
            for (int i = 0; i < 4; ++i)
                if (report->build_ms[k][i] >= 0 && (best < 0 || report->build_ms[k][i] < report->build_ms[k][best]))
                    best = i;
            if (best >= 0) c->tune.build_threads[k] = kCand[best];
            report->chosen.build_threads[k] = c->tune.build_threads[k];
        }
    // narrow phase: the dual-tree kernel plus the per-edge kernel its queue overflows go to
    static const int kQueue[4] = {256, 512, 1024, 1536};
    const TuneTable base = c->tune;
    const int qc = c->last_queue_class;  // the warm-up join's class
    for (int i = 0; i < 4 && st == PZ_OK; ++i) {
        c->tune.narrow_queue[qc] = kQueue[i];
        pzk::prof_reset();
        pzk::prof_enable(true);
        st = run();
        pzk::prof_enable(false);
        double ms = 0, ms2 = 0;
        uint64_t n = 0;
        pzk::prof_query("narrow_dual", &ms, &n);
        pzk::prof_query("narrow", &ms2, &n);
        report->narrow_ms[i] = ms + ms2;
    }
    pzk::prof_reset();
    pzk::prof_enable(was_on);
    c->tune = base;
    if (st != PZ_OK) return st;
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (report->narrow_ms[i] < report->narrow_ms[best]) best = i;
    c->tune.narrow_queue[qc] = kQueue[best];
    // work-queue chunk: tasks a wave takes at a time
    static const int kChunk[4] = {1, 2, 4, 8};
    const TuneTable base2 = c->tune;
    pzk::prof_enable(true);
    for (int i = 0; i < 4 && st == PZ_OK; ++i) {
        c->tune.narrow_chunk[qc] = kChunk[i];
        pzk::prof_reset();
        st = run();
        double ms = 0, ms2 = 0;
        uint64_t n = 0;
        pzk::prof_query("narrow_dual", &ms, &n);
        pzk::prof_query("narrow", &ms2, &n);
        report->chunk_ms[i] = ms + ms2;
    }
    pzk::prof_reset();
    pzk::prof_enable(was_on);
    c->tune = base2;
    if (st != PZ_OK) return st;
    best = 0;
    for (int i = 1; i < 4; ++i)
        if (report->chunk_ms[i] < report->chunk_ms[best]) best = i;
    c->tune.narrow_chunk[qc] = kChunk[best];
    for (int k = 0; k < 2; ++k) {
        report->chosen.narrow_queue[k] = c->tune.narrow_queue[k];
        report->chosen.narrow_chunk[k] = c->tune.narrow_chunk[k];
    }
    report->queue_class = qc;
    return PZ_OK;
}

void pz_ctx_set_graphs(pz_ctx* c, int32_t on) {
    if (c) c->graphs = on != 0;
}

void pz_profile_enable(int32_t on) { pzk::prof_enable(on != 0); }
void pz_profile_reset(void) { pzk::prof_reset(); }
pz_status pz_profile_query(const char* kernel, double* total_ms, uint64_t* launches) {
    if (!kernel || !total_ms || !launches) return fail(PZ_E_INVALID, "pz_profile_query: null argument");
    pzk::prof_query(kernel, total_ms, launches);
    return PZ_OK;
}

// ----------------------------------------------- reference extern "C" helpers
}  // extern "C"

namespace {
// One device workspace per HIP device for the reference-named helpers: grown
// on demand, kept for later calls, returned when the library is unloaded.
// The helpers run on the current device and the null stream and block until
// done, so one mutex serialises them.
struct HelperSpace {
    std::mutex mu;
    std::vector<DevBuf> dev;
    ~HelperSpace() {
        for (DevBuf& b : dev)
            if (b.p) (void)hipFree(b.p);
    }
};
HelperSpace& helper_space() {
    static HelperSpace h;
    return h;
}
// Workspace of at least `bytes` on the current device (caller holds the mutex).
char* helper_ws(size_t bytes) {
    HelperSpace& h = helper_space();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    if ((size_t)dev >= h.dev.size()) h.dev.resize((size_t)dev + 1);
    DevBuf& b = h.dev[(size_t)dev];
    if (b.n < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 20);
        if (hipMalloc(&b.p, want) != hipSuccess) return nullptr;
        b.n = want;
    }
    return static_cast<char*>(b.p);
}
inline size_t up256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

namespace pz {
// Scratch layout of the merge: keys, sorted keys, unique keys (n + 1 u64
// each), then the radix sort's and the unique pass's temporaries.
size_t merge_scratch_bytes(uint64_t n) {
    return up256(3 * (n + 1) * 8) + up256(pzk::radix_sort_temp_bytes(n)) + up256(pzk::unique_temp_bytes(n));
}
// Sorted unique merge of s1[0..n1) and s2[0..n2) into out through `scratch`
// (merge_scratch_bytes(n1 + n2) bytes of device memory) on stream s; the
// sources may alias out (their keys are copied into scratch first).
// Returns the unique count, or -1 on a HIP failure.
long long merge_pairs_into(const pz_pair* s1, uint64_t n1, const pz_pair* s2, uint64_t n2, pz_pair* out,
                           char* scratch, hipStream_t s) {
    const uint64_t n = n1 + n2;
    uint64_t* k = reinterpret_cast<uint64_t*>(scratch);
    uint64_t* k2 = k + (n + 1);
    uint64_t* u = k2 + (n + 1);
    char* tmp = scratch + up256(3 * (n + 1) * 8);
    char* utmp = tmp + up256(pzk::radix_sort_temp_bytes(n));
    pzk::pairs_to_keys(reinterpret_cast<const int32_t*>(s1), n1, k, s);
    pzk::pairs_to_keys(reinterpret_cast<const int32_t*>(s2), n2, k + n1, s);
    const bool alt = pzk::radix_sort_keys(k, k2, n, 0, 64, tmp, s);
    uint64_t* sk = alt ? k2 : k;
    uint64_t* dcount = alt ? k : k2;  // the sort's other buffer is free again: its first slot holds the count
    pzk::unique_u64(sk, n, u, dcount, utmp, s);
    uint64_t cnt = 0;
    if (hipMemcpyAsync(&cnt, dcount, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    pzk::keys_to_pairs(u, cnt, reinterpret_cast<int32_t*>(out), s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return (long long)cnt;
}
}  // namespace pz

extern "C" {

long long merge_and_deduplicate_pairs_gpu(const pz_pair* d1, long long n1, const pz_pair* d2, long long n2,
                                          pz_pair* d_out) {
    if (n1 < 0 || n2 < 0) return -1;
    const uint64_t n = (uint64_t)n1 + (uint64_t)n2;
    if (n == 0) return 0;
    if (!d_out) return -1;
    // a null source with a count: those pairs already sit in the output
    // (mesh_query_deduplication.cu:95-106 copies only non-null sources)
    const pz_pair* s1 = d1 ? d1 : d_out;
    const pz_pair* s2 = d2 ? d2 : d_out + n1;
    std::lock_guard<std::mutex> lk(helper_space().mu);
    char* ws = helper_ws(pz::merge_scratch_bytes(n));
    if (!ws) return -1;
    return pz::merge_pairs_into(s1, (uint64_t)n1, s2, (uint64_t)n2, d_out, ws, nullptr);
}

int compact_hash_table_pairs(const unsigned long long* d_table, unsigned long long size, pz_pair* d_output,
                             int max_output_size) {
    if (!d_table || !d_output || max_output_size < 0) return -1;
    const uint64_t cap = (uint64_t)max_output_size;
    std::lock_guard<std::mutex> lk(helper_space().mu);
    char* ws = helper_ws((cap + 2) * 8);
    if (!ws) return -1;
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(ws);
    uint64_t* keys = reinterpret_cast<uint64_t*>(cnt + 1);
    hipStream_t s = nullptr;
    pzk::fill_ull(cnt, 0ull, 1, s);
    pzk::compact_table(d_table, size, keys, cap, cnt, s);
    unsigned long long total = 0;
    if (hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    const uint64_t n = std::min<uint64_t>(total, cap);
    pzk::keys_to_pairs(keys, n, reinterpret_cast<int32_t*>(d_output), s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return (int)n;
}

long long exclusive_scan_gpu(const int* d_input, long long* d_output, int num_elements) {
    if (num_elements < 0 || (num_elements && (!d_input || !d_output))) return -1;
    const uint64_t n = (uint64_t)num_elements;
    std::lock_guard<std::mutex> lk(helper_space().mu);
    // the scan writes n + 1 entries (the total last); the caller's array holds n
    char* ws = helper_ws(up256((n + 1) * 8) + pzk::scan_temp_bytes(n));
    if (!ws) return -1;
    int64_t* out = reinterpret_cast<int64_t*>(ws);
    char* tmp = ws + up256((n + 1) * 8);
    hipStream_t s = nullptr;
    pzk::exclusive_scan_i32_i64(reinterpret_cast<const int32_t*>(d_input), out, n, tmp, s);
    if (n && hipMemcpyAsync(d_output, out, n * 8, hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
    long long total = -1;
    if (hipMemcpyAsync(&total, out + n, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return total;
}

float pz_estimate_overlap_selectivity_cells(const pz_cell_stats* a, const pz_cell_stats* b, int32_t n,
                                            float cell_volume, float epsilon, float gamma) {
    if (n < 0 || (n && (!a || !b))) return -1.0f;
    if (n == 0) return 0.0f;
    const size_t cb = up256((size_t)n * sizeof(pz_cell_stats));
    std::lock_guard<std::mutex> lk(helper_space().mu);
    char* ws = helper_ws(2 * cb + 256);
    if (!ws) return -1.0f;
    hipStream_t s = nullptr;
    double* total = reinterpret_cast<double*>(ws + 2 * cb);
    if (hipMemcpyAsync(ws, a, (size_t)n * sizeof(pz_cell_stats), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(ws + cb, b, (size_t)n * sizeof(pz_cell_stats), hipMemcpyHostToDevice, s) != hipSuccess)
        return -1.0f;
    pzk::cell_selectivity(reinterpret_cast<const uint32_t*>(ws), reinterpret_cast<const uint32_t*>(ws + cb),
                          (uint32_t)n, cell_volume, epsilon, gamma, total, s);
    double h = 0;
    if (hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return -1.0f;
    if (hipStreamSynchronize(s) != hipSuccess) return -1.0f;
    return (float)h;
}

float pz_estimate_overlap_selectivity(const pz_grid_cell* a, const pz_grid_cell* b, int32_t n, float cell_volume,
                                      float epsilon, float gamma) {
    if (n < 0 || (n && (!a || !b)) || !(cell_volume > 0)) return -1.0f;
    double total = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        if (a[i].touch_count == 0 || b[i].touch_count == 0) continue;
        float comb = a[i].avg_size_mean + b[i].avg_size_mean + epsilon;
        float prob = comb * comb * comb / cell_volume;
        prob *= std::pow(std::sqrt(a[i].vol_ratio * b[i].vol_ratio), gamma);
        prob = std::min(prob, 1.0f);
        total += (double)a[i].touch_count * (double)b[i].touch_count * prob;
    }
    return (float)total;
}

pz_status pz_estimate_grid(const pz_grid_cell* cells_a, uint64_t num_a, const pz_grid_cell* cells_b,
                           uint64_t num_b, float cell_size, float epsilon, float gamma, int32_t alpha_correction,
                           pz_grid_estimate* out) {
    if (!out || (num_a && !cells_a) || (num_b && !cells_b)) return fail(PZ_E_INVALID, "null argument");
    return guarded([&]() -> pz_status {
        estimate_grid(cells_a, num_a, cells_b, num_b, cell_size, epsilon, gamma, alpha_correction != 0, out);
        return PZ_OK;
    });
}

pz_status pz_dataset_estimate_grid(const pz_dataset* a, const pz_dataset* b, float epsilon, float gamma,
                                   int32_t alpha_correction, pz_grid_estimate* out) {
    if (!a || !b || !out) return fail(PZ_E_INVALID, "null argument");
    return pz_estimate_grid(a->grid.data(), a->grid.size(), b->grid.data(), b->grid.size(), a->cell_size, epsilon,
                            gamma, alpha_correction, out);
}

}  // extern "C"
