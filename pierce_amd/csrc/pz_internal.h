// Internal declarations shared by the host translation units (compiled by
// hipcc): context and dataset state, device buffer pool, error plumbing, and
// the per-set statistics both the joins and the size estimator start from.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <initializer_list>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "host/host_mesh.hpp"
#include "kernels/pz_device.h"
#include "kernels/pz_kernels.h"
#include "kernels/pz_kernels_join.h"
#include "pierce_hip.h"

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

// Launch parameters picked per size class (defaults, or measured by pz_autotune).
struct TuneTable {
    int build_threads[3] = {128, 256, 512};  // LDS builder workgroup size for objects <= 128 / 512 / 2048 triangles
    // dual-tree narrow phase: LDS queue entries per wave, for joins whose objects
    // are all LDS-built (<= 2048 triangles) [0] or that index larger ones [1]
    int narrow_queue[2] = {1536, 1536};
    int narrow_chunk[2] = {1, 1};  // tasks a wave takes from the narrow-phase work queue at a time
};

struct pz_ctx {
    int device = 0;
    int cus = 1;                       // compute units of the device (read once at creation)
    hipStream_t s2 = nullptr;          // side stream (default priority) for work that overlaps the main stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void* pinned_est = nullptr;        // host landing area of the side stream's predictions
    size_t pinned_est_n = 0;
    TuneTable tune;
    int last_queue_class = 0;  // narrow-queue class of the last two-level join (what pz_autotune tunes)
    unsigned persistent_blocks = 0;  // grid of the persistent narrow kernel: 8 blocks per CU
    uint64_t pool_epoch = 0;         // bumped whenever a pool buffer moves (captured graphs go stale)
    bool graphs = true;              // capture repeated fixed-shape segments as HIP graphs
    std::vector<struct GraphSeg*> segs;
    hipStream_t s = nullptr;
    std::unordered_map<std::string, DevBuf> bufs;
    void* pinned = nullptr;  // small readbacks
    size_t pinned_n = 0;
    hipEvent_t ev[PZ_MAX_PHASES + 1];
    // generation of each instancing slot (a dataset built into slot k is valid
    // while slot_gen[k] still equals its own generation)
    uint64_t slot_gen[PZ_INSTANCE_SLOTS] = {};
};

struct pz_dataset {
    pz_ctx* ctx = nullptr;
    uint64_t uid = 0;  // unique per upload (keys the captured HIP graphs)
    float* v = nullptr;
    uint32_t* idx = nullptr;
    int32_t* t2o = nullptr;
    float* es = nullptr;
    float* ee = nullptr;
    int32_t* eo = nullptr;
    uint64_t nv = 0, nt = 0, ne = 0;
    int32_t nobj = 0;
    float max_abs = 0.f;
    std::vector<pz_grid_cell> grid;
    float cell_size = 0.f;
    // batched datasets: part p owns objects [part_obj_off[p], part_obj_off[p + 1])
    std::vector<uint64_t> part_obj_off;
    uint32_t* part_of = nullptr;   // device: part id per object
    uint32_t* part_off = nullptr;  // device: part_obj_off as uint32
    // instanced into a context slot: the arrays belong to the context's pool
    int32_t slot = -1;
    uint64_t slot_gen = 0;
    // Objects own disjoint, gap-free vertex runs (found at upload): device
    // per-object first and last triangle, triangle count and last vertex, and
    // the (object, first vertex) chunks the per-join bounds kernel reads.
    bool vruns = false;
    uint32_t* topo = nullptr;  // first[n], last[n], cnt[n], vmax[n], vmin[n] (n = max(nobj, 1))
    uint2* topo_chunks = nullptr;
    uint64_t topo_nchunks = 0;
};

// A dataset whose slot has been reused no longer holds its own geometry.
inline bool dataset_live(const pz_dataset* d) {
    return d->slot < 0 || d->ctx->slot_gen[d->slot] == d->slot_gen;
}

namespace pz {

#define HIP_OK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw PzError(e_ == hipErrorOutOfMemory ? PZ_E_NOMEM : PZ_E_HIP,                           \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                          \
    } while (0)

struct PzError {
    pz_status code;
    std::string msg;
    PzError(pz_status c, std::string m) : code(c), msg(std::move(m)) {}
};

template <class T>
inline T* buf(pz_ctx* c, const std::string& name, size_t count) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    DevBuf& b = c->bufs[name];
    if (b.n < bytes) {
        if (b.p) HIP_OK(hipFree(b.p));
        b.p = nullptr;
        b.n = 0;
        // headroom for growth, at most 1 GiB (large joins would otherwise waste a quarter of HBM)
        size_t want = bytes + std::min<size_t>(bytes / 4, size_t(1) << 30);
        HIP_OK(hipMalloc(&b.p, want));
        b.n = want;
        ++c->pool_epoch;
    }
    return reinterpret_cast<T*>(b.p);
}

inline void* pinned(pz_ctx* c, size_t bytes) {
    if (c->pinned_n < bytes) {
        if (c->pinned) HIP_OK(hipHostFree(c->pinned));
        c->pinned = nullptr;
        HIP_OK(hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
        c->pinned_n = bytes;
    }
    return c->pinned;
}

// Read up to 8 scalars (each from a device address) with one gather kernel
// writing into pinned memory and one synchronisation.
template <class T>
std::vector<T> readback(pz_ctx* c, std::initializer_list<const T*> srcs) {
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    if (srcs.size() > (size_t)pzk::HostGather::kMax) throw PzError(PZ_E_INVALID, "readback of too many values");
    T* h = reinterpret_cast<T*>(pinned(c, srcs.size() * sizeof(T)));
    pzk::HostGather g{};
    for (const T* p : srcs) {
        g.src[g.n] = reinterpret_cast<const uint32_t*>(p);
        g.len[g.n++] = sizeof(T) / 4;
    }
    pzk::gather_to_host(g, reinterpret_cast<uint32_t*>(h), c->s);
    HIP_OK(hipStreamSynchronize(c->s));
    return std::vector<T>(h, h + srcs.size());
}

inline int bits_for(uint64_t n) {  // bits needed to hold values < n
    int b = 0;
    while (b < 64 && (1ull << b) < n) ++b;
    return std::max(b, 1);
}

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class F>
pz_status guarded(F f) {
    try {
        return f();
    } catch (const PzError& e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(PZ_E_NOMEM, "host allocation failed");
    }
}

// per-object statistics of one set (device): ordered-uint boxes (lo xyz, hi xyz),
// triangle counts, first / last triangle index, edge counts
struct SetPrep {
    uint32_t* obox;
    uint32_t* tri_cnt;
    uint32_t* first_tri;
    uint32_t* last_tri;
    uint32_t* edge_cnt;
};
SetPrep prep_set(pz_ctx* c, const pz_dataset* d, const std::string& tag);
// Unique id of a new dataset (keys the captured HIP graphs).
uint64_t next_dataset_uid();
void analyze_topology(pz_ctx* c, pz_dataset* d);

// LBVH over the object boxes of one set (broad-phase target).
struct ObjBvh {
    pzd::Node* nodes = nullptr;  // nullptr: no object index (the broad phase tests every box)
    uint32_t* order = nullptr;
    uint32_t* height = nullptr;
};

// Host-side results of a segment of device work (pool pointers), kept with its graph.
struct SegOut {
    SetPrep pa{}, pb{};
    ObjBvh ob{};
    pzk::ObjGrid og{};
    uint32_t* big[2] = {nullptr, nullptr};  // per-object "built by the big-object path" flags of both sets
};

// predicate | direction bits << 2 of one join description
uint8_t pair_meta_of(const pz_join_desc& d);

// Learned output sizing (mlp.cpp).  One prediction per part (a plain dataset
// is one part): point estimates and calibrated upper bounds of the
// candidate-pair count (object pairs with overlapping boxes) and of the
// result-pair count.
struct SizePrediction {
    double cand, pairs;
    double cand_upper, pairs_upper;
};
// meta_of_part: device predicate | dir << 2 per part (batched) or null (meta for every part).
// Enqueues features, inference and the copy of the outputs to the host on the
// side stream after everything enqueued so far on the main stream; collect
// with mlp_predict_wait.  Returns the number of parts.
uint32_t mlp_predict_async(pz_ctx* c, const pz_mlp* model, const pz_dataset* a, const pz_dataset* b,
                           const SetPrep& pa, const SetPrep& pb, const uint8_t* meta_of_part, uint8_t meta);
std::vector<SizePrediction> mlp_predict_wait(pz_ctx* c, const pz_mlp* model, uint32_t nparts);
// Expected number of object pairs whose boxes meet (plain datasets): the
// 16^3-grid box-pair expectation of the estimator's features, on the main stream.
double candidate_estimate(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, const SetPrep& pa, const SetPrep& pb);

// Sorted unique merge of the reference's merge_and_deduplicate_pairs_gpu
// (capi.cpp): scratch bytes for n input pairs, and the merge of s1[0..n1) and
// s2[0..n2) into out through that scratch on stream s (sources may alias out).
// Returns the unique count, or -1 on a HIP failure.
size_t merge_scratch_bytes(uint64_t n);
long long merge_pairs_into(const pz_pair* s1, uint64_t n1, const pz_pair* s2, uint64_t n2, pz_pair* out,
                           char* scratch, hipStream_t s);
}  // namespace pz
