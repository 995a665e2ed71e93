// Mesh library resident in HBM and the sets assembled from it by instancing
// (pz_library_*), and the read-back of a device-resident dataset
// (pz_dataset_download).
//
// A library keeps one copy of a catalogue of meshes; a query set is a list of
// (library mesh, affine map) instances.  Instantiation writes a batched
// dataset of the same layout pz_dataset_upload_batch builds from host arrays
// (part p = the instances of query p), so every join entry point takes it.
// It moves the set's bytes device to device instead of across PCIe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host/host_mesh.hpp"
#include "pierce_hip.h"
#include "pz_internal.h"

using namespace pz;

struct pz_library {
    pz_ctx* ctx = nullptr;
    float* v = nullptr;
    uint32_t* idx = nullptr;
    float* es = nullptr;
    float* ee = nullptr;
    uint64_t nv = 0, nt = 0, ne = 0;
    // per library object: runs of vertices, triangles and edges, largest |coordinate|
    std::vector<uint32_t> v_off, v_cnt, t_off, t_cnt, e_off, e_cnt;
    std::vector<float> max_abs;
    // every object's triangles reference every vertex of its run: instanced
    // sets then own vertex runs (their topology is known without a scan)
    bool runs_full = true;
};

extern "C" {

pz_status pz_library_create(pz_ctx* c, const pz_mesh_view* v, pz_library** out) {
    if (!c || !v || !out) return fail(PZ_E_INVALID, "pz_library_create: null argument");
    if (v->num_triangles >= (1ull << 31) || v->num_edges >= (1ull << 31) || v->num_vertices >= (1ull << 32))
        return fail(PZ_E_LIMIT, "a library holds at most 2^31 - 1 triangles and edges and 2^32 - 1 vertices");
    if (v->num_triangles && (!v->vertices || !v->indices || !v->triangle_to_object))
        return fail(PZ_E_INVALID, "null geometry array");
    if (v->num_edges && (!v->edge_starts || !v->edge_ends || !v->edge_objects))
        return fail(PZ_E_INVALID, "null edge array");
    return guarded([&]() -> pz_status {
        std::unique_ptr<pz_library> L(new pz_library());
        L->ctx = c;
        L->nv = v->num_vertices;
        L->nt = v->num_triangles;
        L->ne = v->num_edges;
        int32_t nobj = 0;
        for (uint64_t t = 0; t < L->nt; ++t) {
            if (v->triangle_to_object[t] < 0) throw PzError(PZ_E_INVALID, "library triangles need object ids >= 0");
            nobj = std::max(nobj, v->triangle_to_object[t] + 1);
        }
        const size_t n = (size_t)nobj;
        L->v_off.assign(n, 0);
        L->v_cnt.assign(n, 0);
        L->t_off.assign(n, 0);
        L->t_cnt.assign(n, 0);
        L->e_off.assign(n, 0);
        L->e_cnt.assign(n, 0);
        L->max_abs.assign(n, 0.f);
        // triangles: one contiguous run per object, objects in id order
        for (uint64_t t = 0; t < L->nt; ++t) {
            const int32_t o = v->triangle_to_object[t];
            if (L->t_cnt[o] == 0) L->t_off[o] = (uint32_t)t;
            else if (L->t_off[o] + L->t_cnt[o] != t)
                throw PzError(PZ_E_INVALID, "library object " + std::to_string(o) + ": triangles not contiguous");
            ++L->t_cnt[o];
        }
        // vertices: each object's triangles use one run of vertices no other object uses
        std::vector<uint32_t> owner(L->nv, UINT32_MAX);
        std::vector<uint8_t> referenced(L->nv, 0);
        for (size_t o = 0; o < n; ++o) {
            if (!L->t_cnt[o]) continue;
            uint32_t lo = UINT32_MAX, hi = 0;
            for (uint64_t t = L->t_off[o]; t < (uint64_t)L->t_off[o] + L->t_cnt[o]; ++t)
                for (int k = 0; k < 3; ++k) {
                    const uint32_t i = v->indices[3 * t + k];
                    if (i >= L->nv) throw PzError(PZ_E_INVALID, "triangle index beyond the vertex array");
                    lo = std::min(lo, i);
                    hi = std::max(hi, i);
                    referenced[i] = 1;
                }
            for (uint32_t i = lo; i <= hi; ++i) {
                if (owner[i] != UINT32_MAX && owner[i] != (uint32_t)o)
                    throw PzError(PZ_E_INVALID, "library objects share vertices");
                owner[i] = (uint32_t)o;
                L->runs_full = L->runs_full && referenced[i];
            }
            L->v_off[o] = lo;
            L->v_cnt[o] = hi - lo + 1;
            float ma = 0.f;
            for (uint64_t i = 3ull * lo; i < 3ull * (hi + 1ull); ++i) ma = std::max(ma, std::fabs(v->vertices[i]));
            L->max_abs[o] = ma;
        }
        // edges: regrouped per object, in their order within each object
        std::vector<uint64_t> ecount(n + 1, 0);
        for (uint64_t e = 0; e < L->ne; ++e) {
            const int32_t o = v->edge_objects[e];
            if (o < 0 || o >= nobj) throw PzError(PZ_E_INVALID, "library edge of an object without triangles");
            ++ecount[(size_t)o + 1];
        }
        for (size_t o = 0; o < n; ++o) ecount[o + 1] += ecount[o];
        std::vector<float> es(3 * L->ne), ee(3 * L->ne);
        {
            std::vector<uint64_t> fill(ecount.begin(), ecount.end() - 1);
            for (uint64_t e = 0; e < L->ne; ++e) {
                const uint64_t d = fill[v->edge_objects[e]]++;
                std::memcpy(&es[3 * d], &v->edge_starts[3 * e], 12);
                std::memcpy(&ee[3 * d], &v->edge_ends[3 * e], 12);
            }
        }
        for (size_t o = 0; o < n; ++o) {
            L->e_off[o] = (uint32_t)ecount[o];
            L->e_cnt[o] = (uint32_t)(ecount[o + 1] - ecount[o]);
            for (uint64_t i = 3 * ecount[o]; i < 3 * ecount[o + 1]; ++i)
                L->max_abs[o] = std::max(L->max_abs[o], std::max(std::fabs(es[i]), std::fabs(ee[i])));
        }
        HIP_OK(hipSetDevice(c->device));
        auto up = [&](void** dst, const void* src, size_t bytes) {
            HIP_OK(hipMalloc(dst, std::max<size_t>(bytes, 16)));
            if (bytes) HIP_OK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        };
        try {
            up((void**)&L->v, v->vertices, 12 * L->nv);
            up((void**)&L->idx, v->indices, 12 * L->nt);
            up((void**)&L->es, es.data(), 12 * L->ne);
            up((void**)&L->ee, ee.data(), 12 * L->ne);
        } catch (...) {
            pz_library_release(L.release());
            throw;
        }
        *out = L.release();
        return PZ_OK;
    });
}

void pz_library_release(pz_library* L) {
    if (!L) return;
    for (void* p : {(void*)L->v, (void*)L->idx, (void*)L->es, (void*)L->ee})
        if (p) (void)hipFree(p);
    delete L;
}

pz_status pz_library_info(const pz_library* L, uint64_t out[4]) {
    if (!L || !out) return fail(PZ_E_INVALID, "pz_library_info: null argument");
    out[0] = L->nv;
    out[1] = L->nt;
    out[2] = L->ne;
    out[3] = L->t_cnt.size();
    return PZ_OK;
}

pz_status pz_library_instantiate(pz_ctx* c, const pz_library* L, uint32_t nparts, const uint64_t* counts,
                                 const pz_instance* inst, int32_t slot, pz_dataset** out) {
    if (!c || !L || !counts || !out || nparts == 0) return fail(PZ_E_INVALID, "pz_library_instantiate: bad argument");
    if (slot >= PZ_INSTANCE_SLOTS) return fail(PZ_E_INVALID, "pz_library_instantiate: no such instancing slot");
    if (L->ctx != c) return fail(PZ_E_INVALID, "the library belongs to another context");
    uint64_t ninst = 0;
    for (uint32_t p = 0; p < nparts; ++p) ninst += counts[p];
    if (ninst && !inst) return fail(PZ_E_INVALID, "pz_library_instantiate: null instances");
    if (ninst >= (1ull << 31)) return fail(PZ_E_LIMIT, "more than 2^31 - 1 instances in one set");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        const uint64_t nlib = L->t_cnt.size();
        std::vector<pzk::InstanceRec> rec(std::max<uint64_t>(ninst, 1));
        uint64_t nv = 0, nt = 0, ne = 0;
        float ma = 0.f;
        for (uint64_t i = 0; i < ninst; ++i) {
            const pz_instance& x = inst[i];
            if (x.mesh >= nlib) throw PzError(PZ_E_INVALID, "instance of a mesh the library does not hold");
            pzk::InstanceRec& r = rec[i];
            r.sv = L->v_off[x.mesh];
            r.nv = L->v_cnt[x.mesh];
            r.st = L->t_off[x.mesh];
            r.nt = L->t_cnt[x.mesh];
            r.se = L->e_off[x.mesh];
            r.ne = L->e_cnt[x.mesh];
            r.dv = (uint32_t)nv;
            r.dt = (uint32_t)nt;
            r.de = (uint32_t)ne;
            r.obj = (uint32_t)i;
            std::memcpy(r.m, x.xform, sizeof(r.m));
            nv += r.nv;
            nt += r.nt;
            ne += r.ne;
            if (nv >= (1ull << 32) || nt >= (1ull << 31) || ne >= (1ull << 31))
                throw PzError(PZ_E_LIMIT, "an instantiated set is limited to 2^31 - 1 triangles and edges");
            // bound on |A x + t| over the mesh (rounded up for the float evaluation)
            float row = 0.f, tmax = 0.f;
            for (int k = 0; k < 3; ++k) {
                row = std::max(row, std::fabs(x.xform[3 * k]) + std::fabs(x.xform[3 * k + 1]) +
                                        std::fabs(x.xform[3 * k + 2]));
                tmax = std::max(tmax, std::fabs(x.xform[9 + k]));
            }
            const float b = (row * L->max_abs[x.mesh] + tmax) * (1.0f + 0x1p-20f);
            ma = std::max(ma, std::isfinite(b) ? b : INFINITY);
        }
        if (ne == 0) throw PzError(PZ_E_NO_EDGES, "the instantiated set has no edges");
        static_assert(sizeof(pzk::InstanceRec) % 8 == 0, "record layout");
        std::unique_ptr<pz_dataset> d(new pz_dataset());
        d->ctx = c;
        d->uid = next_dataset_uid();
        d->nv = nv;
        d->nt = nt;
        d->ne = ne;
        d->nobj = (int32_t)ninst;
        d->max_abs = ma;
        d->part_obj_off.assign(nparts + 1, 0);
        for (uint32_t p = 0; p < nparts; ++p) d->part_obj_off[p + 1] = d->part_obj_off[p] + counts[p];
        std::vector<uint32_t> part_of(std::max<uint64_t>(ninst, 1), 0u), off(nparts + 1);
        for (uint32_t p = 0; p < nparts; ++p)
            for (uint64_t o = d->part_obj_off[p]; o < d->part_obj_off[p + 1]; ++o) part_of[o] = p;
        for (uint32_t p = 0; p <= nparts; ++p) off[p] = (uint32_t)d->part_obj_off[p];
        // Each instance owns the vertex run and the triangle run it is written to,
        // so the topology the upload's scan finds (first / last triangle, count,
        // last / first vertex per object, and the vertex chunks of the bounds
        // kernel) follows from the records when the library's runs are full.
        const uint32_t K = pzk::run_chunk_vertices();
        const bool runs = L->runs_full && ninst > 0 && nv < (1ull << 32) - K;
        const size_t tn = std::max<uint64_t>(ninst, 1);
        std::vector<uint32_t> topo;
        std::vector<uint2> chunks;
        if (runs) {
            topo.assign(5 * tn, 0u);
            for (uint64_t i = 0; i < ninst; ++i) {
                const pzk::InstanceRec& r = rec[i];
                topo[i] = r.nt ? r.dt : 0xFFFFFFFFu;  // as the scan leaves an object without triangles
                topo[tn + i] = r.nt ? r.dt + r.nt - 1 : 0u;
                topo[2 * tn + i] = r.nt;
                topo[3 * tn + i] = r.nt ? r.dv + r.nv - 1 : 0u;
                topo[4 * tn + i] = r.nt ? r.dv : 0xFFFFFFFFu;
                if (r.nt)
                    for (uint64_t v = r.dv; v < (uint64_t)r.dv + r.nv; v += K)
                        chunks.push_back(make_uint2((uint32_t)i, (uint32_t)v));
            }
        }
        pzk::InstanceRec* drec = nullptr;
        const bool owned = slot < 0;
        try {
            if (owned) {
                HIP_OK(hipMalloc((void**)&d->v, std::max<uint64_t>(12 * nv, 16)));
                HIP_OK(hipMalloc((void**)&d->idx, std::max<uint64_t>(12 * nt, 16)));
                HIP_OK(hipMalloc((void**)&d->t2o, std::max<uint64_t>(4 * nt, 16)));
                HIP_OK(hipMalloc((void**)&d->es, 12 * ne));
                HIP_OK(hipMalloc((void**)&d->ee, 12 * ne));
                HIP_OK(hipMalloc((void**)&d->eo, 4 * ne));
                HIP_OK(hipMalloc((void**)&d->part_of, part_of.size() * 4));
                HIP_OK(hipMalloc((void**)&d->part_off, off.size() * 4));
                HIP_OK(hipMalloc((void**)&drec, rec.size() * sizeof(pzk::InstanceRec)));
                if (runs) {
                    HIP_OK(hipMalloc((void**)&d->topo, 4 * topo.size()));
                    HIP_OK(hipMalloc((void**)&d->topo_chunks, 8 * std::max<size_t>(chunks.size(), 1)));
                }
            } else {  // the context's storage for this slot, grown on demand
                const std::string k = "inst" + std::to_string(slot) + ".";
                d->v = buf<float>(c, k + "v", 3 * nv);
                d->idx = buf<uint32_t>(c, k + "idx", 3 * nt);
                d->t2o = buf<int32_t>(c, k + "t2o", nt);
                d->es = buf<float>(c, k + "es", 3 * ne);
                d->ee = buf<float>(c, k + "ee", 3 * ne);
                d->eo = buf<int32_t>(c, k + "eo", ne);
                d->part_of = buf<uint32_t>(c, k + "part_of", part_of.size());
                d->part_off = buf<uint32_t>(c, k + "part_off", off.size());
                drec = buf<pzk::InstanceRec>(c, k + "rec", rec.size());
                if (runs) {
                    d->topo = buf<uint32_t>(c, k + "topo", topo.size());
                    d->topo_chunks = buf<uint2>(c, k + "topo_chunks", std::max<size_t>(chunks.size(), 1));
                }
                d->slot = slot;
                d->slot_gen = ++c->slot_gen[slot];
            }
            HIP_OK(hipMemcpy(d->part_of, part_of.data(), part_of.size() * 4, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(d->part_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(drec, rec.data(), rec.size() * sizeof(pzk::InstanceRec), hipMemcpyHostToDevice));
            if (runs) {
                HIP_OK(hipMemcpy(d->topo, topo.data(), 4 * topo.size(), hipMemcpyHostToDevice));
                if (!chunks.empty())
                    HIP_OK(hipMemcpy(d->topo_chunks, chunks.data(), 8 * chunks.size(), hipMemcpyHostToDevice));
                d->topo_nchunks = chunks.size();
                d->vruns = true;
            }
            pzk::instantiate(drec, (uint32_t)ninst, L->v, L->idx, L->es, L->ee, d->v, d->idx, d->t2o, d->es, d->ee,
                             d->eo, c->s);
            HIP_OK(hipStreamSynchronize(c->s));
            if (owned) HIP_OK(hipFree(drec));
            drec = nullptr;
        } catch (...) {
            if (drec && owned) (void)hipFree(drec);
            pz_dataset_release(d.release());
            throw;
        }
        *out = d.release();
        return PZ_OK;
    });
}

pz_status pz_dataset_download(const pz_dataset* d, pz_host_mesh** out) {
    if (!d || !out) return fail(PZ_E_INVALID, "pz_dataset_download: null argument");
    if (!dataset_live(d)) return fail(PZ_E_INVALID, "the dataset's instancing slot has been reused");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(d->ctx->device));
        HIP_OK(hipStreamSynchronize(d->ctx->s));
        std::unique_ptr<pz_host_mesh> m(new pz_host_mesh());
        m->vertices.resize(3 * d->nv);
        m->indices.resize(3 * d->nt);
        m->tri_to_object.resize(d->nt);
        m->edge_starts.resize(3 * d->ne);
        m->edge_ends.resize(3 * d->ne);
        m->edge_objects.resize(d->ne);
        auto down = [](void* dst, const void* src, size_t bytes) {
            if (bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        down(m->vertices.data(), d->v, 12 * d->nv);
        down(m->indices.data(), d->idx, 12 * d->nt);
        down(m->tri_to_object.data(), d->t2o, 4 * d->nt);
        down(m->edge_starts.data(), d->es, 12 * d->ne);
        down(m->edge_ends.data(), d->ee, 12 * d->ne);
        down(m->edge_objects.data(), d->eo, 4 * d->ne);
        m->grid = d->grid;
        m->grid_cell_size = d->cell_size;
        m->has_grid = !d->grid.empty();
        *out = m.release();
        return PZ_OK;
    });
}

}  // extern "C"
