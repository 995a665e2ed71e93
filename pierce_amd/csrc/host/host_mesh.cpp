// .pre v4 reader/writer, edge extraction and grid statistics.
//
// File layout (pierce/common/include/BinaryIO.h:14-82): a 56-byte header
// {u32 magic 0x52334442, u32 version 4, u64 numVertices, numIndices,
// numMappings, numEdges, totalTriangles, u8 hasGrid, u8 hasEdges, u8 pad[6]},
// then float3[numVertices], uint3[numIndices], int32[numMappings], and when
// hasEdges: float3 starts[numEdges], float3 ends[numEdges], int32 objs[numEdges];
// when hasGrid: {f32 cellSize, u32 numCells} + numCells * 28-byte cells.
#include "host_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>

namespace {
thread_local std::string g_last_error;

constexpr uint32_t kMagic = 0x52334442u;  // "R3DB"
constexpr uint32_t kVersion = 4;

#pragma pack(push, 1)
struct FileHeader {
    uint32_t magic;
    uint32_t version;
    uint64_t numVertices;
    uint64_t numIndices;
    uint64_t numMappings;
    uint64_t numEdges;
    uint64_t totalTriangles;
    uint8_t hasGrid;
    uint8_t hasEdges;
    uint8_t padding[6];
};
#pragma pack(pop)
static_assert(sizeof(FileHeader) == 56, "header must be 56 bytes");
static_assert(sizeof(pz_grid_cell) == 28, "grid cell must be 28 bytes");
}  // namespace

void pz::set_error(const std::string& msg) { g_last_error = msg; }
pz_status pz::fail(pz_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

extern "C" const char* pz_last_error(void) { return g_last_error.c_str(); }

pz_mesh_view pz_host_mesh::view() const {
    pz_mesh_view v;
    std::memset(&v, 0, sizeof(v));
    v.vertices = vertices.data();
    v.num_vertices = num_vertices();
    v.indices = indices.data();
    v.num_triangles = num_triangles();
    v.triangle_to_object = tri_to_object.data();
    v.edge_starts = edge_starts.empty() ? nullptr : edge_starts.data();
    v.edge_ends = edge_ends.empty() ? nullptr : edge_ends.data();
    v.edge_objects = edge_objects.empty() ? nullptr : edge_objects.data();
    v.num_edges = num_edges();
    v.grid_cells = grid.empty() ? nullptr : grid.data();
    v.num_grid_cells = has_grid ? grid.size() : 0;
    v.grid_cell_size = has_grid ? grid_cell_size : 0.0f;
    return v;
}

extern "C" pz_status pz_host_mesh_view(const pz_host_mesh* mesh, pz_mesh_view* out) {
    if (!mesh || !out) return pz::fail(PZ_E_INVALID, "pz_host_mesh_view: null argument");
    *out = mesh->view();
    return PZ_OK;
}

extern "C" void pz_host_mesh_free(pz_host_mesh* mesh) { delete mesh; }

extern "C" pz_status pz_host_mesh_from_view(const pz_mesh_view* v, pz_host_mesh** out) {
    if (!v || !out) return pz::fail(PZ_E_INVALID, "pz_host_mesh_from_view: null argument");
    auto* m = new pz_host_mesh();
    m->vertices.assign(v->vertices, v->vertices + 3 * v->num_vertices);
    m->indices.assign(v->indices, v->indices + 3 * v->num_triangles);
    m->tri_to_object.assign(v->triangle_to_object, v->triangle_to_object + v->num_triangles);
    if (v->num_edges) {
        m->edge_starts.assign(v->edge_starts, v->edge_starts + 3 * v->num_edges);
        m->edge_ends.assign(v->edge_ends, v->edge_ends + 3 * v->num_edges);
        m->edge_objects.assign(v->edge_objects, v->edge_objects + v->num_edges);
    }
    if (v->grid_cells && v->num_grid_cells) {
        m->grid.assign(v->grid_cells, v->grid_cells + v->num_grid_cells);
        m->has_grid = true;
        m->grid_cell_size = v->grid_cell_size;
    } else if (v->grid_cell_size > 0.0f) {
        m->has_grid = true;
        m->grid_cell_size = v->grid_cell_size;
    }
    *out = m;
    return PZ_OK;
}

template <class T>
static bool read_array(std::ifstream& in, std::vector<T>& dst, uint64_t count) {
    dst.resize(count);
    if (count == 0) return true;
    in.read(reinterpret_cast<char*>(dst.data()), static_cast<std::streamsize>(count * sizeof(T)));
    return static_cast<uint64_t>(in.gcount()) == count * sizeof(T);
}

extern "C" pz_status pz_pre_read(const char* path, pz_host_mesh** out) {
    if (!path || !out) return pz::fail(PZ_E_INVALID, "pz_pre_read: null argument");
    std::ifstream in(path, std::ios::binary);
    if (!in) return pz::fail(PZ_E_IO, std::string("Could not open file for reading: ") + path);
    FileHeader h;
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (in.gcount() != sizeof(h)) return pz::fail(PZ_E_IO, std::string("Truncated header: ") + path);
    if (h.magic != kMagic)
        return pz::fail(PZ_E_IO, std::string("Invalid file format (Magic mismatch). File: ") + path);
    if (h.version != kVersion)
        return pz::fail(PZ_E_IO, "Unsupported binary geometry version " + std::to_string(h.version) +
                                     " in file: " + path + ". Expected version 4. Re-run preprocessing.");
    if (h.numMappings != h.numIndices)
        return pz::fail(PZ_E_IO, std::string("triangle->object map size differs from triangle count: ") + path);
    auto* m = new pz_host_mesh();
    bool ok = read_array(in, m->vertices, 3 * h.numVertices) && read_array(in, m->indices, 3 * h.numIndices) &&
              read_array(in, m->tri_to_object, h.numMappings);
    if (ok && h.hasEdges) {
        ok = read_array(in, m->edge_starts, 3 * h.numEdges) && read_array(in, m->edge_ends, 3 * h.numEdges) &&
             read_array(in, m->edge_objects, h.numEdges);
    }
    if (ok && h.hasGrid) {
        float cs = 0.f;
        uint32_t n = 0;
        in.read(reinterpret_cast<char*>(&cs), 4);
        in.read(reinterpret_cast<char*>(&n), 4);
        ok = in.good() && read_array(in, m->grid, n);
        m->has_grid = true;
        m->grid_cell_size = cs;
    }
    if (!ok) {
        delete m;
        return pz::fail(PZ_E_IO, std::string("Truncated payload: ") + path);
    }
    *out = m;
    return PZ_OK;
}

extern "C" pz_status pz_pre_write(const char* path, const pz_mesh_view* v) {
    if (!path || !v) return pz::fail(PZ_E_INVALID, "pz_pre_write: null argument");
    std::ofstream out(path, std::ios::binary);
    if (!out) return pz::fail(PZ_E_IO, std::string("Could not open file for writing: ") + path);
    FileHeader h;
    std::memset(&h, 0, sizeof(h));
    h.magic = kMagic;
    h.version = kVersion;
    h.numVertices = v->num_vertices;
    h.numIndices = v->num_triangles;
    h.numMappings = v->num_triangles;
    h.numEdges = v->num_edges;
    h.totalTriangles = v->num_triangles;
    // EdgeData::hasEdges() (Geometry.h) is false for an empty edge list.
    h.hasEdges = v->num_edges > 0 ? 1 : 0;
    h.hasGrid = v->grid_cell_size > 0.0f ? 1 : 0;
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    out.write(reinterpret_cast<const char*>(v->vertices), 12 * v->num_vertices);
    out.write(reinterpret_cast<const char*>(v->indices), 12 * v->num_triangles);
    out.write(reinterpret_cast<const char*>(v->triangle_to_object), 4 * v->num_triangles);
    if (h.hasEdges) {
        out.write(reinterpret_cast<const char*>(v->edge_starts), 12 * v->num_edges);
        out.write(reinterpret_cast<const char*>(v->edge_ends), 12 * v->num_edges);
        out.write(reinterpret_cast<const char*>(v->edge_objects), 4 * v->num_edges);
    }
    if (h.hasGrid) {
        float cs = v->grid_cell_size;
        uint32_t n = static_cast<uint32_t>(v->num_grid_cells);
        out.write(reinterpret_cast<const char*>(&cs), 4);
        out.write(reinterpret_cast<const char*>(&n), 4);
        if (n) out.write(reinterpret_cast<const char*>(v->grid_cells), sizeof(pz_grid_cell) * n);
    }
    if (!out) return pz::fail(PZ_E_IO, std::string("Write failed: ") + path);
    return PZ_OK;
}

// ---------------------------------------------------------------- edges
namespace {
struct QPoint {
    int64_t x, y, z;
};
inline bool qless(const QPoint& a, const QPoint& b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    return a.z < b.z;
}
struct EdgeRec {
    QPoint a, b;
    int32_t obj;
    uint64_t order;  // first-seen position (tri * 3 + k)
};
inline bool edge_key_less(const EdgeRec& l, const EdgeRec& r) {
    if (qless(l.a, r.a)) return true;
    if (qless(r.a, l.a)) return false;
    if (qless(l.b, r.b)) return true;
    if (qless(r.b, l.b)) return false;
    return l.obj < r.obj;
}
inline bool edge_key_eq(const EdgeRec& l, const EdgeRec& r) {
    return !edge_key_less(l, r) && !edge_key_less(r, l);
}
inline QPoint quantize(const float* p) {
    const double s = 1e5;
    return {static_cast<int64_t>(std::llround(static_cast<double>(p[0]) * s)),
            static_cast<int64_t>(std::llround(static_cast<double>(p[1]) * s)),
            static_cast<int64_t>(std::llround(static_cast<double>(p[2]) * s))};
}
}  // namespace

// Every triangle contributes its edges (v0,v1), (v1,v2), (v2,v0); the key is
// the lexicographically ordered pair of 1e-5-quantised endpoints plus the
// object id; the first occurrence keeps its float endpoints and orientation;
// output follows ascending key order (the std::map iteration order of
// EdgePreprocessor.cpp:62,103-108).  A stable sort on the key followed by
// keeping the first element of each equal run reproduces that exactly.
void pz::extract_edges_pre_order(const pz_mesh_view& v, std::vector<float>& starts, std::vector<float>& ends,
                                 std::vector<int32_t>& objs) {
    starts.clear();
    ends.clear();
    objs.clear();
    const uint64_t nT = v.num_triangles, nV = v.num_vertices;
    if (nT == 0 || nV == 0) return;
    std::vector<EdgeRec> recs;
    recs.reserve(nT * 3);
    for (uint64_t t = 0; t < nT; ++t) {
        const uint32_t* tri = v.indices + 3 * t;
        if (tri[0] >= nV || tri[1] >= nV || tri[2] >= nV) continue;
        for (int k = 0; k < 3; ++k) {
            uint32_t i0 = tri[k], i1 = tri[(k + 1) % 3];
            QPoint q0 = quantize(v.vertices + 3 * i0), q1 = quantize(v.vertices + 3 * i1);
            if (qless(q1, q0)) std::swap(q0, q1);
            recs.push_back({q0, q1, v.triangle_to_object[t], t * 3 + k});
        }
    }
    std::stable_sort(recs.begin(), recs.end(), edge_key_less);
    for (size_t i = 0; i < recs.size();) {
        size_t j = i + 1;
        while (j < recs.size() && edge_key_eq(recs[i], recs[j])) ++j;
        const EdgeRec& r = recs[i];  // first seen (stable sort keeps input order)
        uint64_t t = r.order / 3;
        int k = static_cast<int>(r.order % 3);
        const uint32_t* tri = v.indices + 3 * t;
        const float* p0 = v.vertices + 3 * tri[k];
        const float* p1 = v.vertices + 3 * tri[(k + 1) % 3];
        starts.insert(starts.end(), p0, p0 + 3);
        ends.insert(ends.end(), p1, p1 + 3);
        objs.push_back(r.obj);
        i = j;
    }
}

extern "C" pz_status pz_host_mesh_extract_edges(pz_host_mesh* mesh) {
    if (!mesh) return pz::fail(PZ_E_INVALID, "pz_host_mesh_extract_edges: null mesh");
    if (mesh->tri_to_object.size() != mesh->num_triangles())
        return pz::fail(PZ_E_INVALID, "triangle->object map size differs from triangle count");
    pz_mesh_view v = mesh->view();
    std::vector<float> s, e;
    std::vector<int32_t> o;
    pz::extract_edges_pre_order(v, s, e, o);
    mesh->edge_starts.swap(s);
    mesh->edge_ends.swap(e);
    mesh->edge_objects.swap(o);
    return PZ_OK;
}

// ----------------------------------------------------------------- grid
// Per object: fp32 AABB over its triangles' vertices; signed volume summed in
// double over triangles relative to the fp32 AABB centre (pierce_preprocess.cpp:77-101);
// ratio = min(1, |vol|/6 / max(aabbVol, 1e-9)); avgSize = (w+h+d)/3; the centre
// cell gets CenterCount++, every cell in [floor(min/cs), floor((max-1e-5)/cs)]
// gets TouchCount++ and sums of avgSize/ratio which are then averaged.
void pz::compute_grid(const pz_mesh_view& v, float cs, std::vector<pz_grid_cell>& out) {
    struct Stats {
        float mn[3] = {1e30f, 1e30f, 1e30f};
        float mx[3] = {-1e30f, -1e30f, -1e30f};
        double vol = 0.0;
        bool seen = false;
    };
    int32_t maxObj = -1;
    for (uint64_t t = 0; t < v.num_triangles; ++t) maxObj = std::max(maxObj, v.triangle_to_object[t]);
    std::vector<Stats> st(static_cast<size_t>(maxObj + 1));
    const uint64_t nV = v.num_vertices;
    auto valid = [&](const uint32_t* tri) { return tri[0] < nV && tri[1] < nV && tri[2] < nV; };
    for (uint64_t t = 0; t < v.num_triangles; ++t) {
        const uint32_t* tri = v.indices + 3 * t;
        int32_t o = v.triangle_to_object[t];
        if (o < 0 || !valid(tri)) continue;
        Stats& s = st[o];
        s.seen = true;
        for (int k = 0; k < 3; ++k) {
            const float* p = v.vertices + 3 * tri[k];
            for (int a = 0; a < 3; ++a) {
                s.mn[a] = std::min(s.mn[a], p[a]);
                s.mx[a] = std::max(s.mx[a], p[a]);
            }
        }
    }
    for (uint64_t t = 0; t < v.num_triangles; ++t) {
        const uint32_t* tri = v.indices + 3 * t;
        int32_t o = v.triangle_to_object[t];
        if (o < 0 || !valid(tri)) continue;
        Stats& s = st[o];
        float c[3];
        for (int a = 0; a < 3; ++a) c[a] = (s.mn[a] + s.mx[a]) * 0.5f;
        const float* p0 = v.vertices + 3 * tri[0];
        const float* p1 = v.vertices + 3 * tri[1];
        const float* p2 = v.vertices + 3 * tri[2];
        float t0[3], t1[3], t2[3];
        for (int a = 0; a < 3; ++a) {
            t0[a] = p0[a] - c[a];
            t1[a] = p1[a] - c[a];
            t2[a] = p2[a] - c[a];
        }
        float cx = t0[1] * t1[2] - t0[2] * t1[1];
        float cy = t0[2] * t1[0] - t0[0] * t1[2];
        float cz = t0[0] * t1[1] - t0[1] * t1[0];
        s.vol += (cx * t2[0] + cy * t2[1] + cz * t2[2]);
    }
    struct Key {
        int32_t x, y, z;
        bool operator<(const Key& o) const {
            if (z != o.z) return z < o.z;
            if (y != o.y) return y < o.y;
            return x < o.x;
        }
    };
    struct Acc {
        uint32_t center = 0, touch = 0;
        float size = 0.f, ratio = 0.f;
    };
    std::map<Key, Acc> cells;
    for (size_t o = 0; o < st.size(); ++o) {
        const Stats& s = st[o];
        if (!s.seen) continue;
        float meshVol = std::fabs(static_cast<float>(s.vol)) / 6.0f;
        float w = s.mx[0] - s.mn[0], h = s.mx[1] - s.mn[1], d = s.mx[2] - s.mn[2];
        float aabbVol = w * h * d;
        if (aabbVol < 1e-9f) aabbVol = 1e-9f;
        float ratio = meshVol / aabbVol;
        if (ratio > 1.0f) ratio = 1.0f;
        float ctr[3];
        for (int a = 0; a < 3; ++a) ctr[a] = (s.mn[a] + s.mx[a]) * 0.5f;
        float avg = (w + h + d) / 3.0f;
        Key ck{(int32_t)std::floor(ctr[0] / cs), (int32_t)std::floor(ctr[1] / cs), (int32_t)std::floor(ctr[2] / cs)};
        cells[ck].center++;
        int32_t lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = (int32_t)std::floor(s.mn[a] / cs);
            hi[a] = (int32_t)std::floor((s.mx[a] - 1e-5f) / cs);
        }
        for (int32_t z = lo[2]; z <= hi[2]; ++z)
            for (int32_t y = lo[1]; y <= hi[1]; ++y)
                for (int32_t x = lo[0]; x <= hi[0]; ++x) {
                    Acc& a = cells[Key{x, y, z}];
                    a.touch++;
                    a.size += avg;
                    a.ratio += ratio;
                }
    }
    out.clear();
    out.reserve(cells.size());
    for (auto& kv : cells) {
        Acc a = kv.second;
        if (a.touch > 0) {
            a.size /= a.touch;
            a.ratio /= a.touch;
        }
        out.push_back({kv.first.x, kv.first.y, kv.first.z, a.center, a.touch, a.size, a.ratio});
    }
}

extern "C" pz_status pz_host_mesh_compute_grid(pz_host_mesh* mesh, float cell_size) {
    if (!mesh) return pz::fail(PZ_E_INVALID, "pz_host_mesh_compute_grid: null mesh");
    if (!(cell_size > 0.0f)) return pz::fail(PZ_E_INVALID, "grid cell size must be > 0");
    pz::compute_grid(mesh->view(), cell_size, mesh->grid);
    mesh->has_grid = true;
    mesh->grid_cell_size = cell_size;
    return PZ_OK;
}
