// Wavefront OBJ reader with the object semantics of the reference's mesh
// loader (pierce/preprocess/src/ObjMeshDatasetLoader.cpp:14-168), which calls
// tinyobjloader (an unpinned conda dependency, not vendored) with
// triangulation on.  The statements that loader's output depends on:
//   v x y z [...]     vertex position (further values ignored)
//   f a b c [d ...]   face; each corner is v, v/vt, v/vt/vn or v//vn.
//                     A positive index counts from 1; a negative one counts
//                     back from the last vertex read so far; 0 is an error.
//                     Faces of more than three corners are split as a
//                     triangle fan (c0, ck, ck+1).
//   o name / g name   start a new shape; a shape without faces is dropped.
// Everything else (vt, vn, vp, usemtl, mtllib, s, l, p, comments) is skipped.
//
// Per shape, as the reference loader does: the shape's distinct vertex
// indices become the object's own vertices; vertices whose index is outside
// the vertex list are left out and the triangles that use them are dropped;
// the object id is the running count of shapes that have faces.  The
// reference orders an object's vertices by std::unordered_set iteration
// (ObjMeshDatasetLoader.cpp:77,111); here they are in ascending file order.
// Only the geometry is the same, so vertex and edge order may differ.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_mesh.hpp"

namespace {

struct Shape {
    std::vector<int64_t> corners;  // resolved 0-based vertex index per triangle corner (3 per triangle)
};

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Parses one face corner's vertex index; returns false on a malformed token.
bool corner_index(const char*& s, int64_t nv, int64_t& out) {
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s, &end, 10);
    if (end == s || errno != 0 || v == 0) return false;
    out = v > 0 ? v - 1 : nv + v;
    s = end;
    while (*s && !is_space(*s)) ++s;  // skip /vt/vn
    return true;
}

}  // namespace

extern "C" pz_status pz_obj_read(const char* path, pz_host_mesh** out) {
    if (!path || !out) return pz::fail(PZ_E_INVALID, "pz_obj_read: null argument");
    *out = nullptr;
    const size_t len = std::strlen(path);
    if (len < 4 || std::strcmp(path + len - 4, ".obj") != 0)
        return pz::fail(PZ_E_IO, std::string("not a .obj file: ") + path);
    std::ifstream in(path);
    if (!in) return pz::fail(PZ_E_IO, std::string("cannot open ") + path);

    std::vector<float> pos;  // 3 per vertex
    std::vector<Shape> shapes(1);
    std::string line;
    uint64_t lineno = 0;
    std::vector<int64_t> face;
    while (std::getline(in, line)) {
        ++lineno;
        const char* s = line.c_str();
        while (is_space(*s)) ++s;
        if (*s == '\0' || *s == '#') continue;
        if (s[0] == 'v' && is_space(s[1])) {
            float xyz[3] = {0.f, 0.f, 0.f};
            const char* p = s + 2;
            for (int k = 0; k < 3; ++k) {
                char* end = nullptr;
                xyz[k] = std::strtof(p, &end);
                if (end == p) return pz::fail(PZ_E_IO, "line " + std::to_string(lineno) + ": bad vertex");
                p = end;
            }
            pos.insert(pos.end(), xyz, xyz + 3);
        } else if (s[0] == 'f' && is_space(s[1])) {
            face.clear();
            const char* p = s + 2;
            const int64_t nv = (int64_t)(pos.size() / 3);
            while (true) {
                while (is_space(*p)) ++p;
                if (*p == '\0' || *p == '#') break;
                int64_t vi;
                if (!corner_index(p, nv, vi))
                    return pz::fail(PZ_E_IO, "line " + std::to_string(lineno) + ": bad face index");
                face.push_back(vi);
            }
            if (face.size() < 3) continue;  // not a polygon
            auto& c = shapes.back().corners;
            for (size_t k = 1; k + 1 < face.size(); ++k) {
                c.push_back(face[0]);
                c.push_back(face[k]);
                c.push_back(face[k + 1]);
            }
        } else if ((s[0] == 'o' || s[0] == 'g') && (is_space(s[1]) || s[1] == '\0')) {
            if (!shapes.back().corners.empty()) shapes.emplace_back();
        }
    }

    auto* m = new pz_host_mesh();
    const int64_t nv = (int64_t)(pos.size() / 3);
    int32_t object = 0;
    std::vector<int64_t> used;
    std::unordered_map<int64_t, uint32_t> local;
    for (const Shape& sh : shapes) {
        if (sh.corners.empty()) continue;
        used.assign(sh.corners.begin(), sh.corners.end());
        std::sort(used.begin(), used.end());
        used.erase(std::unique(used.begin(), used.end()), used.end());
        local.clear();
        for (int64_t g : used) {
            if (g < 0 || g >= nv) continue;
            local[g] = (uint32_t)m->num_vertices();
            m->vertices.insert(m->vertices.end(), pos.begin() + 3 * g, pos.begin() + 3 * g + 3);
        }
        for (size_t t = 0; t + 2 < sh.corners.size(); t += 3) {
            auto a = local.find(sh.corners[t]), b = local.find(sh.corners[t + 1]), c = local.find(sh.corners[t + 2]);
            if (a == local.end() || b == local.end() || c == local.end()) continue;
            m->indices.push_back(a->second);
            m->indices.push_back(b->second);
            m->indices.push_back(c->second);
            m->tri_to_object.push_back(object);
        }
        ++object;
    }
    if (m->vertices.empty()) {
        delete m;
        return pz::fail(PZ_E_IO, std::string("no triangle geometry in ") + path);
    }
    *out = m;
    return PZ_OK;
}

extern "C" pz_status pz_host_mesh_shuffle(pz_host_mesh* mesh, uint64_t seed, float lo, float hi) {
    if (!mesh) return pz::fail(PZ_E_INVALID, "pz_host_mesh_shuffle: null mesh");
    if (!(lo <= hi)) return pz::fail(PZ_E_INVALID, "pz_host_mesh_shuffle: lo > hi");
    // One offset per object, drawn in ascending object id; every vertex an
    // object's triangles use moves once (a vertex shared by two objects
    // moves once per object, as in the reference).
    const uint64_t nt = mesh->num_triangles(), nvtx = mesh->num_vertices();
    int32_t nobj = 0;
    for (int32_t o : mesh->tri_to_object) nobj = std::max(nobj, o + 1);
    std::vector<std::vector<uint32_t>> verts_of(nobj);
    for (uint64_t t = 0; t < nt; ++t) {
        const int32_t o = mesh->tri_to_object[t];
        if (o < 0) continue;
        for (int k = 0; k < 3; ++k) verts_of[o].push_back(mesh->indices[3 * t + k]);
    }
    std::mt19937 rng((uint32_t)(seed ^ (seed >> 32)));
    std::uniform_real_distribution<float> dist(lo, hi);
    for (int32_t o = 0; o < nobj; ++o) {
        auto& vs = verts_of[o];
        if (vs.empty()) continue;
        std::sort(vs.begin(), vs.end());
        vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
        const float d[3] = {dist(rng), dist(rng), dist(rng)};
        for (uint32_t v : vs)
            if (v < nvtx)
                for (int k = 0; k < 3; ++k) mesh->vertices[3 * (uint64_t)v + k] += d[k];
    }
    return PZ_OK;
}
