// Deterministic synthetic dataset generators.
//
// Every mesh draws from its own counter-seeded xoshiro256** stream
// (seed, mesh index), so output does not depend on thread count or host.
// Meshes are closed, consistently oriented 2-manifolds; object id = mesh
// index; vertices and triangles are contiguous per mesh.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>

#include "host_mesh.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
        for (auto& w : s) w = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {  // Box-Muller, one value per call
        double u1 = 1.0 - uniform();  // (0,1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

struct Template {
    std::vector<std::array<double, 3>> v;
    std::vector<std::array<uint32_t, 3>> f;
    std::vector<std::array<uint32_t, 2>> e;  // unique undirected edges, first-seen orientation
};

void build_edges(Template& t) {
    std::map<std::pair<uint32_t, uint32_t>, int> seen;
    t.e.clear();
    for (auto& f : t.f)
        for (int k = 0; k < 3; ++k) {
            uint32_t a = f[k], b = f[(k + 1) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            if (seen.emplace(key, 1).second) t.e.push_back({a, b});
        }
}

// Regular icosahedron subdivided `level` times (shared midpoints), projected to
// the unit sphere.  Faces are oriented outward.
Template make_icosphere(int level) {
    Template t;
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                                {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                   {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                   {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                   {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    auto push = [&](double x, double y, double z) {
        double n = std::sqrt(x * x + y * y + z * z);
        t.v.push_back({x / n, y / n, z / n});
        return static_cast<uint32_t>(t.v.size() - 1);
    };
    for (auto& b : base) push(b[0], b[1], b[2]);
    for (auto& f : faces) t.f.push_back({f[0], f[1], f[2]});
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            auto& A = t.v[a];
            auto& B = t.v[b];
            uint32_t id = push((A[0] + B[0]) * 0.5, (A[1] + B[1]) * 0.5, (A[2] + B[2]) * 0.5);
            mid.emplace(key, id);
            return id;
        };
        std::vector<std::array<uint32_t, 3>> nf;
        nf.reserve(t.f.size() * 4);
        for (auto& f : t.f) {
            uint32_t ab = midpoint(f[0], f[1]), bc = midpoint(f[1], f[2]), ca = midpoint(f[2], f[0]);
            nf.push_back({f[0], ab, ca});
            nf.push_back({f[1], bc, ab});
            nf.push_back({f[2], ca, bc});
            nf.push_back({ab, bc, ca});
        }
        t.f.swap(nf);
    }
    // Orient every face outward (the template is star-shaped about the origin).
    for (auto& f : t.f) {
        auto& a = t.v[f[0]];
        auto& b = t.v[f[1]];
        auto& c = t.v[f[2]];
        double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        double d = n[0] * (a[0] + b[0] + c[0]) + n[1] * (a[1] + b[1] + c[1]) + n[2] * (a[2] + b[2] + c[2]);
        if (d < 0) std::swap(f[1], f[2]);
    }
    build_edges(t);
    return t;
}

// Cube faces of pierce/scripts/generate_test_cubes.py:19-33 (vertex order
// :5-16), unit half-size about the origin.
Template make_cube() {
    Template t;
    const double v[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                            {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    const uint32_t f[12][3] = {{0, 2, 1}, {0, 3, 2}, {4, 5, 6}, {4, 6, 7}, {0, 1, 5}, {0, 5, 4},
                               {1, 2, 6}, {1, 6, 5}, {2, 3, 7}, {2, 7, 6}, {3, 0, 4}, {3, 4, 7}};
    for (auto& p : v) t.v.push_back({p[0], p[1], p[2]});
    for (auto& q : f) t.f.push_back({q[0], q[1], q[2]});
    build_edges(t);
    return t;
}

const Template& icosphere(int level) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<Template>> cache;
    std::lock_guard<std::mutex> g(mu);
    auto& slot = cache[level];
    if (!slot) slot.reset(new Template(make_icosphere(level)));
    return *slot;
}

struct Mat3 {
    double m[3][3];
};

// Uniform random rotation (Shoemake 1992, "Uniform random rotations").
Mat3 random_rotation(Rng& r) {
    double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    double x = a * std::sin(2 * kPi * u2), y = a * std::cos(2 * kPi * u2);
    double z = b * std::sin(2 * kPi * u3), w = b * std::cos(2 * kPi * u3);
    Mat3 R;
    R.m[0][0] = 1 - 2 * (y * y + z * z);
    R.m[0][1] = 2 * (x * y - z * w);
    R.m[0][2] = 2 * (x * z + y * w);
    R.m[1][0] = 2 * (x * y + z * w);
    R.m[1][1] = 1 - 2 * (x * x + z * z);
    R.m[1][2] = 2 * (y * z - x * w);
    R.m[2][0] = 2 * (x * z - y * w);
    R.m[2][1] = 2 * (y * z + x * w);
    R.m[2][2] = 1 - 2 * (x * x + y * y);
    return R;
}

// Plan of one mesh: which template and where its arrays start.
struct MeshPlan {
    const Template* tmpl;
    uint64_t v_off, t_off, e_off;
};

// Fresh ellipsoid vertices: radii U[rmin,rmax]^3, uniform rotation, centre
// U[0,box)^3, optional N(0, jitter * mean radius) per coordinate.
void ellipsoid_vertices(const Template& T, Rng& r, const pz_gen_params& p, float* out) {
    double rad[3];
    for (double& x : rad) x = r.uniform(p.radius_min, p.radius_max);
    Mat3 R = random_rotation(r);
    double c[3];
    for (double& x : c) x = r.uniform(0.0, p.box);
    double sigma = p.jitter * (rad[0] + rad[1] + rad[2]) / 3.0;
    for (size_t i = 0; i < T.v.size(); ++i) {
        double s[3] = {T.v[i][0] * rad[0], T.v[i][1] * rad[1], T.v[i][2] * rad[2]};
        for (int a = 0; a < 3; ++a) {
            double x = R.m[a][0] * s[0] + R.m[a][1] * s[1] + R.m[a][2] * s[2] + c[a];
            if (sigma > 0) x += sigma * r.normal();
            out[3 * i + a] = static_cast<float>(x);
        }
    }
}

void sphere_vertices(const Template& T, Rng& r, const pz_gen_params& p, float* out) {
    double rad = r.uniform(p.radius_min, p.radius_max);
    double c[3];
    for (double& x : c) x = r.uniform(0.0, p.box);
    for (size_t i = 0; i < T.v.size(); ++i)
        for (int a = 0; a < 3; ++a) out[3 * i + a] = static_cast<float>(T.v[i][a] * rad + c[a]);
}

void cube_vertices(const Template& T, Rng& r, const pz_gen_params& p, float* out) {
    double edge = r.uniform(p.radius_min, p.radius_max);
    double c[3];
    for (double& x : c) x = r.uniform(0.0, p.box);
    for (size_t i = 0; i < T.v.size(); ++i)
        for (int a = 0; a < 3; ++a) out[3 * i + a] = static_cast<float>(c[a] + T.v[i][a] * edge * 0.5);
}

int mixed_level(Rng& r, int lmin, int lmax) {
    double total = 0;
    for (int l = lmin; l <= lmax; ++l) total += std::ldexp(1.0, -l);
    double u = r.uniform() * total, acc = 0;
    for (int l = lmin; l <= lmax; ++l) {
        acc += std::ldexp(1.0, -l);
        if (u < acc) return l;
    }
    return lmax;
}

// Level uniform in [lmin, lmax].
int uniform_level(Rng& r, int lmin, int lmax) {
    const int l = lmin + (int)(r.uniform() * (double)(lmax - lmin + 1));
    return std::min(l, lmax);
}

}  // namespace

extern "C" void pz_gen_defaults(int32_t kind, pz_gen_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->edges_mode = PZ_EDGES_GROUPED;
    p->count = 1000;
    p->seed = 42;
    p->level_min = p->level_max = 2;
    p->radius_min = 1.0f;
    p->radius_max = 3.0f;
    p->box = 100.0f;
    p->ring_segments = 16;
    p->step = 2.0f;
    p->branch_prob = 0.05f;
    p->max_branches = 6;
    p->tri_median = 200000;
    p->tri_sigma = 0.5f;
    p->tri_min = 50000;
    p->tri_max = 2000000;
    if (kind == PZ_GEN_SPHERES_MIXED) {
        p->level_min = 1;
        p->level_max = 6;
        p->radius_max = 4.0f;
        p->box = 500.0f;
    } else if (kind == PZ_GEN_CUBES) {
        p->radius_max = 10.0f;
    } else if (kind == PZ_GEN_QUERY_MIX) {
        p->level_min = 1;
        p->level_max = 4;
        p->tri_median = 15000;
        p->tri_sigma = 0.6f;
        p->tri_min = 5000;
        p->tri_max = 50000;
    }
}

static pz_status generate_impl(const pz_gen_params& p, pz_host_mesh& m) {
    const uint64_t n = p.count;
    if (p.level_min < 0 || p.level_max < p.level_min || p.level_max > 8)
        return pz::fail(PZ_E_INVALID, "subdivision levels must satisfy 0 <= min <= max <= 8");
    if (!(p.radius_max >= p.radius_min) || !(p.radius_min > 0) || !(p.box >= 0))
        return pz::fail(PZ_E_INVALID, "invalid radius or box");
    if (p.copy_fraction > 0 && p.kind != PZ_GEN_ELLIPSOIDS)
        return pz::fail(PZ_E_INVALID, "copies are supported for ellipsoid sets only");
    if (p.copy_fraction > 0 && p.level_max != p.level_min)
        return pz::fail(PZ_E_INVALID, "copies need a single subdivision level");

    // Per-mesh template choice (the first draws of each mesh's stream).
    std::vector<MeshPlan> plan(n);
    std::vector<int> level(n, p.level_min);
    const Template* cube = nullptr;
    static const Template cube_t = make_cube();
    if (p.kind == PZ_GEN_CUBES) cube = &cube_t;
    uint64_t nv = 0, nt = 0, ne = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const Template* T;
        if (p.kind == PZ_GEN_CUBES) {
            T = cube;
        } else if (p.kind == PZ_GEN_SPHERES_MIXED) {
            Rng r(p.seed, i);
            level[i] = mixed_level(r, p.level_min, p.level_max);
            T = &icosphere(level[i]);
        } else if (p.level_max > p.level_min) {  // ellipsoids, level uniform in [min, max]
            Rng r(p.seed, i);
            level[i] = uniform_level(r, p.level_min, p.level_max);
            T = &icosphere(level[i]);
        } else {
            T = &icosphere(p.level_min);
        }
        plan[i] = {T, nv, nt, ne};
        nv += T->v.size();
        nt += T->f.size();
        ne += T->e.size();
    }
    if (nv >= (1ull << 32)) return pz::fail(PZ_E_LIMIT, "more than 2^32 vertices in one set");
    m.vertices.assign(3 * nv, 0.f);
    m.indices.assign(3 * nt, 0u);
    m.tri_to_object.assign(nt, 0);
    const bool grouped = p.edges_mode == PZ_EDGES_GROUPED;
    if (grouped) {
        m.edge_starts.assign(3 * ne, 0.f);
        m.edge_ends.assign(3 * ne, 0.f);
        m.edge_objects.assign(ne, 0);
    }

    // Source set of the scaled copies (same shape parameters, its own seed).
    pz_gen_params src = p;
    src.copy_fraction = 0;
    src.seed = p.copy_source_seed;
    src.count = p.copy_source_count;
    const Template& srcT = icosphere(p.level_min);

    pz::parallel_for(n, [&](uint64_t i) {
        const MeshPlan& mp = plan[i];
        const Template& T = *mp.tmpl;
        float* V = m.vertices.data() + 3 * mp.v_off;
        Rng r(p.seed, i);
        if (p.kind == PZ_GEN_SPHERES_MIXED) {
            (void)mixed_level(r, p.level_min, p.level_max);
            sphere_vertices(T, r, p, V);
        } else if (p.kind == PZ_GEN_CUBES) {
            cube_vertices(T, r, p, V);
        } else {
            if (p.level_max > p.level_min) (void)uniform_level(r, p.level_min, p.level_max);
            bool copied = false;
            if (p.copy_fraction > 0) {
                double u = r.uniform();
                if (u < p.copy_fraction && p.copy_source_count > 0) {
                    uint64_t s = static_cast<uint64_t>(r.uniform() * static_cast<double>(p.copy_source_count));
                    if (s >= p.copy_source_count) s = p.copy_source_count - 1;
                    double k = r.uniform(p.copy_scale_min, p.copy_scale_max);
                    Rng rs(src.seed, s);
                    ellipsoid_vertices(srcT, rs, src, V);
                    double c[3] = {0, 0, 0};
                    const size_t nvm = T.v.size();
                    for (size_t j = 0; j < nvm; ++j)
                        for (int a = 0; a < 3; ++a) c[a] += V[3 * j + a];
                    for (double& x : c) x /= static_cast<double>(nvm);
                    for (size_t j = 0; j < nvm; ++j)
                        for (int a = 0; a < 3; ++a)
                            V[3 * j + a] = static_cast<float>(c[a] + k * (static_cast<double>(V[3 * j + a]) - c[a]));
                    copied = true;
                }
            }
            if (!copied) ellipsoid_vertices(T, r, p, V);
        }
        uint32_t* I = m.indices.data() + 3 * mp.t_off;
        for (size_t f = 0; f < T.f.size(); ++f) {
            for (int k = 0; k < 3; ++k) I[3 * f + k] = static_cast<uint32_t>(mp.v_off + T.f[f][k]);
            m.tri_to_object[mp.t_off + f] = static_cast<int32_t>(i);
        }
        if (grouped) {
            for (size_t e = 0; e < T.e.size(); ++e) {
                const float* a = V + 3 * T.e[e][0];
                const float* b = V + 3 * T.e[e][1];
                std::copy(a, a + 3, m.edge_starts.data() + 3 * (mp.e_off + e));
                std::copy(b, b + 3, m.edge_ends.data() + 3 * (mp.e_off + e));
                m.edge_objects[mp.e_off + e] = static_cast<int32_t>(i);
            }
        }
    });
    if (p.edges_mode == PZ_EDGES_PRE) pz::extract_edges_pre_order(m.view(), m.edge_starts, m.edge_ends, m.edge_objects);
    return PZ_OK;
}

pz_status pz_gen_tubes_impl(const pz_gen_params& p, pz_host_mesh& m);

// Append b to a: vertex indices and object ids shifted by a's totals.
static void append_mesh(pz_host_mesh& a, const pz_host_mesh& b, int32_t obj_base) {
    const uint32_t vbase = (uint32_t)a.num_vertices();
    a.vertices.insert(a.vertices.end(), b.vertices.begin(), b.vertices.end());
    for (uint32_t i : b.indices) a.indices.push_back(i + vbase);
    for (int32_t o : b.tri_to_object) a.tri_to_object.push_back(o + obj_base);
    a.edge_starts.insert(a.edge_starts.end(), b.edge_starts.begin(), b.edge_starts.end());
    a.edge_ends.insert(a.edge_ends.end(), b.edge_ends.begin(), b.edge_ends.end());
    for (int32_t o : b.edge_objects) a.edge_objects.push_back(o + obj_base);
}

// Mixed query set: each mesh is an ellipsoid (50%, level uniform in
// [level_min, level_max], radii [radius_min, radius_max]), an axis-aligned
// cube (30%, edge U[1, 10]) or a branching tube (20%, radius U[0.5, 3],
// lognormal triangle counts from the tube fields).  Object ids: ellipsoids,
// then cubes, then tubes.
static pz_status generate_mix(const pz_gen_params& p, pz_host_mesh& m) {
    uint64_t n_kind[3] = {0, 0, 0};
    for (uint64_t i = 0; i < p.count; ++i) {
        Rng r(p.seed ^ 0x6d69787175657279ull, i);
        const double u = r.uniform();
        ++n_kind[u < 0.5 ? 0 : (u < 0.8 ? 1 : 2)];
    }
    const int32_t sub_kind[3] = {PZ_GEN_ELLIPSOIDS, PZ_GEN_CUBES, PZ_GEN_TUBES};
    int32_t obj_base = 0;
    for (int k = 0; k < 3; ++k) {
        if (n_kind[k] == 0) continue;
        pz_gen_params q = p;
        q.kind = sub_kind[k];
        q.count = n_kind[k];
        q.seed = p.seed * 3 + 1 + (uint64_t)k;
        q.edges_mode = p.edges_mode == PZ_EDGES_PRE ? PZ_EDGES_NONE : p.edges_mode;
        q.copy_fraction = 0.f;
        q.degenerate_fraction = 0.f;
        if (k == 1) {
            q.radius_min = 1.f;
            q.radius_max = 10.f;
        } else if (k == 2) {
            q.radius_min = 0.5f;
            q.radius_max = 3.f;
        }
        pz_host_mesh sub;
        pz_status st = k == 2 ? pz_gen_tubes_impl(q, sub) : generate_impl(q, sub);
        if (st != PZ_OK) return st;
        append_mesh(m, sub, obj_base);
        obj_base += (int32_t)n_kind[k];
    }
    if (m.num_vertices() >= (1ull << 32)) return pz::fail(PZ_E_LIMIT, "more than 2^32 vertices in one set");
    if (p.edges_mode == PZ_EDGES_PRE) pz::extract_edges_pre_order(m.view(), m.edge_starts, m.edge_ends, m.edge_objects);
    return PZ_OK;
}

extern "C" pz_status pz_generate(const pz_gen_params* params, pz_host_mesh** out) {
    if (!params || !out) return pz::fail(PZ_E_INVALID, "pz_generate: null argument");
    if (params->kind < PZ_GEN_ELLIPSOIDS || params->kind > PZ_GEN_QUERY_MIX)
        return pz::fail(PZ_E_INVALID, "unknown generator kind");
    auto* m = new pz_host_mesh();
    pz_status st = params->kind == PZ_GEN_TUBES       ? pz_gen_tubes_impl(*params, *m)
                   : params->kind == PZ_GEN_QUERY_MIX ? generate_mix(*params, *m)
                                                      : generate_impl(*params, *m);
    if (st != PZ_OK) {
        delete m;
        return st;
    }
    *out = m;
    return PZ_OK;
}
