// Neuron-like branching tubes: each mesh is a set of capped generalised
// cylinders swept along a random-walk skeleton (step `step`, a branch starts
// at a walk point with probability `branch_prob`, at most `max_branches`
// branches), `ring_segments` vertices per ring, one tube radius per mesh.
// Every tube is a closed, outward-oriented shell; a mesh is the union of its
// shells (branches may interpenetrate).  Triangle counts are lognormal
// (median tri_median, sigma tri_sigma) clipped to [tri_min, tri_max].
// With degenerate_fraction > 0 that fraction of meshes are replaced by copies
// of the source set (copy_source_seed/count): half translated by 1e-6 of the
// mesh extent, half rotated about one face's normal through its centroid so
// that face stays coplanar with the original.
#include <algorithm>
#include <cmath>
#include <cstring>

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
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull + 0x5851F42D4C957F2Dull);
        for (auto& w : s) w = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};
struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 unit(V3 a) {
    double n = std::sqrt(dot(a, a));
    return n > 0 ? (1.0 / n) * a : V3{1, 0, 0};
}
V3 random_dir(Rng& r) { return unit({r.normal(), r.normal(), r.normal()}); }

struct MeshOut {
    std::vector<float> v;
    std::vector<uint32_t> f;
    size_t nv = 0;  // vertex count (kept after v is released)
};

// Unique undirected edges of a face list over nv vertices, ordered by
// (smaller, larger) vertex index, each with the orientation of the first face
// that has it: the order of a stable sort of every face edge by that key,
// obtained here by a counting sort on the smaller index (stable in face order)
// and an insertion sort of each small bucket on the larger one.
void unique_edges(const std::vector<uint32_t>& F, uint32_t nv, std::vector<std::pair<uint32_t, uint32_t>>& out) {
    const size_t ne = F.size();  // 3 per face
    std::vector<uint32_t> start((size_t)nv + 1, 0);
    for (size_t k = 0; k < ne; ++k) {
        const uint32_t a = F[k], b = F[k - k % 3 + (k % 3 + 1) % 3];
        ++start[std::min(a, b) + 1];
    }
    for (uint32_t v = 0; v < nv; ++v) start[v + 1] += start[v];
    std::vector<std::pair<uint32_t, uint32_t>> bucket(ne);  // (a, b) as in the face
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (size_t k = 0; k < ne; ++k) {
        const uint32_t a = F[k], b = F[k - k % 3 + (k % 3 + 1) % 3];
        bucket[fill[std::min(a, b)]++] = {a, b};
    }
    out.clear();
    out.reserve(ne / 2 + 1);
    for (uint32_t v = 0; v < nv; ++v) {
        auto* lo = bucket.data() + start[v];
        auto* hi = bucket.data() + start[v + 1];
        auto larger = [](const std::pair<uint32_t, uint32_t>& e) { return std::max(e.first, e.second); };
        for (auto* p = lo + 1; p < hi; ++p)  // stable insertion sort on the larger index
            for (auto* q = p; q > lo && larger(*(q - 1)) > larger(*q); --q) std::swap(*(q - 1), *q);
        for (auto* p = lo; p < hi; ++p)
            if (p == lo || larger(*p) != larger(*(p - 1))) out.push_back(*p);
    }
}

// Emit one capped tube through skeleton points P (>= 2 points).
void emit_tube(const std::vector<V3>& P, double radius, int seg, MeshOut& out) {
    const size_t k = P.size();
    const uint32_t base = static_cast<uint32_t>(out.v.size() / 3);
    V3 t0 = unit(P[1] - P[0]);
    V3 ref = std::fabs(t0.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    V3 n = unit(cross(t0, ref));
    for (size_t i = 0; i < k; ++i) {
        V3 t = i == 0 ? unit(P[1] - P[0]) : (i + 1 == k ? unit(P[k - 1] - P[k - 2]) : unit(P[i + 1] - P[i - 1]));
        n = unit(n - dot(n, t) * t);  // rotation-minimising frame by projection
        V3 b = cross(t, n);
        for (int s = 0; s < seg; ++s) {
            double a = 2.0 * kPi * s / seg;
            V3 q = P[i] + radius * (std::cos(a) * n + std::sin(a) * b);
            out.v.push_back(static_cast<float>(q.x));
            out.v.push_back(static_cast<float>(q.y));
            out.v.push_back(static_cast<float>(q.z));
        }
    }
    auto ring = [&](size_t i, int s) { return base + static_cast<uint32_t>(i * seg + ((s % seg) + seg) % seg); };
    // Side quads, oriented outward: ring i -> i+1 along +t, angle increases about +t.
    for (size_t i = 0; i + 1 < k; ++i)
        for (int s = 0; s < seg; ++s) {
            uint32_t a = ring(i, s), b = ring(i, s + 1), c = ring(i + 1, s + 1), d = ring(i + 1, s);
            out.f.insert(out.f.end(), {a, b, c, a, c, d});
        }
    // Caps: fan around the centre points.
    uint32_t c0 = static_cast<uint32_t>(out.v.size() / 3);
    out.v.insert(out.v.end(), {static_cast<float>(P[0].x), static_cast<float>(P[0].y), static_cast<float>(P[0].z)});
    uint32_t c1 = c0 + 1;
    out.v.insert(out.v.end(),
                 {static_cast<float>(P[k - 1].x), static_cast<float>(P[k - 1].y), static_cast<float>(P[k - 1].z)});
    for (int s = 0; s < seg; ++s) {
        out.f.insert(out.f.end(), {c0, ring(0, s + 1), ring(0, s)});
        out.f.insert(out.f.end(), {c1, ring(k - 1, s), ring(k - 1, s + 1)});
    }
}

// Orient each closed shell outward: flip all its faces if the signed volume is negative.
void orient_shell(MeshOut& m, size_t f_begin) {
    double vol = 0;
    for (size_t i = f_begin; i < m.f.size(); i += 3) {
        const float* a = &m.v[3 * m.f[i]];
        const float* b = &m.v[3 * m.f[i + 1]];
        const float* c = &m.v[3 * m.f[i + 2]];
        V3 A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]};
        vol += dot(A, cross(B, C));
    }
    if (vol < 0)
        for (size_t i = f_begin; i < m.f.size(); i += 3) std::swap(m.f[i + 1], m.f[i + 2]);
}

MeshOut make_tube_mesh(const pz_gen_params& p, uint64_t seed, uint64_t idx) {
    Rng r(seed, idx);
    const int seg = std::max(3, p.ring_segments);
    double ln = std::log(static_cast<double>(std::max<int64_t>(p.tri_median, 1))) + p.tri_sigma * r.normal();
    int64_t target = static_cast<int64_t>(std::llround(std::exp(ln)));
    target = std::min(std::max(target, p.tri_min), std::max(p.tri_max, p.tri_min));
    double radius = r.uniform(p.radius_min, p.radius_max);
    V3 c{r.uniform(0, p.box), r.uniform(0, p.box), r.uniform(0, p.box)};
    // Steps: each step adds 2*seg triangles; each tube also has 2*seg cap triangles.
    int64_t steps_total = std::max<int64_t>(1, (target - 2 * seg) / (2 * seg));
    MeshOut out;
    struct Branch {
        V3 start, dir;
    };
    std::vector<Branch> pending{{c, random_dir(r)}};
    int branches = 0;
    int64_t remaining = steps_total;
    while (!pending.empty() && remaining > 0) {
        Branch br = pending.back();
        pending.pop_back();
        int64_t len = pending.empty() && branches >= p.max_branches ? remaining
                                                                    : std::max<int64_t>(1, remaining / 2);
        if (len > remaining) len = remaining;
        std::vector<V3> P{br.start};
        V3 d = br.dir;
        for (int64_t s = 0; s < len; ++s) {
            d = unit(d + 0.35 * random_dir(r));
            P.push_back(P.back() + p.step * d);
            if (branches < p.max_branches && r.uniform() < p.branch_prob) {
                ++branches;
                pending.push_back({P.back(), unit(d + random_dir(r))});
            }
        }
        remaining -= len;
        size_t f0 = out.f.size();
        emit_tube(P, radius, seg, out);
        orient_shell(out, f0);
        if (pending.empty() && remaining > 0) pending.push_back({P.back(), unit(d + random_dir(r))});
    }
    return out;
}

void apply_degenerate_copy(MeshOut& m, Rng& r, double scale_hint) {
    if (r.uniform() < 0.5) {
        V3 d = random_dir(r);
        double off = 1e-6 * scale_hint;
        for (size_t i = 0; i < m.v.size(); i += 3) {
            m.v[i] = static_cast<float>(m.v[i] + off * d.x);
            m.v[i + 1] = static_cast<float>(m.v[i + 1] + off * d.y);
            m.v[i + 2] = static_cast<float>(m.v[i + 2] + off * d.z);
        }
        return;
    }
    // Rotate about the normal of a random face through that face's centroid.
    size_t nf = m.f.size() / 3;
    size_t f = static_cast<size_t>(r.uniform() * nf) % nf;
    const float* a = &m.v[3 * m.f[3 * f]];
    const float* b = &m.v[3 * m.f[3 * f + 1]];
    const float* c = &m.v[3 * m.f[3 * f + 2]];
    V3 A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]};
    V3 ctr = (1.0 / 3.0) * (A + B + C);
    V3 n = unit(cross(B - A, C - A));
    double th = r.uniform(0.1, 2.0 * kPi - 0.1), cs = std::cos(th), sn = std::sin(th);
    for (size_t i = 0; i < m.v.size(); i += 3) {
        V3 q = V3{m.v[i], m.v[i + 1], m.v[i + 2]} - ctr;
        V3 rq = cs * q + sn * cross(n, q) + (dot(n, q) * (1 - cs)) * n;  // Rodrigues
        V3 w = ctr + rq;
        m.v[i] = static_cast<float>(w.x);
        m.v[i + 1] = static_cast<float>(w.y);
        m.v[i + 2] = static_cast<float>(w.z);
    }
}
}  // namespace

pz_status pz_gen_tubes_impl(const pz_gen_params& p, pz_host_mesh& m) {
    if (p.ring_segments < 3 || !(p.step > 0) || !(p.radius_min > 0) || p.radius_max < p.radius_min)
        return pz::fail(PZ_E_INVALID, "invalid tube parameters");
    const uint64_t n = p.count;
    std::vector<MeshOut> meshes(n);
    pz::parallel_for(n, [&](uint64_t i) {
        Rng r(p.seed ^ 0xA5A5A5A5ull, i);
        if (p.degenerate_fraction > 0 && p.copy_source_count > 0 && r.uniform() < p.degenerate_fraction) {
            uint64_t s = static_cast<uint64_t>(r.uniform() * static_cast<double>(p.copy_source_count));
            if (s >= p.copy_source_count) s = p.copy_source_count - 1;
            pz_gen_params src = p;
            src.degenerate_fraction = 0;
            meshes[i] = make_tube_mesh(src, p.copy_source_seed, s);
            apply_degenerate_copy(meshes[i], r, p.box);
        } else {
            meshes[i] = make_tube_mesh(p, p.seed, i);
        }
    });
    uint64_t nv = 0, nt = 0;
    std::vector<uint64_t> voff(n), toff(n);
    for (uint64_t i = 0; i < n; ++i) {
        voff[i] = nv;
        toff[i] = nt;
        nv += meshes[i].v.size() / 3;
        nt += meshes[i].f.size() / 3;
    }
    if (nv >= (1ull << 32)) return pz::fail(PZ_E_LIMIT, "more than 2^32 vertices in one set");
    m.vertices.resize(3 * nv);
    m.indices.resize(3 * nt);
    m.tri_to_object.resize(nt);
    pz::parallel_for(n, [&](uint64_t i) {
        meshes[i].nv = meshes[i].v.size() / 3;
        std::copy(meshes[i].v.begin(), meshes[i].v.end(), m.vertices.begin() + 3 * voff[i]);
        for (size_t k = 0; k < meshes[i].f.size(); ++k)
            m.indices[3 * toff[i] + k] = static_cast<uint32_t>(voff[i] + meshes[i].f[k]);
        std::fill(m.tri_to_object.begin() + toff[i], m.tri_to_object.begin() + toff[i] + meshes[i].f.size() / 3,
                  static_cast<int32_t>(i));
        MeshOut().v.swap(meshes[i].v);
    });
    // Edges: every triangle edge once per mesh (each undirected edge is shared by two faces).
    if (p.edges_mode == PZ_EDGES_PRE) {
        pz::extract_edges_pre_order(m.view(), m.edge_starts, m.edge_ends, m.edge_objects);
    } else if (p.edges_mode == PZ_EDGES_GROUPED) {
        std::vector<uint64_t> eoff(n + 1, 0);
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> per(n);
        pz::parallel_for(n, [&](uint64_t i) {
            unique_edges(meshes[i].f, (uint32_t)(meshes[i].nv), per[i]);
            std::vector<uint32_t>().swap(meshes[i].f);
        });
        for (uint64_t i = 0; i < n; ++i) eoff[i + 1] = eoff[i] + per[i].size();
        m.edge_starts.resize(3 * eoff[n]);
        m.edge_ends.resize(3 * eoff[n]);
        m.edge_objects.resize(eoff[n]);
        pz::parallel_for(n, [&](uint64_t i) {
            for (size_t k = 0; k < per[i].size(); ++k) {
                const float* a = &m.vertices[3 * (voff[i] + per[i][k].first)];
                const float* b = &m.vertices[3 * (voff[i] + per[i][k].second)];
                std::copy(a, a + 3, m.edge_starts.begin() + 3 * (eoff[i] + k));
                std::copy(b, b + 3, m.edge_ends.begin() + 3 * (eoff[i] + k));
                m.edge_objects[eoff[i] + k] = static_cast<int32_t>(i);
            }
        });
    }
    return PZ_OK;
}
