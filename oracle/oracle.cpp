// ORACLE - test infrastructure only.  Nothing in the product links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker / CPU baseline.
//
// A CPU restatement of the reference's join semantics with exact predicates:
//   - overlap(a, b): some closed triangle of a meets some closed triangle of b,
//     the predicate of Face (baselines/face/src/overlap_query.cpp:159-177,
//     CGAL EPICK all_intersected_primitives on Triangle_3), evaluated as
//     "an edge of one triangle meets the other triangle, either way round";
//   - the edge form Pierce traces (pierce/query/src/applications/
//     pierce_overlap_main.cpp:119-141): edges of each set against triangles of
//     the other, both directions;
//   - containment(A, B) = not overlap and the first vertex of B's first
//     triangle strictly inside A (pierce/query/src/applications/
//     pierce_containment_main.cpp:152-169,336-347);
//   - intersection(a, b) = overlap or either first vertex inside the other
//     (pierce/query/src/applications/pierce_intersection_main.cpp:103-155).
// Point-in-mesh is +Z ray parity (pierce/query/src/cuda/mesh_containment.cu:179-293)
// made exact: the query point is perturbed symbolically by (eps, eps^2) in x, y
// (simulation of simplicity) so the ray never meets an edge or vertex.
//
// Exactness: every fp32 input is an integer multiple of 2^-149, so scaling by
// 2^149 makes all coordinates integers below 2^278 and the determinants
// integers below 2^840, evaluated here in 28-limb signed big integers after
// a double-precision filter (Shewchuk 1997, "Adaptive Precision Floating-Point
// Arithmetic and Fast Robust Geometric Predicates", orient2d/orient3d error
// bound A).
//
// Degenerate (zero-area) triangles meet nothing (CGAL requires non-degenerate
// triangles); every generator here emits non-degenerate ones.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------- big integers
constexpr int NL = 28;  // 896-bit magnitude
struct Big {
    int sign = 0;  // -1, 0, +1
    uint32_t m[NL] = {0};
};

int cmp_mag(const Big& a, const Big& b) {
    for (int i = NL - 1; i >= 0; --i)
        if (a.m[i] != b.m[i]) return a.m[i] < b.m[i] ? -1 : 1;
    return 0;
}
void add_mag(const Big& a, const Big& b, Big& r) {
    uint64_t c = 0;
    for (int i = 0; i < NL; ++i) {
        c += static_cast<uint64_t>(a.m[i]) + b.m[i];
        r.m[i] = static_cast<uint32_t>(c);
        c >>= 32;
    }
}
void sub_mag(const Big& a, const Big& b, Big& r) {  // |a| >= |b|
    int64_t br = 0;
    for (int i = 0; i < NL; ++i) {
        int64_t d = static_cast<int64_t>(a.m[i]) - b.m[i] - br;
        br = d < 0;
        r.m[i] = static_cast<uint32_t>(d + (br << 32));
    }
}
Big add(const Big& a, const Big& b) {
    Big r;
    if (a.sign == 0) return b;
    if (b.sign == 0) return a;
    if (a.sign == b.sign) {
        add_mag(a, b, r);
        r.sign = a.sign;
        return r;
    }
    int c = cmp_mag(a, b);
    if (c == 0) return Big();
    if (c > 0) {
        sub_mag(a, b, r);
        r.sign = a.sign;
    } else {
        sub_mag(b, a, r);
        r.sign = b.sign;
    }
    return r;
}
Big neg(Big a) {
    a.sign = -a.sign;
    return a;
}
Big sub(const Big& a, const Big& b) { return add(a, neg(b)); }
Big mul(const Big& a, const Big& b) {
    Big r;
    if (a.sign == 0 || b.sign == 0) return r;
    uint64_t acc[2 * NL] = {0};
    int na = NL, nb = NL;
    while (na > 0 && a.m[na - 1] == 0) --na;
    while (nb > 0 && b.m[nb - 1] == 0) --nb;
    for (int i = 0; i < na; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < nb; ++j) {
            uint64_t t = static_cast<uint64_t>(a.m[i]) * b.m[j] + acc[i + j] + carry;
            acc[i + j] = t & 0xFFFFFFFFull;
            carry = t >> 32;
        }
        int k = i + nb;
        while (carry) {
            uint64_t t = acc[k] + carry;
            acc[k] = t & 0xFFFFFFFFull;
            carry = t >> 32;
            ++k;
        }
    }
    for (int i = 0; i < NL; ++i) r.m[i] = static_cast<uint32_t>(acc[i]);
    r.sign = a.sign * b.sign;
    return r;
}
// f * 2^149 as an exact integer.
Big from_float(float f) {
    Big r;
    if (f == 0.0f) return r;
    double d = std::ldexp(static_cast<double>(f), 149);  // exact: |d| < 2^277, 24-bit mantissa
    int e;
    double fr = std::frexp(std::fabs(d), &e);            // |d| = fr * 2^e, fr in [0.5, 1)
    uint64_t mant = static_cast<uint64_t>(std::ldexp(fr, 53));  // exact 53-bit integer
    int shift = e - 53;
    if (shift < 0) {
        mant >>= -shift;  // exact: d is an integer
        shift = 0;
    }
    // place mant << shift
    int limb = shift / 32, bit = shift % 32;
    unsigned __int128 v = static_cast<unsigned __int128>(mant) << bit;
    for (int k = 0; k < 4 && limb + k < NL; ++k) r.m[limb + k] = static_cast<uint32_t>(v >> (32 * k));
    r.sign = f < 0 ? -1 : 1;
    return r;
}

// ---------------------------------------------------------------- predicates
const double kEps = std::ldexp(1.0, -53);
const double kO3dErrA = (7.0 + 56.0 * kEps) * kEps;
const double kCcwErrA = (3.0 + 16.0 * kEps) * kEps;

int orient3d_exact(const float* a, const float* b, const float* c, const float* d) {
    Big dx = from_float(d[0]), dy = from_float(d[1]), dz = from_float(d[2]);
    Big adx = sub(from_float(a[0]), dx), ady = sub(from_float(a[1]), dy), adz = sub(from_float(a[2]), dz);
    Big bdx = sub(from_float(b[0]), dx), bdy = sub(from_float(b[1]), dy), bdz = sub(from_float(b[2]), dz);
    Big cdx = sub(from_float(c[0]), dx), cdy = sub(from_float(c[1]), dy), cdz = sub(from_float(c[2]), dz);
    Big t1 = mul(adz, sub(mul(bdx, cdy), mul(cdx, bdy)));
    Big t2 = mul(bdz, sub(mul(cdx, ady), mul(adx, cdy)));
    Big t3 = mul(cdz, sub(mul(adx, bdy), mul(bdx, ady)));
    return add(add(t1, t2), t3).sign;
}

// Sign of the determinant |a-d; b-d; c-d|: positive when d lies below the
// plane of a, b, c seen counter-clockwise from above (Shewchuk's convention).
int orient3d(const float* a, const float* b, const float* c, const float* d) {
    double adx = (double)a[0] - d[0], bdx = (double)b[0] - d[0], cdx = (double)c[0] - d[0];
    double ady = (double)a[1] - d[1], bdy = (double)b[1] - d[1], cdy = (double)c[1] - d[1];
    double adz = (double)a[2] - d[2], bdz = (double)b[2] - d[2], cdz = (double)c[2] - d[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy, cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    double perm = (std::fabs(bdxcdy) + std::fabs(cdxbdy)) * std::fabs(adz) +
                  (std::fabs(cdxady) + std::fabs(adxcdy)) * std::fabs(bdz) +
                  (std::fabs(adxbdy) + std::fabs(bdxady)) * std::fabs(cdz);
    double err = kO3dErrA * perm;
    if (det > err) return 1;
    if (-det > err) return -1;
    return orient3d_exact(a, b, c, d);
}

// orient2d on coordinates (u, v) = (p[i], p[j]).
int orient2d_exact(const float* a, const float* b, const float* c, int i, int j) {
    Big acx = sub(from_float(a[i]), from_float(c[i])), bcy = sub(from_float(b[j]), from_float(c[j]));
    Big acy = sub(from_float(a[j]), from_float(c[j])), bcx = sub(from_float(b[i]), from_float(c[i]));
    return sub(mul(acx, bcy), mul(acy, bcx)).sign;
}
int orient2d(const float* a, const float* b, const float* c, int i, int j) {
    double detl = ((double)a[i] - c[i]) * ((double)b[j] - c[j]);
    double detr = ((double)a[j] - c[j]) * ((double)b[i] - c[i]);
    double det = detl - detr, sum;
    if (detl > 0) {
        if (detr <= 0) return det > 0 ? 1 : (det < 0 ? -1 : 0);
        sum = detl + detr;
    } else if (detl < 0) {
        if (detr >= 0) return det > 0 ? 1 : (det < 0 ? -1 : 0);
        sum = -detl - detr;
    } else {
        return det > 0 ? 1 : (det < 0 ? -1 : 0);
    }
    double err = kCcwErrA * sum;
    if (det >= err && det != 0) return 1;
    if (-det >= err && det != 0) return -1;
    return orient2d_exact(a, b, c, i, j);
}

inline bool between(float a, float b, float x) { return std::min(a, b) <= x && x <= std::max(a, b); }
// p collinear with segment ab (in the (i, j) projection): is it on the closed segment?
inline bool on_seg2(const float* a, const float* b, const float* p, int i, int j) {
    return between(a[i], b[i], p[i]) && between(a[j], b[j], p[j]);
}
// Closed segments pq and ab in the (i, j) projection (textbook orientation test).
bool seg_seg2(const float* p, const float* q, const float* a, const float* b, int i, int j) {
    int d1 = orient2d(a, b, p, i, j), d2 = orient2d(a, b, q, i, j);
    int d3 = orient2d(p, q, a, i, j), d4 = orient2d(p, q, b, i, j);
    if (d1 * d2 < 0 && d3 * d4 < 0) return true;
    if (d1 == 0 && on_seg2(a, b, p, i, j)) return true;
    if (d2 == 0 && on_seg2(a, b, q, i, j)) return true;
    if (d3 == 0 && on_seg2(p, q, a, i, j)) return true;
    if (d4 == 0 && on_seg2(p, q, b, i, j)) return true;
    return false;
}
// Point in closed triangle (non-degenerate in the projection).
bool pt_in_tri2(const float* p, const float* a, const float* b, const float* c, int i, int j) {
    int s1 = orient2d(a, b, p, i, j), s2 = orient2d(b, c, p, i, j), s3 = orient2d(c, a, p, i, j);
    bool has_neg = s1 < 0 || s2 < 0 || s3 < 0, has_pos = s1 > 0 || s2 > 0 || s3 > 0;
    return !(has_neg && has_pos);
}

// Closed segment pq against closed triangle abc.
bool seg_tri(const float* p, const float* q, const float* a, const float* b, const float* c) {
    int op = orient3d(a, b, c, p), oq = orient3d(a, b, c, q);
    if (op * oq > 0) return false;
    if (op == 0 && oq == 0) {
        // Coplanar: project on a coordinate plane where the triangle keeps its area.
        static const int P[3][2] = {{0, 1}, {1, 2}, {2, 0}};
        int k = 0;
        for (; k < 3; ++k)
            if (orient2d(a, b, c, P[k][0], P[k][1]) != 0) break;
        if (k == 3) return false;  // degenerate triangle
        int i = P[k][0], j = P[k][1];
        if (pt_in_tri2(p, a, b, c, i, j) || pt_in_tri2(q, a, b, c, i, j)) return true;
        return seg_seg2(p, q, a, b, i, j) || seg_seg2(p, q, b, c, i, j) || seg_seg2(p, q, c, a, i, j);
    }
    // The line pq crosses the plane once, at a point of the closed segment.
    int s1 = orient3d(p, q, a, b), s2 = orient3d(p, q, b, c), s3 = orient3d(p, q, c, a);
    bool has_neg = s1 < 0 || s2 < 0 || s3 < 0, has_pos = s1 > 0 || s2 > 0 || s3 > 0;
    if (has_neg && has_pos) return false;
    if (s1 == 0 && s2 == 0 && s3 == 0) return false;  // degenerate triangle
    return true;
}

bool tri_degenerate(const float* a, const float* b, const float* c) {
    return orient2d(a, b, c, 0, 1) == 0 && orient2d(a, b, c, 1, 2) == 0 && orient2d(a, b, c, 2, 0) == 0;
}

bool tri_tri(const float* t1[3], const float* t2[3]) {
    if (tri_degenerate(t1[0], t1[1], t1[2]) || tri_degenerate(t2[0], t2[1], t2[2])) return false;
    for (int k = 0; k < 3; ++k)
        if (seg_tri(t1[k], t1[(k + 1) % 3], t2[0], t2[1], t2[2])) return true;
    for (int k = 0; k < 3; ++k)
        if (seg_tri(t2[k], t2[(k + 1) % 3], t1[0], t1[1], t1[2])) return true;
    return false;
}

// Sign of orient2d_xy(u, v, p') with p' = p + (eps, eps^2): f(p) + eps*(u.y - v.y) + eps^2*(v.x - u.x).
int orient2d_sos(const float* u, const float* v, const float* p) {
    int s = orient2d(u, v, p, 0, 1);
    if (s) return s;
    if (u[1] != v[1]) return u[1] > v[1] ? 1 : -1;
    if (v[0] != u[0]) return v[0] > u[0] ? 1 : -1;
    return 0;
}
// Does the +Z ray from the perturbed p cross triangle abc above p?
bool ray_cross(const float* p, const float* a, const float* b, const float* c) {
    int s = orient2d(a, b, c, 0, 1);
    if (s == 0) return false;  // vertical triangle: never met by the perturbed ray
    int e1 = orient2d_sos(a, b, p), e2 = orient2d_sos(b, c, p), e3 = orient2d_sos(c, a, p);
    if (!(e1 == e2 && e2 == e3)) return false;
    int o = orient3d(a, b, c, p);
    return o != 0 && o == s;
}

// ---------------------------------------------------------------- meshes
struct Mesh {
    const float* v;
    uint64_t nv;
    const uint32_t* idx;
    uint64_t nt;
    const int32_t* t2o;
    const float* es;
    const float* ee;
    const int32_t* eo;
    uint64_t ne;
    int32_t nobj = 0;
    std::vector<uint64_t> tri_off, tri_list;  // CSR triangles by object
    std::vector<uint64_t> edge_off, edge_list;
    std::vector<float> obox;                  // 6 per object (min xyz, max xyz)
    std::vector<int64_t> first_tri;           // first triangle index per object (-1 none)

    const float* vert(uint64_t t, int k) const { return v + 3 * idx[3 * t + k]; }
    bool valid(uint64_t t) const {
        return idx[3 * t] < nv && idx[3 * t + 1] < nv && idx[3 * t + 2] < nv && t2o[t] >= 0;
    }
    void prepare() {
        nobj = 0;
        for (uint64_t t = 0; t < nt; ++t) nobj = std::max(nobj, t2o[t] + 1);
        for (uint64_t e = 0; e < ne; ++e) nobj = std::max(nobj, eo[e] + 1);
        tri_off.assign(nobj + 1, 0);
        first_tri.assign(nobj, -1);
        obox.assign(6 * (size_t)nobj, 0.f);
        for (int32_t o = 0; o < nobj; ++o)
            for (int a = 0; a < 3; ++a) {
                obox[6 * o + a] = INFINITY;
                obox[6 * o + 3 + a] = -INFINITY;
            }
        for (uint64_t t = 0; t < nt; ++t)
            if (valid(t)) {
                int32_t o = t2o[t];
                tri_off[o + 1]++;
                if (first_tri[o] < 0) first_tri[o] = (int64_t)t;
                for (int k = 0; k < 3; ++k)
                    for (int a = 0; a < 3; ++a) {
                        float x = vert(t, k)[a];
                        obox[6 * o + a] = std::min(obox[6 * o + a], x);
                        obox[6 * o + 3 + a] = std::max(obox[6 * o + 3 + a], x);
                    }
            }
        for (int32_t o = 0; o < nobj; ++o) tri_off[o + 1] += tri_off[o];
        tri_list.assign(tri_off[nobj], 0);
        std::vector<uint64_t> cur(tri_off.begin(), tri_off.end() - 1);
        for (uint64_t t = 0; t < nt; ++t)
            if (valid(t)) tri_list[cur[t2o[t]]++] = t;
        edge_off.assign(nobj + 1, 0);
        for (uint64_t e = 0; e < ne; ++e)
            if (eo[e] >= 0) edge_off[eo[e] + 1]++;
        for (int32_t o = 0; o < nobj; ++o) edge_off[o + 1] += edge_off[o];
        edge_list.assign(edge_off[nobj], 0);
        std::vector<uint64_t> ec(edge_off.begin(), edge_off.end() - 1);
        for (uint64_t e = 0; e < ne; ++e)
            if (eo[e] >= 0) edge_list[ec[eo[e]]++] = e;
    }
};

// Median-split BVH over primitive boxes (leaves of <= 4 primitives).
struct Bvh {
    struct Node {
        float lo[3], hi[3];
        int64_t left, right;  // children, or -1
        uint64_t begin, end;  // item range for leaves
    };
    struct Item {
        uint64_t id;
        float b[6];
    };
    std::vector<Node> nodes;
    std::vector<Item> items;

    // ids: the primitives to index; boxfn(id, float[6]) gives each box.
    template <class BoxFn>
    void build(const std::vector<uint64_t>& ids, BoxFn boxfn) {
        items.resize(ids.size());
        for (size_t i = 0; i < ids.size(); ++i) {
            items[i].id = ids[i];
            boxfn(ids[i], items[i].b);
        }
        nodes.clear();
        if (!items.empty()) rec(0, items.size());
    }
    int64_t rec(uint64_t b, uint64_t e) {
        Node nd;
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = INFINITY;
            nd.hi[a] = -INFINITY;
        }
        for (uint64_t i = b; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                nd.lo[a] = std::min(nd.lo[a], items[i].b[a]);
                nd.hi[a] = std::max(nd.hi[a], items[i].b[3 + a]);
            }
        nd.left = nd.right = -1;
        nd.begin = b;
        nd.end = e;
        int64_t id = (int64_t)nodes.size();
        nodes.push_back(nd);
        if (e - b <= 4) return id;
        int ax = 0;
        float ext = -1;
        for (int a = 0; a < 3; ++a)
            if (nd.hi[a] - nd.lo[a] > ext) {
                ext = nd.hi[a] - nd.lo[a];
                ax = a;
            }
        uint64_t mid = (b + e) / 2;
        std::nth_element(items.begin() + b, items.begin() + mid, items.begin() + e, [&](const Item& x, const Item& y) {
            return x.b[ax] + x.b[3 + ax] < y.b[ax] + y.b[3 + ax];
        });
        int64_t l = rec(b, mid), r = rec(mid, e);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
    // Visit primitives whose closed boxes meet [lo, hi]; f returns false to stop.
    template <class F>
    void query(const float* lo, const float* hi, F f) const {
        if (nodes.empty()) return;
        int64_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node& nd = nodes[stack[--sp]];
            if (nd.lo[0] > hi[0] || nd.hi[0] < lo[0] || nd.lo[1] > hi[1] || nd.hi[1] < lo[1] || nd.lo[2] > hi[2] ||
                nd.hi[2] < lo[2])
                continue;
            if (nd.left < 0) {
                for (uint64_t i = nd.begin; i < nd.end; ++i) {
                    const float* bx = items[i].b;
                    if (bx[0] > hi[0] || bx[3] < lo[0] || bx[1] > hi[1] || bx[4] < lo[1] || bx[2] > hi[2] ||
                        bx[5] < lo[2])
                        continue;
                    if (!f(items[i].id)) return;
                }
            } else {
                stack[sp++] = nd.left;
                stack[sp++] = nd.right;
            }
        }
    }
};

void tri_box(const Mesh& m, uint64_t t, float* bx) {
    for (int a = 0; a < 3; ++a) {
        float x0 = m.vert(t, 0)[a], x1 = m.vert(t, 1)[a], x2 = m.vert(t, 2)[a];
        bx[a] = std::min(x0, std::min(x1, x2));
        bx[3 + a] = std::max(x0, std::max(x1, x2));
    }
}
void build_tri_bvh(const Mesh& m, Bvh& b) {
    std::vector<uint64_t> valid;
    valid.reserve(m.nt);
    for (uint64_t t = 0; t < m.nt; ++t)
        if (m.valid(t)) valid.push_back(t);
    b.build(valid, [&](uint64_t t, float* bx) { tri_box(m, t, bx); });
}

struct PairSet {
    std::vector<int32_t> data;  // (a, b) interleaved, sorted, unique
};

void finish_pairs(std::vector<std::vector<std::pair<int32_t, int32_t>>>& per, PairSet& out) {
    std::vector<std::pair<int32_t, int32_t>> all;
    for (auto& v : per) all.insert(all.end(), v.begin(), v.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    out.data.resize(2 * all.size());
    for (size_t i = 0; i < all.size(); ++i) {
        out.data[2 * i] = all[i].first;
        out.data[2 * i + 1] = all[i].second;
    }
}

int nthreads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

// Face: for each A object, B objects with a triangle meeting one of its triangles.
void face_join(const Mesh& A, const Mesh& B, const Bvh& bvhB, PairSet& out, double* query_us) {
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(nthreads());
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel
    {
        std::vector<int32_t> mark(B.nobj, -1);
        auto& mine = per[thread_id()];
#pragma omp for schedule(dynamic, 16)
        for (int32_t a = 0; a < A.nobj; ++a) {
            for (uint64_t k = A.tri_off[a]; k < A.tri_off[a + 1]; ++k) {
                uint64_t ta = A.tri_list[k];
                float bx[6];
                tri_box(A, ta, bx);
                const float* t1[3] = {A.vert(ta, 0), A.vert(ta, 1), A.vert(ta, 2)};
                bvhB.query(bx, bx + 3, [&](uint64_t tb) {
                    int32_t ob = B.t2o[tb];
                    if (mark[ob] == a) return true;
                    const float* t2[3] = {B.vert(tb, 0), B.vert(tb, 1), B.vert(tb, 2)};
                    if (tri_tri(t1, t2)) {
                        mark[ob] = a;
                        mine.push_back({a, ob});
                    }
                    return true;
                });
            }
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (query_us) *query_us = std::chrono::duration<double, std::micro>(t1 - t0).count();
    finish_pairs(per, out);
}

// Edges of S against triangles of T; pairs reported as (S obj, T obj) or swapped.
void edge_dir(const Mesh& S, const Mesh& T, const Bvh& bvhT, bool swap,
              std::vector<std::vector<std::pair<int32_t, int32_t>>>& per) {
#pragma omp parallel
    {
        std::vector<int32_t> mark(T.nobj, -1);
        auto& mine = per[thread_id()];
#pragma omp for schedule(dynamic, 16)
        for (int32_t s = 0; s < S.nobj; ++s) {
            for (uint64_t k = S.edge_off[s]; k < S.edge_off[s + 1]; ++k) {
                uint64_t e = S.edge_list[k];
                const float* p = S.es + 3 * e;
                const float* q = S.ee + 3 * e;
                float lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(p[a], q[a]);
                    hi[a] = std::max(p[a], q[a]);
                }
                bvhT.query(lo, hi, [&](uint64_t t) {
                    int32_t o = T.t2o[t];
                    if (mark[o] == s) return true;
                    const float *a = T.vert(t, 0), *b = T.vert(t, 1), *c = T.vert(t, 2);
                    if (seg_tri(p, q, a, b, c)) {
                        mark[o] = s;
                        mine.push_back(swap ? std::make_pair(o, s) : std::make_pair(s, o));
                    }
                    return true;
                });
            }
        }
    }
}

bool point_in_object(const Mesh& M, const Bvh& bvh, int32_t obj, const float* p) {
    float lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], INFINITY};
    bool odd = false;
    bvh.query(lo, hi, [&](uint64_t t) {
        if (M.t2o[t] != obj) return true;
        if (ray_cross(p, M.vert(t, 0), M.vert(t, 1), M.vert(t, 2))) odd = !odd;
        return true;
    });
    return odd;
}

bool boxes_meet(const float* a, const float* b) {
    return !(a[0] > b[3] || a[3] < b[0] || a[1] > b[4] || a[4] < b[1] || a[2] > b[5] || a[5] < b[2]);
}

// Object pairs whose closed boxes meet (brute force over a sort on x).
std::vector<std::pair<int32_t, int32_t>> box_pairs(const Mesh& A, const Mesh& B) {
    std::vector<std::pair<int32_t, int32_t>> out;
    std::vector<int32_t> ob;
    for (int32_t b = 0; b < B.nobj; ++b)
        if (B.first_tri[b] >= 0) ob.push_back(b);
    std::sort(ob.begin(), ob.end(), [&](int32_t x, int32_t y) { return B.obox[6 * x] < B.obox[6 * y]; });
    std::vector<float> bx0(ob.size());
    for (size_t i = 0; i < ob.size(); ++i) bx0[i] = B.obox[6 * ob[i]];
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(nthreads());
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t a = 0; a < A.nobj; ++a) {
        if (A.first_tri[a] < 0) continue;
        const float* ba = &A.obox[6 * a];
        size_t end = std::upper_bound(bx0.begin(), bx0.end(), ba[3]) - bx0.begin();
        for (size_t i = 0; i < end; ++i)
            if (boxes_meet(ba, &B.obox[6 * ob[i]])) per[thread_id()].push_back({a, ob[i]});
    }
    for (auto& v : per) out.insert(out.end(), v.begin(), v.end());
    std::sort(out.begin(), out.end());
    return out;
}

bool has_pair(const std::vector<int32_t>& sorted_pairs, int32_t a, int32_t b) {
    size_t lo = 0, hi = sorted_pairs.size() / 2;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        std::pair<int32_t, int32_t> m{sorted_pairs[2 * mid], sorted_pairs[2 * mid + 1]};
        if (m < std::make_pair(a, b))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < sorted_pairs.size() / 2 && sorted_pairs[2 * lo] == a && sorted_pairs[2 * lo + 1] == b;
}

const float* first_vertex(const Mesh& M, int32_t o) { return M.vert((uint64_t)M.first_tri[o], 0); }


// TOUCH (baselines/face/src/touch_query.cpp:125-511): STR tree with fanout 2
// over mesh1 object boxes; each mesh2 object is assigned to the deepest node
// on its descent where more than one child box meets its box (or to the leaf
// it reaches); candidates are (leaf objects below the node) x (assigned
// objects) with meeting boxes; refinement asks whether any triangle of the
// mesh1 object meets the mesh2 object (per-object triangle BVH).  Count only.
struct StrNode {
    float lo[3], hi[3];
    std::vector<int> kids;     // child node ids (internal)
    std::vector<int32_t> objs; // mesh1 objects (leaf)
    std::vector<int32_t> assigned;
    bool leaf = false;
};

struct StrTree {
    std::vector<StrNode> nodes;
    int root = -1;
    const Mesh* M = nullptr;

    void extend(StrNode& n, const float* lo, const float* hi) {
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = std::min(n.lo[a], lo[a]);
            n.hi[a] = std::max(n.hi[a], hi[a]);
        }
    }
    static void empty_box(StrNode& n) {
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = INFINITY;
            n.hi[a] = -INFINITY;
        }
    }
    // Sort-tile-recursive packing of items (boxes given by boxfn) into groups of `fan`.
    template <class BoxFn>
    std::vector<std::vector<int>> pack(std::vector<int> items, int fan, BoxFn boxfn) {
        std::vector<std::vector<int>> groups;
        size_t P = (items.size() + fan - 1) / fan;
        int Sx = (int)std::ceil(std::cbrt((double)P));
        int Sy = (int)std::ceil(std::sqrt((double)P / Sx));
        auto ctr = [&](int it, int ax) {
            const float* b = boxfn(it);
            return b[ax] + b[3 + ax];
        };
        auto sort_ax = [&](size_t b, size_t e, int ax) {
            std::sort(items.begin() + b, items.begin() + e, [&](int x, int y) { return ctr(x, ax) < ctr(y, ax); });
        };
        sort_ax(0, items.size(), 0);
        size_t ix = 0;
        int sx = Sx;
        while (ix < items.size()) {
            size_t ex = std::min(items.size(), ix + (items.size() - ix + sx - 1) / sx);
            sort_ax(ix, ex, 1);
            size_t iy = ix;
            int sy = Sy;
            while (iy < ex) {
                size_t ey = std::min(ex, iy + (ex - iy + sy - 1) / sy);
                sort_ax(iy, ey, 2);
                for (size_t iz = iy; iz < ey;) {
                    std::vector<int> g;
                    for (int k = 0; k < fan && iz < ey; ++k, ++iz) g.push_back(items[iz]);
                    groups.push_back(g);
                }
                --sy;
                iy = ey;
            }
            --sx;
            ix = ex;
        }
        return groups;
    }
    void build(const Mesh& A, int fan) {
        M = &A;
        std::vector<int> objs;
        for (int32_t o = 0; o < A.nobj; ++o)
            if (A.first_tri[o] >= 0) objs.push_back(o);
        if (objs.empty()) return;
        auto obox = [&](int o) { return &A.obox[6 * (size_t)o]; };
        std::vector<int> level;
        for (auto& g : pack(objs, fan, obox)) {
            StrNode n;
            empty_box(n);
            n.leaf = true;
            for (int o : g) {
                n.objs.push_back(o);
                extend(n, obox(o), obox(o) + 3);
            }
            level.push_back((int)nodes.size());
            nodes.push_back(n);
        }
        while (level.size() > 1) {
            std::vector<float> nb;
            auto nbox = [&](int id) {
                static thread_local float tmp[6];
                for (int a = 0; a < 3; ++a) {
                    tmp[a] = nodes[id].lo[a];
                    tmp[3 + a] = nodes[id].hi[a];
                }
                return tmp;
            };
            std::vector<int> next;
            for (auto& g : pack(level, fan, nbox)) {
                StrNode n;
                empty_box(n);
                for (int id : g) {
                    n.kids.push_back(id);
                    extend(n, nodes[id].lo, nodes[id].hi);
                }
                next.push_back((int)nodes.size());
                nodes.push_back(n);
            }
            level.swap(next);
        }
        root = level.front();
    }
    static bool meet(const float* lo, const float* hi, const StrNode& n) {
        return !(lo[0] > n.hi[0] || hi[0] < n.lo[0] || lo[1] > n.hi[1] || hi[1] < n.lo[1] || lo[2] > n.hi[2] ||
                 hi[2] < n.lo[2]);
    }
    void assign(const Mesh& B) {
        if (root < 0) return;
        for (int32_t b = 0; b < B.nobj; ++b) {
            if (B.first_tri[b] < 0) continue;
            const float* lo = &B.obox[6 * (size_t)b];
            const float* hi = lo + 3;
            int p = root;
            if (!meet(lo, hi, nodes[p])) continue;
            bool assigned = false;
            while (!nodes[p].leaf) {
                int cnt = 0, last = -1;
                for (int k : nodes[p].kids)
                    if (meet(lo, hi, nodes[k])) {
                        ++cnt;
                        last = k;
                    }
                if (cnt == 0) break;
                if (cnt > 1) {
                    nodes[p].assigned.push_back(b);
                    assigned = true;
                    break;
                }
                p = last;
            }
            if (!assigned && nodes[p].leaf) nodes[p].assigned.push_back(b);
        }
    }
    void leaves_below(int n, std::vector<int>& out) const {
        if (nodes[n].leaf) {
            out.push_back(n);
            return;
        }
        for (int k : nodes[n].kids) leaves_below(k, out);
    }
};
}  // namespace

// ------------------------------------------------------------------- C API
extern "C" {

struct or_mesh_view {
    const float* vertices;
    uint64_t num_vertices;
    const uint32_t* indices;
    uint64_t num_triangles;
    const int32_t* triangle_to_object;
    const float* edge_starts;
    const float* edge_ends;
    const int32_t* edge_objects;
    uint64_t num_edges;
};

struct or_pairs {
    PairSet set;
    std::vector<uint8_t> types;
    double query_us = 0;
};

static Mesh make_mesh(const or_mesh_view* v) {
    Mesh m;
    m.v = v->vertices;
    m.nv = v->num_vertices;
    m.idx = v->indices;
    m.nt = v->num_triangles;
    m.t2o = v->triangle_to_object;
    m.es = v->edge_starts;
    m.ee = v->edge_ends;
    m.eo = v->edge_objects;
    m.ne = v->num_edges;
    m.prepare();
    return m;
}

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int or_get_threads(void) { return nthreads(); }

int or_orient3d(const float* a, const float* b, const float* c, const float* d) { return orient3d(a, b, c, d); }
int or_orient3d_exact(const float* a, const float* b, const float* c, const float* d) {
    return orient3d_exact(a, b, c, d);
}
int or_orient2d(const float* a, const float* b, const float* c, int i, int j) { return orient2d(a, b, c, i, j); }
int or_seg_tri(const float* p, const float* q, const float* a, const float* b, const float* c) {
    return seg_tri(p, q, a, b, c) ? 1 : 0;
}
int or_tri_tri(const float* t1, const float* t2) {
    const float* a[3] = {t1, t1 + 3, t1 + 6};
    const float* b[3] = {t2, t2 + 3, t2 + 6};
    return tri_tri(a, b) ? 1 : 0;
}
int or_ray_cross(const float* p, const float* a, const float* b, const float* c) {
    return ray_cross(p, a, b, c) ? 1 : 0;
}

int64_t or_pairs_count(const or_pairs* r) { return r ? (int64_t)(r->set.data.size() / 2) : -1; }
const int32_t* or_pairs_data(const or_pairs* r) { return r ? r->set.data.data() : nullptr; }
const uint8_t* or_pairs_types(const or_pairs* r) { return r && !r->types.empty() ? r->types.data() : nullptr; }
double or_pairs_query_us(const or_pairs* r) { return r ? r->query_us : 0.0; }
void or_pairs_free(or_pairs* r) { delete r; }

// Face semantics (triangle against triangle).
or_pairs* or_face_overlap(const or_mesh_view* a, const or_mesh_view* b) {
    Mesh A = make_mesh(a), B = make_mesh(b);
    Bvh bvhB;
    build_tri_bvh(B, bvhB);
    auto* r = new or_pairs();
    face_join(A, B, bvhB, r->set, &r->query_us);
    return r;
}

// Pierce edge semantics; direction 0 both, 1 A-edges vs B-triangles, 2 B-edges vs A-triangles.
or_pairs* or_edge_overlap(const or_mesh_view* a, const or_mesh_view* b, int direction) {
    Mesh A = make_mesh(a), B = make_mesh(b);
    Bvh bvhA, bvhB;
    build_tri_bvh(A, bvhA);
    build_tri_bvh(B, bvhB);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(nthreads());
    auto t0 = std::chrono::steady_clock::now();
    if (direction == 0 || direction == 1) edge_dir(A, B, bvhB, false, per);
    if (direction == 0 || direction == 2) edge_dir(B, A, bvhA, true, per);
    auto* r = new or_pairs();
    r->query_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    finish_pairs(per, r->set);
    return r;
}

// predicate 1: intersection, 2: containment (B inside A); include_overlap adds
// overlap pairs with type 1 for containment.  direction (intersection only,
// pierce_intersection_main.cpp:103-155): 0 both, 1 A's edges against B's
// triangles and A's first vertices in B's objects, 2 the reverse; containment
// always uses both edge directions (pierce_containment_main.cpp:385-550).
or_pairs* or_predicate_join_dir(const or_mesh_view* a, const or_mesh_view* b, int predicate, int include_overlap,
                                int direction) {
    Mesh A = make_mesh(a), B = make_mesh(b);
    Bvh bvhA, bvhB;
    build_tri_bvh(A, bvhA);
    build_tri_bvh(B, bvhB);
    auto t0 = std::chrono::steady_clock::now();
    if (predicate != 1) direction = 0;
    const bool d12 = direction != 2, d21 = direction != 1;
    PairSet ov;
    {
        std::vector<std::vector<std::pair<int32_t, int32_t>>> per(nthreads());
        if (d12) edge_dir(A, B, bvhB, false, per);
        if (d21) edge_dir(B, A, bvhA, true, per);
        finish_pairs(per, ov);
    }
    auto cands = box_pairs(A, B);
    std::vector<uint8_t> keep(cands.size(), 0), type(cands.size(), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)cands.size(); ++i) {
        int32_t x = cands[i].first, y = cands[i].second;
        bool o = has_pair(ov.data, x, y);
        if (predicate == 1) {
            keep[i] = o || (d12 && point_in_object(B, bvhB, y, first_vertex(A, x))) ||
                      (d21 && point_in_object(A, bvhA, x, first_vertex(B, y)));
        } else {
            if (o) {
                keep[i] = include_overlap ? 1 : 0;
                type[i] = 1;
            } else {
                keep[i] = point_in_object(A, bvhA, x, first_vertex(B, y));
            }
        }
    }
    auto* r = new or_pairs();
    for (size_t i = 0; i < cands.size(); ++i)
        if (keep[i]) {
            r->set.data.push_back(cands[i].first);
            r->set.data.push_back(cands[i].second);
            if (predicate == 2 && include_overlap) r->types.push_back(type[i]);
        }
    r->query_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

or_pairs* or_predicate_join(const or_mesh_view* a, const or_mesh_view* b, int predicate, int include_overlap) {
    return or_predicate_join_dir(a, b, predicate, include_overlap, 0);
}

// Point-in-object parity for one point (tests).
int or_point_in_object(const or_mesh_view* m, int32_t obj, const float* p) {
    Mesh M = make_mesh(m);
    Bvh bvh;
    build_tri_bvh(M, bvh);
    return point_in_object(M, bvh, obj, p) ? 1 : 0;
}

// Containment-ray tracking (pierce/query/src/cuda/mesh_intersection_anyhit.cu:46-240,
// read out by pierce/query/src/utilities/PairHitTracking.h:30-117): for every
// source object with triangles, the +Z ray from the first vertex of its first
// triangle against every triangle of T, brute force; hits[s * T.nobj + t] =
// triangles of object t crossed.  Returns 0.
int or_ray_target_hits(const or_mesh_view* s, const or_mesh_view* t, uint32_t* hits) {
    Mesh S = make_mesh(s), T = make_mesh(t);
    const uint64_t nt = (uint64_t)std::max<int32_t>(T.nobj, 0);
    std::fill(hits, hits + (uint64_t)std::max<int32_t>(S.nobj, 0) * nt, 0u);
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t o = 0; o < S.nobj; ++o) {
        if (S.first_tri[o] < 0) continue;
        const float* p = first_vertex(S, o);
        for (uint64_t k = 0; k < T.nt; ++k) {
            if (!T.valid(k)) continue;
            if (ray_cross(p, T.vert(k, 0), T.vert(k, 1), T.vert(k, 2))) ++hits[(uint64_t)o * nt + T.t2o[k]];
        }
    }
    return 0;
}

// Brute force: every triangle pair (tiny inputs only).
or_pairs* or_bruteforce_overlap(const or_mesh_view* a, const or_mesh_view* b) {
    Mesh A = make_mesh(a), B = make_mesh(b);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(nthreads());
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ta = 0; ta < (int64_t)A.nt; ++ta) {
        if (!A.valid(ta)) continue;
        const float* t1[3] = {A.vert(ta, 0), A.vert(ta, 1), A.vert(ta, 2)};
        for (uint64_t tb = 0; tb < B.nt; ++tb) {
            if (!B.valid(tb)) continue;
            const float* t2[3] = {B.vert(tb, 0), B.vert(tb, 1), B.vert(tb, 2)};
            if (tri_tri(t1, t2)) per[thread_id()].push_back({A.t2o[ta], B.t2o[tb]});
        }
    }
    auto* r = new or_pairs();
    finish_pairs(per, r->set);
    return r;
}

// Number of (edge, triangle) pairs, both directions, with meeting closed fp32 boxes.
uint64_t or_count_candidate_tests(const or_mesh_view* a, const or_mesh_view* b) {
    Mesh A = make_mesh(a), B = make_mesh(b);
    Bvh bvhA, bvhB;
    build_tri_bvh(A, bvhA);
    build_tri_bvh(B, bvhB);
    std::atomic<uint64_t> total{0};
    auto count_dir = [&](const Mesh& S, const Bvh& bvhT) {
#pragma omp parallel
        {
            uint64_t local = 0;
#pragma omp for schedule(dynamic, 4096)
            for (int64_t e = 0; e < (int64_t)S.ne; ++e) {
                if (S.eo[e] < 0) continue;
                const float* p = S.es + 3 * e;
                const float* q = S.ee + 3 * e;
                float lo[3], hi[3];
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(p[k], q[k]);
                    hi[k] = std::max(p[k], q[k]);
                }
                bvhT.query(lo, hi, [&](uint64_t) {
                    ++local;
                    return true;
                });
            }
            total += local;
        }
    };
    count_dir(A, bvhB);
    count_dir(B, bvhA);
    return total.load();
}


// TOUCH restatement: number of overlapping object pairs; *query_us gets the
// candidate generation + refinement time (touch_query.cpp:403-503).
int64_t or_touch_overlap_count(const or_mesh_view* a, const or_mesh_view* b, double* query_us) {
    Mesh A = make_mesh(a), B = make_mesh(b);
    // per-object triangle BVHs of mesh2
    std::vector<Bvh> trees(B.nobj);
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t o = 0; o < B.nobj; ++o) {
        if (B.tri_off[o] == B.tri_off[o + 1]) continue;
        std::vector<uint64_t> sub(B.tri_list.begin() + B.tri_off[o], B.tri_list.begin() + B.tri_off[o + 1]);
        trees[o].build(sub, [&](uint64_t t, float* bx) { tri_box(B, t, bx); });
    }
    StrTree str;
    str.build(A, 2);
    str.assign(B);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> work;
    for (size_t n = 0; n < str.nodes.size(); ++n)
        if (!str.nodes[n].assigned.empty()) work.push_back((int)n);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(nthreads());
#pragma omp parallel for schedule(dynamic)
    for (int64_t w = 0; w < (int64_t)work.size(); ++w) {
        const StrNode& nd = str.nodes[work[w]];
        std::vector<int> lv;
        str.leaves_below(work[w], lv);
        for (int l : lv)
            for (int32_t ob : nd.assigned) {
                const float* blo = &B.obox[6 * (size_t)ob];
                if (!StrTree::meet(blo, blo + 3, str.nodes[l])) continue;
                for (int32_t oa : str.nodes[l].objs)
                    if (boxes_meet(&A.obox[6 * (size_t)oa], blo)) per[thread_id()].push_back({oa, ob});
            }
    }
    std::vector<std::pair<int32_t, int32_t>> cands;
    for (auto& v : per) cands.insert(cands.end(), v.begin(), v.end());
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t i = 0; i < (int64_t)cands.size(); ++i) {
        int32_t oa = cands[i].first, ob = cands[i].second;
        bool hit = false;
        for (uint64_t k = A.tri_off[oa]; k < A.tri_off[oa + 1] && !hit; ++k) {
            uint64_t ta = A.tri_list[k];
            float bx[6];
            tri_box(A, ta, bx);
            const float* t1[3] = {A.vert(ta, 0), A.vert(ta, 1), A.vert(ta, 2)};
            trees[ob].query(bx, bx + 3, [&](uint64_t tb) {
                const float* t2[3] = {B.vert(tb, 0), B.vert(tb, 1), B.vert(tb, 2)};
                if (tri_tri(t1, t2)) {
                    hit = true;
                    return false;
                }
                return true;
            });
        }
        if (hit) ++total;
    }
    if (query_us)
        *query_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return total;
}

}  // extern "C"
