// Exact geometric predicates for fp32 inputs on gfx950.
//
// Each orientation sign is decided by the cheapest stage that can prove it:
//   1. fp32 evaluation with Shewchuk's forward error bound A
//      (J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
//      Robust Geometric Predicates", DCG 18, 1997, section 4.3) plus an
//      absolute 2^-100 term that covers gradual underflow when every
//      coordinate difference is below 2^40 (the caller guarantees this by
//      passing allow_f32 only for such inputs);
//   2. the same evaluation in fp64 with the fp64 bound (no underflow or
//      overflow is possible for fp32 inputs);
//   3. exact evaluation: the determinant expanded into monomials of fp32
//      inputs, each monomial formed exactly as a (hi, lo) pair of doubles via
//      FMA, accumulated in a 1024-bit fixed-point integer whose range covers
//      every product of three fp32 values (bit positions -500 .. 385).
// Contraction is disabled so every rounded operation matches the analysis.
#pragma once

#include "pz_device.h"

namespace pzd {

// Escalation record of one composite test: 0 = all fp32, 1 = some fp64, 2 = some exact.
struct Esc {
    int level;
    __device__ void bump(int l) { level = l > level ? l : level; }
};

// (7 + 56 eps) eps rounded up, eps = 2^-24 and 2^-53.
#define PZ_O3D_A32 0x1.c0001p-22f
#define PZ_O3D_A64 0x1.c000000000008p-51
// (3 + 16 eps) eps, eps = 2^-53, rounded up.
#define PZ_CCW_A64 0x1.8000000000008p-52

struct ExactAcc {
    static constexpr int kLimbs = 32;
    static constexpr int kEmin = -500;  // lowest bit position representable
    long long L[kLimbs];
    __device__ void clear() {
        for (int i = 0; i < kLimbs; ++i) L[i] = 0;
    }
    // Add a (normal or zero) double exactly.
    __device__ void add(double x) {
        if (x == 0.0) return;
        uint64_t bits = (uint64_t)__double_as_longlong(x);
        bool neg = bits >> 63;
        int E = (int)((bits >> 52) & 0x7FF);
        uint64_t M = (bits & 0xFFFFFFFFFFFFFull) | (1ull << 52);
        int pos = E - 1075 - kEmin;  // >= 0 for every value formed here
        int li = pos >> 5, sh = pos & 31;
        uint64_t lo = M << sh;
        uint64_t hi = sh ? (M >> (64 - sh)) : 0ull;
        long long d0 = (long long)(lo & 0xFFFFFFFFull), d1 = (long long)(lo >> 32), d2 = (long long)hi;
        if (neg) {
            d0 = -d0;
            d1 = -d1;
            d2 = -d2;
        }
        L[li] += d0;
        L[li + 1] += d1;
        L[li + 2] += d2;
    }
    __device__ int sign() {
        for (int i = 0; i < kLimbs - 1; ++i) {
            long long carry = L[i] >> 32;  // arithmetic shift: floor division
            L[i] -= carry * 4294967296LL;
            L[i + 1] += carry;
        }
        if (L[kLimbs - 1] != 0) return L[kLimbs - 1] > 0 ? 1 : -1;
        for (int i = kLimbs - 2; i >= 0; --i)
            if (L[i] != 0) return 1;
        return 0;
    }
};

__device__ __forceinline__ void acc_prod3(ExactAcc& A, float u, float v, float w, bool negate) {
#pragma clang fp contract(off)
    double uv = (double)u * (double)v;  // exact: 48-bit product
    double hi = uv * (double)w;
    double lo = fma(uv, (double)w, -hi);  // exact remainder
    if (negate) {
        hi = -hi;
        lo = -lo;
    }
    A.add(hi);
    A.add(lo);
}

// Adds s * det|p; q; r| (rows).
__device__ __forceinline__ void acc_det3(ExactAcc& A, const float3& p, const float3& q, const float3& r, bool neg) {
    acc_prod3(A, p.x, q.y, r.z, neg);
    acc_prod3(A, p.x, q.z, r.y, !neg);
    acc_prod3(A, p.y, q.x, r.z, !neg);
    acc_prod3(A, p.y, q.z, r.x, neg);
    acc_prod3(A, p.z, q.x, r.y, neg);
    acc_prod3(A, p.z, q.y, r.x, !neg);
}

// |a-d; b-d; c-d| = det(a,b,c) - det(d,b,c) - det(a,d,c) - det(a,b,d).
__device__ __noinline__ int orient3d_exact(float3 a, float3 b, float3 c, float3 d) {
    ExactAcc A;
    A.clear();
    acc_det3(A, a, b, c, false);
    acc_det3(A, d, b, c, true);
    acc_det3(A, a, d, c, true);
    acc_det3(A, a, b, d, true);
    return A.sign();
}

// Out of line: in fp32-filtered traversal this stage is rare, and keeping its
// doubles out of the caller's register budget keeps traversal occupancy up.
__device__ __noinline__ int orient3d_f64(float3 a, float3 b, float3 c, float3 d, Esc& esc) {
#pragma clang fp contract(off)
    double adx = (double)a.x - d.x, bdx = (double)b.x - d.x, cdx = (double)c.x - d.x;
    double ady = (double)a.y - d.y, bdy = (double)b.y - d.y, cdy = (double)c.y - d.y;
    double adz = (double)a.z - d.z, bdz = (double)b.z - d.z, cdz = (double)c.z - d.z;
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy, cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                  (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    double err = PZ_O3D_A64 * perm;
    if (det > err) return 1;
    if (-det > err) return -1;
    esc.bump(2);
    return orient3d_exact(a, b, c, d);
}

// Sign of |a-d; b-d; c-d|: positive when d lies below the plane of a, b, c
// seen counter-clockwise from above.
template <bool AllowF32>
__device__ __forceinline__ int orient3d(const float3& a, const float3& b, const float3& c, const float3& d,
                                        bool allow_f32, Esc& esc) {
#pragma clang fp contract(off)
    if (AllowF32 && allow_f32) {
        float adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
        float ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
        float adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
        float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy, cdxady = cdx * ady, adxcdy = adx * cdy;
        float adxbdy = adx * bdy, bdxady = bdx * ady;
        float det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
        float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * fabsf(adz) + (fabsf(cdxady) + fabsf(adxcdy)) * fabsf(bdz) +
                     (fabsf(adxbdy) + fabsf(bdxady)) * fabsf(cdz);
        float err = PZ_O3D_A32 * perm + 0x1p-100f;
        if (det > err) return 1;
        if (-det > err) return -1;
        esc.bump(1);
    }
    return orient3d_f64(a, b, c, d, esc);
}

// fp32 stage alone: the sign when the filter proves it, 0 when it cannot
// (the caller decides such cases with the exact predicates).  Valid only
// where orient3d's allow_f32 holds.
__device__ __forceinline__ int orient3d_f32(const float3& a, const float3& b, const float3& c, const float3& d) {
#pragma clang fp contract(off)
    float adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    float ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    float adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy, cdxady = cdx * ady, adxcdy = adx * cdy;
    float adxbdy = adx * bdy, bdxady = bdx * ady;
    float det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * fabsf(adz) + (fabsf(cdxady) + fabsf(adxcdy)) * fabsf(bdz) +
                 (fabsf(adxbdy) + fabsf(bdxady)) * fabsf(cdz);
    float err = PZ_O3D_A32 * perm + 0x1p-100f;
    return det > err ? 1 : (-det > err ? -1 : 0);
}

// Naive fp32 sign (no filter): what an unguarded fp32 test would decide.
__device__ __forceinline__ int orient3d_naive(const float3& a, const float3& b, const float3& c, const float3& d) {
#pragma clang fp contract(off)
    float adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    float ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    float adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    float det = adz * (bdx * cdy - cdx * bdy) + bdz * (cdx * ady - adx * cdy) + cdz * (adx * bdy - bdx * ady);
    return det > 0.f ? 1 : (det < 0.f ? -1 : 0);
}

__device__ __forceinline__ float cget(const float3& p, int i) { return i == 0 ? p.x : (i == 1 ? p.y : p.z); }

__device__ __noinline__ int orient2d_exact(float3 a, float3 b, float3 c, int i, int j) {
    ExactAcc A;
    A.clear();
    // det2(a,b) - det2(c,b) - det2(a,c), det2(p,q) = p_i q_j - p_j q_i (products exact in fp64)
    float ai = cget(a, i), aj = cget(a, j), bi = cget(b, i), bj = cget(b, j), ci = cget(c, i), cj = cget(c, j);
    A.add((double)ai * bj);
    A.add(-((double)aj * bi));
    A.add(-((double)ci * bj));
    A.add((double)cj * bi);
    A.add(-((double)ai * cj));
    A.add((double)aj * ci);
    return A.sign();
}

// orient2d of a, b, c projected on coordinates (i, j); fp64 filter then exact.
__device__ __forceinline__ int orient2d(const float3& a, const float3& b, const float3& c, int i, int j, Esc& esc) {
#pragma clang fp contract(off)
    double detl = ((double)cget(a, i) - cget(c, i)) * ((double)cget(b, j) - cget(c, j));
    double detr = ((double)cget(a, j) - cget(c, j)) * ((double)cget(b, i) - cget(c, i));
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
    double err = PZ_CCW_A64 * sum;
    if (det >= err && det != 0) return 1;
    if (-det >= err && det != 0) return -1;
    esc.bump(2);
    return orient2d_exact(a, b, c, i, j);
}

__device__ __forceinline__ bool between(float a, float b, float x) { return fminf(a, b) <= x && x <= fmaxf(a, b); }

__device__ __forceinline__ bool seg_seg2(const float3& p, const float3& q, const float3& a, const float3& b, int i,
                                         int j, Esc& esc) {
    int d1 = orient2d(a, b, p, i, j, esc), d2 = orient2d(a, b, q, i, j, esc);
    int d3 = orient2d(p, q, a, i, j, esc), d4 = orient2d(p, q, b, i, j, esc);
    if (d1 * d2 < 0 && d3 * d4 < 0) return true;
    if (d1 == 0 && between(cget(a, i), cget(b, i), cget(p, i)) && between(cget(a, j), cget(b, j), cget(p, j)))
        return true;
    if (d2 == 0 && between(cget(a, i), cget(b, i), cget(q, i)) && between(cget(a, j), cget(b, j), cget(q, j)))
        return true;
    if (d3 == 0 && between(cget(p, i), cget(q, i), cget(a, i)) && between(cget(p, j), cget(q, j), cget(a, j)))
        return true;
    if (d4 == 0 && between(cget(p, i), cget(q, i), cget(b, i)) && between(cget(p, j), cget(q, j), cget(b, j)))
        return true;
    return false;
}

__device__ __forceinline__ bool pt_in_tri2(const float3& p, const float3& a, const float3& b, const float3& c, int i,
                                           int j, Esc& esc) {
    int s1 = orient2d(a, b, p, i, j, esc), s2 = orient2d(b, c, p, i, j, esc), s3 = orient2d(c, a, p, i, j, esc);
    bool neg = s1 < 0 || s2 < 0 || s3 < 0, pos = s1 > 0 || s2 > 0 || s3 > 0;
    return !(neg && pos);
}

// Coplanar closed segment / closed triangle test.
__device__ __noinline__ bool seg_tri_coplanar(float3 p, float3 q, float3 a, float3 b, float3 c, Esc* escp) {
    Esc esc = *escp;
    int i = 0, j = 1;
    if (orient2d(a, b, c, 0, 1, esc) != 0) {
        i = 0;
        j = 1;
    } else if (orient2d(a, b, c, 1, 2, esc) != 0) {
        i = 1;
        j = 2;
    } else if (orient2d(a, b, c, 2, 0, esc) != 0) {
        i = 2;
        j = 0;
    } else {
        *escp = esc;
        return false;  // degenerate triangle
    }
    bool hit = pt_in_tri2(p, a, b, c, i, j, esc) || pt_in_tri2(q, a, b, c, i, j, esc) ||
               seg_seg2(p, q, a, b, i, j, esc) || seg_seg2(p, q, b, c, i, j, esc) || seg_seg2(p, q, c, a, i, j, esc);
    *escp = esc;
    return hit;
}

// Closed segment pq against closed triangle abc, exact.
template <bool AllowF32>
__device__ __forceinline__ bool seg_tri(const float3& p, const float3& q, const float3& a, const float3& b,
                                        const float3& c, bool allow_f32, Esc& esc) {
    int op = orient3d<AllowF32>(a, b, c, p, allow_f32, esc);
    int oq = orient3d<AllowF32>(a, b, c, q, allow_f32, esc);
    if (op * oq > 0) return false;
    if (op == 0 && oq == 0) return seg_tri_coplanar(p, q, a, b, c, &esc);
    int s1 = orient3d<AllowF32>(p, q, a, b, allow_f32, esc);
    int s2 = orient3d<AllowF32>(p, q, b, c, allow_f32, esc);
    if (s1 * s2 < 0) return false;
    int s3 = orient3d<AllowF32>(p, q, c, a, allow_f32, esc);
    if (s1 * s3 < 0 || s2 * s3 < 0) return false;
    return !(s1 == 0 && s2 == 0 && s3 == 0);
}

// seg_tri decided by the fp32 filter alone: 1 = intersects, 0 = does not,
// -1 = undecided.  A certain answer agrees with seg_tri: a miss needs two
// proven opposite signs (or both endpoints proven on one side), which decide
// seg_tri whatever the unproven signs are; a hit needs every sign proven.
__device__ __forceinline__ int seg_tri_f32(const float3& p, const float3& q, const float3& a, const float3& b,
                                           const float3& c) {
    const int op = orient3d_f32(a, b, c, p), oq = orient3d_f32(a, b, c, q);
    if (op != 0 && op == oq) return 0;
    if (op == 0 || oq == 0) return -1;
    const int s1 = orient3d_f32(p, q, a, b), s2 = orient3d_f32(p, q, b, c);
    if (s1 != 0 && s2 != 0 && s1 != s2) return 0;
    const int s3 = orient3d_f32(p, q, c, a);
    if (s3 != 0 && ((s1 != 0 && s1 != s3) || (s2 != 0 && s2 != s3))) return 0;
    return (s1 != 0 && s2 != 0 && s3 != 0) ? 1 : -1;
}

// What unguarded fp32 arithmetic would decide (for the misclassification count).
__device__ __noinline__ bool seg_tri_naive(float3 p, float3 q, float3 a, float3 b, float3 c) {
    int op = orient3d_naive(a, b, c, p), oq = orient3d_naive(a, b, c, q);
    if (op * oq > 0) return false;
    if (op == 0 && oq == 0) {
        Esc e{0};
        return seg_tri_coplanar(p, q, a, b, c, &e);
    }
    int s1 = orient3d_naive(p, q, a, b), s2 = orient3d_naive(p, q, b, c), s3 = orient3d_naive(p, q, c, a);
    bool neg = s1 < 0 || s2 < 0 || s3 < 0, pos = s1 > 0 || s2 > 0 || s3 > 0;
    return !(neg && pos) && !(s1 == 0 && s2 == 0 && s3 == 0);
}

// orient2d_xy(u, v, p') with p' = p + (eps, eps^2):
// f(p) + eps (u.y - v.y) + eps^2 (v.x - u.x)  (simulation of simplicity).
__device__ __forceinline__ int orient2d_sos(const float3& u, const float3& v, const float3& p, Esc& esc) {
    int s = orient2d(u, v, p, 0, 1, esc);
    if (s) return s;
    if (u.y != v.y) return u.y > v.y ? 1 : -1;
    if (v.x != u.x) return v.x > u.x ? 1 : -1;
    return 0;
}

// Does the +Z ray from the symbolically perturbed p cross triangle abc above p?
__device__ __forceinline__ bool ray_cross(const float3& p, const float3& a, const float3& b, const float3& c,
                                          Esc& esc) {
    int s = orient2d(a, b, c, 0, 1, esc);
    if (s == 0) return false;
    int e1 = orient2d_sos(a, b, p, esc);
    int e2 = orient2d_sos(b, c, p, esc);
    if (e1 != e2) return false;
    int e3 = orient2d_sos(c, a, p, esc);
    if (e3 != e1) return false;
    int o = orient3d<false>(a, b, c, p, false, esc);
    return o != 0 && o == s;
}

}  // namespace pzd
