// Learned output sizing on gfx950.
//
//   features : per part (query) of a join, 32 scale-free numbers from the
//              object boxes of both sets: counts, triangle totals, a closed-form
//              uniform estimate, and a 16-bin histogram of per-cell expected
//              box-pair overlaps on a 16^3 grid over the part's scene.
//   inference: 32 -> 128 -> 128 -> 2 MLP, bf16 operands on the matrix cores
//              (v_mfma_f32_16x16x32_bf16), fp32 accumulation, one wavefront
//              per 16 queries, activations re-laid out through LDS.
//   training : fp32 forward + backward per sample, batch gradient with a fixed
//              summation order (bit-reproducible), AdamW update; all launch
//              parameters that change per step live in device memory so an
//              epoch can be captured once as a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <math.h>

#include "pz_device.h"
#include "pz_kernels.h"
#include "pz_kernels_est.h"

namespace {

using pzd::f2ord;
using pzd::ord2f;
using namespace pzk;

constexpr int kBlock = 256;
inline unsigned grid1(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ features
struct PartAcc {
    uint32_t lo[3], hi[3];        // ordered-uint box of the set's objects in the part
    uint32_t n, max_tris;         // objects with triangles, largest triangle count
    unsigned long long tris;      // triangle total
    unsigned long long ext[3];    // box extents in cells of the part grid, 2^-16 fixed point
    unsigned long long n_in_r;    // objects whose box meets the other set's box
};
struct CellAcc {
    uint32_t cnt, pad;
    unsigned long long ext[3];    // summed extents (cells, 2^-16 fixed point)
};
constexpr int kHistSlots = kFeatBins + 2;  // bins, total estimate, cells holding both sets
constexpr float kFix = 65536.0f;

struct Scratch {
    PartAcc* acc;
    CellAcc* cells;
    unsigned long long* hist;
};
Scratch carve(void* p, uint32_t nparts) {
    char* c = static_cast<char*>(p);
    Scratch s;
    s.acc = reinterpret_cast<PartAcc*>(c);
    c += ((sizeof(PartAcc) * 2 * nparts + 255) / 256) * 256;
    s.cells = reinterpret_cast<CellAcc*>(c);
    c += sizeof(CellAcc) * 2 * (size_t)kFeatCells * nparts;
    s.hist = reinterpret_cast<unsigned long long*>(c);
    return s;
}

struct Frame {  // union box U, intersection R and cell size of one part
    float ulo[3], uhi[3], rlo[3], rhi[3], cs[3];
    bool r_empty;
};
#pragma clang fp contract(off)
__device__ inline Frame part_frame(const PartAcc& A, const PartAcc& B) {
    Frame f;
    f.r_empty = false;
    for (int k = 0; k < 3; ++k) {
        float alo = ord2f(A.lo[k]), ahi = ord2f(A.hi[k]), blo = ord2f(B.lo[k]), bhi = ord2f(B.hi[k]);
        f.ulo[k] = fminf(alo, blo);
        f.uhi[k] = fmaxf(ahi, bhi);
        f.rlo[k] = fmaxf(alo, blo);
        f.rhi[k] = fminf(ahi, bhi);
        if (f.rlo[k] > f.rhi[k]) f.r_empty = true;
        f.cs[k] = fmaxf((f.uhi[k] - f.ulo[k]) / (float)kFeatGrid, 1e-30f);
    }
    return f;
}

__global__ void k_feat_init(PartAcc* acc, uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    PartAcc a;
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = 0xFFFFFFFFu;
        a.hi[k] = 0u;
        a.ext[k] = 0ull;
    }
    a.n = a.max_tris = 0;
    a.tris = a.n_in_r = 0ull;
    acc[i] = a;
}

// One object per lane.  The objects of one part are contiguous ids, so the
// lanes of a wave holding one part form one run: a segmented wave scan
// (combining only lanes of the same part) leaves each run's totals in its
// last lane, which adds them to the part with one atomic per field.  A part
// of tens of objects costs a few atomics instead of one per object on one
// address.  Lanes past the last object carry no part.
constexpr uint32_t kNoPart = 0xFFFFFFFFu;

template <class T, class Op>
__device__ inline T seg_scan(T v, uint32_t part, Op op) {
    const int lane = __lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        const T u = __shfl_up(v, off, 64);
        const uint32_t pu = (uint32_t)__shfl_up((int)part, off, 64);
        if (lane >= off && pu == part) v = op(v, u);
    }
    return v;
}
__device__ inline bool seg_last(uint32_t part) {
    const int lane = __lane_id();
    const uint32_t next = (uint32_t)__shfl_down((int)part, 1, 64);
    return part != kNoPart && (lane == 63 || next != part);
}
struct MinOp {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return min(a, b); }
};
struct MaxOp {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return max(a, b); }
};
struct SumOp {
    template <class T>
    __device__ T operator()(T a, T b) const { return a + b; }
};

// A wave walks `chunk` objects, 64 per step.  Runs that end inside a step
// are added to their part at once; the run reaching lane 63 may go on in the
// next step, so its totals are carried (wave-uniform) and seeded into lane 0
// of the next step when that step starts with the same part.
inline int feat_chunk(int32_t nobj) { return nobj > (1 << 16) ? 256 : 64; }
inline unsigned feat_grid(int32_t nobj) {
    const uint64_t waves = ((uint64_t)nobj + feat_chunk(nobj) - 1) / feat_chunk(nobj);
    return (unsigned)((waves * 64 + kBlock - 1) / kBlock);
}

// Both sets in one launch: blocks [0, blocks0) walk set 0, the rest set 1.
struct FeatSets {
    const uint32_t* obox[2];
    const uint32_t* tri[2];
    const uint32_t* part_of[2];
    int32_t nobj[2];
    int chunk[2];
    unsigned blocks0;
};

// Box, counts and triangle totals of each set per part.
__global__ __launch_bounds__(kBlock) void k_feat_bounds(FeatSets fs, PartAcc* __restrict__ acc) {
    const int set = blockIdx.x < fs.blocks0 ? 0 : 1;
    const uint32_t* __restrict__ obox = fs.obox[set];
    const uint32_t* __restrict__ tri = fs.tri[set];
    const uint32_t* __restrict__ part_of = fs.part_of[set];
    const int32_t nobj = fs.nobj[set];
    const int chunk = fs.chunk[set];
    const unsigned blk = blockIdx.x - (set ? fs.blocks0 : 0u);
    const int lane = __lane_id();
    const int64_t begin = (int64_t)((blk * kBlock + threadIdx.x) / 64) * chunk;
    const int64_t end = min(begin + (int64_t)chunk, (int64_t)nobj);
    uint32_t cp = kNoPart, clo[3], chi[3], cn = 0, ctm = 0;  // carried run (wave-uniform)
    unsigned long long cts = 0;
    auto add = [&](uint32_t p, const uint32_t* l, const uint32_t* h, uint32_t n, unsigned long long ts, uint32_t tm) {
        PartAcc& A = acc[2 * p + set];
        for (int k = 0; k < 3; ++k) {
            atomicMin(&A.lo[k], l[k]);
            atomicMax(&A.hi[k], h[k]);
        }
        atomicAdd(&A.n, n);
        atomicAdd(&A.tris, ts);
        atomicMax(&A.max_tris, tm);
    };
    for (int64_t base = begin; base < end; base += 64) {
        const int64_t o = base + lane;
        const uint32_t t = o < end ? tri[o] : 0u;
        const bool valid = t > 0;
        // every object keeps its part (objects without triangles add
        // identities), so a part is one run of lanes
        const uint32_t p = o < end ? (part_of ? part_of[o] : 0u) : kNoPart;
        uint32_t lo[3], hi[3], n = valid ? 1u : 0u, tm = t;
        unsigned long long ts = t;
        for (int k = 0; k < 3; ++k) {
            lo[k] = valid ? obox[6ll * o + k] : 0xFFFFFFFFu;
            hi[k] = valid ? obox[6ll * o + 3 + k] : 0u;
        }
        const uint32_t p0 = (uint32_t)__shfl((int)p, 0, 64);
        if (cp != kNoPart) {
            if (cp == p0) {
                if (lane == 0) {
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = min(lo[k], clo[k]);
                        hi[k] = max(hi[k], chi[k]);
                    }
                    n += cn;
                    ts += cts;
                    tm = max(tm, ctm);
                }
            } else if (lane == 0) {
                add(cp, clo, chi, cn, cts, ctm);
            }
            cp = kNoPart;
        }
        for (int k = 0; k < 3; ++k) {
            lo[k] = seg_scan(lo[k], p, MinOp{});
            hi[k] = seg_scan(hi[k], p, MaxOp{});
        }
        n = seg_scan(n, p, SumOp{});
        ts = seg_scan(ts, p, SumOp{});
        tm = seg_scan(tm, p, MaxOp{});
        if (seg_last(p) && lane != 63) add(p, lo, hi, n, ts, tm);
        cp = (uint32_t)__shfl((int)p, 63, 64);
        if (cp != kNoPart) {
            for (int k = 0; k < 3; ++k) {
                clo[k] = (uint32_t)__shfl((int)lo[k], 63, 64);
                chi[k] = (uint32_t)__shfl((int)hi[k], 63, 64);
            }
            cn = (uint32_t)__shfl((int)n, 63, 64);
            cts = __shfl(ts, 63, 64);
            ctm = (uint32_t)__shfl((int)tm, 63, 64);
        }
    }
    if (cp != kNoPart && lane == 0) add(cp, clo, chi, cn, cts, ctm);
}

// Per-cell counts and extents (cell of the box centre), per-part extent sums
// and objects meeting the other set's box, for both sets.
__global__ __launch_bounds__(kBlock) void k_feat_cells(FeatSets fs, PartAcc* __restrict__ acc,
                                                       CellAcc* __restrict__ cells) {
    const int set = blockIdx.x < fs.blocks0 ? 0 : 1;
    const uint32_t* __restrict__ obox = fs.obox[set];
    const uint32_t* __restrict__ tri = fs.tri[set];
    const uint32_t* __restrict__ part_of = fs.part_of[set];
    const int32_t nobj = fs.nobj[set];
    const int chunk = fs.chunk[set];
    const unsigned blk = blockIdx.x - (set ? fs.blocks0 : 0u);
    const int lane = __lane_id();
    const int64_t begin = (int64_t)((blk * kBlock + threadIdx.x) / 64) * chunk;
    const int64_t end = min(begin + (int64_t)chunk, (int64_t)nobj);
    uint32_t cp = kNoPart;  // carried run (wave-uniform)
    unsigned long long cq[3] = {0ull, 0ull, 0ull}, cr = 0ull;
    auto add = [&](uint32_t p, const unsigned long long* q, unsigned long long r) {
        PartAcc& A = acc[2 * p + set];
        for (int k = 0; k < 3; ++k) atomicAdd(&A.ext[k], q[k]);
        atomicAdd(&A.n_in_r, r);
    };
    for (int64_t base = begin; base < end; base += 64) {
        const int64_t o = base + lane;
        bool valid = o < end && tri[o] > 0;
        const uint32_t p = o < end ? (part_of ? part_of[o] : 0u) : kNoPart;
        unsigned long long q[3] = {0ull, 0ull, 0ull}, r = 0ull;
        if (valid) {
            const PartAcc& A = acc[2 * p];
            const PartAcc& B = acc[2 * p + 1];
            valid = A.n > 0 && B.n > 0;
            if (valid) {
                const Frame f = part_frame(A, B);
                int cell[3];
                bool meets = !f.r_empty;
                for (int k = 0; k < 3; ++k) {
                    const float lo = ord2f(obox[6ll * o + k]), hi = ord2f(obox[6ll * o + 3 + k]);
                    const float c = (lo + hi) * 0.5f;
                    const float u = (c - f.ulo[k]) / f.cs[k];
                    cell[k] = min(max((int)floorf(u), 0), kFeatGrid - 1);
                    const float e = fminf((hi - lo) / f.cs[k], kFix);
                    q[k] = (unsigned long long)(e * kFix + 0.5f);
                    if (lo > f.rhi[k] || hi < f.rlo[k]) meets = false;
                }
                r = meets ? 1ull : 0ull;
                CellAcc& C =
                    cells[(2ull * p + set) * kFeatCells + (cell[2] * kFeatGrid + cell[1]) * kFeatGrid + cell[0]];
                atomicAdd(&C.cnt, 1u);
                for (int k = 0; k < 3; ++k) atomicAdd(&C.ext[k], q[k]);
            }
        }
        const uint32_t p0 = (uint32_t)__shfl((int)p, 0, 64);
        if (cp != kNoPart) {
            if (cp == p0) {
                if (lane == 0) {
                    for (int k = 0; k < 3; ++k) q[k] += cq[k];
                    r += cr;
                }
            } else if (lane == 0) {
                add(cp, cq, cr);
            }
            cp = kNoPart;
        }
        for (int k = 0; k < 3; ++k) q[k] = seg_scan(q[k], p, SumOp{});
        r = seg_scan(r, p, SumOp{});
        if (seg_last(p) && lane != 63) add(p, q, r);
        cp = (uint32_t)__shfl((int)p, 63, 64);
        if (cp != kNoPart) {
            for (int k = 0; k < 3; ++k) cq[k] = __shfl(q[k], 63, 64);
            cr = __shfl(r, 63, 64);
        }
    }
    if (cp != kNoPart && lane == 0) add(cp, cq, cr);
}

// Probability that two intervals of half-width-sum d (in cell units) whose
// centres fall uniformly in one unit cell overlap: P(|x - y| <= d) = d(2 - d).
__device__ inline double overlap_prob(double d) { return d >= 1.0 ? 1.0 : d * (2.0 - d); }

// Expected overlapping box pairs per cell; histogram over log2 bins.  A block
// covers 256 cells of one part (4096 cells per part): it sums its bins in LDS
// and adds each non-zero bin to the part's histogram once.
__global__ __launch_bounds__(kBlock) void k_feat_hist(const CellAcc* __restrict__ cells, uint32_t nparts,
                                                      unsigned long long* __restrict__ hist) {
    static_assert(kFeatCells % kBlock == 0, "a block stays inside one part");
    __shared__ unsigned long long bins[kHistSlots];
    if (threadIdx.x < kHistSlots) bins[threadIdx.x] = 0ull;
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t p = ((uint64_t)blockIdx.x * kBlock) / kFeatCells;
    if (i < (uint64_t)nparts * kFeatCells) {
        const uint64_t c = i % kFeatCells;
        const CellAcc& a = cells[(2 * p) * kFeatCells + c];
        const CellAcc& b = cells[(2 * p + 1) * kFeatCells + c];
        if (a.cnt != 0 && b.cnt != 0) {
            double est = (double)a.cnt * (double)b.cnt;
            for (int k = 0; k < 3; ++k) {
                const double ea = (double)a.ext[k] / a.cnt / kFix, eb = (double)b.ext[k] / b.cnt / kFix;
                est *= overlap_prob(0.5 * (ea + eb));
            }
            if (est > 0) {
                const int bin = min(max((int)floor(log2(est)) + 4, 0), kFeatBins - 1);
                const unsigned long long qv = (unsigned long long)(fmin(est, 0x1p40) * kFix + 0.5);
                atomicAdd(&bins[bin], qv);
                atomicAdd(&bins[kFeatBins], qv);
            }
            atomicAdd(&bins[kFeatBins + 1], 1ull);
        }
    }
    __syncthreads();
    if (threadIdx.x < kHistSlots && p < nparts && bins[threadIdx.x] != 0ull)
        atomicAdd(&hist[p * kHistSlots + threadIdx.x], bins[threadIdx.x]);
}

__global__ void k_feat_finalize(const PartAcc* __restrict__ acc, const unsigned long long* __restrict__ hist,
                                uint32_t nparts, const uint8_t* __restrict__ meta_of_part, uint8_t meta,
                                float* __restrict__ X) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= nparts) return;
    const PartAcc& A = acc[2 * p];
    const PartAcc& B = acc[2 * p + 1];
    float f[kMlpIn];
    for (int k = 0; k < kMlpIn; ++k) f[k] = 0.f;
    f[0] = log2f(1.f + (float)A.n);
    f[1] = log2f(1.f + (float)B.n);
    f[2] = (float)log2(1.0 + (double)A.tris);
    f[3] = (float)log2(1.0 + (double)B.tris);
    f[4] = log2f(1.f + (float)A.max_tris);
    f[5] = log2f(1.f + (float)B.max_tris);
    const int pred = (meta_of_part ? meta_of_part[p] : meta) & 3;
    if (pred <= 2) f[6 + pred] = 1.f;
    f[9] = (float)log2(1.0 + (double)A.n_in_r);
    f[10] = (float)log2(1.0 + (double)B.n_in_r);
    if (A.n > 0 && B.n > 0) {
        const Frame fr = part_frame(A, B);
        double euni = 0;
        if (!fr.r_empty) {
            euni = (double)A.n_in_r * (double)B.n_in_r;
            for (int k = 0; k < 3; ++k) {
                const double ea = (double)A.ext[k] / A.n / kFix * fr.cs[k];
                const double eb = (double)B.ext[k] / B.n / kFix * fr.cs[k];
                const double side = (double)fr.rhi[k] - (double)fr.rlo[k];
                euni *= side > 0 ? overlap_prob(0.5 * (ea + eb) / side) : 1.0;
            }
        }
        f[11] = (float)log2(1.0 + euni);
        const unsigned long long* h = hist + (uint64_t)p * kHistSlots;
        f[12] = (float)log2(1.0 + (double)h[kFeatBins] / kFix);
        for (int b = 0; b < kFeatBins; ++b) f[13 + b] = (float)log2(1.0 + (double)h[b] / kFix);
        f[29] = (float)log2(1.0 + (double)h[kFeatBins + 1]);
        double ga = 1, gb = 1;
        for (int k = 0; k < 3; ++k) {
            ga *= (double)A.ext[k] / A.n / kFix;
            gb *= (double)B.ext[k] / B.n / kFix;
        }
        f[30] = (float)log2(cbrt(ga) + 0x1p-16);
        f[31] = (float)log2(cbrt(gb) + 0x1p-16);
    }
    for (int k = 0; k < kMlpIn; ++k) X[(uint64_t)p * kMlpIn + k] = f[k];
}
#pragma clang fp contract(on)

__global__ void k_part_pair_counts(const uint64_t* __restrict__ pairs, uint64_t np,
                                   const uint32_t* __restrict__ part_of_a, unsigned long long* __restrict__ counts) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < np;
    const uint32_t p = valid ? part_of_a[(uint32_t)(pairs[i] >> 32)] : 0u;
    const uint64_t vm = __ballot(valid);
    if (vm == 0) return;
    const int leader = __ffsll((long long)vm) - 1;
    const uint32_t p0 = (uint32_t)__shfl((int)p, leader, 64);
    if (__ballot(valid && p == p0) == vm) {
        if (__lane_id() == leader) atomicAdd(&counts[p0], (unsigned long long)__popcll(vm));
    } else if (valid) {
        atomicAdd(&counts[p], 1ull);
    }
}

__global__ void k_zero_u64(unsigned long long* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ----------------------------------------------------------------- inference
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (NaN -> quiet NaN).
__device__ inline uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ void k_mlp_pack_bf16(const float* __restrict__ P, uint16_t* __restrict__ w16) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    constexpr int n1 = kMlpHid * kMlpIn, n2 = kMlpHid * kMlpHid;
    if (i < n1) {
        w16[i] = to_bf16(P[kOffW1 + i]);
    } else if (i < n1 + n2) {
        w16[i] = to_bf16(P[kOffW2 + (i - n1)]);
    } else if (i < kW16Elems) {
        const int j = i - n1 - n2, row = j / kMlpHid, col = j % kMlpHid;
        w16[i] = row < kMlpOut ? to_bf16(P[kOffW3 + row * kMlpHid + col]) : (uint16_t)0;
    }
}

constexpr int kLdsStride = kMlpHid + 8;  // bf16 elements per activation row in LDS

__device__ inline bf16x8 load8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(p)); }

// One wavefront = 16 query rows.  MFMA 16x16x32 bf16 operand layout (lane l,
// r = l % 16, g = l / 16): A holds row r, k = 8g..8g+7; B holds column r,
// k = 8g..8g+7; the f32x4 result holds rows 4g..4g+3 of column r.
__global__ __launch_bounds__(64) void k_mlp_infer(const float* __restrict__ X, uint32_t m,
                                                  const float* __restrict__ norm, const uint16_t* __restrict__ w16,
                                                  const float* __restrict__ P, float* __restrict__ Y) {
    __shared__ __align__(16) uint16_t H[16 * kLdsStride];
    const int l = threadIdx.x, r = l & 15, g = l >> 4;
    const uint32_t row = blockIdx.x * 16 + r;
    const uint16_t* W1 = w16;
    const uint16_t* W2 = w16 + kMlpHid * kMlpIn;
    const uint16_t* W3 = W2 + kMlpHid * kMlpHid;

    u16x8 xa;
    for (int i = 0; i < 8; ++i) {
        const int k = 8 * g + i;
        const float x = row < m ? (X[(uint64_t)row * kMlpIn + k] - norm[k]) * norm[kMlpIn + k] : 0.f;
        xa[i] = to_bf16(x);
    }
    const bf16x8 a1 = __builtin_bit_cast(bf16x8, xa);
    // layer 1: K = 32, one MFMA per 16-column tile
    for (int t = 0; t < kMlpHid / 16; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, load8(W1 + (16 * t + r) * kMlpIn + 8 * g), acc, 0, 0, 0);
        const float bias = P[kOffB1 + 16 * t + r];
        for (int i = 0; i < 4; ++i) H[(4 * g + i) * kLdsStride + 16 * t + r] = to_bf16(fmaxf(acc[i] + bias, 0.f));
    }
    __syncthreads();
    bf16x8 a2[4];
    for (int kc = 0; kc < 4; ++kc) a2[kc] = load8(&H[r * kLdsStride + 32 * kc + 8 * g]);
    __syncthreads();
    // layer 2: K = 128
    for (int t = 0; t < kMlpHid / 16; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < 4; ++kc)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[kc], load8(W2 + (16 * t + r) * kMlpHid + 32 * kc + 8 * g),
                                                          acc, 0, 0, 0);
        const float bias = P[kOffB2 + 16 * t + r];
        for (int i = 0; i < 4; ++i) H[(4 * g + i) * kLdsStride + 16 * t + r] = to_bf16(fmaxf(acc[i] + bias, 0.f));
    }
    __syncthreads();
    for (int kc = 0; kc < 4; ++kc) a2[kc] = load8(&H[r * kLdsStride + 32 * kc + 8 * g]);
    // layer 3: 2 outputs in a zero-padded 16-column tile
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < 4; ++kc)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[kc], load8(W3 + r * kMlpHid + 32 * kc + 8 * g), acc, 0, 0, 0);
    if (r < kMlpOut) {
        const float bias = P[kOffB3 + r];
        for (int i = 0; i < 4; ++i) {
            const uint32_t rr = blockIdx.x * 16 + 4 * g + i;
            if (rr < m) Y[(uint64_t)rr * kMlpOut + r] = acc[i] + bias;
        }
    }
}

// ------------------------------------------------------------------ training
// Work area of a batch of B samples (floats): XS [B][32] normalised inputs,
// H1/H2 [B][128] activations, DZ1/DZ2 [B][128] pre-activation deltas,
// DY [B][2] output deltas, LOSS [B].
struct Work {
    float *xs, *h1, *h2, *dz1, *dz2, *dy, *loss;
};
__host__ __device__ inline Work work_of(float* w, uint32_t B) {
    Work k;
    k.xs = w;
    k.h1 = k.xs + (size_t)B * kMlpIn;
    k.h2 = k.h1 + (size_t)B * kMlpHid;
    k.dz1 = k.h2 + (size_t)B * kMlpHid;
    k.dz2 = k.dz1 + (size_t)B * kMlpHid;
    k.dy = k.dz2 + (size_t)B * kMlpHid;
    k.loss = k.dy + (size_t)B * kMlpOut;
    return k;
}

// Sum over the 128 threads of a block in a fixed order.
__device__ inline float block_sum128(float v, float* red) {
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1];
}

__global__ __launch_bounds__(kMlpHid) void k_mlp_fwd_bwd(const float* __restrict__ X, const float* __restrict__ T,
                                                          const uint32_t* __restrict__ perm,
                                                          const uint32_t* __restrict__ ctr, uint32_t steps_per_epoch,
                                                          uint32_t B, const float* __restrict__ norm,
                                                          const float* __restrict__ P, float* __restrict__ work,
                                                          float* __restrict__ Y) {
    __shared__ float xs[kMlpIn], h1[kMlpHid], dz2[kMlpHid], red[2];
    const uint32_t b = blockIdx.x;
    const uint32_t step = ctr ? ctr[0] % steps_per_epoch : 0u;
    const uint32_t row = perm ? perm[(uint64_t)step * B + b] : b;
    const int j = threadIdx.x;
    Work w = work_of(work, B);
    if (j < kMlpIn) {
        const float x = (X[(uint64_t)row * kMlpIn + j] - norm[j]) * norm[kMlpIn + j];
        xs[j] = x;
        w.xs[(uint64_t)b * kMlpIn + j] = x;
    }
    __syncthreads();
    float z1 = P[kOffB1 + j];
    for (int k = 0; k < kMlpIn; ++k) z1 = fmaf(P[kOffW1 + j * kMlpIn + k], xs[k], z1);
    const float a1 = fmaxf(z1, 0.f);
    h1[j] = a1;
    w.h1[(uint64_t)b * kMlpHid + j] = a1;
    __syncthreads();
    float z2 = P[kOffB2 + j];
    for (int k = 0; k < kMlpHid; ++k) z2 = fmaf(P[kOffW2 + j * kMlpHid + k], h1[k], z2);
    const float a2 = fmaxf(z2, 0.f);
    w.h2[(uint64_t)b * kMlpHid + j] = a2;
    const float y0 = P[kOffB3 + 0] + block_sum128(P[kOffW3 + j] * a2, red);
    const float y1 = P[kOffB3 + 1] + block_sum128(P[kOffW3 + kMlpHid + j] * a2, red);
    if (Y && j == 0) {
        Y[(uint64_t)row * kMlpOut + 0] = y0;
        Y[(uint64_t)row * kMlpOut + 1] = y1;
    }
    if (!T) return;
    const float e0 = y0 - T[(uint64_t)row * kMlpOut + 0], e1 = y1 - T[(uint64_t)row * kMlpOut + 1];
    const float dy0 = 2.f * e0 / (float)B, dy1 = 2.f * e1 / (float)B;
    if (j == 0) {
        w.dy[(uint64_t)b * kMlpOut + 0] = dy0;
        w.dy[(uint64_t)b * kMlpOut + 1] = dy1;
        w.loss[b] = e0 * e0 + e1 * e1;
    }
    const float d2 = z2 > 0.f ? fmaf(P[kOffW3 + j], dy0, P[kOffW3 + kMlpHid + j] * dy1) : 0.f;
    dz2[j] = d2;
    w.dz2[(uint64_t)b * kMlpHid + j] = d2;
    __syncthreads();
    float d1 = 0.f;
    if (z1 > 0.f)
        for (int i = 0; i < kMlpHid; ++i) d1 = fmaf(P[kOffW2 + i * kMlpHid + j], dz2[i], d1);
    w.dz1[(uint64_t)b * kMlpHid + j] = d1;
}

// One thread per parameter: gradient = sum over b (ascending) of delta x input.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(const float* __restrict__ work, uint32_t B,
                                                     float* __restrict__ P, float* __restrict__ G, int adam,
                                                     float* __restrict__ M, float* __restrict__ V,
                                                     const float* __restrict__ hyper,
                                                     const uint32_t* __restrict__ ctr) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kMlpParams) return;
    Work w = work_of(const_cast<float*>(work), B);
    const float* delta;
    const float* in = nullptr;
    int dstride, istride, di, ii = 0;
    if (p < kOffB1) {
        delta = w.dz1, dstride = kMlpHid, di = p / kMlpIn;
        in = w.xs, istride = kMlpIn, ii = p % kMlpIn;
    } else if (p < kOffW2) {
        delta = w.dz1, dstride = kMlpHid, di = p - kOffB1, istride = 0;
    } else if (p < kOffB2) {
        delta = w.dz2, dstride = kMlpHid, di = (p - kOffW2) / kMlpHid;
        in = w.h1, istride = kMlpHid, ii = (p - kOffW2) % kMlpHid;
    } else if (p < kOffW3) {
        delta = w.dz2, dstride = kMlpHid, di = p - kOffB2, istride = 0;
    } else if (p < kOffB3) {
        delta = w.dy, dstride = kMlpOut, di = (p - kOffW3) / kMlpHid;
        in = w.h2, istride = kMlpHid, ii = (p - kOffW3) % kMlpHid;
    } else {
        delta = w.dy, dstride = kMlpOut, di = p - kOffB3, istride = 0;
    }
    float g = 0.f;
    if (in) {
        for (uint32_t b = 0; b < B; ++b) g = fmaf(delta[(uint64_t)b * dstride + di], in[(uint64_t)b * istride + ii], g);
    } else {
        for (uint32_t b = 0; b < B; ++b) g += delta[(uint64_t)b * dstride + di];
    }
    if (!adam) {
        G[p] = g;
        return;
    }
    const float lr = hyper[0], b1 = hyper[1], b2 = hyper[2], eps = hyper[3], wd = hyper[4];
    const float t = (float)(ctr[1] + 1u);
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    const float mh = m / (1.f - powf(b1, t));
    const float vh = v / (1.f - powf(b2, t));
    P[p] = P[p] - lr * (mh / (sqrtf(vh) + eps) + wd * P[p]);
}

__global__ void k_mlp_step_end(const float* __restrict__ work, uint32_t B, float* __restrict__ step_loss,
                               uint32_t* __restrict__ ctr) {
    if (threadIdx.x != 0) return;
    const Work w = work_of(const_cast<float*>(work), B);
    float s = 0.f;
    for (uint32_t b = 0; b < B; ++b) s += w.loss[b];
    step_loss[ctr[1]] = s / (float)B;
    ctr[0] += 1u;
    ctr[1] += 1u;
}

// ------------------------------------------------------ grid-cell estimate
// Per matched cell (pierce/query/src/cuda/estimated_overlap.cu:8-41, in the
// same float operations): TA * TB * min(1, (sA + sB + eps)^3 / V * sqrt(rA rB)^gamma),
// 0 when either touch count is 0.  Cells are GridCell words {CenterCount,
// TouchCount, AvgSizeMean, VolRatio}.  One block sums the terms in double:
// each thread over a fixed stride, then a fixed tree, so the total does not
// depend on scheduling.
#pragma clang fp contract(off)
constexpr int kCellBlock = 1024;
__global__ __launch_bounds__(kCellBlock) void k_cell_selectivity(const uint32_t* __restrict__ ga,
                                                                 const uint32_t* __restrict__ gb, uint32_t n,
                                                                 float cell_volume, float eps, float gamma,
                                                                 double* __restrict__ total) {
    __shared__ double part[kCellBlock];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += kCellBlock) {
        const uint32_t ta = ga[4ull * i + 1], tb = gb[4ull * i + 1];
        if (ta == 0u || tb == 0u) continue;
        const float sa = __uint_as_float(ga[4ull * i + 2]), sb = __uint_as_float(gb[4ull * i + 2]);
        const float ra = __uint_as_float(ga[4ull * i + 3]), rb = __uint_as_float(gb[4ull * i + 3]);
        const float comb = sa + sb + eps;
        float prob = comb * comb * comb / cell_volume;
        prob *= powf(sqrtf(ra * rb), gamma);
        prob = fminf(prob, 1.0f);
        acc += (double)((float)ta * (float)tb * prob);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kCellBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = part[0];
}
#pragma clang fp contract(on)

}  // namespace

namespace pzk {

void cell_selectivity(const uint32_t* cells_a, const uint32_t* cells_b, uint32_t n, float cell_volume, float eps,
                      float gamma, double* total, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_selectivity, dim3(1), dim3(kCellBlock), 0, s, cells_a, cells_b, n, cell_volume, eps,
                       gamma, total);
}


size_t features_scratch_bytes(uint32_t nparts) {
    return ((sizeof(PartAcc) * 2 * nparts + 255) / 256) * 256 + sizeof(CellAcc) * 2 * (size_t)kFeatCells * nparts +
           sizeof(unsigned long long) * kHistSlots * nparts;
}

void join_features(const uint32_t* oboxA, const uint32_t* triA, int32_t nA, const uint32_t* partA,
                   const uint32_t* oboxB, const uint32_t* triB, int32_t nB, const uint32_t* partB, uint32_t nparts,
                   const uint8_t* meta_of_part, uint8_t meta, void* scratch, float* X, hipStream_t s) {
    if (nparts == 0) return;
    Scratch sc = carve(scratch, nparts);
    KernelTimer kt_("features", s);
    hipLaunchKernelGGL(k_feat_init, dim3(grid1(2ull * nparts)), dim3(kBlock), 0, s, sc.acc, 2u * nparts);
    const uint64_t zero_n = (sizeof(CellAcc) * 2 * (uint64_t)kFeatCells * nparts) / 8 + (uint64_t)kHistSlots * nparts;
    hipLaunchKernelGGL(k_zero_u64, dim3(grid1(zero_n)), dim3(kBlock), 0, s,
                       reinterpret_cast<unsigned long long*>(sc.cells), zero_n);
    const unsigned gA = nA > 0 ? feat_grid(nA) : 0u, gB = nB > 0 ? feat_grid(nB) : 0u;
    const FeatSets fs{{oboxA, oboxB}, {triA, triB}, {partA, partB}, {std::max(nA, 0), std::max(nB, 0)},
                      {feat_chunk(nA), feat_chunk(nB)}, gA};
    if (gA + gB > 0) {
        hipLaunchKernelGGL(k_feat_bounds, dim3(gA + gB), dim3(kBlock), 0, s, fs, sc.acc);
        hipLaunchKernelGGL(k_feat_cells, dim3(gA + gB), dim3(kBlock), 0, s, fs, sc.acc, sc.cells);
    }
    hipLaunchKernelGGL(k_feat_hist, dim3(grid1((uint64_t)nparts * kFeatCells)), dim3(kBlock), 0, s, sc.cells, nparts,
                       sc.hist);
    hipLaunchKernelGGL(k_feat_finalize, dim3(grid1(nparts)), dim3(kBlock), 0, s, sc.acc, sc.hist, nparts,
                       meta_of_part, meta, X);
}

void part_pair_counts(const uint64_t* pairs, uint64_t np, const uint32_t* part_of_a, uint32_t nparts,
                      unsigned long long* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(grid1(nparts)), dim3(kBlock), 0, s, counts, (uint64_t)nparts);
    if (np)
        hipLaunchKernelGGL(k_part_pair_counts, dim3(grid1(np)), dim3(kBlock), 0, s, pairs, np, part_of_a, counts);
}

void mlp_pack_bf16(const float* params, uint16_t* w16, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack_bf16, dim3(grid1(kW16Elems)), dim3(kBlock), 0, s, params, w16);
}

void mlp_infer_bf16(const float* X, uint32_t m, const float* norm, const uint16_t* w16, const float* params,
                    float* Y, hipStream_t s) {
    if (m == 0) return;
    KernelTimer kt_("mlp_infer", s);
    hipLaunchKernelGGL(k_mlp_infer, dim3((m + 15) / 16), dim3(64), 0, s, X, m, norm, w16, params, Y);
}

size_t mlp_work_floats(uint32_t batch) {
    return (size_t)batch * (kMlpIn + 4 * kMlpHid + kMlpOut + 1);
}

void mlp_forward_backward(const float* X, const float* T, const uint32_t* perm, const uint32_t* ctr,
                          uint32_t steps_per_epoch, uint32_t batch, const float* norm, const float* params,
                          float* work, float* Y, hipStream_t s) {
    if (batch == 0) return;
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(batch), dim3(kMlpHid), 0, s, X, T, perm, ctr,
                       steps_per_epoch ? steps_per_epoch : 1u, batch, norm, params, work, Y);
}

void mlp_grad(const float* work, uint32_t batch, float* params, float* grad, bool adam, float* m, float* v,
              const float* hyper, const uint32_t* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(grid1(kMlpParams)), dim3(kBlock), 0, s, work, batch, params, grad,
                       adam ? 1 : 0, m, v, hyper, ctr);
}

void mlp_step_end(const float* work, uint32_t batch, float* step_loss, uint32_t* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_step_end, dim3(1), dim3(64), 0, s, work, batch, step_loss, ctr);
}

}  // namespace pzk
