// Instancing from a resident mesh library: every instance copies one library
// mesh's vertices, triangles and edges into a set under an affine map
// x' = A x + t, evaluated as ((a0 x + a1 y) + a2 z) + t0 with every product
// and sum rounded (no fused multiply-add), so a host evaluation in the same
// order reproduces the coordinates bit for bit.
#include <hip/hip_runtime.h>

#include "pz_kernels.h"

namespace {

constexpr int kInstThreads = 256;

#pragma clang fp contract(off)
__device__ __forceinline__ void map3(const float* __restrict__ m, const float* __restrict__ p, float* __restrict__ q) {
    const float x = p[0], y = p[1], z = p[2];
    for (int r = 0; r < 3; ++r) q[r] = ((m[3 * r] * x + m[3 * r + 1] * y) + m[3 * r + 2] * z) + m[9 + r];
}

// One workgroup per instance; each of its three runs (vertices, triangles,
// edges) is read and written contiguously.
__global__ __launch_bounds__(kInstThreads) void k_instantiate(
    const pzk::InstanceRec* __restrict__ inst, const float* __restrict__ lv, const uint32_t* __restrict__ lidx,
    const float* __restrict__ les, const float* __restrict__ lee, float* __restrict__ v, uint32_t* __restrict__ idx,
    int32_t* __restrict__ t2o, float* __restrict__ es, float* __restrict__ ee, int32_t* __restrict__ eo) {
    const pzk::InstanceRec r = inst[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < r.nv; i += kInstThreads)
        map3(r.m, lv + 3ull * (r.sv + i), v + 3ull * (r.dv + i));
    for (uint32_t i = threadIdx.x; i < r.nt; i += kInstThreads) {
        const uint64_t s = 3ull * (r.st + i), d = 3ull * (r.dt + i);
        for (int k = 0; k < 3; ++k) idx[d + k] = lidx[s + k] - r.sv + r.dv;
        t2o[r.dt + i] = (int32_t)r.obj;
    }
    for (uint32_t i = threadIdx.x; i < r.ne; i += kInstThreads) {
        map3(r.m, les + 3ull * (r.se + i), es + 3ull * (r.de + i));
        map3(r.m, lee + 3ull * (r.se + i), ee + 3ull * (r.de + i));
        eo[r.de + i] = (int32_t)r.obj;
    }
}
#pragma clang fp contract(on)

}  // namespace

namespace pzk {

void instantiate(const InstanceRec* inst, uint32_t n, const float* lv, const uint32_t* lidx, const float* les,
                 const float* lee, float* v, uint32_t* idx, int32_t* t2o, float* es, float* ee, int32_t* eo,
                 hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_instantiate, dim3(n), dim3(kInstThreads), 0, s, inst, lv, lidx, les, lee, v, idx, t2o,
                           es, ee, eo);
}

}  // namespace pzk
