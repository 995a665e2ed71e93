// merge_and_deduplicate_pairs_gpu_noalloc: the reference-named merge with its
// scratch in a fixed device array instead of a workspace that grows, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
// the reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "pierce_hip.h"
#include "pz_internal.h"

namespace {
// One fixed array per device, made when the library's code object is loaded
// on that device and kept until the process ends.  Host code never uses its
// name as an address: each device's copy is located once with
// hipGetSymbolAddress, with that device current.
constexpr size_t kNoAllocBytes = size_t(40) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct NoAllocScratch {
    std::mutex mu;
    std::vector<char*> ptr;  // per device, null until located
};
NoAllocScratch& noalloc_scratch() {
    static NoAllocScratch s;
    return s;
}
}  // namespace

extern "C" long long merge_and_deduplicate_pairs_gpu_noalloc(const pz_pair* d1, long long n1, const pz_pair* d2,
                                                             long long n2, pz_pair* d_out) {
    if (n1 < 0 || n2 < 0) {
        (void)pz::fail(PZ_E_INVALID, "merge_and_deduplicate_pairs_gpu_noalloc: negative count");
        return -1;
    }
    const uint64_t n = (uint64_t)n1 + (uint64_t)n2;
    if (n > (uint64_t)PZ_NOALLOC_MAX_PAIRS) {
        (void)pz::fail(PZ_E_LIMIT, "merge_and_deduplicate_pairs_gpu_noalloc: " + std::to_string(n) +
                                       " pairs exceed the fixed scratch limit of " +
                                       std::to_string(PZ_NOALLOC_MAX_PAIRS));
        return -1;
    }
    if (n == 0) return 0;
    if (!d_out) {
        (void)pz::fail(PZ_E_INVALID, "merge_and_deduplicate_pairs_gpu_noalloc: null output");
        return -1;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        (void)pz::fail(PZ_E_NODEVICE, "merge_and_deduplicate_pairs_gpu_noalloc: no current HIP device");
        return -1;
    }
    NoAllocScratch& ns = noalloc_scratch();
    std::lock_guard<std::mutex> lk(ns.mu);
    if ((size_t)dev >= ns.ptr.size()) ns.ptr.resize((size_t)dev + 1, nullptr);
    if (!ns.ptr[(size_t)dev]) {  // first call on this device: locate its copy of the array
        if (pz::merge_scratch_bytes(PZ_NOALLOC_MAX_PAIRS) > kNoAllocBytes) {
            (void)pz::fail(PZ_E_LIMIT, "merge_and_deduplicate_pairs_gpu_noalloc: scratch smaller than its limit needs");
            return -1;
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)pz::fail(PZ_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return -1;
        }
        ns.ptr[(size_t)dev] = static_cast<char*>(p);
    }
    // Every scratch word the merge reads was written earlier in the same call,
    // so the previous call's contents need no clearing.
    const pz_pair* s1 = d1 ? d1 : d_out;
    const pz_pair* s2 = d2 ? d2 : d_out + n1;
    const long long r = pz::merge_pairs_into(s1, (uint64_t)n1, s2, (uint64_t)n2, d_out, ns.ptr[(size_t)dev], nullptr);
    if (r < 0) (void)pz::fail(PZ_E_HIP, "merge_and_deduplicate_pairs_gpu_noalloc: HIP failure");
    return r;
}
