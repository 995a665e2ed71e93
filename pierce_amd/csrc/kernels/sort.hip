// Scan, radix sort, unique: the data-movement primitives of the join.
//
// Radix sort: LSD, 8-bit digits, tiles of 256 threads x 16 keys.  Each wave
// ranks its keys stably with a ballot-based digit match (8 ballots give the
// lanes holding the same digit) and wave-private LDS counters, so no
// block-wide sort network is needed; wave totals are prefixed across the
// block and offset by the digit-major global scan of tile histograms.
#include "pz_device.h"
#include "pz_kernels.h"

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / PZ_WAVE;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096
constexpr int kRadix = 256;

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = __lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Block-wide exclusive scan of per-thread values; returns the thread's
// exclusive prefix and writes the block total to *total.
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T* total) {
    __shared__ T wsum[kScanThreads / 64];
    const int lane = __lane_id(), w = threadIdx.x / 64;
    T inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    T off = 0, tot = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) {
        T s = wsum[i];
        if (i < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

template <class Tin, class Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const Tin* __restrict__ in, uint64_t n,
                                                              Tout* __restrict__ sums) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    Tout s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) s += (Tout)in[base + i];
    Tout tot;
    (void)block_excl_scan<Tout>(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class Tin, class Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const Tin* __restrict__ in, uint64_t n,
                                                             Tout* __restrict__ out, const Tout* __restrict__ offs,
                                                             int write_total) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    Tout v[kScanItems];
    Tout s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = base + i < n ? (Tout)in[base + i] : (Tout)0;
        s += v[i];
    }
    Tout tot;
    Tout run = block_excl_scan<Tout>(s, &tot) + (offs ? offs[blockIdx.x] : (Tout)0);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) out[n] = run;
}

// Exclusive scan of `in` (n items) into out[0..n]; recursion over tile sums.
template <class Tin, class Tout>
void scan_impl(const Tin* in, Tout* out, uint64_t n, char* temp, hipStream_t s) {
    if (n == 0) {
        hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3(1), dim3(kScanThreads), 0, s, in, 0, out,
                           (const Tout*)nullptr, 1);
        return;
    }
    uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3(1), dim3(kScanThreads), 0, s, in, n, out,
                           (const Tout*)nullptr, 1);
        return;
    }
    Tout* sums = reinterpret_cast<Tout*>(temp);
    Tout* sums_scan = sums + nb;
    char* rest = temp + ((2 * nb + 1) * sizeof(Tout) + 255) / 256 * 256;
    hipLaunchKernelGGL((k_scan_reduce<Tin, Tout>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, sums);
    scan_impl<Tout, Tout>(sums, sums_scan, nb, rest, s);
    hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, out,
                       (const Tout*)sums_scan, 1);
}

size_t scan_temp_rec(uint64_t n, size_t elem) {
    uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) return 0;
    return ((2 * nb + 1) * elem + 255) / 256 * 256 + scan_temp_rec(nb, elem);
}

// ----------------------------------------------------------------- radix
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t* __restrict__ keys, uint64_t n, int shift,
                                                             uint32_t* __restrict__ counts, uint32_t nblocks) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        uint64_t idx = base + (uint64_t)i * kSortThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    counts[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <bool HasVals>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t* __restrict__ keys,
                                                                const uint32_t* __restrict__ vals, uint64_t n,
                                                                int shift, const uint32_t* __restrict__ offsets,
                                                                uint32_t nblocks, uint64_t* __restrict__ okeys,
                                                                uint32_t* __restrict__ ovals) {
    __shared__ uint32_t cnt[kSortWaves][kRadix];
    const int lane = __lane_id(), w = threadIdx.x / 64;
    for (int i = threadIdx.x; i < kSortWaves * kRadix; i += kSortThreads) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t wbase = (uint64_t)blockIdx.x * kSortTile + (uint64_t)w * (kSortItems * 64);
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t local[kSortItems];
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        uint64_t idx = wbase + (uint64_t)i * 64 + lane;
        bool valid = idx < n;
        k[i] = valid ? keys[idx] : 0ull;
        if (HasVals) v[i] = valid ? vals[idx] : 0u;
        uint32_t d = (uint32_t)(k[i] >> shift) & 0xFF;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t base = valid ? cnt[w][d] : 0u;
        if (valid && (peers & lt) == 0) cnt[w][d] = base + (uint32_t)__popcll(peers);
        local[i] = base + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    // Exclusive prefix of wave counts per digit, then add the global tile offset.
    {
        uint32_t d = threadIdx.x;
        uint32_t run = offsets[(uint64_t)d * nblocks + blockIdx.x];
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ++ww) {
            uint32_t t = cnt[ww][d];
            cnt[ww][d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        uint64_t idx = wbase + (uint64_t)i * 64 + lane;
        if (idx < n) {
            uint32_t d = (uint32_t)(k[i] >> shift) & 0xFF;
            uint32_t pos = cnt[w][d] + local[i];
            okeys[pos] = k[i];
            if (HasVals) ovals[pos] = v[i];
        }
    }
}

// Up to one radix tile of keys: a bitonic sort in the LDS of one workgroup
// (one launch instead of three per 8-bit pass).  Keys are below 2^63, so the
// padding key ~0 sorts last; equal keys may leave in any order.
constexpr int kSmallSortThreads = 1024;
template <bool HasVals>
__global__ __launch_bounds__(kSmallSortThreads) void k_sort_small(uint64_t* __restrict__ keys,
                                                                  uint32_t* __restrict__ vals, uint32_t n) {
    __shared__ uint64_t K[kSortTile];
    __shared__ uint32_t V[HasVals ? kSortTile : 1];
    uint32_t npow = 2;
    while (npow < n) npow <<= 1;
    for (uint32_t i = threadIdx.x; i < npow; i += kSmallSortThreads) {
        K[i] = i < n ? keys[i] : ~0ull;
        if (HasVals) V[i] = i < n ? vals[i] : 0u;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= npow; size <<= 1)
        for (uint32_t st = size >> 1; st > 0; st >>= 1) {
            for (uint32_t i = threadIdx.x; i < npow / 2; i += kSmallSortThreads) {
                const uint32_t a = 2 * i - (i & (st - 1)), b = a + st;
                const bool up = (a & size) == 0;
                const uint64_t x = K[a], y = K[b];
                if ((x > y) == up) {
                    K[a] = y;
                    K[b] = x;
                    if (HasVals) {
                        const uint32_t t = V[a];
                        V[a] = V[b];
                        V[b] = t;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < n; i += kSmallSortThreads) {
        keys[i] = K[i];
        if (HasVals) vals[i] = V[i];
    }
}

// Tiles of kSortTile keys, each sorted by one workgroup in LDS (the bitonic
// network of k_sort_small).  With dn the key count is read from device
// memory, and a count above cap leaves the keys as they are (the caller then
// sorts with the count it has read back).
template <bool HasVals>
__global__ __launch_bounds__(kSmallSortThreads) void k_sort_tiles(uint64_t* __restrict__ keys,
                                                                  uint32_t* __restrict__ vals, uint64_t n,
                                                                  const unsigned long long* __restrict__ dn,
                                                                  uint64_t cap) {
    __shared__ uint64_t K[kSortTile];
    __shared__ uint32_t V[HasVals ? kSortTile : 1];
    if (dn) {
        n = *dn;
        if (n > cap) return;  // workgroup-uniform
    }
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    if (base >= n) return;  // workgroup-uniform
    const uint32_t len = (uint32_t)min((uint64_t)kSortTile, n - base);
    uint32_t npow = 2;
    while (npow < len) npow <<= 1;
    for (uint32_t i = threadIdx.x; i < npow; i += kSmallSortThreads) {
        K[i] = i < len ? keys[base + i] : ~0ull;
        if (HasVals) V[i] = i < len ? vals[base + i] : 0u;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= npow; size <<= 1)
        for (uint32_t st = size >> 1; st > 0; st >>= 1) {
            for (uint32_t i = threadIdx.x; i < npow / 2; i += kSmallSortThreads) {
                const uint32_t a = 2 * i - (i & (st - 1)), b = a + st;
                const bool up = (a & size) == 0;
                const uint64_t x = K[a], y = K[b];
                if ((x > y) == up) {
                    K[a] = y;
                    K[b] = x;
                    if (HasVals) {
                        const uint32_t t = V[a];
                        V[a] = V[b];
                        V[b] = t;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < len; i += kSmallSortThreads) {
        keys[base + i] = K[i];
        if (HasVals) vals[base + i] = V[i];
    }
}

// Thread per key of sorted tiles: its rank in the whole is its place in its
// tile plus, in every other tile, the keys below it (tiles before its own also
// count equal keys, so equal keys keep distinct places).
template <bool HasVals>
__global__ __launch_bounds__(256) void k_merge_tiles(const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals, uint64_t n,
                                                     uint64_t* __restrict__ okeys, uint32_t* __restrict__ ovals,
                                                     const unsigned long long* __restrict__ dn, uint64_t cap) {
    if (dn) {
        n = *dn;
        if (n > cap) return;
    }
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t t = i / kSortTile, ntiles = (n + kSortTile - 1) / kSortTile;
    const uint64_t x = keys[i];
    uint64_t pos = i - t * kSortTile;
    for (uint64_t j = 0; j < ntiles; ++j) {
        if (j == t) continue;
        uint64_t lo = j * kSortTile, hi = min(n, lo + kSortTile);  // first key of tile j not before x
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            const uint64_t y = keys[mid];
            if (j < t ? y <= x : y < x)
                lo = mid + 1;
            else
                hi = mid;
        }
        pos += lo - j * kSortTile;
    }
    okeys[pos] = x;
    if (HasVals) ovals[pos] = vals[i];
}

template <bool HasVals>
bool radix_impl(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint64_t n, int begin_bit,
                int end_bit, void* temp, hipStream_t s) {
    if (n <= 1 || end_bit <= begin_bit) return false;
    if (n <= (uint64_t)kSortTile && begin_bit == 0 && end_bit < 64) {  // keys < 2^end_bit: one workgroup
        hipLaunchKernelGGL((k_sort_small<HasVals>), dim3(1), dim3(kSmallSortThreads), 0, s, keys, vals, (uint32_t)n);
        return false;
    }
    uint32_t nb = (uint32_t)((n + kSortTile - 1) / kSortTile);
    uint64_t ncounts = (uint64_t)kRadix * nb;
    uint32_t* counts = reinterpret_cast<uint32_t*>(temp);
    uint32_t* offs = counts + ncounts;
    char* scan_tmp = reinterpret_cast<char*>(offs + ncounts + 1);
    scan_tmp = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(scan_tmp) + 255) & ~uintptr_t(255));
    bool alt = false;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        uint64_t* ki = alt ? keys_alt : keys;
        uint64_t* ko = alt ? keys : keys_alt;
        uint32_t* vi = alt ? vals_alt : vals;
        uint32_t* vo = alt ? vals : vals_alt;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, ki, n, shift, counts, nb);
        scan_impl<uint32_t, uint32_t>(counts, offs, ncounts, scan_tmp, s);
        pzk::KernelTimer kt_("radix_scatter", s);
        hipLaunchKernelGGL((k_radix_scatter<HasVals>), dim3(nb), dim3(kSortThreads), 0, s, ki, vi, n, shift, offs,
                           nb, ko, vo);
        alt = !alt;
    }
    return alt;
}

__global__ void k_unique_flags(const uint64_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ flags) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || in[i] != in[i - 1]) ? 1u : 0u;
}
__global__ void k_unique_scatter(const uint64_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ flags,
                                 const uint64_t* __restrict__ pos, uint64_t* __restrict__ out,
                                 uint64_t* __restrict__ count) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) out[pos[i]] = in[i];
    if (i == 0) *count = pos[n];
}

__global__ void k_keys_to_pairs(const uint64_t* __restrict__ keys, uint64_t n, int2* __restrict__ pairs) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint64_t k = keys[i];
        pairs[i] = make_int2((int32_t)(uint32_t)(k >> 32), (int32_t)(uint32_t)k);
    }
}
__global__ void k_pairs_to_keys(const int2* __restrict__ pairs, uint64_t n, uint64_t* __restrict__ keys) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = pzd::pack_pair(pairs[i].x, pairs[i].y);
}
template <class T>
__global__ void k_fill(T* p, T v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned grid1(uint64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace pzk {

namespace {
struct ProfState {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t next = 0;
    struct Rec {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::map<std::string, std::pair<double, uint64_t>> totals;
    std::mutex mu;
};
ProfState& prof_state() {
    static ProfState st;
    return st;
}
void prof_drain(ProfState& st) {
    for (auto& r : st.pending) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& t = st.totals[r.name];
            t.first += ms;
            t.second += 1;
        }
    }
    st.pending.clear();
    st.next = 0;
}
}  // namespace

void prof_enable(bool on) {
    auto& st = prof_state();
    std::lock_guard<std::mutex> g(st.mu);
    st.on = on;
}
bool prof_enabled() { return prof_state().on; }
hipEvent_t prof_event() {
    auto& st = prof_state();
    std::lock_guard<std::mutex> g(st.mu);
    if (st.next == st.pool.size()) {
        if (st.next >= 4096) prof_drain(st);  // bounded pool: fold completed pairs
        if (st.next == st.pool.size()) {
            hipEvent_t e = nullptr;
            (void)hipEventCreate(&e);
            st.pool.push_back(e);
        }
    }
    return st.pool[st.next++];
}
void prof_record(const char* name, hipEvent_t a, hipEvent_t b) {
    auto& st = prof_state();
    std::lock_guard<std::mutex> g(st.mu);
    st.pending.push_back({name, a, b});
}
void prof_query(const char* name, double* total_ms, uint64_t* count) {
    auto& st = prof_state();
    std::lock_guard<std::mutex> g(st.mu);
    prof_drain(st);
    auto it = st.totals.find(name);
    *total_ms = it == st.totals.end() ? 0.0 : it->second.first;
    *count = it == st.totals.end() ? 0 : it->second.second;
}
void prof_reset() {
    auto& st = prof_state();
    std::lock_guard<std::mutex> g(st.mu);
    prof_drain(st);
    st.totals.clear();
}

size_t scan_temp_bytes(uint64_t n) { return scan_temp_rec(n, 8) + 256; }

void exclusive_scan_u32_u64(const uint32_t* in, uint64_t* out, uint64_t n, void* temp, hipStream_t s) {
    scan_impl<uint32_t, uint64_t>(in, out, n, reinterpret_cast<char*>(temp), s);
}
void exclusive_scan_u32_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, hipStream_t s) {
    scan_impl<uint32_t, uint32_t>(in, out, n, reinterpret_cast<char*>(temp), s);
}
void exclusive_scan_i32_i64(const int32_t* in, int64_t* out, uint64_t n, void* temp, hipStream_t s) {
    scan_impl<int32_t, int64_t>(in, out, n, reinterpret_cast<char*>(temp), s);
}

size_t radix_sort_temp_bytes(uint64_t n) {
    uint64_t nb = (n + kSortTile - 1) / kSortTile;
    uint64_t ncounts = (uint64_t)kRadix * (nb ? nb : 1);
    return (2 * ncounts + 1) * 4 + 512 + scan_temp_rec(ncounts, 4);
}

bool radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint64_t n,
                      int begin_bit, int end_bit, void* temp, hipStream_t s) {
    return radix_impl<true>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, temp, s);
}
bool radix_sort_keys(uint64_t* keys, uint64_t* keys_alt, uint64_t n, int begin_bit, int end_bit, void* temp,
                     hipStream_t s) {
    return radix_impl<false>(keys, nullptr, keys_alt, nullptr, n, begin_bit, end_bit, temp, s);
}

bool sort_tiles(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint64_t n, hipStream_t s) {
    if (n <= 1) return false;
    const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
    if (vals)
        hipLaunchKernelGGL((k_sort_tiles<true>), dim3((unsigned)ntiles), dim3(kSmallSortThreads), 0, s, keys, vals, n,
                           (const unsigned long long*)nullptr, (uint64_t)0);
    else
        hipLaunchKernelGGL((k_sort_tiles<false>), dim3((unsigned)ntiles), dim3(kSmallSortThreads), 0, s, keys, vals,
                           n, (const unsigned long long*)nullptr, (uint64_t)0);
    if (ntiles == 1) return false;
    if (vals)
        hipLaunchKernelGGL((k_merge_tiles<true>), dim3(grid1(n)), dim3(256), 0, s, keys, vals, n, keys_alt, vals_alt,
                           (const unsigned long long*)nullptr, (uint64_t)0);
    else
        hipLaunchKernelGGL((k_merge_tiles<false>), dim3(grid1(n)), dim3(256), 0, s, keys, vals, n, keys_alt,
                           vals_alt, (const unsigned long long*)nullptr, (uint64_t)0);
    return true;
}
void sort_tiles_counted(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                        const unsigned long long* dn, uint64_t cap, hipStream_t s) {
    if (cap == 0) return;
    const uint64_t ntiles = (cap + kSortTile - 1) / kSortTile;
    if (vals) {
        hipLaunchKernelGGL((k_sort_tiles<true>), dim3((unsigned)ntiles), dim3(kSmallSortThreads), 0, s, keys, vals,
                           (uint64_t)0, dn, cap);
        hipLaunchKernelGGL((k_merge_tiles<true>), dim3(grid1(cap)), dim3(256), 0, s, keys, vals, (uint64_t)0, keys_alt,
                           vals_alt, dn, cap);
    } else {
        hipLaunchKernelGGL((k_sort_tiles<false>), dim3((unsigned)ntiles), dim3(kSmallSortThreads), 0, s, keys, vals,
                           (uint64_t)0, dn, cap);
        hipLaunchKernelGGL((k_merge_tiles<false>), dim3(grid1(cap)), dim3(256), 0, s, keys, vals, (uint64_t)0,
                           keys_alt, vals_alt, dn, cap);
    }
}
void sort_tile_counted(uint64_t* keys, uint32_t* vals, const unsigned long long* dn, hipStream_t s) {
    if (vals)
        hipLaunchKernelGGL((k_sort_tiles<true>), dim3(1), dim3(kSmallSortThreads), 0, s, keys, vals, (uint64_t)0, dn,
                           (uint64_t)kSortTile);
    else
        hipLaunchKernelGGL((k_sort_tiles<false>), dim3(1), dim3(kSmallSortThreads), 0, s, keys, vals, (uint64_t)0, dn,
                           (uint64_t)kSortTile);
}

__global__ void k_keys_to_pairs_counted(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                        const unsigned long long* __restrict__ dn, uint64_t cap,
                                        int2* __restrict__ pairs, uint32_t* __restrict__ ovals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = *dn;
    if (n <= cap && i < n) {
        const uint64_t k = keys[i];
        pairs[i] = make_int2((int32_t)(uint32_t)(k >> 32), (int32_t)(uint32_t)k);
        if (vals) ovals[i] = vals[i];
    }
}
void keys_to_pairs_counted(const uint64_t* keys, const uint32_t* vals, const unsigned long long* dn, uint64_t cap,
                           int32_t* pairs, uint32_t* ovals, hipStream_t s) {
    if (cap)
        hipLaunchKernelGGL(k_keys_to_pairs_counted, dim3(grid1(cap)), dim3(256), 0, s, keys, vals, dn, cap,
                           reinterpret_cast<int2*>(pairs), ovals);
}

size_t unique_temp_bytes(uint64_t n) { return (n * 4 + 255) / 256 * 256 + (n + 1) * 8 + 256 + scan_temp_bytes(n); }

void unique_u64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* d_count, void* temp, hipStream_t s) {
    if (n == 0) {
        hipLaunchKernelGGL((k_fill<uint64_t>), dim3(1), dim3(64), 0, s, d_count, (uint64_t)0, (uint64_t)1);
        return;
    }
    char* t = reinterpret_cast<char*>(temp);
    uint32_t* flags = reinterpret_cast<uint32_t*>(t);
    uint64_t* pos = reinterpret_cast<uint64_t*>(t + (n * 4 + 255) / 256 * 256);
    char* st = reinterpret_cast<char*>(pos + n + 1);
    st = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(st) + 255) & ~uintptr_t(255));
    hipLaunchKernelGGL(k_unique_flags, dim3(grid1(n)), dim3(256), 0, s, in, n, flags);
    scan_impl<uint32_t, uint64_t>(flags, pos, n, st, s);
    hipLaunchKernelGGL(k_unique_scatter, dim3(grid1(n)), dim3(256), 0, s, in, n, flags, pos, out, d_count);
}

void keys_to_pairs(const uint64_t* keys, uint64_t n, int32_t* pairs, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_keys_to_pairs, dim3(grid1(n)), dim3(256), 0, s, keys, n, reinterpret_cast<int2*>(pairs));
}
void pairs_to_keys(const int32_t* pairs, uint64_t n, uint64_t* keys, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_pairs_to_keys, dim3(grid1(n)), dim3(256), 0, s, reinterpret_cast<const int2*>(pairs), n,
                           keys);
}
__global__ void k_gather_to_host(HostGather g, uint32_t* __restrict__ dst) {
    uint32_t at = 0;
    for (uint32_t k = 0; k < g.n; ++k) {
        for (uint32_t i = threadIdx.x; i < g.len[k]; i += blockDim.x) dst[at + i] = g.src[k][i];
        at += g.len[k];
    }
}
void gather_to_host(const HostGather& g, uint32_t* host_dst, hipStream_t s) {
    if (g.n) hipLaunchKernelGGL(k_gather_to_host, dim3(1), dim3(64), 0, s, g, host_dst);
}
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL((k_fill<uint32_t>), dim3(grid1(n)), dim3(256), 0, s, p, v, n);
}
void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL((k_fill<uint64_t>), dim3(grid1(n)), dim3(256), 0, s, p, v, n);
}

}  // namespace pzk
