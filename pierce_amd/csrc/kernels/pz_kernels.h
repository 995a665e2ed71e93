// Host-callable launchers of the gfx950 kernels (internal interface between
// the C ABI implementation and the .hip translation units).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pzk {

// ------------------------------------------------------------- scan / sort
// Exclusive scan; out has n + 1 entries, out[n] = total.  temp: scan_temp_bytes(n).
size_t scan_temp_bytes(uint64_t n);
void exclusive_scan_u32_u64(const uint32_t* in, uint64_t* out, uint64_t n, void* temp, hipStream_t s);
void exclusive_scan_u32_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, hipStream_t s);
void exclusive_scan_i32_i64(const int32_t* in, int64_t* out, uint64_t n, void* temp, hipStream_t s);

// Stable LSD radix sort of (u64 key, u32 value) on bits [begin_bit, end_bit).
// Sorted data ends in (keys, vals) or (keys_alt, vals_alt): returns true for alt.
size_t radix_sort_temp_bytes(uint64_t n);
bool radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint64_t n,
                      int begin_bit, int end_bit, void* temp, hipStream_t s);
// Keys only (values ignored).
bool radix_sort_keys(uint64_t* keys, uint64_t* keys_alt, uint64_t n, int begin_bit, int end_bit, void* temp,
                     hipStream_t s);

// Remove adjacent duplicates of a sorted u64 array into out; count -> *d_count (u64, device).
void unique_u64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* d_count, void* temp, hipStream_t s);
size_t unique_temp_bytes(uint64_t n);

// Unpack (a << 32 | b) keys into pairs {a, b}.
void keys_to_pairs(const uint64_t* keys, uint64_t n, int32_t* pairs, hipStream_t s);
// Keys below 2^63 sorted as tiles of 4,096 (one workgroup each) merged by
// rank: two launches up to kSortTilesMax keys; true when the sorted keys are
// in keys_alt / vals_alt (more than one tile).  vals may be null.
constexpr uint64_t kSortTileKeys = 4096, kSortTilesMax = 16 * kSortTileKeys;
bool sort_tiles(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint64_t n, hipStream_t s);
// One tile sorted in place with the key count read from *dn (nothing done
// when it exceeds 4,096), and those keys as pairs with their values copied
// beside them (nothing when above cap; vals may be null).  pairs and ovals
// may be pinned host memory: the kernel writes the result there directly.
void sort_tile_counted(uint64_t* keys, uint32_t* vals, const unsigned long long* dn, hipStream_t s);
// sort_tiles with the key count read from *dn (at most cap, else nothing is
// done); the sorted keys and values are always in keys_alt / vals_alt.
void sort_tiles_counted(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                        const unsigned long long* dn, uint64_t cap, hipStream_t s);
void keys_to_pairs_counted(const uint64_t* keys, const uint32_t* vals, const unsigned long long* dn, uint64_t cap,
                           int32_t* pairs, uint32_t* ovals, hipStream_t s);
void pairs_to_keys(const int32_t* pairs, uint64_t n, uint64_t* keys, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// Up to 8 device ranges of 32-bit words, copied back to back into host-pinned
// memory by one small kernel (one launch instead of one copy per range).
struct HostGather {
    static constexpr int kMax = 8;
    const uint32_t* src[kMax];
    uint32_t len[kMax];  // words
    uint32_t n;
};
void gather_to_host(const HostGather& g, uint32_t* host_dst, hipStream_t s);

// Optional per-kernel timing: when enabled, launches bracketed by KernelTimer
// record a HIP event pair on their stream; totals are collected after the
// stream is synchronised.
void prof_enable(bool on);
bool prof_enabled();
void prof_record(const char* name, hipEvent_t start, hipEvent_t stop);
hipEvent_t prof_event();
// Sum (ms) and count of completed launches of `name` since the last reset.
void prof_query(const char* name, double* total_ms, uint64_t* count);
void prof_reset();

struct KernelTimer {
    const char* name;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(const char* n, hipStream_t st) : name(n), s(st) {
        if (prof_enabled()) {
            a = prof_event();
            b = prof_event();
            (void)hipEventRecord(a, s);
        }
    }
    ~KernelTimer() {
        if (a) {
            (void)hipEventRecord(b, s);
            prof_record(name, a, b);
        }
    }
};
void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);

// ---------------------------------------------------------- instancing
// One instance of a library mesh in a set (instance.hip): the library's runs
// of vertices, triangles and edges, where they go in the set, the object id
// they get, and the map x' = A x + t (A row-major in m[0..8], t in m[9..11]).
struct InstanceRec {
    uint32_t sv, nv, st, nt, se, ne;  // library: first vertex / count, first triangle / count, first edge / count
    uint32_t dv, dt, de;              // set: first vertex, first triangle, first edge
    uint32_t obj;                     // object id in the set
    float m[12];
};
void instantiate(const InstanceRec* inst, uint32_t n, const float* lv, const uint32_t* lidx, const float* les,
                 const float* lee, float* v, uint32_t* idx, int32_t* t2o, float* es, float* ee, int32_t* eo,
                 hipStream_t s);

}  // namespace pzk
