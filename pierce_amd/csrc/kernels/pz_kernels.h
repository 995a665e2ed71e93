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
void pairs_to_keys(const int32_t* pairs, uint64_t n, uint64_t* keys, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);

}  // namespace pzk
