// Launchers of the BVH, join and plain-path kernels.
#pragma once

#include "pz_device.h"
#include "pz_kernels.h"

namespace pzk {

// --- LBVH (bvh.hip)
// range_end[i] (n entries): the other end of internal node i's sorted range.
void lbvh_karras(const uint64_t* sorted_keys, uint64_t n, const uint32_t* seg_off, int seg_shift, pzd::Node* nodes,
                 uint32_t* parent_int, uint32_t* parent_leaf, uint32_t* counters, uint32_t* range_end,
                 hipStream_t s);
// Boxes bottom-up from the leaf records in sorted order (gather_leaf_tris
// runs first); max_depth gets the largest segment height.
void lbvh_refit_leaves(const uint64_t* sorted_keys, uint64_t n, const uint32_t* seg_off, int seg_shift,
                       const pzd::LeafTri* leaves, pzd::Node* nodes, const uint32_t* parent_int,
                       const uint32_t* parent_leaf, uint32_t* counters, const uint32_t* range_end,
                       uint32_t* max_depth, hipStream_t s);
void lbvh_refit_ord_boxes(const uint64_t* sorted_keys, const uint32_t* order, uint64_t n, const uint32_t* seg_off,
                          int seg_shift, const uint32_t* obox, pzd::Node* nodes, const uint32_t* parent_int,
                          const uint32_t* parent_leaf, uint32_t* counters, const uint32_t* range_end,
                          uint32_t* max_depth, hipStream_t s);
void gather_leaf_tris(const uint32_t* order, uint64_t n, const float* verts, const uint32_t* idx, const int32_t* t2o,
                      pzd::LeafTri* out, hipStream_t s);

}  // namespace pzk
