// Launchers of the join kernels (join.hip).
#pragma once

#include "pz_kernels_bvh.h"

namespace pzd {
// Device view of one set's per-object acceleration data for a two-level join.
struct SetView {
    const BNode* nodes;        // per-object BVHs (own-box nodes)
    const LeafTri* leaves;
    const uint32_t* seg_off;   // per needed object: first sorted triangle (n + 1 entries)
    const EdgeRec* edges;      // edges grouped per needed object
    const uint32_t* eseg_off;  // per needed object: first edge (n + 1 entries)
    const uint32_t* loc;       // object id -> needed-object index
    const uint32_t* obox;      // ordered-uint boxes by object id (6 per object)
};
// One task of the narrow phase's work queue: candidate pair `pair` still to
// be walked below a (mesh1 node, mesh2 node) pair, given by both nodes' child
// links (links as in BNode; a side that is a single leaf has x1 / y1 =
// 0x7FFFFFFF).
struct DualTask {
    uint32_t pair;
    int32_t x0, x1, y0, y1;
    uint32_t pad[3];
};
// One set's geometry as the pair sieve reads it: the dataset's arrays and the
// per-object triangle ranges and boxes of the join.
struct SieveSet {
    const float* v;
    const uint32_t* idx;
    const uint32_t* first_tri;
    const uint32_t* tri_cnt;
    const uint32_t* last_tri;
    const uint32_t* obox;
    const uint32_t* vmin;  // objects' vertex runs [vmin, vmax] (null: no runs)
    const uint32_t* vmax;
};
}  // namespace pzd

namespace pzk {

// Device counters of one join (u64 each): 0 tests, 1 decided in fp64,
// 2 decided exactly, 3 fp32 misclassified, 4 edges traced, 5 hash-set slots
// probed, 6 probed slots holding another pair, 7 parity rays cast, 8 triangles
// they crossed, 9 rays crossing at least one, 10 rays ending inside, 11 their
// ray-triangle tests, 12 dual-tree walks, 13 candidate pairs proven touching,
// 14 candidate pairs the pair sieve decided, 15 triangles those pairs' objects
// hold (what the sieve read for them).
constexpr int kStatSlots = 16;
constexpr int kStatParityRays = 7, kStatWalks = 12, kStatTouching = 13, kStatSieved = 14, kStatSieveTris = 15;

using pzd::DualTask;
using pzd::EdgeRec;
using pzd::LeafTri;
using pzd::BNode;
using pzd::Node;
using pzd::SetView;
using pzd::SieveSet;

// Stack depth for a BVH of the given height (0: unsupported).
int stack_for_height(uint32_t h);

void object_stats(const float* verts, uint64_t nv, const uint32_t* idx, const int32_t* t2o, uint64_t nt, int32_t nobj,
                  uint32_t* obox, uint32_t* tri_cnt, uint32_t* first_tri, uint32_t* last_tri, hipStream_t s);

// Topology of a dataset, once at upload: per object the first and last
// triangle, the triangle count and the least and greatest vertex index; per
// vertex the referencing object (owner, nv entries; a vertex referenced by
// two objects sets *clash); *referenced = vertices some object references.
void topology_scan(const uint32_t* idx, const int32_t* t2o, uint64_t nt, uint64_t nv, int32_t nobj, uint32_t* first,
                   uint32_t* last, uint32_t* cnt, uint32_t* vmin, uint32_t* vmax, int32_t* owner, uint32_t* clash,
                   unsigned long long* referenced, hipStream_t s);
// Vertices per chunk of object_bounds_runs.
uint32_t run_chunk_vertices();
// Per-object boxes of a dataset whose objects own disjoint, gap-free vertex
// runs: chunks[c] = (object, first vertex) of at most run_chunk_vertices()
// vertices of one run, which ends at vmax[object].
void object_bounds_runs(const float* verts, const uint2* chunks, uint64_t nch, const uint32_t* vmax, int32_t nobj,
                        uint32_t* obox, hipStream_t s);

// --- small-object LDS builder (build_small.hip)
// Objects with at most kSmallMax triangles stored contiguously are built
// entirely in LDS by one workgroup each; the rest use the global pipeline.
// The LDS builder runs one launch per size class so each launch sizes its
// LDS (and picks its workgroup size) for the class.
constexpr uint32_t kSmallMax = 2048;
constexpr int kSmallClasses = 3;
constexpr uint32_t kClassMax[kSmallClasses] = {128, 512, 2048};
// big[o] flags; cls[k * (nobj + 1) + o] flags per class; totals[5 + k] = max
// triangle count of class k (atomicMax: zero totals first).
void need_split(const uint32_t* need, const uint32_t* tri_cnt, const uint32_t* first_tri, const uint32_t* last_tri,
                int32_t nobj, uint32_t* big, uint32_t* cls, uint32_t* totals, hipStream_t s);
// loc[o]: big objects first (loc_big), then class 0, 1, 2 in id order
// (loc_cls: exclusive scans of the class flags, nobj + 1 entries each);
// totals[0] = big objects, totals[2 + k] = objects of class k.
void merge_locs(const uint32_t* big, const uint32_t* loc_big, const uint32_t* loc_cls, int32_t nobj, uint32_t* loc,
                uint32_t* totals, hipStream_t s);
// totals[1] = seg_off[totals[0]] (triangles of the big objects).
void big_prefix(const uint32_t* seg_off, uint32_t* totals, hipStream_t s);
// Dynamic LDS bytes of one workgroup for objects of at most cap triangles.
size_t build_small_lds(uint32_t cap);
// One workgroup of `threads` (128, 256, 512 or 1024) per object l in
// [l_begin, l_begin + count), all of at most cap triangles.  false: bad cap or threads.
bool build_small(const float* verts, const uint32_t* idx, const int32_t* obj_of_loc,
                 const uint32_t* first_tri, const uint32_t* tri_cnt, const uint32_t* obox, const uint32_t* seg_off,
                 uint32_t l_begin, uint32_t count, uint32_t cap, int threads, pzd::BNode* nodes, pzd::LeafTri* leaves,
                 uint32_t* height, const char* timer, hipStream_t s);
// Big objects (segments [0, nbig)) are built in chunks of kChunk consecutive
// triangles, each chunk by the LDS builder, joined by a balanced top tree per
// object whose K - 1 nodes take the node slots seg_off[l] + j * kChunk
// (j < K - 1; the chunks' own nodes start one slot later), so the object's
// root stays at its segment's first node.  The chunks are runs of each
// object's own triangle order (sorted in LDS), or of a Morton order of the
// set's big objects (keys / order of the global pipeline's sort) when some
// are not contiguous or an object's own order is not spatially coherent.
// Chunk / workgroup sizes measured on 2,000 vs 2,000 tubes with own-order
// chunks: 2,048 / 1,024 build 38.9 ms + walk 31.0 ms, 1,024 / 512 30.9 +
// 32.0 ms, 512 / 256 28.5 + 34.5 ms, 1,024 / 256 34.2 + 31.9 ms, 1,024 /
// 1,024 44.9 + 32.0 ms.
constexpr uint32_t kChunk = 1024;   // triangles per chunk (<= kSmallMax)
constexpr int kChunkThreads = 512;  // one chunk build's workgroup
struct ChunkRoot {
    float lo[3], hi[3];
    int32_t link, pad;  // as in Node: >= 0 node, < 0 leaf ~index
};
// Scratch of build_big_chunked: chunk counts and offsets (nbig + 1 each) and
// two ChunkRoot arrays of big_chunk_bound entries.
uint64_t big_chunk_bound(uint64_t nbig, uint64_t nbig_tri);
// sorted_keys = (segment << mbits) | the top mbits of the 30-bit Morton code,
// order = the triangle at each sorted position.  Null sorted_keys: every big
// object stores its triangles contiguously (from first_tri), its chunks are
// runs of its own triangle order, each sorted in LDS, and the objects whose
// order is not spatially coherent (chunk boxes against the object's box in
// obox) are counted into *incoherent: their trees are valid but loose, and
// the caller rebuilds such a set from the sorted order.
void build_big_chunked(const float* verts, const uint32_t* idx, const int32_t* obj_of_loc, const uint32_t* seg_off,
                       const uint32_t* first_tri, const uint32_t* obox, uint32_t nbig, uint64_t nbig_tri,
                       const uint64_t* sorted_keys, int mbits, const uint32_t* order, uint32_t* chunk_cnt,
                       uint32_t* chunk_off, void* scan_tmp, ChunkRoot* ent0, ChunkRoot* ent1, pzd::BNode* nodes,
                       pzd::LeafTri* leaves, uint32_t* height, uint32_t* incoherent, hipStream_t s);
void edge_count(const int32_t* eo, uint64_t ne, int32_t nobj, uint32_t* edge_cnt, hipStream_t s);
void scene_box(const uint32_t* obox, int32_t nobj, uint32_t* scene, hipStream_t s);
// keys = (part_of[o] << 30) | Morton of the box centre (part 0 when part_of is null)
void object_keys(const uint32_t* obox, int32_t nobj, const uint32_t* scene, const uint32_t* part_of, uint64_t* keys,
                 uint32_t* vals, hipStream_t s);
// Broad phase: mesh1 object boxes against the mesh2 object LBVH (stack of 64).
// With parts (batched joins), mesh1 object a only meets the mesh2 objects of
// its part: sorted positions [boff[part_of_a[a]], boff[part + 1]).
// nodesB == nullptr: no LBVH; every mesh2 object of the part is tested
// (object ids [boff[part], boff[part + 1]), or all nB without parts).
void broad_count(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, const Node* nodesB, const uint32_t* orderB,
                 int32_t nB, const uint32_t* part_of_a, const uint32_t* boff, uint32_t* cnt, hipStream_t s);
void broad_write(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, const Node* nodesB, const uint32_t* orderB,
                 int32_t nB, const uint32_t* part_of_a, const uint32_t* boff, const uint64_t* off, uint64_t* out,
                 hipStream_t s);
// Pairs into out[0, cap) with *cursor = their number (which may exceed cap);
// cnt (nA), off (nA + 1) and scan_tmp (scan_temp_bytes(nA)) are scratch.
void broad_append(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, const Node* nodesB,
                  const uint32_t* orderB, int32_t nB, const uint32_t* part_of_a, const uint32_t* boff, uint64_t* out,
                  unsigned long long* cursor, uint64_t cap, uint32_t* cnt, uint64_t* off, void* scan_tmp,
                  hipStream_t s);
// Broad phase over a hashed grid of the mesh2 object boxes (non-batched sets):
// cells are three times the boxes' mean extent per axis, hashed into nb (a power of two)
// buckets; boxes over more than 27 cells are kept in `big` and tested by every
// query.  counts and start hold nb + 1 words, entries grid_entry_bound(nB) (ebox twice that),
// big nB, nbig 1 word, stats 4 floats.  Same outputs as the broad_* calls.
struct ObjGrid {
    uint32_t* counts;
    uint32_t* start;
    uint4* entries;
    float4* ebox;  // two per entry: the box of the entry's object
    uint32_t* big;
    uint32_t* nbig;
    float* stats;
    uint32_t nb;
};
uint64_t grid_entry_bound(int32_t nB);
void grid_build(const uint32_t* oboxB, int32_t nB, const ObjGrid& g, void* scan_tmp, hipStream_t s);
void grid_count(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, int32_t nB, const ObjGrid& g, uint32_t* cnt,
                hipStream_t s);
void grid_write(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, int32_t nB, const ObjGrid& g,
                const uint64_t* off, uint64_t* out, hipStream_t s);
void grid_append(const uint32_t* oboxA, int32_t nA, const uint32_t* oboxB, int32_t nB, const ObjGrid& g, uint64_t* out,
                 unsigned long long* cursor, uint64_t cap, hipStream_t s);
// What pair_prep zeroes besides writing meta: per-pair u32 flag arrays of
// np + 1 words, single u64 counters and a block of u32 words (null: none).
struct PairPrepZero {
    static constexpr int kFlags = 4, kCounters = 6;
    uint32_t* flags[kFlags];
    unsigned long long* counters[kCounters];
    uint32_t* words;
    uint32_t nwords;
};
// pair_meta's meta and the zeroes of z, in one launch.
void pair_prep(const uint64_t* pairs, uint64_t np, const uint32_t* part_of_a, const uint8_t* meta_of_part,
               uint8_t uniform, uint8_t* meta, const PairPrepZero& z, hipStream_t s);
// meta[p] = predicate | dir_mask << 2, from the pair's part or `uniform`.
void pair_meta(const uint64_t* pairs, uint64_t np, const uint32_t* part_of_a, const uint8_t* meta_of_part,
               uint8_t uniform, uint8_t* meta, hipStream_t s);
void mark_needed(const uint64_t* pairs, uint64_t np, uint32_t* needA, uint32_t* needB, hipStream_t s);
void needed_layout(const uint32_t* need, const uint32_t* loc, int32_t nobj, const uint32_t* tri_cnt,
                   const uint32_t* edge_cnt, int32_t* obj_of_loc, uint32_t* seg_tri_cnt, uint32_t* seg_edge_cnt,
                   hipStream_t s);
void tri_keys_two_level(const float* verts, uint64_t nv, const uint32_t* idx, const int32_t* t2o, uint64_t nt,
                        int32_t nobj, const uint32_t* need, const uint32_t* loc, const uint32_t* obox,
                        const uint32_t* seg_off, const uint32_t* first_tri, const uint32_t* last_tri,
                        const uint32_t* tri_cnt, uint32_t* cursor, uint64_t* keys, uint32_t* vals, int mbits,
                        hipStream_t s);
void tri_keys_global(const float* verts, const uint32_t* idx, uint64_t nt, const uint32_t* scene, uint64_t* keys,
                     uint32_t* vals, hipStream_t s);
void edge_scatter(const float* es, const float* ee, const int32_t* eo, uint64_t ne, int32_t nobj,
                  const uint32_t* need, const uint32_t* loc, const uint32_t* eseg_off, uint32_t* cursor,
                  EdgeRec* out, hipStream_t s);
void task_counts(const uint64_t* pairs, uint64_t np, const uint32_t* locA, const uint32_t* esegA,
                 const uint32_t* locB, const uint32_t* esegB, const uint8_t* meta, uint32_t* counts, hipStream_t s);
bool narrow(const uint64_t* pairs, uint64_t np, const uint64_t* task_off, uint64_t ntasks, const SetView& A,
            const SetView& B, uint32_t* found, const uint8_t* meta, bool f32, int allow_f32, uint32_t height,
            unsigned long long* stats, hipStream_t s);
// One round of the dual-tree narrow phase as a persistent work queue.
// Waves fetch `chunk` tasks at a time from *next_task (zeroed by the caller)
// until ntasks are taken.  tasks_in == nullptr: task i is candidate pair i
// from the two object roots.  A wave whose LDS queue (`queue` node pairs:
// 256, 512, 1024 or 1536) would overflow moves its oldest half to tasks_out
// (reserved through *nout, at most out_cap) for the next round; if that is
// full too, fallback[p] = 1 and the pair must be re-run with narrow().
// With f32 (and allow_f32) the walk uses the fp32 filter alone: leaf pairs
// it cannot decide are appended to esc_list (pair, mesh1 leaf, mesh2 leaf;
// *esc_count zeroed by the caller, at most esc_cap kept, a pair that finds
// the list full goes to fallback) for escalate().
// `blocks` = persistent grid size.  false: bad queue length.
bool narrow_dual(const uint64_t* pairs, const SetView& A, const SetView& B, uint32_t* found, uint32_t* fallback,
                 const uint8_t* meta, bool f32, int allow_f32, int queue, const DualTask* tasks_in, uint64_t ntasks,
                 unsigned long long* next_task, int chunk, DualTask* tasks_out, unsigned long long* nout,
                 uint64_t out_cap, uint4* esc_list, unsigned long long* esc_count, uint64_t esc_cap, unsigned blocks,
                 unsigned long long* stats, hipStream_t s);
// Decides the leaf pairs narrow_dual left open with the fp64 and exact
// predicates: found[pair] = 1 on a hit (entries min(*count, cap)).
void escalate(const uint4* list, const unsigned long long* count, uint64_t cap, const SetView& A, const SetView& B,
              const uint8_t* meta, uint32_t* found, unsigned long long* stats, hipStream_t s);
// fallback[tasks[i].pair] = 1 for every task (work the rounds did not finish).
void mark_task_pairs(const DualTask* tasks, uint64_t n, uint32_t* fallback, hipStream_t s);
// Waves per block of narrow_dual.
int narrow_dual_waves();
// Compact pairs with flag[p] != 0 into out (count -> *count).
void compact_flagged(const uint64_t* pairs, uint64_t np, const uint32_t* flag, uint64_t* out, uint32_t* out_index,
                     unsigned long long* count, hipStream_t s);
// found[index[i]] |= sub_found[i]
void scatter_found(const uint32_t* sub_found, const uint32_t* index, uint64_t n, uint32_t* found, hipStream_t s);
bool parity(const uint64_t* pairs, uint64_t np, const uint32_t* found, const SetView& A, const SetView& B,
            const float* vA, const uint32_t* iA, const uint32_t* ftA, const float* vB, const uint32_t* iB,
            const uint32_t* ftB, const uint8_t* meta, uint32_t* inside, uint32_t height,
            unsigned long long* stats, hipStream_t s);
// Containment-ray tracking: for each source object with triangles, the +Z
// ray from its first vertex through the set-wide BVH (nodes, leaves, ntri
// triangles, height) of the other set.  Per source: ids/hits[src *
// max_targets + i] for the first `count` objects met, `overflow` hits on
// objects beyond max_targets, iters = 1 if a ray was cast.  false: the BVH is
// too tall for the traversal stack.
bool track_rays(const float* vS, const uint32_t* iS, const uint32_t* ftS, const uint32_t* tcS, int32_t nsrc,
                const Node* nodes, const LeafTri* leaves, uint64_t ntri, uint32_t height, int max_targets,
                int32_t* ids, uint32_t* hits, uint32_t* count, uint32_t* overflow, uint32_t* iters,
                unsigned long long* stats, hipStream_t s);
// Set-wide containment rays -> pair keys (see k_parity_keys): keys[off[s] + i]
// for each entry of each source's list, ~0 for an even hit count.
void parity_keys(const int32_t* ids, const uint32_t* hits, const uint32_t* count, const uint64_t* off, int32_t nsrc,
                 int max_targets, bool src_is_mesh2, uint64_t* keys, hipStream_t s);
// Appends the keys of `in` (other than ~0) that the sorted unique `ref` lacks,
// with type `type` (count: device cursor).
void key_difference(const uint64_t* in, uint64_t n, const uint64_t* ref, uint64_t nref, uint64_t* out,
                    uint32_t* out_type, uint32_t type, unsigned long long* count, hipStream_t s);
void select_pairs(const uint64_t* pairs, uint64_t np, const uint32_t* found, const uint32_t* inside,
                  const uint8_t* meta, int include_overlap, uint64_t* out_keys, uint32_t* out_type,
                  unsigned long long* count, unsigned long long* touching, hipStream_t s);
bool trace_global(int mode, const float* es, const float* ee, const int32_t* eo, uint64_t ne, const Node* nodes,
                  const LeafTri* leaves, uint64_t ntri, uint32_t height, int swap, bool f32, int allow_f32,
                  uint32_t* counts, const uint64_t* off, uint64_t* out, unsigned long long* table, uint64_t mask,
                  uint32_t* overflow, unsigned long long* stats, hipStream_t s);
void compact_table(const unsigned long long* table, uint64_t size, uint64_t* out, uint64_t cap,
                   unsigned long long* count, hipStream_t s);
void fill_ull(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s);

// --- pair sieve (sieve.hip)
// Narrow phase (and parity rays) of candidate pairs whose objects hold at most
// 8,192 contiguous triangles, one workgroup per pair, without per-object
// trees: found[p] = 1 for a proven hit, defer[p] = 1 for a pair left to the
// tree path (nothing else written for it).  With f32 (and allow_f32) open triangle pairs go to esc as (pair,
// mesh1 triangle, mesh2 triangle) for escalate_tris.
// run_desc holds 5 np uint4 (the vertex-run sieve's per-pair descriptors), gwork
// np pair indices, ctr 8 words, zeroed by the caller; blocks = compute units.
// rays (rays_cap entries) and nrays: the parity-ray candidates the first pass
// records for sieve_parity; ray_overflow (np flags) marks pairs some of whose
// candidates did not fit.
void pair_sieve(const uint64_t* pairs, uint64_t np, const SieveSet& A, const SieveSet& B, const uint8_t* meta,
                bool f32, int allow_f32, uint32_t* found, uint32_t* defer, uint4* run_desc, uint32_t* gwork,
                uint32_t* ctr, uint2* rays, unsigned long long* nrays, uint64_t rays_cap, uint32_t* ray_overflow,
                uint4* esc, unsigned long long* esc_count, uint64_t esc_cap, unsigned blocks,
                unsigned long long* stats, hipStream_t s);
// Parity rays (parity()'s semantics) of the pairs the sieve decided that do
// not touch, from the candidates its first pass recorded; run after
// escalate_tris.
// cross: 2 np counters, zeroed.
void sieve_parity(const uint64_t* pairs, uint64_t np, const SieveSet& A, const SieveSet& B, const uint8_t* meta,
                  const uint32_t* found, const uint32_t* defer, const uint2* rays, const unsigned long long* nrays,
                  uint64_t rays_cap, const uint32_t* overflow, uint32_t* cross, uint32_t* inside,
                  unsigned long long* stats, hipStream_t s);
void escalate_tris(const uint4* list, const unsigned long long* count, uint64_t cap, const SieveSet& A,
                   const SieveSet& B, const uint8_t* meta, uint32_t* found, unsigned long long* stats, hipStream_t s);

}  // namespace pzk
