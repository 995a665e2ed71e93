/*
 * pierce_hip.h - C ABI of the MI355X (gfx950) spatial-join engine.
 *
 * Plain C types only: pointers, sizes, enums as int32.  Device work runs on
 * the context's HIP stream; every call blocks until its results are on the
 * host unless stated otherwise.  No call exits the process: failures return a
 * negative pz_status and pz_last_error() describes the most recent failure of
 * the calling thread.
 *
 * Reference interfaces these entry points replace (paths relative to the
 * Pierce source tree):
 *   - dataset load/upload: pierce/query/src/runtime/GeometryIO.cpp:6-20,
 *     pierce/query/src/geometry/GeometryUploader.cpp:12-34,
 *     pierce/query/src/geometry/PrecomputedEdgeData.cpp:5-29
 *   - index build: pierce/query/src/optix/OptixAccelerationStructure.cpp:22-68
 *   - query launches: MeshOverlapEdgesLauncher.h:49-51,
 *     MeshIntersectionLauncher.h:17-22, MeshContainmentLauncher.h:26-30
 *     (all under pierce/query/src/raytracing/)
 *   - extern "C" helpers: pierce/query/src/cuda/mesh_query_deduplication.h:6-18,
 *     pierce/query/src/cuda/scan_utils.h:7-9
 *   - selectivity estimate: pierce/query/src/cuda/estimated_overlap.h:6-13
 */
#ifndef PIERCE_HIP_H
#define PIERCE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef int32_t pz_status;
#define PZ_OK            0
#define PZ_E_INVALID    -1  /* bad argument */
#define PZ_E_HIP        -2  /* HIP runtime error */
#define PZ_E_IO         -3  /* file missing / malformed / wrong version */
#define PZ_E_OVERFLOW   -4  /* a sized buffer was too small (recoverable) */
#define PZ_E_NOMEM      -5  /* device or host allocation failed */
#define PZ_E_NO_EDGES   -6  /* dataset lacks the precomputed edge payload */
#define PZ_E_LIMIT      -7  /* input beyond a documented engine limit */
#define PZ_E_NODEVICE   -8  /* no HIP device available */

/* Most recent failure message of the calling thread ("" if none). */
const char* pz_last_error(void);

/* ------------------------------------------------------------- host types */
/* One result pair; layout of MeshQueryResult (pierce/common/include/common.h:34-37). */
typedef struct {
    int32_t object_id_mesh1;
    int32_t object_id_mesh2;
} pz_pair;

/* Sparse grid cell of the .pre payload: GridCell (pierce/common/include/GridCell.h:5-10)
 * preceded by its int3 index (SparseGridEntry, pierce/common/include/Geometry.h:15-18). */
typedef struct {
    int32_t ix, iy, iz;
    uint32_t center_count;
    uint32_t touch_count;
    float avg_size_mean;
    float vol_ratio;
} pz_grid_cell;

/* Borrowed view of one dataset: the .pre v4 payload
 * (pierce/common/include/BinaryIO.h:14-151).  All arrays are AoS exactly as in
 * the file: float3 vertices, uint3 indices, int32 triangle->object map, float3
 * edge starts/ends and int32 edge source objects. */
typedef struct {
    const float*    vertices;        /* 3 * num_vertices */
    uint64_t        num_vertices;
    const uint32_t* indices;         /* 3 * num_triangles */
    uint64_t        num_triangles;
    const int32_t*  triangle_to_object; /* num_triangles */
    const float*    edge_starts;     /* 3 * num_edges (may be NULL if num_edges == 0) */
    const float*    edge_ends;       /* 3 * num_edges */
    const int32_t*  edge_objects;    /* num_edges */
    uint64_t        num_edges;
    const pz_grid_cell* grid_cells;  /* num_grid_cells, may be NULL */
    uint64_t        num_grid_cells;
    float           grid_cell_size;  /* 0 if no grid */
} pz_mesh_view;

/* Host-owned dataset (generated or read from a .pre file). */
typedef struct pz_host_mesh pz_host_mesh;

pz_status pz_host_mesh_view(const pz_host_mesh* mesh, pz_mesh_view* out);
void      pz_host_mesh_free(pz_host_mesh* mesh);

/* Read a .pre v4 file (pierce/common/include/BinaryIO.h:84-151).  Wrong magic or
 * version -> PZ_E_IO (the reference returns empty geometry and prints an error). */
pz_status pz_pre_read(const char* path, pz_host_mesh** out);
/* Write a .pre v4 file byte-compatible with BinaryIO.h:32-82. */
pz_status pz_pre_write(const char* path, const pz_mesh_view* mesh);

/* Read a Wavefront .obj file into a host mesh, one object per shape (`o` /
 * `g` group with faces), as pierce/preprocess/src/ObjMeshDatasetLoader.cpp:14-168
 * loads it through tinyobjloader with triangulation: polygon faces become
 * triangle fans, negative face indices count back from the last vertex,
 * shapes without faces are dropped, vertices outside the vertex list drop the
 * triangles that use them.  An object's vertices are stored in ascending file
 * order (the reference's order is std::unordered_set iteration order).
 * No edges and no grid are attached.  Not a .obj path / unreadable /
 * malformed `v` or `f` line / no triangles -> PZ_E_IO. */
pz_status pz_obj_read(const char* path, pz_host_mesh** out);
/* Translate every object by its own offset drawn uniformly from [lo, hi]^3
 * (pierce/preprocess/src/pierce_preprocess.cpp:243-283, which uses [-15, 15]
 * and an unseeded generator); deterministic for a given seed. */
pz_status pz_host_mesh_shuffle(pz_host_mesh* mesh, uint64_t seed, float lo, float hi);

/* Copy a view into a new host mesh (owned arrays). */
pz_status pz_host_mesh_from_view(const pz_mesh_view* view, pz_host_mesh** out);

/* Replace the mesh's edge payload with the unique per-object edges of
 * EdgePreprocessor::extractEdges (pierce/preprocess/src/EdgePreprocessor.cpp:44-111):
 * 1e-5 quantised endpoint keys, first-seen float endpoints, key order. */
pz_status pz_host_mesh_extract_edges(pz_host_mesh* mesh);
/* Replace the mesh's grid payload with generateGridStats
 * (pierce/preprocess/src/pierce_preprocess.cpp:49-167); cells in ascending
 * (z, y, x) order, objects visited in ascending id order. */
pz_status pz_host_mesh_compute_grid(pz_host_mesh* mesh, float cell_size);

/* ------------------------------------------------------------- generators */
#define PZ_GEN_ELLIPSOIDS     0  /* icosphere ellipsoids at one subdivision level */
#define PZ_GEN_SPHERES_MIXED  1  /* spheres, level l in [level_min, level_max] with P(l) ~ 2^-l */
#define PZ_GEN_CUBES          2  /* axis-aligned cubes, 12 triangles */
#define PZ_GEN_TUBES          3  /* branching generalised-cylinder tubes */
#define PZ_GEN_QUERY_MIX      4  /* per mesh: 50% ellipsoids (level uniform in [level_min, level_max]),
                                    30% cubes (edge U[1, 10]), 20% tubes (radius U[0.5, 3]) */

#define PZ_EDGES_NONE     0
#define PZ_EDGES_GROUPED  1  /* each mesh's unique edges, meshes in id order (fast) */
#define PZ_EDGES_PRE      2  /* EdgePreprocessor key order over the whole set (a .pre file's order) */

typedef struct {
    int32_t  kind;            /* PZ_GEN_* */
    int32_t  edges_mode;      /* PZ_EDGES_* */
    uint64_t count;           /* number of meshes */
    uint64_t seed;
    int32_t  level_min;       /* icosphere subdivision (ellipsoids: uniform in [min, max]) */
    int32_t  level_max;
    float    radius_min;      /* ellipsoid radii / sphere radius / cube edge / tube radius */
    float    radius_max;
    float    box;             /* centres uniform in [0, box)^3 */
    float    jitter;          /* vertex jitter sigma as a fraction of the mean radius */
    /* scaled copies of another generated set (containment workloads) */
    float    copy_fraction;
    float    copy_scale_min;
    float    copy_scale_max;
    uint64_t copy_source_count;
    uint64_t copy_source_seed;
    /* tubes */
    int64_t  tri_median;      /* lognormal median of triangle counts */
    float    tri_sigma;       /* lognormal sigma */
    int64_t  tri_min;
    int64_t  tri_max;
    int32_t  ring_segments;
    float    step;
    float    branch_prob;
    int32_t  max_branches;
    float    degenerate_fraction; /* fraction replaced by near-coplanar/offset copies of the source set */
    int32_t  reserved[4];
} pz_gen_params;

/* Fill *p with the defaults of a kind (ellipsoids at level 2, radii [1,3], box 100). */
void      pz_gen_defaults(int32_t kind, pz_gen_params* p);
pz_status pz_generate(const pz_gen_params* params, pz_host_mesh** out);

/* ---------------------------------------------------------------- context */
typedef struct pz_ctx pz_ctx;
typedef struct pz_dataset pz_dataset;

/* Number of visible HIP devices (0 when none; never fails). */
int32_t   pz_device_count(void);
pz_status pz_ctx_create(int32_t device, pz_ctx** out);
void      pz_ctx_destroy(pz_ctx* ctx);
/* Repeated joins of the same datasets replay their fixed-shape segments
 * (object statistics, object index) from HIP graphs captured on the second
 * join; on by default.  No reference counterpart. */
void      pz_ctx_set_graphs(pz_ctx* ctx, int32_t on);
/* Free and total HBM of the context's device (sizing from free memory,
 * pierce/query/src/applications/pierce_overlap_main.cpp:474-489). */
pz_status pz_ctx_mem_info(pz_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* Return the context's cached device work buffers (they grow to the largest
 * join seen and are kept for the next one) and drop its captured graphs.
 * Uploaded datasets are not touched; datasets instanced into a slot of the
 * context (pz_library_instantiate) lose their arrays.  For alternating very large joins whose work
 * buffers do not fit in HBM together.  No reference counterpart. */
pz_status pz_ctx_trim(pz_ctx* ctx);

/* Upload one dataset to HBM (GeometryUploader::upload + PrecomputedEdgeData).
 * A dataset without edges is rejected with PZ_E_NO_EDGES
 * (pierce/query/src/runtime/GeometryIO.cpp:11-20). */
pz_status pz_dataset_upload(pz_ctx* ctx, const pz_mesh_view* view, pz_dataset** out);
pz_status pz_dataset_load_pre(pz_ctx* ctx, const char* path, pz_dataset** out);
void      pz_dataset_release(pz_dataset* ds);
/* out[0..4] = vertices, triangles, edges, objects (max id + 1), grid cells */
pz_status pz_dataset_info(const pz_dataset* ds, uint64_t out[5]);
/* Layout facts found at upload.  PZ_DATASET_VERTEX_RUNS: every object owns a
 * disjoint, gap-free run of vertices (ordered by object id) referenced by its
 * triangles alone, so a join computes the object boxes from those runs (12 B
 * per vertex) rather than from every triangle's corners.  Diagnostics; no
 * reference counterpart (the reference's GAS build reads every triangle). */
#define PZ_DATASET_VERTEX_RUNS 1u
pz_status pz_dataset_flags(const pz_dataset* ds, uint32_t* flags);

/* ------------------------------------------------------------------- join */
#define PZ_PRED_OVERLAP       0  /* surfaces intersect (pierce_overlap*) */
#define PZ_PRED_INTERSECTION  1  /* overlap or one solid contains the other (pierce_intersection) */
#define PZ_PRED_CONTAINMENT   2  /* mesh2 object strictly inside mesh1 object, no contact (pierce_containment) */

#define PZ_DIR_BOTH       0
#define PZ_DIR_1TO2       1  /* mesh1 edges against mesh2 triangles only */
#define PZ_DIR_2TO1       2

#define PZ_STRATEGY_TWO_LEVEL 0  /* object broad phase + per-object BVHs, early exit per pair */
#define PZ_STRATEGY_GLOBAL    1  /* every edge traced through a set-wide BVH (reference shape) */

#define PZ_PREC_FP32_FILTERED 0  /* fp32 filter, fp64 filter, exact fallback */
#define PZ_PREC_FP64          1  /* fp64 filter, exact fallback */

#define PZ_SIZING_TWO_PASS 0  /* exact count pass, scan, write pass */
#define PZ_SIZING_ESTIMATE 1  /* grid selectivity estimate (estimated_overlap.cu:29-40) */
#define PZ_SIZING_LEARNED  2  /* MLP estimate (needs pz_join_desc.model); the buffer holds the
                                 prediction raised by the model's calibrated margin */
#define PZ_SIZING_FIXED    3  /* capacity given by pz_join_desc.capacity */

#define PZ_DEDUP_SORT 0  /* GLOBAL strategy: raw hits sorted + uniqued */
#define PZ_DEDUP_HASH 1  /* GLOBAL strategy: open-addressing hash set, then compaction */

typedef struct pz_mlp pz_mlp;

typedef struct {
    int32_t predicate;
    int32_t direction;          /* overlap only */
    int32_t strategy;
    int32_t precision;
    int32_t sizing;
    int32_t dedup;              /* GLOBAL strategy only */
    int32_t include_overlap_pairs; /* containment: also return overlap pairs (type 1) */
    int32_t max_hits_per_edge;  /* 0 = unlimited (the reference caps at overlap_max_iterations) */
    uint64_t capacity;          /* FIXED: buffer slots; ESTIMATE/LEARNED: 0 = computed */
    float    load_factor;       /* hash sets: slots = expected / load_factor (0 -> 0.5) */
    float    margin;            /* ESTIMATE: buffers hold expected * (1 + margin) (0 -> 0.25) */
    float    gamma;             /* grid estimator exponent (0 -> 0.8) */
    float    epsilon;           /* grid estimator size epsilon (0 -> 0.001) */
    int32_t  no_alpha_correction;
    int32_t  keep_device_result; /* reserved, must be 0 */
    const pz_mlp* model;        /* PZ_SIZING_LEARNED */
    int32_t  narrow_mode;       /* two-level: 0 = pair sieve (candidate pairs of objects with at most
                                   8,192 contiguous triangles, decided in LDS without trees), the
                                   other pairs through per-object trees as in 2; 1 = per-object
                                   trees, per-edge kernel only; 2 = per-object trees, dual-tree walk
                                   with per-edge fallback */
    int32_t  broad_mode;        /* two-level: 0 = auto (box scan for <= 4096 mesh2 objects, else the
                                   hashed grid for a single join and the object LBVH walk for a
                                   batch), 1 = object LBVH walk, 2 = box scan (every mesh1 box
                                   against every mesh2 box), 3 = hashed grid of the mesh2 boxes
                                   (single joins only; pz_join_batch rejects it) */
    int32_t  reserved[6];
} pz_join_desc;

/* Fill *d with the defaults: overlap, both directions, TWO_LEVEL, FP32_FILTERED, TWO_PASS. */
void pz_join_desc_defaults(pz_join_desc* d);

typedef struct {
    char   name[48];
    double ms;
} pz_phase;

#define PZ_MAX_PHASES 24

typedef struct {
    uint64_t num_pairs;
    pz_pair* pairs;             /* host, library-owned; sorted by (mesh1, mesh2), unique */
    uint8_t* pair_types;        /* containment + include_overlap_pairs: 0 = containment, 1 = overlap; else NULL */
    uint64_t num_candidates;    /* TWO_LEVEL: object pairs with overlapping boxes; GLOBAL: raw edge hits */
    uint64_t capacity;          /* slots of the sized buffer actually used */
    uint64_t predicted;         /* sizing prediction of the sized count (0 for TWO_PASS): candidate
                                   object pairs (TWO_LEVEL) or unique pairs (GLOBAL) */
    uint64_t predicted_pairs;   /* LEARNED: predicted result pairs */
    int32_t  overflowed;        /* 1: the estimate-sized buffer overflowed; the exact count was used */
    int32_t  num_phases;
    uint64_t tests_total;       /* primitive predicate evaluations (segment-triangle and parity) */
    uint64_t tests_fp64;        /* decided by the fp64 filter */
    uint64_t tests_exact;       /* decided by exact arithmetic */
    uint64_t tests_fp32_misclassified; /* escalated tests an unguarded fp32 evaluation gets wrong */
    uint64_t edges_traced;      /* edge traversals started */
    uint64_t stack_overflows;   /* always 0: a BVH too tall for the traversal stack fails the join with
                                   PZ_E_LIMIT before any traversal runs */
    uint64_t index_triangles;   /* triangles indexed by the per-object (two-level) or set-wide BVHs */
    uint64_t index_edges;       /* edges regrouped per object (two-level) */
    uint64_t tasks_spilled;     /* narrow-phase subtree tasks handed from one work-queue round to the next */
    int32_t  narrow_rounds;     /* work-queue rounds of the narrow phase */
    int32_t  graph_launches;    /* fixed-shape segments replayed from captured HIP graphs */
    uint64_t hash_accesses;     /* GLOBAL with a hash set: slots probed (0 elsewhere) */
    uint64_t hash_contentions;  /* probed slots already holding another pair
                                   (pierce_overlap_main.cpp:676-686 reports both per run) */
    /* Walk counters (the quantities pierce_intersection --enable-profiling-stats reports,
     * pierce/query/src/cuda/mesh_intersection.h:7-22): */
    uint64_t narrow_fallback_pairs; /* two-level: candidate pairs the dual-tree walk handed to the
                                       per-edge kernel (queue or list full); a normal, exact outcome */
    uint64_t walks_started;     /* narrow-phase walks: dual-tree tasks taken + per-edge traversals */
    uint64_t pairs_touching;    /* candidate pairs proven to have touching surfaces */
    uint64_t parity_rays;       /* +Z parity rays cast (one per pair and direction tested) */
    uint64_t parity_crossings;  /* triangles crossed by those rays */
    uint64_t parity_rays_crossing; /* rays that crossed at least one triangle */
    uint64_t parity_inside;     /* rays that crossed an odd number of triangles (point inside) */
    uint64_t parity_tests;      /* ray-triangle tests of those rays (part of tests_total) */
    uint64_t sieve_pairs;       /* two-level, narrow_mode 0: candidate pairs the pair sieve decided
                                   (the others went to the per-object trees) */
    uint64_t sieve_triangles;   /* triangles of those pairs' objects, per pair (what the sieve read) */
    uint32_t index_height;      /* tallest triangle BVH the join built (at most 63, else PZ_E_LIMIT) */
    uint32_t chunk_sorted_sets; /* two-level: sets (0-2) whose objects over 2,048 triangles were cut into
                                   chunks in a Morton-sorted order instead of their own triangle order:
                                   some were not contiguous, or an object's own order was not spatially
                                   coherent (its chunks' boxes summed to over 4 times its box's area) */
    double   ms_total;          /* device-resident geometry -> sorted host pair set */
    pz_phase phases[PZ_MAX_PHASES];
} pz_join_result;

pz_status pz_join(pz_ctx* ctx, const pz_dataset* mesh1, const pz_dataset* mesh2,
                  const pz_join_desc* desc, pz_join_result* out);
void      pz_join_result_release(pz_join_result* r);

/* Read a device-resident dataset back into a new host mesh (its vertices,
 * triangles, object map, edges and grid). */
pz_status pz_dataset_download(const pz_dataset* ds, pz_host_mesh** out);

/* ------------------------------------------------------------------ batches */
/* One dataset holding many independent sets ("parts") concatenated: part p's
 * object ids, vertex indices and edge objects are shifted by the totals of the
 * parts before it.  Used by pz_join_batch; no reference counterpart (the
 * reference runs one join per process). */
pz_status pz_dataset_upload_batch(pz_ctx* ctx, uint32_t nparts, const pz_mesh_view* views, pz_dataset** out);
/* Batched join: query q joins part q of mesh1 with part q of mesh2 in one
 * pass of the two-level pipeline (one object subtree per part, one set of
 * launches for all queries).  descs[q].predicate / direction /
 * include_overlap_pairs apply to query q; every other field is taken from
 * descs[0] (strategy must be TWO_LEVEL).  results[q] gets query q's pairs with
 * part-local object ids; summary (optional) gets the batch's timings and
 * counters (no pairs).  results[q].num_candidates is query q's candidate
 * object-pair count; with LEARNED sizing results[q].predicted /
 * predicted_pairs / capacity hold query q's prediction and its sized share. */
pz_status pz_join_batch(pz_ctx* ctx, const pz_dataset* mesh1, const pz_dataset* mesh2, uint32_t nqueries,
                        const pz_join_desc* descs, pz_join_result* results, pz_join_result* summary);

/* ------------------------------------------------------------ mesh library */
/* A catalogue of meshes resident in HBM from which query sets are assembled
 * by instancing: a set is a list of (library mesh, affine map) instances.
 * Streams of queries over sets that reuse a bounded set of shapes move each
 * set device to device instead of across PCIe.  No reference counterpart
 * (the reference loads every set from its own .pre file). */
typedef struct pz_library pz_library;
typedef struct {
    uint32_t mesh;       /* library object id */
    uint32_t pad;
    float    xform[12];  /* x' = A x + t: A row-major in xform[0..8], t in xform[9..11] */
} pz_instance;
/* Every object of the view must own one contiguous run of triangles and one
 * of vertices (as pz_generate and .pre files from the reference's
 * preprocessor store them); edges are regrouped per object.  Object ids >= 0. */
pz_status pz_library_create(pz_ctx* ctx, const pz_mesh_view* view, pz_library** out);
void      pz_library_release(pz_library* lib);
/* out[0..3] = vertices, triangles, edges, objects */
pz_status pz_library_info(const pz_library* lib, uint64_t out[4]);
/* A batched dataset of nparts parts, as pz_dataset_upload_batch makes: part p
 * holds instances [counts[0] + .. + counts[p-1], .. + counts[p]) in order as
 * its objects.  Vertices and edge endpoints are mapped in fp32 as
 * ((a0 x + a1 y) + a2 z) + t0 per row, every operation rounded (no fused
 * multiply-add), so the same evaluation on the host gives the same bits.
 * slot < 0: the dataset owns its arrays.  slot in [0, PZ_INSTANCE_SLOTS): its
 * arrays are the context's storage for that slot, kept and grown across
 * calls (no allocation once large enough), and the next instantiation into
 * the same slot takes them over: joins then reject the older dataset. */
#define PZ_INSTANCE_SLOTS 8
pz_status pz_library_instantiate(pz_ctx* ctx, const pz_library* lib, uint32_t nparts, const uint64_t* counts,
                                 const pz_instance* instances, int32_t slot, pz_dataset** out);

/* --------------------------------------------------- learned output sizing */
/* A 32-128-128-2 MLP predicts, per join (or per query of a batch),
 * y0 = log2(1 + candidate object pairs) and y1 = log2(1 + result pairs) from
 * 32 features of the two sets' object boxes.  Inference runs in bf16 on the
 * matrix cores; training runs on the GPU in fp32 (forward, backward, AdamW).
 * Replaces the grid estimate of
 * pierce/query/src/applications/pierce_overlap_main.cpp:376-501 as the sizing
 * step; an overflow still falls back to the exact two-pass count. */
#define PZ_MLP_FEATURES 32
#define PZ_MLP_OUTPUTS  2
#define PZ_MLP_PARAMS   20994  /* W1[128][32] b1[128] W2[128][128] b2[128] W3[2][128] b3[2] */
#define PZ_MLP_NORM     64     /* feature mean[32], 1/std[32] */

typedef struct {
    int32_t  epochs;          /* passes over the samples */
    int32_t  batch;           /* samples per step (clamped to the sample count) */
    float    lr;              /* AdamW learning rate, cosine-decayed to lr_final */
    float    lr_final;
    float    weight_decay;    /* decoupled decay (0 = none) */
    int32_t  use_graph;       /* 1: capture one epoch as a HIP graph and replay it */
    uint64_t seed;            /* shuffling */
    float    margin_quantile; /* quantile of (target - prediction) over the validation samples used
                                 as the log2 sizing margin (plus 0.05) */
    float    validation_fraction; /* samples held out of training to calibrate the margins (a seeded
                                     random subset; with fewer than 8 the training samples are used) */
    int32_t  reserved[4];
} pz_mlp_train_desc;

/* 150 epochs, batch 64, lr 2e-3 -> 1e-4, quantile 0.95, validation 0.2 */
void      pz_mlp_train_desc_defaults(pz_mlp_train_desc* d);
/* New model on the context's device: He-uniform weights from seed, identity
 * normalisation (set from the data by the first pz_mlp_train). */
pz_status pz_mlp_create(pz_ctx* ctx, uint64_t seed, pz_mlp** out);
void      pz_mlp_destroy(pz_mlp* model);
pz_status pz_mlp_save(const pz_mlp* model, const char* path);
pz_status pz_mlp_load(pz_ctx* ctx, const char* path, pz_mlp** out);
/* Any output may be NULL: params[PZ_MLP_PARAMS], norm[PZ_MLP_NORM], margin[PZ_MLP_OUTPUTS]. */
pz_status pz_mlp_get_params(const pz_mlp* model, float* params, float* norm, float* margin);
pz_status pz_mlp_set_params(pz_mlp* model, const float* params, const float* norm, const float* margin);

/* Features of one join (out[32]) or of every query of a batch (out[nq][32]). */
pz_status pz_join_features(pz_ctx* ctx, const pz_dataset* mesh1, const pz_dataset* mesh2, int32_t predicate,
                           float* out);
pz_status pz_batch_features(pz_ctx* ctx, const pz_dataset* mesh1, const pz_dataset* mesh2, uint32_t nqueries,
                            const int32_t* predicates, float* out);

/* Y[n][2] from X[n][32] (host arrays): bf16 matrix-core inference (the path the
 * joins use) or the fp32 forward of the training kernels. */
pz_status pz_mlp_predict(pz_ctx* ctx, const pz_mlp* model, const float* X, uint64_t n, float* Y);
pz_status pz_mlp_forward(pz_ctx* ctx, const pz_mlp* model, const float* X, uint64_t n, float* Y);
/* Loss = mean over samples of sum over outputs of (y - t)^2 and its gradient
 * grad[PZ_MLP_PARAMS] over all n samples as one batch (fp32). */
pz_status pz_mlp_gradients(pz_ctx* ctx, const pz_mlp* model, const float* X, const float* T, uint64_t n,
                           float* loss, float* grad);
/* Train on (X[n][32], T[n][2]) with T = log2(1 + exact counts) from two-pass
 * joins; epoch_loss (optional) gets the mean step loss of each epoch.  Then
 * calibrates the sizing margins on the validation samples. */
pz_status pz_mlp_train(pz_ctx* ctx, pz_mlp* model, const float* X, const float* T, uint64_t n,
                       const pz_mlp_train_desc* desc, float* epoch_loss);

/* ---------------------------------------------------------------- tuning */
/* Launch parameters picked per size class.  The per-object LDS index builder
 * runs one launch per size class of objects (at most 128, 512 and 2048
 * triangles); each class has its own workgroup size.  No reference
 * counterpart (OptiX picks its own build parameters). */
typedef struct {
    int32_t build_threads[3];   /* 128, 256, 512 or 1024 */
    int32_t narrow_queue[2];    /* dual-tree narrow phase: node pairs queued in LDS per wave
                                   (256, 512, 1024 or 1536; overflowing pairs go to the per-edge kernel),
                                   [0] for joins whose objects all have <= 2048 triangles, [1] otherwise */
    int32_t narrow_chunk[2];    /* work-queue tasks a wave takes at a time (1..64), same two classes */
    int32_t reserved[9];
} pz_tune;

typedef struct {
    pz_tune chosen;
    double  build_ms[3][4];     /* class x candidate {128, 256, 512, 1024}: device ms of that class's
                                   launches in one join; < 0 when the workload has no object of the class */
    double  narrow_ms[4];       /* queue {256, 512, 1024, 1536}: device ms of the narrow phase kernels */
    double  chunk_ms[4];        /* chunk {1, 2, 4, 8} at the chosen queue: device ms of the narrow phase */
    int32_t joins_run;
    int32_t queue_class;        /* which narrow_queue entry this workload tuned */
    int32_t reserved[6];
} pz_tune_report;

pz_status pz_tune_get(const pz_ctx* ctx, pz_tune* out);
pz_status pz_tune_set(pz_ctx* ctx, const pz_tune* in);
/* Autotuner: after one warm-up join, runs the join once per candidate
 * workgroup size and times every size class's builder launches with HIP
 * events, keeping the fastest per class; then once per queue length and once
 * per work-queue chunk, keeping the fastest narrow phase for the workload's
 * class.  The choices stay in the context (13 joins in all).  The tuned
 * parameters belong to the per-object trees, so a desc with narrow_mode 0
 * (pair sieve) is tuned with narrow_mode 2 (every pair through the trees). */
pz_status pz_autotune(pz_ctx* ctx, const pz_dataset* mesh1, const pz_dataset* mesh2, const pz_join_desc* desc,
                      pz_tune_report* report);

/* Number of (edge, triangle) pairs over both directions whose closed fp32
 * boxes overlap (the BVH-independent candidate-test count). */
pz_status pz_count_candidate_tests(pz_ctx* ctx, const pz_dataset* mesh1, const pz_dataset* mesh2,
                                   uint64_t* out);

/* Containment-ray tracking of pierce_intersection --enable-tracking.
 * Replaces the per-source tracking buffers that the any-hit program fills
 * (pierce/query/src/cuda/mesh_intersection_anyhit.cu:46-240) and that
 * pierce/query/src/utilities/PairHitTracking.h:30-117 and
 * ContainmentTracking.h:22-100 copy out.  For each object s of `source` with
 * triangles, the +Z ray from the first vertex of its first triangle is cast
 * through every triangle of `target`; a triangle it crosses is one hit on
 * that triangle's object.  Host outputs (all sized by source's object count):
 *   target_ids/target_hits[s * max_targets + i]: the objects hit an odd
 *     number of times (those that contain the point), ascending ids, with
 *     their hit counts; unused slots -1 / 0;
 *   candidate_count[s]: distinct objects hit (at most max_targets);
 *   candidate_overflow[s]: hits on objects beyond the first max_targets met;
 *   iterations[s]: 1 if a ray was cast, 0 for an object without triangles.
 * 1 <= max_targets <= 4096 (the reference tracks 256). */
pz_status pz_track_containment_rays(pz_ctx* ctx, const pz_dataset* source, const pz_dataset* target,
                                    int32_t max_targets, int32_t* target_ids, uint32_t* target_hits,
                                    uint32_t* candidate_count, uint32_t* candidate_overflow,
                                    uint32_t* iterations);

/* ------------------------------------------------------------- profiling */
/* Per-kernel device time from HIP events recorded on the launching stream
 * (the phase timer of pierce/query/src/timer.cpp, at kernel granularity).
 * Kernel names: object_stats, edge_count, tri_keys, edge_scatter, karras,
 * refit, gather_leaves, radix_scatter, narrow, parity, trace_global. */
void      pz_profile_enable(int32_t on);
void      pz_profile_reset(void);
pz_status pz_profile_query(const char* kernel, double* total_ms, uint64_t* launches);

/* ------------------------------------------------ reference extern "C" helpers */
/* These keep the reference's conventions: the current HIP device, the null
 * stream, caller-owned device buffers, counts returned by value, and a block
 * until the result is on the device.  Their scratch comes from one workspace
 * per device that grows on demand and is kept for later calls (freed when
 * the library is unloaded); calls from several host threads are serialised. */
/* pierce/query/src/cuda/mesh_query_deduplication.h:7-11
 * (mesh_query_deduplication.cu:85-111): the pairs of d_results1 followed by
 * those of d_results2 are written to d_merged_output (sized >= n1 + n2) as the
 * sorted unique set.  A null d_results1 (d_results2) with a non-zero count
 * means those pairs are already at d_merged_output (d_merged_output + n1),
 * the in-place form pierce_overlap_two_pass_main.cpp:137 uses.  Returns the
 * number of unique pairs (0 for no input), sorted by (mesh1, mesh2) (the
 * reference's u64 reinterpretation orders by (mesh2, mesh1); the set is the
 * same), or -1 on failure. */
long long merge_and_deduplicate_pairs_gpu(const pz_pair* d_results1, long long num_results1,
                                          const pz_pair* d_results2, long long num_results2,
                                          pz_pair* d_merged_output);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): merge_and_deduplicate_pairs_gpu with the same arguments,
 * in-place forms and result, whose scratch is a fixed device array of the
 * library (40 MiB per device, held from load to process exit) instead of the
 * growable workspace.  It calls no allocating routine.  More than
 * PZ_NOALLOC_MAX_PAIRS input pairs (n1 + n2) -> -1 and pz_last_error() says
 * why.  Every call on one device uses that device's one array: calls that
 * overlap on one device race (serialise them). */
#define PZ_NOALLOC_MAX_PAIRS 1048576
long long merge_and_deduplicate_pairs_gpu_noalloc(const pz_pair* d_results1, long long num_results1,
                                                  const pz_pair* d_results2, long long num_results2,
                                                  pz_pair* d_merged_output);
/* mesh_query_deduplication.h:13-16: device hash table of (id1 << 32 | id2) keys,
 * empty = ~0ull; writes at most max_output_size pairs; returns the count written. */
int compact_hash_table_pairs(const unsigned long long* d_hash_table, unsigned long long table_size,
                             pz_pair* d_output, int max_output_size);
/* scan_utils.h:7-9: exclusive scan of int -> long long on the device; returns the total. */
long long exclusive_scan_gpu(const int* d_input, long long* d_output, int num_elements);

/* estimated_overlap.h:6-13 (host inputs): sum over matched cells of the grid
 * estimate, evaluated on the host; -1.0f on failure.  Takes the .pre file's
 * 28-byte cells (index + GridCell). */
float pz_estimate_overlap_selectivity(const pz_grid_cell* cells_a, const pz_grid_cell* cells_b,
                                      int32_t num_cells, float cell_volume, float epsilon, float gamma);

/* The 16-byte GridCell of pierce/common/include/GridCell.h:5-10, field for field. */
typedef struct {
    uint32_t center_count;   /* CenterCount */
    uint32_t touch_count;    /* TouchCount */
    float    avg_size_mean;  /* AvgSizeMean */
    float    vol_ratio;      /* VolRatio */
} pz_cell_stats;

/* estimateOverlapSelectivity of pierce/query/src/cuda/estimated_overlap.h:6-13
 * (estimated_overlap.cu:8-77) on GridCell-layout host arrays: the per-cell
 * terms and their sum are evaluated on the current HIP device (the sum in
 * double, in a fixed order), and the total is returned as float; -1.0f on
 * failure, as the reference returns.  A C++ definition of the reference's
 * own declaration can forward here with the arrays reinterpreted (the
 * layouts are the same). */
float pz_estimate_overlap_selectivity_cells(const pz_cell_stats* h_grid_a, const pz_cell_stats* h_grid_b,
                                            int32_t num_cells, float cell_volume, float epsilon, float gamma);

/* Whole-set grid estimate with its intermediates: the values the reference's
 * verbose estimation block prints (pierce/query/src/applications/pierce_overlap_main.cpp:376-471,
 * pierce_intersection_main.cpp:213-311, pierce_containment_main.cpp:65-135). */
typedef struct {
    int32_t has_grid;          /* 0 when either side has no cells or cell_size <= 0 */
    int32_t matched_cells;     /* cells of b whose index is also a cell of a */
    double  raw_pairs;         /* per-cell estimate summed over the matches */
    float   avg_size[2];       /* touch-weighted mean AvgSizeMean of a, b (0 if untouched) */
    float   avg_vol_ratio[2];  /* touch-weighted mean VolRatio of a, b (1 if untouched) */
    float   effective_size[2]; /* avg_size * cbrt(avg_vol_ratio) */
    float   alpha;             /* max(1, (e_a + e_b)^3 / cell_size^3); 1 without correction */
    int64_t estimated_pairs;   /* (int64)(raw_pairs / alpha) */
} pz_grid_estimate;

/* Host-only: needs no device.  cell_size is a's (the reference takes mesh1's). */
pz_status pz_estimate_grid(const pz_grid_cell* cells_a, uint64_t num_a, const pz_grid_cell* cells_b,
                           uint64_t num_b, float cell_size, float epsilon, float gamma, int32_t alpha_correction,
                           pz_grid_estimate* out);
/* The same on two loaded datasets' grids; this is the estimate PZ_SIZING_ESTIMATE sizes with. */
pz_status pz_dataset_estimate_grid(const pz_dataset* a, const pz_dataset* b, float epsilon, float gamma,
                                   int32_t alpha_correction, pz_grid_estimate* out);

#ifdef __cplusplus
}
#endif
#endif /* PIERCE_HIP_H */
