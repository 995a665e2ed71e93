"""Python mirror of the pierce_hip C ABI (ctypes over libpierce_hip.so).

The engine is the native library: this module only marshals numpy arrays in
and out.  There is no Python or CPU fallback for the join: without the built
library, or without a HIP device, every join call raises.

Reference interfaces mirrored here: the ``.pre`` v4 payload
(pierce/common/include/BinaryIO.h:14-151), MeshQueryResult pairs
(pierce/common/include/common.h:34-37) and the four query applications'
predicates (pierce/query/src/applications/*_main.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpierce_hip.so")
BIN_DIR = os.path.join(_HERE, "bin")

# ---------------------------------------------------------------- constants
PZ_OK = 0
PZ_E_INVALID, PZ_E_HIP, PZ_E_IO, PZ_E_OVERFLOW, PZ_E_NOMEM, PZ_E_NO_EDGES, PZ_E_LIMIT, PZ_E_NODEVICE = (
    -1, -2, -3, -4, -5, -6, -7, -8)

GEN_ELLIPSOIDS, GEN_SPHERES_MIXED, GEN_CUBES, GEN_TUBES, GEN_QUERY_MIX = 0, 1, 2, 3, 4
EDGES_NONE, EDGES_GROUPED, EDGES_PRE = 0, 1, 2

PRED_OVERLAP, PRED_INTERSECTION, PRED_CONTAINMENT = 0, 1, 2
DIR_BOTH, DIR_1TO2, DIR_2TO1 = 0, 1, 2
STRATEGY_TWO_LEVEL, STRATEGY_GLOBAL = 0, 1
PREC_FP32_FILTERED, PREC_FP64 = 0, 1
SIZING_TWO_PASS, SIZING_ESTIMATE, SIZING_LEARNED, SIZING_FIXED = 0, 1, 2, 3
DEDUP_SORT, DEDUP_HASH = 0, 1
MAX_PHASES = 24
NOALLOC_MAX_PAIRS = 1 << 20  # PZ_NOALLOC_MAX_PAIRS


class PierceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pierce_hip error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------- structs
class GridCell(C.Structure):
    _fields_ = [("ix", C.c_int32), ("iy", C.c_int32), ("iz", C.c_int32), ("center_count", C.c_uint32),
                ("touch_count", C.c_uint32), ("avg_size_mean", C.c_float), ("vol_ratio", C.c_float)]


GRID_DTYPE = np.dtype([("ix", "<i4"), ("iy", "<i4"), ("iz", "<i4"), ("center_count", "<u4"),
                       ("touch_count", "<u4"), ("avg_size_mean", "<f4"), ("vol_ratio", "<f4")])


class CellStats(C.Structure):
    """GridCell of pierce/common/include/GridCell.h:5-10 (16 bytes)."""
    _fields_ = [("center_count", C.c_uint32), ("touch_count", C.c_uint32), ("avg_size_mean", C.c_float),
                ("vol_ratio", C.c_float)]


CELL_DTYPE = np.dtype([("center_count", "<u4"), ("touch_count", "<u4"), ("avg_size_mean", "<f4"),
                       ("vol_ratio", "<f4")])


class GridEstimate(C.Structure):
    _fields_ = [("has_grid", C.c_int32), ("matched_cells", C.c_int32), ("raw_pairs", C.c_double),
                ("avg_size", C.c_float * 2), ("avg_vol_ratio", C.c_float * 2), ("effective_size", C.c_float * 2),
                ("alpha", C.c_float), ("estimated_pairs", C.c_int64)]

    def as_dict(self) -> dict:
        return dict(has_grid=bool(self.has_grid), matched_cells=self.matched_cells, raw_pairs=self.raw_pairs,
                    avg_size=tuple(self.avg_size), avg_vol_ratio=tuple(self.avg_vol_ratio),
                    effective_size=tuple(self.effective_size), alpha=self.alpha,
                    estimated_pairs=self.estimated_pairs)


class MeshView(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint64), ("indices", C.c_void_p),
                ("num_triangles", C.c_uint64), ("triangle_to_object", C.c_void_p), ("edge_starts", C.c_void_p),
                ("edge_ends", C.c_void_p), ("edge_objects", C.c_void_p), ("num_edges", C.c_uint64),
                ("grid_cells", C.c_void_p), ("num_grid_cells", C.c_uint64), ("grid_cell_size", C.c_float)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("edges_mode", C.c_int32), ("count", C.c_uint64), ("seed", C.c_uint64),
                ("level_min", C.c_int32), ("level_max", C.c_int32), ("radius_min", C.c_float),
                ("radius_max", C.c_float), ("box", C.c_float), ("jitter", C.c_float), ("copy_fraction", C.c_float),
                ("copy_scale_min", C.c_float), ("copy_scale_max", C.c_float), ("copy_source_count", C.c_uint64),
                ("copy_source_seed", C.c_uint64), ("tri_median", C.c_int64), ("tri_sigma", C.c_float),
                ("tri_min", C.c_int64), ("tri_max", C.c_int64), ("ring_segments", C.c_int32), ("step", C.c_float),
                ("branch_prob", C.c_float), ("max_branches", C.c_int32), ("degenerate_fraction", C.c_float),
                ("reserved", C.c_int32 * 4)]


class JoinDesc(C.Structure):
    _fields_ = [("predicate", C.c_int32), ("direction", C.c_int32), ("strategy", C.c_int32),
                ("precision", C.c_int32), ("sizing", C.c_int32), ("dedup", C.c_int32),
                ("include_overlap_pairs", C.c_int32), ("max_hits_per_edge", C.c_int32), ("capacity", C.c_uint64),
                ("load_factor", C.c_float), ("margin", C.c_float), ("gamma", C.c_float), ("epsilon", C.c_float),
                ("no_alpha_correction", C.c_int32), ("keep_device_result", C.c_int32), ("model", C.c_void_p),
                ("narrow_mode", C.c_int32), ("broad_mode", C.c_int32), ("reserved", C.c_int32 * 6)]


class Phase(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("ms", C.c_double)]


class JoinResult(C.Structure):
    _fields_ = [("num_pairs", C.c_uint64), ("pairs", C.c_void_p), ("pair_types", C.c_void_p),
                ("num_candidates", C.c_uint64), ("capacity", C.c_uint64), ("predicted", C.c_uint64),
                ("predicted_pairs", C.c_uint64), ("overflowed", C.c_int32), ("num_phases", C.c_int32), ("tests_total", C.c_uint64),
                ("tests_fp64", C.c_uint64), ("tests_exact", C.c_uint64), ("tests_fp32_misclassified", C.c_uint64),
                ("edges_traced", C.c_uint64), ("stack_overflows", C.c_uint64), ("index_triangles", C.c_uint64),
                ("index_edges", C.c_uint64), ("tasks_spilled", C.c_uint64), ("narrow_rounds", C.c_int32),
                ("graph_launches", C.c_int32), ("hash_accesses", C.c_uint64), ("hash_contentions", C.c_uint64),
                ("narrow_fallback_pairs", C.c_uint64), ("walks_started", C.c_uint64), ("pairs_touching", C.c_uint64),
                ("parity_rays", C.c_uint64), ("parity_crossings", C.c_uint64), ("parity_rays_crossing", C.c_uint64),
                ("parity_inside", C.c_uint64), ("parity_tests", C.c_uint64), ("sieve_pairs", C.c_uint64),
                ("sieve_triangles", C.c_uint64),
                ("index_height", C.c_uint32),
                ("chunk_sorted_sets", C.c_uint32), ("ms_total", C.c_double),
                ("phases", Phase * MAX_PHASES)]


MLP_FEATURES, MLP_OUTPUTS, MLP_PARAMS, MLP_NORM = 32, 2, 20994, 64


class MlpTrainDesc(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("weight_decay", C.c_float), ("use_graph", C.c_int32), ("seed", C.c_uint64),
                ("margin_quantile", C.c_float), ("validation_fraction", C.c_float), ("reserved", C.c_int32 * 4)]


_FP = C.POINTER(C.c_float)


class Tune(C.Structure):
    _fields_ = [("build_threads", C.c_int32 * 3), ("narrow_queue", C.c_int32 * 2), ("narrow_chunk", C.c_int32 * 2),
                ("reserved", C.c_int32 * 9)]


class TuneReport(C.Structure):
    _fields_ = [("chosen", Tune), ("build_ms", (C.c_double * 4) * 3), ("narrow_ms", C.c_double * 4),
                ("chunk_ms", C.c_double * 4),
                ("joins_run", C.c_int32), ("queue_class", C.c_int32), ("reserved", C.c_int32 * 6)]


TUNE_CANDIDATES = (128, 256, 512, 1024)
QUEUE_CANDIDATES = (256, 512, 1024, 1536)
CHUNK_CANDIDATES = (1, 2, 4, 8)

# Every function declared in include/pierce_hip.h: (name, restype, argtypes).
_SIGS = [
    ("pz_last_error", C.c_char_p, []),
    ("pz_host_mesh_view", C.c_int32, [C.c_void_p, C.POINTER(MeshView)]),
    ("pz_host_mesh_free", None, [C.c_void_p]),
    ("pz_pre_read", C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("pz_obj_read", C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("pz_host_mesh_shuffle", C.c_int32, [C.c_void_p, C.c_uint64, C.c_float, C.c_float]),
    ("pz_pre_write", C.c_int32, [C.c_char_p, C.POINTER(MeshView)]),
    ("pz_host_mesh_from_view", C.c_int32, [C.POINTER(MeshView), C.POINTER(C.c_void_p)]),
    ("pz_host_mesh_extract_edges", C.c_int32, [C.c_void_p]),
    ("pz_host_mesh_compute_grid", C.c_int32, [C.c_void_p, C.c_float]),
    ("pz_gen_defaults", None, [C.c_int32, C.POINTER(GenParams)]),
    ("pz_generate", C.c_int32, [C.POINTER(GenParams), C.POINTER(C.c_void_p)]),
    ("pz_device_count", C.c_int32, []),
    ("pz_ctx_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("pz_ctx_destroy", None, [C.c_void_p]),
    ("pz_ctx_set_graphs", None, [C.c_void_p, C.c_int32]),
    ("pz_ctx_mem_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("pz_ctx_trim", C.c_int32, [C.c_void_p]),
    ("pz_dataset_upload", C.c_int32, [C.c_void_p, C.POINTER(MeshView), C.POINTER(C.c_void_p)]),
    ("pz_dataset_load_pre", C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("pz_dataset_release", None, [C.c_void_p]),
    ("pz_dataset_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("pz_dataset_flags", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("pz_join_desc_defaults", None, [C.POINTER(JoinDesc)]),
    ("pz_join", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(JoinDesc), C.POINTER(JoinResult)]),
    ("pz_join_result_release", None, [C.POINTER(JoinResult)]),
    ("pz_count_candidate_tests", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    ("pz_track_containment_rays", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pz_dataset_upload_batch", C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(MeshView), C.POINTER(C.c_void_p)]),
    ("pz_dataset_download", C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("pz_library_create", C.c_int32, [C.c_void_p, C.POINTER(MeshView), C.POINTER(C.c_void_p)]),
    ("pz_library_release", None, [C.c_void_p]),
    ("pz_library_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("pz_library_instantiate", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                           C.c_int32, C.POINTER(C.c_void_p)]),
    ("pz_join_batch", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(JoinDesc),
                                  C.POINTER(JoinResult), C.POINTER(JoinResult)]),
    ("pz_profile_enable", None, [C.c_int32]),
    ("pz_profile_reset", None, []),
    ("pz_profile_query", C.c_int32, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("merge_and_deduplicate_pairs_gpu", C.c_longlong,
     [C.c_void_p, C.c_longlong, C.c_void_p, C.c_longlong, C.c_void_p]),
    ("merge_and_deduplicate_pairs_gpu_noalloc", C.c_longlong,
     [C.c_void_p, C.c_longlong, C.c_void_p, C.c_longlong, C.c_void_p]),
    ("compact_hash_table_pairs", C.c_int, [C.c_void_p, C.c_ulonglong, C.c_void_p, C.c_int]),
    ("exclusive_scan_gpu", C.c_longlong, [C.c_void_p, C.c_void_p, C.c_int]),
    ("pz_estimate_overlap_selectivity", C.c_float,
     [C.POINTER(GridCell), C.POINTER(GridCell), C.c_int32, C.c_float, C.c_float, C.c_float]),
    ("pz_estimate_overlap_selectivity_cells", C.c_float,
     [C.POINTER(CellStats), C.POINTER(CellStats), C.c_int32, C.c_float, C.c_float, C.c_float]),
    ("pz_estimate_grid", C.c_int32,
     [C.POINTER(GridCell), C.c_uint64, C.POINTER(GridCell), C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_int32,
      C.POINTER(GridEstimate)]),
    ("pz_dataset_estimate_grid", C.c_int32,
     [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_int32, C.POINTER(GridEstimate)]),
    ("pz_tune_get", C.c_int32, [C.c_void_p, C.POINTER(Tune)]),
    ("pz_tune_set", C.c_int32, [C.c_void_p, C.POINTER(Tune)]),
    ("pz_autotune", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(JoinDesc), C.POINTER(TuneReport)]),
    ("pz_mlp_train_desc_defaults", None, [C.POINTER(MlpTrainDesc)]),
    ("pz_mlp_create", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("pz_mlp_destroy", None, [C.c_void_p]),
    ("pz_mlp_save", C.c_int32, [C.c_void_p, C.c_char_p]),
    ("pz_mlp_load", C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("pz_mlp_get_params", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pz_mlp_set_params", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pz_join_features", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    ("pz_batch_features", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("pz_mlp_predict", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("pz_mlp_forward", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("pz_mlp_gradients", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, _FP, C.c_void_p]),
    ("pz_mlp_train", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                 C.POINTER(MlpTrainDesc), C.c_void_p]),
]

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libpierce_hip.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PierceError(PZ_E_INVALID, f"native library not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list:
    return [n for n, _, _ in _SIGS]


def _check(st: int) -> None:
    if st != PZ_OK:
        msg = lib().pz_last_error()
        raise PierceError(st, msg.decode() if msg else "")


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None or a.size == 0 else a.ctypes.data


# ---------------------------------------------------------------- host meshes
@dataclass
class Mesh:
    """One dataset's .pre payload as numpy arrays (AoS, as in the file)."""
    vertices: np.ndarray            # (nV, 3) float32
    indices: np.ndarray             # (nT, 3) uint32
    triangle_to_object: np.ndarray  # (nT,) int32
    edge_starts: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    edge_ends: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    edge_objects: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    grid: np.ndarray = field(default_factory=lambda: np.zeros((0,), GRID_DTYPE))
    grid_cell_size: float = 0.0

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, np.uint32).reshape(-1, 3)
        self.triangle_to_object = np.ascontiguousarray(self.triangle_to_object, np.int32).reshape(-1)
        self.edge_starts = np.ascontiguousarray(self.edge_starts, np.float32).reshape(-1, 3)
        self.edge_ends = np.ascontiguousarray(self.edge_ends, np.float32).reshape(-1, 3)
        self.edge_objects = np.ascontiguousarray(self.edge_objects, np.int32).reshape(-1)
        self.grid = np.ascontiguousarray(self.grid, GRID_DTYPE).reshape(-1)

    @property
    def num_objects(self) -> int:
        m = -1
        if self.triangle_to_object.size:
            m = max(m, int(self.triangle_to_object.max()))
        if self.edge_objects.size:
            m = max(m, int(self.edge_objects.max()))
        return m + 1

    def view(self) -> MeshView:
        v = MeshView()
        v.vertices = _ptr(self.vertices)
        v.num_vertices = self.vertices.shape[0]
        v.indices = _ptr(self.indices)
        v.num_triangles = self.indices.shape[0]
        v.triangle_to_object = _ptr(self.triangle_to_object)
        v.edge_starts = _ptr(self.edge_starts)
        v.edge_ends = _ptr(self.edge_ends)
        v.edge_objects = _ptr(self.edge_objects)
        v.num_edges = self.edge_objects.shape[0]
        v.grid_cells = _ptr(self.grid)
        v.num_grid_cells = self.grid.shape[0]
        v.grid_cell_size = float(self.grid_cell_size)
        return v


class _HostMeshOwner:
    """Keeps a library-owned host mesh alive while numpy arrays view it."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            lib().pz_host_mesh_free(self.h)
        except Exception:
            pass


def _mesh_from_handle(h: C.c_void_p) -> Mesh:
    """Arrays view the library's host mesh without a copy (large sets would
    otherwise be held twice); the mesh is freed when the last view goes."""
    L = lib()
    v = MeshView()
    _check(L.pz_host_mesh_view(h, C.byref(v)))
    owner = _HostMeshOwner(h)

    def arr(p, n, dt, shape):
        if n == 0 or not p:
            return np.zeros(tuple(0 if x == -1 else x for x in shape), dt)
        buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(p)
        buf.owner = owner
        return np.frombuffer(buf, dt).reshape(shape)

    m = Mesh(arr(v.vertices, 3 * v.num_vertices, np.float32, (-1, 3)),
             arr(v.indices, 3 * v.num_triangles, np.uint32, (-1, 3)),
             arr(v.triangle_to_object, v.num_triangles, np.int32, (-1,)),
             arr(v.edge_starts, 3 * v.num_edges, np.float32, (-1, 3)),
             arr(v.edge_ends, 3 * v.num_edges, np.float32, (-1, 3)),
             arr(v.edge_objects, v.num_edges, np.int32, (-1,)),
             arr(v.grid_cells, v.num_grid_cells, GRID_DTYPE, (-1,)),
             float(v.grid_cell_size))
    return m


def gen_params(kind: int = GEN_ELLIPSOIDS, **kw) -> GenParams:
    p = GenParams()
    lib().pz_gen_defaults(kind, C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def generate(kind: int = GEN_ELLIPSOIDS, **kw) -> Mesh:
    """Deterministic synthetic dataset (see pz_gen_params in include/pierce_hip.h)."""
    p = gen_params(kind, **kw)
    h = C.c_void_p()
    _check(lib().pz_generate(C.byref(p), C.byref(h)))
    return _mesh_from_handle(h)


def read_pre(path: str) -> Mesh:
    h = C.c_void_p()
    _check(lib().pz_pre_read(path.encode(), C.byref(h)))
    return _mesh_from_handle(h)


def read_obj(path: str) -> Mesh:
    """A Wavefront .obj file, one object per shape with faces
    (pierce/preprocess/src/ObjMeshDatasetLoader.cpp:14-168); no edges, no grid."""
    h = C.c_void_p()
    _check(lib().pz_obj_read(path.encode(), C.byref(h)))
    return _mesh_from_handle(h)


def write_obj(path: str, mesh: Mesh) -> None:
    """Writes one `o object<i>` shape per object (triangles of object i,
    1-based indices into the shared vertex list)."""
    tri = mesh.indices
    t2o = mesh.triangle_to_object
    with open(path, "w") as f:
        for x, y, z in mesh.vertices.tolist():
            f.write(f"v {x!r} {y!r} {z!r}\n")
        order = np.argsort(t2o, kind="stable")
        cur = None
        for t in order.tolist():
            o = int(t2o[t])
            if o != cur:
                f.write(f"o object{o}\n")
                cur = o
            a, b, c = (int(i) + 1 for i in tri[t])
            f.write(f"f {a} {b} {c}\n")


def shuffle(mesh: Mesh, seed: int, lo: float = -15.0, hi: float = 15.0) -> Mesh:
    """Every object moved by its own uniform offset in [lo, hi]^3
    (pierce/preprocess/src/pierce_preprocess.cpp:243-283)."""
    return _with_handle(mesh, lambda h: lib().pz_host_mesh_shuffle(h, seed, lo, hi))


def write_pre(path: str, mesh: Mesh) -> None:
    v = mesh.view()
    _check(lib().pz_pre_write(path.encode(), C.byref(v)))


def _with_handle(mesh: Mesh, fn) -> Mesh:
    v = mesh.view()
    h = C.c_void_p()
    _check(lib().pz_host_mesh_from_view(C.byref(v), C.byref(h)))
    try:
        _check(fn(h))
    except Exception:
        lib().pz_host_mesh_free(h)
        raise
    return _mesh_from_handle(h)


def extract_edges(mesh: Mesh) -> Mesh:
    """EdgePreprocessor::extractEdges semantics (pierce/preprocess/src/EdgePreprocessor.cpp:44-111)."""
    return _with_handle(mesh, lambda h: lib().pz_host_mesh_extract_edges(h))


def compute_grid(mesh: Mesh, cell_size: float) -> Mesh:
    """generateGridStats semantics (pierce/preprocess/src/pierce_preprocess.cpp:49-167)."""
    return _with_handle(mesh, lambda h: lib().pz_host_mesh_compute_grid(h, C.c_float(cell_size)))


def estimate_overlap_selectivity(cells_a: np.ndarray, cells_b: np.ndarray, cell_volume: float,
                                 epsilon: float = 0.001, gamma: float = 0.8) -> float:
    a = np.ascontiguousarray(cells_a, GRID_DTYPE)
    b = np.ascontiguousarray(cells_b, GRID_DTYPE)
    return float(lib().pz_estimate_overlap_selectivity(
        a.ctypes.data_as(C.POINTER(GridCell)), b.ctypes.data_as(C.POINTER(GridCell)), len(a),
        cell_volume, epsilon, gamma))


def estimate_overlap_selectivity_cells(cells_a: np.ndarray, cells_b: np.ndarray, cell_volume: float,
                                       epsilon: float = 0.001, gamma: float = 0.8) -> float:
    """estimateOverlapSelectivity (pierce/query/src/cuda/estimated_overlap.h:6-13) on
    16-byte GridCell arrays, evaluated on the current HIP device."""
    a = np.ascontiguousarray(cells_a, CELL_DTYPE)
    b = np.ascontiguousarray(cells_b, CELL_DTYPE)
    return float(lib().pz_estimate_overlap_selectivity_cells(
        a.ctypes.data_as(C.POINTER(CellStats)), b.ctypes.data_as(C.POINTER(CellStats)), len(a), cell_volume,
        epsilon, gamma))


def estimate_grid(cells_a: np.ndarray, cells_b: np.ndarray, cell_size: float, epsilon: float = 0.001,
                  gamma: float = 0.8, alpha_correction: bool = True) -> dict:
    """Whole-set grid estimate and the intermediates the reference's verbose
    block prints (pierce/query/src/applications/pierce_overlap_main.cpp:376-471)."""
    a = np.ascontiguousarray(cells_a, GRID_DTYPE)
    b = np.ascontiguousarray(cells_b, GRID_DTYPE)
    out = GridEstimate()
    _check(lib().pz_estimate_grid(a.ctypes.data_as(C.POINTER(GridCell)), len(a),
                                  b.ctypes.data_as(C.POINTER(GridCell)), len(b), cell_size, epsilon, gamma,
                                  1 if alpha_correction else 0, C.byref(out)))
    return out.as_dict()


# ---------------------------------------------------------------- device
def device_count() -> int:
    return int(lib().pz_device_count())


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().pz_ctx_create(device, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().pz_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graphs(self, on: bool) -> None:
        lib().pz_ctx_set_graphs(self._h, 1 if on else 0)

    def mem_info(self) -> tuple:
        """(free, total) bytes of the context's device."""
        f, t = C.c_uint64(), C.c_uint64()
        _check(lib().pz_ctx_mem_info(self._h, C.byref(f), C.byref(t)))
        return int(f.value), int(t.value)

    def trim(self) -> None:
        """Free the cached device work buffers (pz_ctx_trim)."""
        _check(lib().pz_ctx_trim(self._h))

    def upload(self, mesh: Mesh) -> "Dataset":
        return Dataset(self, mesh)

    def load_pre(self, path: str) -> "Dataset":
        return Dataset(self, None, path)

    def upload_batch(self, meshes) -> "Dataset":
        """One dataset whose part p is meshes[p] (for join_batch)."""
        return Dataset(self, None, None, parts=list(meshes))

    def library(self, mesh: Mesh) -> "Library":
        """A mesh catalogue resident on the device (pz_library_create)."""
        return Library(self, mesh)


class Dataset:
    def __init__(self, ctx: Context, mesh: Optional[Mesh], path: Optional[str] = None, parts=None):
        self.ctx = ctx
        self._h = C.c_void_p()
        if parts is not None:
            views = (MeshView * len(parts))(*[m.view() for m in parts])
            _check(lib().pz_dataset_upload_batch(ctx._h, len(parts), views, C.byref(self._h)))
        elif mesh is not None:
            v = mesh.view()
            _check(lib().pz_dataset_upload(ctx._h, C.byref(v), C.byref(self._h)))
        else:
            _check(lib().pz_dataset_load_pre(ctx._h, path.encode(), C.byref(self._h)))

    def info(self) -> dict:
        out = (C.c_uint64 * 5)()
        _check(lib().pz_dataset_info(self._h, out))
        return dict(vertices=out[0], triangles=out[1], edges=out[2], objects=out[3], grid_cells=out[4])

    @property
    def vertex_runs(self) -> bool:
        """True when every object owns a disjoint, gap-free run of vertices
        (found at upload): joins then bound objects from those runs."""
        f = C.c_uint32(0)
        _check(lib().pz_dataset_flags(self._h, C.byref(f)))
        return bool(f.value & 1)

    def estimate_grid(self, other: "Dataset", epsilon: float = 0.001, gamma: float = 0.8,
                      alpha_correction: bool = True) -> dict:
        """estimate_grid over this dataset's grid (a) and other's (b)."""
        out = GridEstimate()
        _check(lib().pz_dataset_estimate_grid(self._h, other._h, epsilon, gamma, 1 if alpha_correction else 0,
                                              C.byref(out)))
        return out.as_dict()

    def download(self) -> Mesh:
        """The device-resident arrays read back (pz_dataset_download)."""
        h = C.c_void_p()
        _check(lib().pz_dataset_download(self._h, C.byref(h)))
        return _mesh_from_handle(h)

    @classmethod
    def _adopt(cls, ctx: "Context", h: C.c_void_p) -> "Dataset":
        d = cls.__new__(cls)
        d.ctx = ctx
        d._h = h
        return d

    def close(self):
        if self._h:
            lib().pz_dataset_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# One instance of a library mesh: mesh id, pad, x' = A x + t (A row-major, then t).
INSTANCE_DTYPE = np.dtype([("mesh", "<u4"), ("pad", "<u4"), ("xform", "<f4", (12,))])


class Library:
    """A mesh catalogue in HBM; sets are built from it by instancing."""

    def __init__(self, ctx: Context, mesh: Mesh):
        self.ctx = ctx
        self._h = C.c_void_p()
        v = mesh.view()
        _check(lib().pz_library_create(ctx._h, C.byref(v), C.byref(self._h)))

    def info(self) -> dict:
        out = (C.c_uint64 * 4)()
        _check(lib().pz_library_info(self._h, out))
        return dict(vertices=out[0], triangles=out[1], edges=out[2], objects=out[3])

    def instantiate(self, counts, instances: np.ndarray, slot: int = -1) -> Dataset:
        """Batched dataset: part p holds the next counts[p] instances (structured
        INSTANCE_DTYPE array) as its objects (pz_library_instantiate).  slot >= 0
        builds it in the context's reusable storage for that slot (valid until
        the next instantiation into the same slot)."""
        cnt = np.ascontiguousarray(counts, np.uint64).reshape(-1)
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE).reshape(-1)
        if int(cnt.sum()) != inst.shape[0]:
            raise ValueError("counts must add up to the number of instances")
        h = C.c_void_p()
        _check(lib().pz_library_instantiate(self.ctx._h, self._h, cnt.shape[0], cnt.ctypes.data,
                                            inst.ctypes.data if inst.size else None, int(slot), C.byref(h)))
        return Dataset._adopt(self.ctx, h)

    def close(self):
        if self._h:
            lib().pz_library_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class JoinOutput:
    pairs: np.ndarray          # (n, 2) int32 sorted by (mesh1, mesh2)
    types: Optional[np.ndarray]
    num_candidates: int
    capacity: int
    predicted: int
    overflowed: bool
    tests_total: int
    tests_fp64: int
    tests_exact: int
    tests_fp32_misclassified: int
    edges_traced: int
    index_triangles: int
    index_edges: int
    ms_total: float
    phases: list
    predicted_pairs: int = 0
    tasks_spilled: int = 0
    narrow_rounds: int = 0
    graph_launches: int = 0
    hash_accesses: int = 0
    hash_contentions: int = 0
    narrow_fallback_pairs: int = 0
    walks_started: int = 0
    pairs_touching: int = 0
    parity_rays: int = 0
    parity_crossings: int = 0
    parity_rays_crossing: int = 0
    parity_inside: int = 0
    stack_overflows: int = 0
    parity_tests: int = 0
    index_height: int = 0
    sieve_pairs: int = 0
    sieve_triangles: int = 0
    chunk_sorted_sets: int = 0


def join_desc(**kw) -> JoinDesc:
    d = JoinDesc()
    lib().pz_join_desc_defaults(C.byref(d))
    for k, v in kw.items():
        if not hasattr(d, k):
            raise KeyError(k)
        setattr(d, k, _field(v))
    return d


def _field(v):
    return v._h.value if isinstance(v, Model) else v


def _output(r: JoinResult) -> JoinOutput:
    n = int(r.num_pairs)
    if n and r.pairs:
        buf = (C.c_int32 * (2 * n)).from_address(r.pairs)
        pairs = np.frombuffer(buf, np.int32).reshape(n, 2).copy()
    else:
        pairs = np.zeros((0, 2), np.int32)
    types = None
    if r.pair_types:
        types = np.frombuffer((C.c_uint8 * max(n, 1)).from_address(r.pair_types), np.uint8)[:n].copy()
    phases = [(r.phases[i].name.decode(), r.phases[i].ms) for i in range(r.num_phases)]
    return JoinOutput(pairs, types, int(r.num_candidates), int(r.capacity), int(r.predicted), bool(r.overflowed),
                      int(r.tests_total), int(r.tests_fp64), int(r.tests_exact), int(r.tests_fp32_misclassified),
                      int(r.edges_traced), int(r.index_triangles), int(r.index_edges), float(r.ms_total), phases,
                      int(r.predicted_pairs), int(r.tasks_spilled), int(r.narrow_rounds), int(r.graph_launches),
                      int(r.hash_accesses), int(r.hash_contentions), int(r.narrow_fallback_pairs),
                      int(r.walks_started), int(r.pairs_touching), int(r.parity_rays), int(r.parity_crossings),
                      int(r.parity_rays_crossing), int(r.parity_inside), int(r.stack_overflows),
                      int(r.parity_tests), int(r.index_height), int(r.sieve_pairs),
                      int(r.sieve_triangles), int(r.chunk_sorted_sets))


def join_batch(ctx: Context, a: Dataset, b: Dataset, queries, **common):
    """Batched join of part q of a with part q of b for every query q.

    queries: list of dicts of per-query pz_join_desc fields (predicate, direction,
    include_overlap_pairs); common: fields shared by the batch.  Returns
    (list of JoinOutput with part-local ids, summary JoinOutput)."""
    nq = len(queries)
    descs = (JoinDesc * nq)()
    for q, kw in enumerate(queries):
        lib().pz_join_desc_defaults(C.byref(descs[q]))
        for k, v in list(common.items()) + list(kw.items()):
            setattr(descs[q], k, _field(v))
    res = (JoinResult * nq)()
    summ = JoinResult()
    _check(lib().pz_join_batch(ctx._h, a._h, b._h, nq, descs, res, C.byref(summ)))
    try:
        outs = [_output(res[q]) for q in range(nq)]
        return outs, _output(summ)
    finally:
        for q in range(nq):
            lib().pz_join_result_release(C.byref(res[q]))


def join(ctx: Context, a: Dataset, b: Dataset, **kw) -> JoinOutput:
    """Run one join on the GPU; keyword arguments set pz_join_desc fields."""
    d = join_desc(**kw)
    r = JoinResult()
    _check(lib().pz_join(ctx._h, a._h, b._h, C.byref(d), C.byref(r)))
    try:
        return _output(r)
    finally:
        lib().pz_join_result_release(C.byref(r))


def profile_enable(on: bool = True) -> None:
    lib().pz_profile_enable(1 if on else 0)


def profile_reset() -> None:
    lib().pz_profile_reset()


def profile_query(kernel: str) -> tuple:
    """(total ms, launches) of one kernel since the last reset."""
    ms, n = C.c_double(), C.c_uint64()
    _check(lib().pz_profile_query(kernel.encode(), C.byref(ms), C.byref(n)))
    return float(ms.value), int(n.value)


def count_candidate_tests(ctx: Context, a: Dataset, b: Dataset) -> int:
    out = C.c_uint64()
    _check(lib().pz_count_candidate_tests(ctx._h, a._h, b._h, C.byref(out)))
    return int(out.value)


def track_containment_rays(ctx: Context, source: Dataset, target: Dataset, max_targets: int = 256) -> dict:
    """Per-source containment-ray tracking (pierce_intersection --enable-tracking;
    pierce/query/src/cuda/mesh_intersection_anyhit.cu:46-240).

    Returns arrays over source objects: ``target_ids`` / ``target_hits``
    (n, max_targets): objects crossed an odd number of times, ascending, -1 / 0
    padded; ``candidate_count``: distinct objects crossed; ``candidate_overflow``:
    hits beyond the first max_targets objects; ``iterations``: 1 if a ray was cast.
    """
    n = int(source.info()["objects"])
    ids = np.empty((n, max_targets), np.int32)
    hits = np.empty((n, max_targets), np.uint32)
    cnt, ovf, it = (np.empty(n, np.uint32) for _ in range(3))
    _check(lib().pz_track_containment_rays(ctx._h, source._h, target._h, int(max_targets), _ptr(ids), _ptr(hits),
                                           _ptr(cnt), _ptr(ovf), _ptr(it)))
    return dict(target_ids=ids, target_hits=hits, candidate_count=cnt, candidate_overflow=ovf, iterations=it)


# ------------------------------------------------------------ learned sizing
def _f32(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def train_desc(**kw) -> MlpTrainDesc:
    d = MlpTrainDesc()
    lib().pz_mlp_train_desc_defaults(C.byref(d))
    for k, v in kw.items():
        if not hasattr(d, k):
            raise KeyError(k)
        setattr(d, k, v)
    return d


class Model:
    """Selectivity MLP on a context's device (pz_mlp)."""

    def __init__(self, ctx: Context, seed: int = 1, path: Optional[str] = None):
        self.ctx = ctx
        self._h = C.c_void_p()
        if path is None:
            _check(lib().pz_mlp_create(ctx._h, seed, C.byref(self._h)))
        else:
            _check(lib().pz_mlp_load(ctx._h, path.encode(), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().pz_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def save(self, path: str) -> None:
        _check(lib().pz_mlp_save(self._h, path.encode()))

    def get(self):
        """(params[20994], norm[64], margin[2]) as float32 arrays."""
        p = np.zeros(MLP_PARAMS, np.float32)
        n = np.zeros(MLP_NORM, np.float32)
        m = np.zeros(MLP_OUTPUTS, np.float32)
        _check(lib().pz_mlp_get_params(self._h, _ptr(p), _ptr(n), _ptr(m)))
        return p, n, m

    def set(self, params=None, norm=None, margin=None) -> None:
        p = None if params is None else _f32(params, (MLP_PARAMS,))
        n = None if norm is None else _f32(norm, (MLP_NORM,))
        m = None if margin is None else _f32(margin, (MLP_OUTPUTS,))
        _check(lib().pz_mlp_set_params(self._h, _ptr(p), _ptr(n), _ptr(m)))

    def _run(self, fn, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_FEATURES)
        Y = np.zeros((X.shape[0], MLP_OUTPUTS), np.float32)
        _check(fn(self.ctx._h, self._h, _ptr(X), X.shape[0], _ptr(Y)))
        return Y

    def predict(self, X) -> np.ndarray:
        """bf16 matrix-core inference: Y[n][2] = log2(1 + counts)."""
        return self._run(lib().pz_mlp_predict, X)

    def forward(self, X) -> np.ndarray:
        """fp32 forward of the training kernels."""
        return self._run(lib().pz_mlp_forward, X)

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_FEATURES)
        T = _f32(T, (X.shape[0], MLP_OUTPUTS))
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib().pz_mlp_gradients(self.ctx._h, self._h, _ptr(X), _ptr(T), X.shape[0], C.byref(loss), _ptr(g)))
        return float(loss.value), g

    def train(self, X, T, **kw) -> np.ndarray:
        """Train on the GPU; returns the mean loss of each epoch."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_FEATURES)
        T = _f32(T, (X.shape[0], MLP_OUTPUTS))
        d = train_desc(**kw)
        losses = np.zeros(max(d.epochs, 1), np.float32)
        _check(lib().pz_mlp_train(self.ctx._h, self._h, _ptr(X), _ptr(T), X.shape[0], C.byref(d), _ptr(losses)))
        return losses


def join_features(ctx: Context, a: Dataset, b: Dataset, predicate: int = PRED_OVERLAP) -> np.ndarray:
    out = np.zeros(MLP_FEATURES, np.float32)
    _check(lib().pz_join_features(ctx._h, a._h, b._h, predicate, _ptr(out)))
    return out


def batch_features(ctx: Context, a: Dataset, b: Dataset, predicates) -> np.ndarray:
    preds = np.ascontiguousarray(predicates, dtype=np.int32)
    out = np.zeros((preds.shape[0], MLP_FEATURES), np.float32)
    _check(lib().pz_batch_features(ctx._h, a._h, b._h, preds.shape[0], _ptr(preds), _ptr(out)))
    return out


def log_counts(counts) -> np.ndarray:
    """Training target of a count: log2(1 + count)."""
    return np.log2(1.0 + np.asarray(counts, dtype=np.float64)).astype(np.float32)


# ------------------------------------------------------------------- tuning
def tune_get(ctx: Context) -> dict:
    t = Tune()
    _check(lib().pz_tune_get(ctx._h, C.byref(t)))
    return {"build_threads": list(t.build_threads), "narrow_queue": list(t.narrow_queue),
            "narrow_chunk": list(t.narrow_chunk)}


def tune_set(ctx: Context, build_threads, narrow_queue=(1536, 1536), narrow_chunk=(1, 1)) -> None:
    t = Tune()
    t.build_threads[:] = list(build_threads)
    t.narrow_queue[:] = list(narrow_queue)
    t.narrow_chunk[:] = list(narrow_chunk)
    _check(lib().pz_tune_set(ctx._h, C.byref(t)))


def autotune(ctx: Context, a: Dataset, b: Dataset, **kw) -> dict:
    """Pick the LDS builder's workgroup size per size class on this workload."""
    d = join_desc(**kw)
    r = TuneReport()
    _check(lib().pz_autotune(ctx._h, a._h, b._h, C.byref(d), C.byref(r)))
    return {"build_threads": list(r.chosen.build_threads), "narrow_queue": list(r.chosen.narrow_queue),
            "narrow_chunk": list(r.chosen.narrow_chunk), "queue_class": int(r.queue_class),
            "chunk_ms": {str(CHUNK_CANDIDATES[i]): round(r.chunk_ms[i], 4) for i in range(4)},
            "joins_run": int(r.joins_run),
            "narrow_ms": {str(QUEUE_CANDIDATES[i]): round(r.narrow_ms[i], 4) for i in range(4)},
            "build_ms": {f"class{k}": {str(TUNE_CANDIDATES[i]): (None if r.build_ms[k][i] < 0
                                                              else round(r.build_ms[k][i], 4))
                                      for i in range(4)} for k in range(3)}}
