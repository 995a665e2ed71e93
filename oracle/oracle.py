"""ORACLE - test infrastructure only (tests/, __graft_entry__.smoke(),
bench.py's cpu_baseline leg).  The product never imports this module.

ctypes wrapper of liboracle.so (oracle/oracle.cpp): exact CPU restatement of
the join semantics (Face/TOUCH baselines, baselines/face/src/overlap_query.cpp:75-188)
used as the checker of the GPU path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrMeshView(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint64), ("indices", C.c_void_p),
                ("num_triangles", C.c_uint64), ("triangle_to_object", C.c_void_p), ("edge_starts", C.c_void_p),
                ("edge_ends", C.c_void_p), ("edge_objects", C.c_void_p), ("num_edges", C.c_uint64)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        for name in ("or_orient3d", "or_orient3d_exact"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [fp] * 4
        L.or_orient2d.restype = C.c_int
        L.or_orient2d.argtypes = [fp, fp, fp, C.c_int, C.c_int]
        L.or_seg_tri.restype = C.c_int
        L.or_seg_tri.argtypes = [fp] * 5
        L.or_tri_tri.restype = C.c_int
        L.or_tri_tri.argtypes = [fp, fp]
        L.or_ray_cross.restype = C.c_int
        L.or_ray_cross.argtypes = [fp] * 4
        L.or_pairs_count.restype = C.c_int64
        L.or_pairs_count.argtypes = [C.c_void_p]
        L.or_pairs_data.restype = C.c_void_p
        L.or_pairs_data.argtypes = [C.c_void_p]
        L.or_pairs_types.restype = C.c_void_p
        L.or_pairs_types.argtypes = [C.c_void_p]
        L.or_pairs_query_us.restype = C.c_double
        L.or_pairs_query_us.argtypes = [C.c_void_p]
        L.or_pairs_free.restype = None
        L.or_pairs_free.argtypes = [C.c_void_p]
        mv = C.POINTER(OrMeshView)
        L.or_face_overlap.restype = C.c_void_p
        L.or_face_overlap.argtypes = [mv, mv]
        L.or_edge_overlap.restype = C.c_void_p
        L.or_edge_overlap.argtypes = [mv, mv, C.c_int]
        L.or_predicate_join.restype = C.c_void_p
        L.or_predicate_join.argtypes = [mv, mv, C.c_int, C.c_int]
        L.or_predicate_join_dir.restype = C.c_void_p
        L.or_predicate_join_dir.argtypes = [mv, mv, C.c_int, C.c_int, C.c_int]
        L.or_bruteforce_overlap.restype = C.c_void_p
        L.or_bruteforce_overlap.argtypes = [mv, mv]
        L.or_point_in_object.restype = C.c_int
        L.or_point_in_object.argtypes = [mv, C.c_int32, fp]
        L.or_ray_target_hits.restype = C.c_int
        L.or_ray_target_hits.argtypes = [mv, mv, C.POINTER(C.c_uint32)]
        L.or_count_candidate_tests.restype = C.c_uint64
        L.or_count_candidate_tests.argtypes = [mv, mv]
        L.or_set_threads.restype = None
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_get_threads.argtypes = []
        L.or_touch_overlap_count.restype = C.c_int64
        L.or_touch_overlap_count.argtypes = [mv, mv, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _f(a) -> C.POINTER(C.c_float):
    a = np.ascontiguousarray(a, np.float32)
    return a.ctypes.data_as(C.POINTER(C.c_float)), a


class _View:
    """Keeps the arrays alive while the view is in use."""

    def __init__(self, mesh):
        self.arrs = [np.ascontiguousarray(mesh.vertices, np.float32), np.ascontiguousarray(mesh.indices, np.uint32),
                     np.ascontiguousarray(mesh.triangle_to_object, np.int32),
                     np.ascontiguousarray(mesh.edge_starts, np.float32),
                     np.ascontiguousarray(mesh.edge_ends, np.float32),
                     np.ascontiguousarray(mesh.edge_objects, np.int32)]
        v, i, t, es, ee, eo = self.arrs
        self.v = OrMeshView(v.ctypes.data, v.shape[0], i.ctypes.data, i.shape[0], t.ctypes.data,
                            es.ctypes.data if es.size else None, ee.ctypes.data if ee.size else None,
                            eo.ctypes.data if eo.size else None, eo.shape[0])


def _take(h) -> tuple:
    L = lib()
    n = L.or_pairs_count(h)
    if n > 0:
        data = np.frombuffer((C.c_int32 * (2 * n)).from_address(L.or_pairs_data(h)), np.int32).reshape(n, 2).copy()
    else:
        data = np.zeros((0, 2), np.int32)
    tp = L.or_pairs_types(h)
    types = np.frombuffer((C.c_uint8 * max(n, 1)).from_address(tp), np.uint8)[:n].copy() if tp else None
    us = L.or_pairs_query_us(h)
    L.or_pairs_free(h)
    return data, types, us


def orient3d(a, b, c, d) -> int:
    args = [_f(x) for x in (a, b, c, d)]
    return lib().or_orient3d(*[p for p, _ in args])


def orient3d_exact(a, b, c, d) -> int:
    args = [_f(x) for x in (a, b, c, d)]
    return lib().or_orient3d_exact(*[p for p, _ in args])


def orient2d(a, b, c, i=0, j=1) -> int:
    args = [_f(x) for x in (a, b, c)]
    return lib().or_orient2d(*[p for p, _ in args], i, j)


def seg_tri(p, q, a, b, c) -> bool:
    args = [_f(x) for x in (p, q, a, b, c)]
    return bool(lib().or_seg_tri(*[x for x, _ in args]))


def tri_tri(t1, t2) -> bool:
    a, _a = _f(np.asarray(t1, np.float32).reshape(9))
    b, _b = _f(np.asarray(t2, np.float32).reshape(9))
    return bool(lib().or_tri_tri(a, b))


def ray_cross(p, a, b, c) -> bool:
    args = [_f(x) for x in (p, a, b, c)]
    return bool(lib().or_ray_cross(*[x for x, _ in args]))


def face_overlap(m1, m2, with_time=False):
    """Face semantics: sorted unique (a, b) with some triangle pair meeting."""
    v1, v2 = _View(m1), _View(m2)
    pairs, _, us = _take(lib().or_face_overlap(C.byref(v1.v), C.byref(v2.v)))
    return (pairs, us) if with_time else pairs


def edge_overlap(m1, m2, direction=0, with_time=False):
    v1, v2 = _View(m1), _View(m2)
    pairs, _, us = _take(lib().or_edge_overlap(C.byref(v1.v), C.byref(v2.v), direction))
    return (pairs, us) if with_time else pairs


def predicate_join(m1, m2, predicate, include_overlap=False, direction=0):
    """predicate 1 = intersection, 2 = containment; returns (pairs, types).
    direction (intersection only): 0 both, 1 mesh1 edges and points against
    mesh2, 2 the reverse."""
    v1, v2 = _View(m1), _View(m2)
    pairs, types, _ = _take(lib().or_predicate_join_dir(C.byref(v1.v), C.byref(v2.v), predicate,
                                                        int(include_overlap), int(direction)))
    return pairs, types


def bruteforce_overlap(m1, m2):
    v1, v2 = _View(m1), _View(m2)
    pairs, _, _ = _take(lib().or_bruteforce_overlap(C.byref(v1.v), C.byref(v2.v)))
    return pairs


def point_in_object(m, obj: int, p) -> bool:
    v = _View(m)
    pp, _k = _f(p)
    return bool(lib().or_point_in_object(C.byref(v.v), obj, pp))


def ray_target_hits(src, tgt) -> np.ndarray:
    """(src objects, tgt objects) uint32: triangles of each target object that
    the +Z ray from the first vertex of each source object crosses (brute
    force; pierce/query/src/cuda/mesh_intersection_anyhit.cu:46-240)."""
    vs, vt = _View(src), _View(tgt)
    out = np.zeros((src.num_objects, tgt.num_objects), np.uint32)
    lib().or_ray_target_hits(C.byref(vs.v), C.byref(vt.v), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def grid_estimate(cells_a, cells_b, cell_size: float, epsilon: float = 0.001, gamma: float = 0.8,
                  alpha_correction: bool = True) -> dict:
    """Whole-set selectivity estimate over two sparse grids (structured arrays
    with ix, iy, iz, touch_count, avg_size_mean, vol_ratio), in float32 as the
    reference computes it.  Cell matching by index, a's last entry winning:
    pierce/query/src/applications/pierce_overlap_main.cpp:399-414; per-cell term
    pierce/query/src/cuda/estimated_overlap.cu:8-41; touch-weighted global means
    pierce/query/src/applications/pierce_containment_main.cpp:34-63; effective
    sizes, alpha and the final division pierce_overlap_main.cpp:430-448."""
    f = np.float32
    if len(cells_a) == 0 or len(cells_b) == 0 or not cell_size > 0:
        return dict(has_grid=False, estimated_pairs=0)
    ma = {}
    for c in cells_a:
        ma[(int(c["ix"]), int(c["iy"]), int(c["iz"]))] = c
    cv = f(cell_size) * f(cell_size) * f(cell_size)
    raw, matched = 0.0, 0
    for c in cells_b:
        a = ma.get((int(c["ix"]), int(c["iy"]), int(c["iz"])))
        if a is None:
            continue
        matched += 1
        if a["touch_count"] == 0 or c["touch_count"] == 0:
            continue
        s = f(a["avg_size_mean"]) + f(c["avg_size_mean"]) + f(epsilon)
        prob = f(s * s * s / cv)
        prob = f(prob * f(np.power(np.sqrt(f(a["vol_ratio"]) * f(c["vol_ratio"])), f(gamma))))
        raw += float(a["touch_count"]) * float(c["touch_count"]) * float(min(prob, f(1.0)))

    def mean(cells, key, empty):
        t = w = 0
        for c in cells:
            if c["touch_count"] > 0:
                t += float(c[key]) * int(c["touch_count"])
                w += int(c["touch_count"])
        return f(t / w) if w else f(empty)

    size = [mean(cells_a, "avg_size_mean", 0.0), mean(cells_b, "avg_size_mean", 0.0)]
    ratio = [mean(cells_a, "vol_ratio", 1.0), mean(cells_b, "vol_ratio", 1.0)]
    eff = [f(size[i] * f(np.cbrt(ratio[i]))) for i in range(2)]
    comb = f(eff[0] + eff[1])
    alpha = f(1.0)
    if alpha_correction:
        alpha = max(f(1.0), f(comb * comb * comb / max(cv, f(1e-9))))
    raw32 = f(raw)
    return dict(has_grid=True, matched_cells=matched, raw_pairs=float(raw32), avg_size=tuple(size),
                avg_vol_ratio=tuple(ratio), effective_size=tuple(eff), alpha=float(alpha),
                estimated_pairs=int(f(raw32 / alpha)))


def count_candidate_tests(m1, m2) -> int:
    v1, v2 = _View(m1), _View(m2)
    return int(lib().or_count_candidate_tests(C.byref(v1.v), C.byref(v2.v)))


def touch_overlap_count(m1, m2):
    """TOUCH restatement: number of overlapping pairs and its query time (us)."""
    v1, v2 = _View(m1), _View(m2)
    us = C.c_double()
    n = lib().or_touch_overlap_count(C.byref(v1.v), C.byref(v2.v), C.byref(us))
    return int(n), float(us.value)


def set_threads(n: int) -> None:
    lib().or_set_threads(n)


def get_threads() -> int:
    return int(lib().or_get_threads())
