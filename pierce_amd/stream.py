"""Synthetic join-query streams and training data for the learned sizing.

A query is (set A, set B, predicate).  Both sets mix the three mesh kinds of
pz_generate's QUERY_MIX generator:
- ellipsoids at subdivision 1-4;
- 12-triangle axis-aligned cubes;
- branching tubes.

Set sizes are lognormal.  Each query draws its own scene density, and B is
shifted against A, so the selectivity varies over several orders of
magnitude.  Training targets are log2(1 + exact counts) from the batched
two-pass join: candidate object pairs and result pairs.
"""
from __future__ import annotations

import numpy as np

from . import (GEN_CUBES, GEN_ELLIPSOIDS, GEN_QUERY_MIX, GEN_TUBES, INSTANCE_DTYPE, PRED_CONTAINMENT,
               PRED_INTERSECTION, PRED_OVERLAP, SIZING_TWO_PASS, Context, Mesh, batch_features, generate, join_batch,
               log_counts)


def translate(mesh: Mesh, offset) -> Mesh:
    """Copy of mesh moved by offset (grid statistics dropped)."""
    o = np.asarray(offset, np.float32).reshape(1, 3)
    return Mesh(mesh.vertices + o, mesh.indices, mesh.triangle_to_object, mesh.edge_starts + o, mesh.edge_ends + o,
                mesh.edge_objects)


def make_query(seed: int, median: float = 40.0, sigma: float = 0.9, max_meshes: int = 400,
               tris=(500, 1500, 4000), density=(4.0, 14.0), predicate=None):
    """(A, B, predicate) of one query; tris = (min, median, max) tube triangles."""
    rng = np.random.default_rng(seed)
    nA, nB = (int(np.clip(round(rng.lognormal(np.log(median), sigma)), 2, max_meshes)) for _ in range(2))
    side = float((nA + nB) ** (1.0 / 3.0) * rng.uniform(*density))
    kw = dict(box=side, tri_min=tris[0], tri_median=tris[1], tri_max=tris[2])
    A = generate(GEN_QUERY_MIX, count=nA, seed=2 * seed + 1, **kw)
    B = generate(GEN_QUERY_MIX, count=nB, seed=2 * seed + 2, **kw)
    B = translate(B, rng.uniform(-0.4, 0.4, 3) * side)
    pred = int(rng.integers(0, 3)) if predicate is None else int(predicate)
    return A, B, pred


def make_stream(n: int, seed0: int = 0, **kw):
    return [make_query(seed0 + i, **kw) for i in range(n)]


def run_stream(ctx: Context, queries, batch: int = 32, with_features: bool = False, **join_kw):
    """Run the queries as batched joins of `batch` queries each.

    Returns (outputs per query, batch summaries, features or None)."""
    outs, summaries, feats = [], [], []
    for i in range(0, len(queries), batch):
        chunk = queries[i:i + batch]
        da = ctx.upload_batch([q[0] for q in chunk])
        db = ctx.upload_batch([q[1] for q in chunk])
        try:
            preds = [q[2] for q in chunk]
            if with_features:
                feats.append(batch_features(ctx, da, db, preds))
            o, s = join_batch(ctx, da, db, [dict(predicate=p) for p in preds], **join_kw)
            outs.extend(o)
            summaries.append(s)
        finally:
            da.close()
            db.close()
    return outs, summaries, (np.concatenate(feats) if with_features else None)


def training_data(ctx: Context, queries, batch: int = 32):
    """(X [n][32], T [n][2], outputs) with T from the exact two-pass join."""
    outs, _, X = run_stream(ctx, queries, batch, with_features=True, sizing=SIZING_TWO_PASS)
    T = np.stack([log_counts([o.num_candidates for o in outs]), log_counts([len(o.pairs) for o in outs])], axis=1)
    return X, T, outs


__all__ = ["translate", "make_query", "make_stream", "run_stream", "training_data", "PRED_OVERLAP",
           "PRED_INTERSECTION", "PRED_CONTAINMENT", "concat_meshes", "centre_objects", "stream_catalogue",
           "random_rotations", "stream_query", "library_batches", "run_library_stream", "instantiate_host"]


# ----------------------------------------------------------------------------
# Query streams over a mesh library: every set is a list of instances of a
# resident catalogue of meshes (pz_library_*), so a long stream never crosses
# PCIe with its geometry.

def concat_meshes(meshes):
    """One mesh holding the given meshes as consecutive objects (ids shifted)."""
    v, idx, t2o, es, ee, eo = [], [], [], [], [], []
    nv = 0
    no = 0
    for m in meshes:
        v.append(m.vertices)
        idx.append(m.indices + np.uint32(nv))
        t2o.append(m.triangle_to_object + np.int32(no))
        es.append(m.edge_starts)
        ee.append(m.edge_ends)
        eo.append(m.edge_objects + np.int32(no))
        nv += m.vertices.shape[0]
        no += m.num_objects
    return Mesh(np.concatenate(v), np.concatenate(idx), np.concatenate(t2o), np.concatenate(es), np.concatenate(ee),
                np.concatenate(eo))


def centre_objects(m: Mesh) -> Mesh:
    """Each object moved so that its box centre is the origin (fp32).  Objects
    must own consecutive runs of vertices (as the generators store them)."""
    vobj = np.empty(m.vertices.shape[0], np.int64)
    vobj[m.indices.reshape(-1)] = np.repeat(m.triangle_to_object.astype(np.int64), 3)
    if np.any(np.diff(vobj) < 0):
        raise ValueError("objects do not own consecutive vertex runs")
    starts = np.flatnonzero(np.r_[True, vobj[1:] != vobj[:-1]])
    lo = np.minimum.reduceat(m.vertices, starts, axis=0)
    hi = np.maximum.reduceat(m.vertices, starts, axis=0)
    c = np.zeros((m.num_objects, 3), np.float32)
    c[vobj[starts]] = ((lo + hi) * np.float32(0.5)).astype(np.float32)
    return Mesh(m.vertices - c[vobj], m.indices, m.triangle_to_object, m.edge_starts - c[m.edge_objects],
                m.edge_ends - c[m.edge_objects], m.edge_objects)


def stream_catalogue(seed: int = 5, ellipsoids_per_level: int = 256, tubes: int = 512,
                     tube_tris=(5000, 15000, 50000)):
    """The catalogue of the library stream: ellipsoids at subdivision 1-4
    (radii U[1, 3], random orientation), one unit cube, and branching tubes of
    tube_tris = (min, median, max) triangles (lognormal, sigma 0.6), each
    centred at the origin.  Returns (mesh, ids) with ids the first object of
    each kind."""
    parts, ids = [], {}
    for lv in range(1, 5):
        ids[f"ell{lv}"] = sum(p.num_objects for p in parts)
        parts.append(generate(GEN_ELLIPSOIDS, count=ellipsoids_per_level, seed=seed * 100 + lv, level_min=lv,
                              level_max=lv, radius_min=1.0, radius_max=3.0, box=0.0))
    ids["cube"] = sum(p.num_objects for p in parts)
    parts.append(generate(GEN_CUBES, count=1, seed=0, radius_min=1.0, radius_max=1.0, box=0.0))
    ids["tube"] = sum(p.num_objects for p in parts)
    parts.append(centre_objects(generate(GEN_TUBES, count=tubes, seed=seed, box=0.0, tri_min=tube_tris[0],
                                         tri_median=tube_tris[1], tri_max=tube_tris[2], tri_sigma=0.6)))
    ids["ellipsoids_per_level"] = ellipsoids_per_level
    ids["tubes"] = tubes
    return concat_meshes(parts), ids


def random_rotations(rng, n):
    """n uniform random rotation matrices (Shoemake's quaternion method), (n, 3, 3) float64."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w, x, y, z = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3), \
        b * np.cos(2 * np.pi * u3)
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def stream_query(q: int, ids: dict, seed0: int = 0, median: float = 500.0, sigma: float = 0.9,
                 max_meshes: int = 50000, spacing=(30.0, 120.0), predicate=None):
    """One query of the library stream: (instances of A, instances of B, predicate).

    Set sizes lognormal(median, sigma) clipped to [1, max_meshes]; per mesh 50%
    ellipsoids (subdivision uniform in 1-4, random orientation), 30% axis-aligned
    cubes with edge U[1, 10], 20% tubes (random orientation); centres uniform in
    [0, L)^3 with L = s * (nA + nB)^(1/3) and the spacing s uniform in `spacing`
    per query."""
    rng = np.random.default_rng([seed0, q])
    n = [int(np.clip(round(rng.lognormal(np.log(median), sigma)), 1, max_meshes)) for _ in range(2)]
    pred = int(rng.integers(0, 3)) if predicate is None else int(predicate)
    side = float(rng.uniform(*spacing) * (n[0] + n[1]) ** (1.0 / 3.0))
    sets = []
    E, T = ids["ellipsoids_per_level"], ids["tubes"]
    for k in n:
        kind = rng.choice(3, size=k, p=[0.5, 0.3, 0.2])
        lvl = rng.integers(1, 5, size=k)
        ell = np.array([ids["ell1"], ids["ell2"], ids["ell3"], ids["ell4"]])[lvl - 1] + rng.integers(0, E, size=k)
        mesh = np.where(kind == 0, ell, np.where(kind == 1, ids["cube"], ids["tube"] + rng.integers(0, T, size=k)))
        A = random_rotations(rng, k)
        edge = rng.uniform(1.0, 10.0, size=k)
        A[kind == 1] = np.eye(3)[None] * edge[kind == 1, None, None]
        inst = np.zeros(k, INSTANCE_DTYPE)
        inst["mesh"] = mesh.astype(np.uint32)
        inst["xform"][:, :9] = A.reshape(k, 9).astype(np.float32)
        inst["xform"][:, 9:] = rng.uniform(0.0, side, size=(k, 3)).astype(np.float32)
        sets.append(inst)
    return sets[0], sets[1], pred


def library_batches(queries, tri_per_object, max_tris: int = 120_000_000, max_queries: int = 256):
    """Consecutive runs of queries whose sets hold at most max_tris triangles a side."""
    out, cur, ta, tb = [], [], 0, 0
    for q in queries:
        a = int(tri_per_object[q[0]["mesh"]].sum())
        b = int(tri_per_object[q[1]["mesh"]].sum())
        if cur and (ta + a > max_tris or tb + b > max_tris or len(cur) >= max_queries):
            out.append(cur)
            cur, ta, tb = [], 0, 0
        cur.append(q)
        ta += a
        tb += b
    if cur:
        out.append(cur)
    return out


def run_library_stream(ctx: Context, library, queries, tri_per_object, with_features: bool = False,
                       max_tris: int = 120_000_000, **join_kw):
    """Run a library stream as batched joins.  Returns (outputs per query,
    batch summaries, features or None); instancing is outside the joins' times."""
    outs, summaries, feats = [], [], []
    for chunk in library_batches(queries, tri_per_object, max_tris):
        # two reusable instancing slots: no device allocation once they are large enough
        da = library.instantiate([len(q[0]) for q in chunk], np.concatenate([q[0] for q in chunk]), slot=0)
        db = library.instantiate([len(q[1]) for q in chunk], np.concatenate([q[1] for q in chunk]), slot=1)
        try:
            preds = [q[2] for q in chunk]
            if with_features:
                feats.append(batch_features(ctx, da, db, preds))
            o, s = join_batch(ctx, da, db, [dict(predicate=p) for p in preds], **join_kw)
            outs.extend(o)
            summaries.append(s)
        finally:
            da.close()
            db.close()
    return outs, summaries, (np.concatenate(feats) if with_features else None)


def instantiate_host(lib_mesh, counts, inst):
    """The parts pz_library_instantiate builds, in numpy: per instance the
    library object's vertices, triangles and edges (edges in their order within
    the object), mapped as ((a0 x + a1 y) + a2 z) + t0 per row with every fp32
    product and sum rounded.  Returns one Mesh per part (part-local ids)."""
    t2o = lib_mesh.triangle_to_object
    f = np.float32

    def runs(o):  # an object id without triangles has an empty run (lo > hi)
        tri = lib_mesh.indices[t2o == o]
        if tri.shape[0] == 0:
            return tri, 0, -1
        lo, hi = int(tri.min()), int(tri.max())
        return tri, lo, hi

    def mapped(A, t, p):
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        return np.stack([((f(A[r, 0]) * x + f(A[r, 1]) * y) + f(A[r, 2]) * z) + f(t[r]) for r in range(3)], 1)

    parts, k = [], 0
    for c in counts:
        v, idx, to, es, ee, eo = [], [], [], [], [], []
        nv = 0
        for j in range(int(c)):
            x = inst[k]
            k += 1
            o = int(x["mesh"])
            A = x["xform"][:9].reshape(3, 3).astype(np.float32)
            t = x["xform"][9:].astype(np.float32)
            tri, lo, hi = runs(o)
            v.append(mapped(A, t, lib_mesh.vertices[lo:hi + 1]))
            idx.append(tri - np.uint32(lo) + np.uint32(nv))
            to.append(np.full(len(tri), j, np.int32))
            sel = lib_mesh.edge_objects == o
            es.append(mapped(A, t, lib_mesh.edge_starts[sel]))
            ee.append(mapped(A, t, lib_mesh.edge_ends[sel]))
            eo.append(np.full(int(sel.sum()), j, np.int32))
            nv += hi - lo + 1
        cat = (lambda a, shape: np.concatenate(a) if a else np.zeros(shape, np.float32))
        parts.append(Mesh(cat(v, (0, 3)), np.concatenate(idx) if idx else np.zeros((0, 3), np.uint32),
                             np.concatenate(to) if to else np.zeros(0, np.int32), cat(es, (0, 3)), cat(ee, (0, 3)),
                             np.concatenate(eo) if eo else np.zeros(0, np.int32)))
    return parts
