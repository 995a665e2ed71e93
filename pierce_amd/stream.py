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

from . import (GEN_QUERY_MIX, PRED_CONTAINMENT, PRED_INTERSECTION, PRED_OVERLAP, SIZING_TWO_PASS, Context, Mesh,
               batch_features, generate, join_batch, log_counts)


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
           "PRED_INTERSECTION", "PRED_CONTAINMENT"]
