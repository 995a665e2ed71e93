"""Object bounds from vertex runs against the per-triangle bounds kernel.

A dataset whose objects own disjoint, gap-free vertex runs (found at upload)
has its object boxes computed from those runs; any other layout goes through
the per-triangle kernel.  Either way the join must return the oracle's pairs
(bit-exact sorted unique pairs), and the layout flag must say which path ran.
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _spheres(seed):
    # levels 1..5: the 10,242-vertex spheres span several 4,096-vertex chunks
    return pz.generate(pz.GEN_SPHERES_MIXED, count=60, seed=seed, level_min=1, level_max=5, radius_min=1.0,
                       radius_max=4.0, box=30.0)


def _remesh(m, vertices, indices, t2o=None):
    t2o = m.triangle_to_object if t2o is None else t2o
    return pz.extract_edges(pz.Mesh(np.ascontiguousarray(vertices, np.float32),
                                    np.ascontiguousarray(indices, np.uint32), np.ascontiguousarray(t2o, np.int32)))


def _permuted(m, seed):
    p = np.random.default_rng(seed).permutation(m.vertices.shape[0])
    inv = np.empty_like(p)
    inv[p] = np.arange(p.shape[0])
    return _remesh(m, m.vertices[p], inv[m.indices])


def _hole(m):
    # an unreferenced vertex inside object 0's run
    k = int(m.indices[m.triangle_to_object == 0].min()) + 1
    v = np.insert(m.vertices, k, np.float32([500.0, 500.0, 500.0]), axis=0)
    idx = m.indices.astype(np.int64)
    idx[idx >= k] += 1
    return _remesh(m, v, idx)


def _shared(m):
    # object 1's first triangle takes a corner of object 0
    idx = m.indices.copy()
    t = int(np.flatnonzero(m.triangle_to_object == 1)[0])
    idx[t, 0] = m.indices[m.triangle_to_object == 0].max()
    return _remesh(m, m.vertices, idx)


def _tail(m):
    # unreferenced vertices after every run
    v = np.concatenate([m.vertices, np.full((5, 3), -700.0, np.float32)])
    return _remesh(m, v, m.indices)


def _ragged(m):
    # ids 3, 10, 17, ... own no triangle (their vertices are then unreferenced
    # and outside every run of the other objects)
    keep = (m.triangle_to_object % 7) != 3
    return _remesh(m, m.vertices, m.indices[keep], m.triangle_to_object[keep])


CASES = [("runs", lambda m, s: m, True), ("permuted", _permuted, False), ("hole", lambda m, s: _hole(m), False),
         ("shared", lambda m, s: _shared(m), False), ("tail", lambda m, s: _tail(m), True),
         ("ragged", lambda m, s: _ragged(m), True)]


@pytest.mark.parametrize("name,make,runs", CASES, ids=[c[0] for c in CASES])
def test_object_bounds_layouts(ctx, name, make, runs):
    a, b = make(_spheres(61), 1), make(_spheres(62), 2)
    da, db = ctx.upload(a), ctx.upload(b)
    print(f"[layout] {name}: vertex runs {da.vertex_runs}/{db.vertex_runs}", flush=True)
    assert da.vertex_runs == runs and db.vertex_runs == runs
    exp = orc.face_overlap(a, b)
    assert len(exp) > 0
    np.testing.assert_array_equal(pz.join(ctx, da, db).pairs, exp)
    np.testing.assert_array_equal(pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL).pairs, exp)
    if name in ("runs", "permuted"):
        for pred in (pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT):
            np.testing.assert_array_equal(pz.join(ctx, da, db, predicate=pred).pairs,
                                          orc.predicate_join(a, b, pred)[0])


def test_runs_and_permuted_give_the_same_candidates(ctx):
    a, b = _spheres(63), _spheres(64)
    ap, bp = _permuted(a, 3), _permuted(b, 4)
    da, db, dap, dbp = ctx.upload(a), ctx.upload(b), ctx.upload(ap), ctx.upload(bp)
    assert da.vertex_runs and not dap.vertex_runs
    r1, r2 = pz.join(ctx, da, db), pz.join(ctx, dap, dbp)
    # the object boxes are equal, so the broad phase finds the same candidates
    assert r1.num_candidates == r2.num_candidates
    np.testing.assert_array_equal(r1.pairs, r2.pairs)
