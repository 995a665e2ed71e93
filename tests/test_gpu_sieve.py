"""The pair sieve (narrow_mode 0) against the oracle and the per-object trees.

The sieve decides candidate pairs of objects with at most 8,192 contiguous
triangles in LDS, without trees; pairs beyond its caps go to the per-object
trees (narrow_mode 2).  Both must return the oracle's pairs, and on inputs
where no walk stops early they test the same triangle pairs and cast the same
parity rays, so the counters agree too.
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from tests.meshes import containment_sets

pytestmark = pytest.mark.gpu

PREDS = (pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT)


def _expected(a, b, pred):
    return orc.face_overlap(a, b) if pred == pz.PRED_OVERLAP else orc.predicate_join(a, b, pred)[0]


def test_small_objects_all_sieved(ctx):
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=600, seed=42, box=80.0, jitter=0.02)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=600, seed=43, box=80.0, jitter=0.02)
    da, db = ctx.upload(a), ctx.upload(b)
    for pred in PREDS:
        out = pz.join(ctx, da, db, predicate=pred)
        print(f"[sieve small] predicate {pred}: {out.sieve_pairs} of {out.num_candidates} sieved", flush=True)
        assert out.num_candidates > 0 and out.sieve_pairs == out.num_candidates
        assert out.index_triangles == 0  # no per-object tree was built
        np.testing.assert_array_equal(out.pairs, _expected(a, b, pred))


def test_large_objects_deferred_to_trees(ctx):
    # levels 1..5: the 20,480-triangle spheres exceed the sieve's object cap
    a = pz.generate(pz.GEN_SPHERES_MIXED, count=60, seed=71, level_min=1, level_max=5, radius_min=1.0,
                    radius_max=4.0, box=30.0)
    b = pz.generate(pz.GEN_SPHERES_MIXED, count=60, seed=72, level_min=1, level_max=5, radius_min=1.0,
                    radius_max=4.0, box=30.0)
    assert np.bincount(a.triangle_to_object).max() > 8192
    da, db = ctx.upload(a), ctx.upload(b)
    for pred in PREDS:
        out = pz.join(ctx, da, db, predicate=pred)
        print(f"[sieve large] predicate {pred}: {out.sieve_pairs} of {out.num_candidates} sieved", flush=True)
        assert 0 < out.sieve_pairs < out.num_candidates
        exp = _expected(a, b, pred)
        np.testing.assert_array_equal(out.pairs, exp)
        np.testing.assert_array_equal(pz.join(ctx, da, db, predicate=pred, narrow_mode=2).pairs, exp)


def test_many_kept_triangles_deferred(ctx):
    # a 5,120-triangle sphere and its copy scaled by 1 + 2^-10 about the centre:
    # every triangle of both meets the overlap box (more than the sieve keeps)
    s = pz.generate(pz.GEN_ELLIPSOIDS, count=1, seed=5, level_min=4, level_max=4, radius_min=2.0, radius_max=2.0,
                    box=0.0)
    t = pz.Mesh(s.vertices * np.float32(1 + 2.0 ** -10), s.indices, s.triangle_to_object)
    a, b = pz.extract_edges(s), pz.extract_edges(t)
    da, db = ctx.upload(a), ctx.upload(b)
    for pred in PREDS:
        out = pz.join(ctx, da, db, predicate=pred)
        assert out.num_candidates == 1 and out.sieve_pairs == 0
        np.testing.assert_array_equal(out.pairs, _expected(a, b, pred))


def _diamonds(centres):
    """Edge-2 cubes turned 45 degrees about z, bottom at z = 0 and top at z = 2."""
    from tests.meshes import cube_mesh

    c = np.float32(np.sqrt(0.5))
    m = cube_mesh([np.array([0, 0, 1], np.float32)] * len(centres), [2.0] * len(centres))
    v = m.vertices.copy()
    x, y = v[:, 0].copy(), v[:, 1].copy()
    v[:, 0], v[:, 1] = (x - y) * c, (x + y) * c
    for o, (cx, cy) in enumerate(centres):
        sel = np.unique(m.indices[m.triangle_to_object == o])
        v[sel, 0] += np.float32(cx)
        v[sel, 1] += np.float32(cy)
    return pz.extract_edges(pz.Mesh(v, m.indices, m.triangle_to_object))


def test_escalations_match_the_trees(ctx):
    """Interleaved grids of turned cubes: coplanar top and bottom faces with
    gaps between them, so triangle boxes meet, surfaces do not, and the fp32
    filter leaves the coplanar pairs open.  No walk stops early, so the sieve
    and the trees test the same triangle pairs and escalate the same ones."""
    grid = [(4.0 * i, 4.0 * j) for i in range(8) for j in range(8)]
    a, b = _diamonds(grid), _diamonds([(x + 2, y + 2) for x, y in grid])
    assert len(orc.face_overlap(a, b)) == 0
    da, db = ctx.upload(a), ctx.upload(b)
    s, t = pz.join(ctx, da, db), pz.join(ctx, da, db, narrow_mode=2)
    print(f"[sieve escalations] sieve {s.tests_total}/{s.tests_fp64 + s.tests_exact}, "
          f"trees {t.tests_total}/{t.tests_fp64 + t.tests_exact}", flush=True)
    assert len(s.pairs) == 0 and len(t.pairs) == 0
    assert s.sieve_pairs == s.num_candidates > 0
    assert s.tests_fp64 + s.tests_exact > 0
    assert (s.tests_total, s.tests_fp64, s.tests_exact) == (t.tests_total, t.tests_fp64, t.tests_exact)


@pytest.mark.parametrize("direction", [pz.DIR_BOTH, pz.DIR_1TO2, pz.DIR_2TO1])
@pytest.mark.parametrize("pred", [pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT])
def test_parity_rays_match_the_trees(ctx, pred, direction):
    a, b = containment_sets()
    da, db = ctx.upload(a), ctx.upload(b)
    s = pz.join(ctx, da, db, predicate=pred, direction=direction)
    t = pz.join(ctx, da, db, predicate=pred, direction=direction, narrow_mode=2)
    exp = orc.predicate_join(a, b, pred, direction=direction)[0]
    np.testing.assert_array_equal(s.pairs, exp)
    np.testing.assert_array_equal(t.pairs, exp)
    assert s.sieve_pairs == s.num_candidates
    key = lambda o: (o.parity_rays, o.parity_crossings, o.parity_rays_crossing, o.parity_inside, o.pairs_touching)
    print(f"[sieve parity] predicate {pred} direction {direction}: {key(s)}", flush=True)
    assert s.parity_rays > 0 and key(s) == key(t)
