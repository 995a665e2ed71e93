"""Batched joins: one pass over many independent (mesh1 part, mesh2 part)
queries must equal the per-query joins and the oracle, query by query."""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from tests.meshes import containment_sets

pytestmark = pytest.mark.gpu


def query_sets(nq, seed=100):
    rng = np.random.default_rng(seed)
    A, B = [], []
    for q in range(nq):
        n = int(rng.integers(1, 120))
        box = float(rng.uniform(8, 40))
        lvl = int(rng.integers(0, 3))
        a = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=1000 + q, level_min=lvl, level_max=lvl, radius_min=0.5,
                        radius_max=2.0, box=box)
        m = int(rng.integers(1, 120))
        if q % 3 == 0:
            b = pz.generate(pz.GEN_ELLIPSOIDS, count=m, seed=5000 + q, level_min=lvl, level_max=lvl, radius_min=0.5,
                            radius_max=2.0, box=box, copy_fraction=0.3, copy_scale_min=0.3, copy_scale_max=0.8,
                            copy_source_count=n, copy_source_seed=1000 + q)
        elif q % 3 == 1:
            b = pz.generate(pz.GEN_CUBES, count=m, seed=5000 + q, radius_min=1.0, radius_max=6.0, box=box)
        else:
            b = pz.generate(pz.GEN_SPHERES_MIXED, count=m, seed=5000 + q, level_min=1, level_max=3, radius_min=0.5,
                            radius_max=2.0, box=box)
        A.append(a)
        B.append(b)
    return A, B


def expected(a, b, pred, include_overlap=False):
    if pred == pz.PRED_OVERLAP:
        return orc.face_overlap(a, b), None
    return orc.predicate_join(a, b, pred, include_overlap)


@pytest.mark.parametrize("sizing", [pz.SIZING_TWO_PASS, pz.SIZING_FIXED])
def test_batch_matches_oracle_per_query(ctx, sizing):
    nq = 24
    A, B = query_sets(nq)
    preds = [q % 3 for q in range(nq)]
    print(f"[batch] {nq} queries sizing={sizing}", flush=True)
    da, db = ctx.upload_batch(A), ctx.upload_batch(B)
    outs, summ = pz.join_batch(ctx, da, db, [dict(predicate=p) for p in preds], sizing=sizing, capacity=64)
    for q in range(nq):
        exp, _ = expected(A[q], B[q], preds[q])
        np.testing.assert_array_equal(outs[q].pairs, exp, err_msg=f"query {q} predicate {preds[q]}")
    assert summ.num_candidates >= sum(len(o.pairs) for o in outs)


def test_batch_equals_single_joins(ctx):
    nq = 10
    A, B = query_sets(nq, seed=7)
    da, db = ctx.upload_batch(A), ctx.upload_batch(B)
    queries = [dict(predicate=pz.PRED_CONTAINMENT, include_overlap_pairs=q % 2) for q in range(nq)]
    outs, _ = pz.join_batch(ctx, da, db, queries)
    for q in range(nq):
        sa, sb = ctx.upload(A[q]), ctx.upload(B[q])
        single = pz.join(ctx, sa, sb, **queries[q])
        np.testing.assert_array_equal(outs[q].pairs, single.pairs)
        if q % 2:
            np.testing.assert_array_equal(outs[q].types, single.types)
        else:
            assert outs[q].types is None


def test_batch_of_one_is_the_join(ctx):
    a, b = containment_sets()
    da, db = ctx.upload_batch([a]), ctx.upload_batch([b])
    outs, _ = pz.join_batch(ctx, da, db, [dict(predicate=pz.PRED_INTERSECTION)])
    np.testing.assert_array_equal(outs[0].pairs, orc.predicate_join(a, b, 1)[0])
