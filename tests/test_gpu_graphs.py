"""Repeated joins replay their fixed-shape segments from captured HIP graphs.

Each replay is checked against the eager result, with other joins (eager
work and synchronisation) between the replays, and again after a larger join
has moved the context's buffers (the graphs must be dropped and re-captured).
The segments are the object statistics of both sets, the layout of the
objects that occur in candidate pairs (about 40 small launches whose shapes
depend only on the object counts) and, when the broad phase walks one (WALK),
the mesh2 object index: three graph launches per join through the
per-object trees (narrow_mode 2).  With the box scan there is no object index
and two; the pair sieve (narrow_mode 0) needs no layout, so one segment fewer.
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


WALK = dict(broad_mode=1, narrow_mode=2)


def _sets(count, seed, box):
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=count, seed=seed, box=box, jitter=0.02)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=count, seed=seed + 1, box=box, jitter=0.02)
    return a, b


def test_replays_match_eager_with_interleaved_work(ctx):
    a, b = _sets(1000, 42, 100.0)
    c, d = _sets(300, 7, 40.0)
    da, db, dc, dd = ctx.upload(a), ctx.upload(b), ctx.upload(c), ctx.upload(d)
    ctx.set_graphs(False)
    want_ab = pz.join(ctx, da, db, **WALK)
    want_cd = pz.join(ctx, dc, dd, predicate=pz.PRED_INTERSECTION)
    np.testing.assert_array_equal(want_ab.pairs, orc.face_overlap(a, b))
    ctx.set_graphs(True)
    launches = []
    for i in range(5):
        out = pz.join(ctx, da, db, **WALK)
        np.testing.assert_array_equal(out.pairs, want_ab.pairs, err_msg=f"replay {i}")
        assert out.num_candidates == want_ab.num_candidates
        launches.append(out.graph_launches)
        other = pz.join(ctx, dc, dd, predicate=pz.PRED_INTERSECTION)  # eager or its own graphs, in between
        np.testing.assert_array_equal(other.pairs, want_cd.pairs, err_msg=f"interleaved {i}")
    # first join of each key runs eagerly; the second captures; later ones replay
    assert launches[0] == 0 and all(n == 3 for n in launches[2:]), launches


def test_buffer_growth_drops_graphs(ctx):
    a, b = _sets(500, 11, 70.0)
    big_a, big_b = _sets(4000, 13, 160.0)
    da, db = ctx.upload(a), ctx.upload(b)
    want = orc.face_overlap(a, b)
    for _ in range(3):
        np.testing.assert_array_equal(pz.join(ctx, da, db, **WALK).pairs, want)
    dba, dbb = ctx.upload(big_a), ctx.upload(big_b)
    np.testing.assert_array_equal(pz.join(ctx, dba, dbb, **WALK).pairs, orc.face_overlap(big_a, big_b))
    first = pz.join(ctx, da, db, **WALK)  # buffers moved: runs eagerly and re-records
    np.testing.assert_array_equal(first.pairs, want)
    assert first.graph_launches == 0
    for _ in range(3):
        out = pz.join(ctx, da, db, **WALK)
        np.testing.assert_array_equal(out.pairs, want)
    assert out.graph_launches == 3
    for _ in range(3):  # box scan: statistics and layout segments only
        out = pz.join(ctx, da, db, broad_mode=2, narrow_mode=2)
        np.testing.assert_array_equal(out.pairs, want)
    assert out.graph_launches == 2


def test_sieve_replays_statistics_and_object_index(ctx):
    a, b = _sets(600, 31, 80.0)
    da, db = ctx.upload(a), ctx.upload(b)
    want = orc.face_overlap(a, b)
    for _ in range(4):
        out = pz.join(ctx, da, db, broad_mode=1)
        np.testing.assert_array_equal(out.pairs, want)
        assert out.sieve_pairs == out.num_candidates
    assert out.graph_launches == 2


def test_graphs_with_learned_sizing_side_stream(ctx):
    a, b = _sets(800, 21, 90.0)
    da, db = ctx.upload(a), ctx.upload(b)
    want = orc.face_overlap(a, b)
    m = pz.Model(ctx, seed=3)
    rng = np.random.default_rng(0)
    X = rng.normal(0, 1, (64, pz.MLP_FEATURES)).astype(np.float32)
    m.train(X, np.abs(rng.normal(8, 2, (64, 2))).astype(np.float32), epochs=5)
    for _ in range(4):
        out = pz.join(ctx, da, db, sizing=pz.SIZING_LEARNED, model=m, **WALK)
        np.testing.assert_array_equal(out.pairs, want)
    assert out.graph_launches == 3
    m.close()
