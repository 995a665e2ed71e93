"""Learned output sizing end to end.

The model is trained on the GPU on (features, exact two-pass counts) of a
query stream; held-out queries then check three things:
1. what it gains: its result-count predictions are closer to the exact counts
   than the reference's grid estimate
   (pierce/query/src/cuda/estimated_overlap.cu:8-77 with the alpha correction
   of pierce_overlap_main.cpp:190-219), and learned-sized batches mostly skip
   the counting pass;
2. parity: learned-sized joins return exactly the two-pass pair sets;
3. fallback: an overflowing prediction falls back to the exact two-pass count.
"""
import time

import numpy as np
import pytest

import pierce_amd as pz
from pierce_amd.stream import make_stream, run_stream, training_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(ctx):
    t0 = time.perf_counter()
    queries = make_stream(768, seed0=1000)
    X, T, _ = training_data(ctx, queries, batch=48)
    t1 = time.perf_counter()
    m = pz.Model(ctx, seed=4)
    losses = m.train(X, T)
    print(f"[learned] {len(queries)} training queries in {t1 - t0:.1f} s, trained in {time.perf_counter() - t1:.1f} s,"
          f" loss {losses[0]:.3f} -> {losses[-1]:.4f}, margins {m.get()[2]}", flush=True)
    yield m, X, T, losses
    m.close()


def test_training_fits_the_stream(trained):
    _, X, T, losses = trained
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0]
    assert T[:, 0].max() > 8  # the stream spans several orders of magnitude of candidates


def test_gain_over_the_grid_estimate(ctx, trained):
    m = trained[0]
    held = make_stream(48, seed0=50000, predicate=pz.PRED_OVERLAP)
    X, T, _ = training_data(ctx, held, batch=24)
    Y = m.predict(X)
    err_mlp = np.abs(Y[:, 1] - T[:, 1])
    est = []
    for A, B, _ in held:
        da = ctx.upload(pz.compute_grid(A, 5.0))
        db = ctx.upload(pz.compute_grid(B, 5.0))
        est.append(pz.join(ctx, da, db, sizing=pz.SIZING_ESTIMATE).predicted_pairs)
        da.close()
        db.close()
    err_grid = np.abs(pz.log_counts(est) - T[:, 1])
    print(f"[learned] median |log2 error| of result pairs: mlp {np.median(err_mlp):.3f}, "
          f"grid estimate {np.median(err_grid):.3f}; candidates: mlp {np.median(np.abs(Y[:, 0] - T[:, 0])):.3f}",
          flush=True)
    assert np.median(err_mlp) < np.median(err_grid)
    assert np.median(err_mlp) < 1.0
    assert np.median(np.abs(Y[:, 0] - T[:, 0])) < 1.0


def test_learned_sizing_matches_two_pass(ctx, trained):
    m = trained[0]
    held = make_stream(96, seed0=60000)
    t0 = time.perf_counter()
    exact, s_exact, _ = run_stream(ctx, held, batch=32, sizing=pz.SIZING_TWO_PASS)
    t1 = time.perf_counter()
    learned, s_learned, _ = run_stream(ctx, held, batch=32, sizing=pz.SIZING_LEARNED, model=m)
    t2 = time.perf_counter()
    for q, (e, l) in enumerate(zip(exact, learned)):
        np.testing.assert_array_equal(l.pairs, e.pairs, err_msg=f"query {q}")
        assert l.num_candidates == e.num_candidates
    covered = np.mean([l.capacity >= e.num_candidates for e, l in zip(exact, learned)])
    overflowed = sum(s.overflowed for s in s_learned)
    print(f"[learned] per-query capacity covers the exact candidates for {covered:.0%} of queries; "
          f"{overflowed}/{len(s_learned)} batches overflowed; wall {t1 - t0:.3f} s two-pass vs "
          f"{t2 - t1:.3f} s learned", flush=True)
    assert covered >= 0.8
    assert overflowed <= len(s_learned) // 2


def test_overflow_falls_back_to_two_pass(ctx, trained):
    m = trained[0]
    p, norm, margin = m.get()
    low = pz.Model(ctx, seed=1)
    q = p.copy()
    q[-2:] = -40.0  # outputs clamp to 0: every sized buffer gets the 64-slot floor
    low.set(params=q, norm=norm, margin=np.zeros(2, np.float32))
    held = make_stream(8, seed0=70000, median=80)
    exact, _, _ = run_stream(ctx, held, batch=8, sizing=pz.SIZING_TWO_PASS)
    got, summ, _ = run_stream(ctx, held, batch=8, sizing=pz.SIZING_LEARNED, model=low)
    assert summ[0].overflowed
    for e, g in zip(exact, got):
        np.testing.assert_array_equal(g.pairs, e.pairs)
    low.close()


@pytest.mark.parametrize("strategy", [pz.STRATEGY_TWO_LEVEL, pz.STRATEGY_GLOBAL])
def test_learned_single_join(ctx, trained, strategy):
    m = trained[0]
    (A, B, _), = make_stream(1, seed0=80000, median=120)
    da, db = ctx.upload(A), ctx.upload(B)
    want = pz.join(ctx, da, db, strategy=strategy, sizing=pz.SIZING_TWO_PASS)
    got = pz.join(ctx, da, db, strategy=strategy, sizing=pz.SIZING_LEARNED, model=m)
    np.testing.assert_array_equal(got.pairs, want.pairs)
    assert got.predicted > 0 and got.predicted_pairs > 0
