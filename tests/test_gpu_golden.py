"""GPU joins against the golden brute-force fixtures (tests/golden), for every
predicate, strategy and precision, single and batched."""
import numpy as np
import pytest

import pierce_amd as pz
from tests.golden_io import CASES, PREDICATES, expected, meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uploaded(ctx):
    out = {}
    for case in CASES:
        a, b = meshes(case)
        out[case] = (a, b, ctx.upload(a), ctx.upload(b))
    return out


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("pred", list(PREDICATES))
@pytest.mark.parametrize("precision", [pz.PREC_FP32_FILTERED, pz.PREC_FP64])
@pytest.mark.parametrize("narrow_mode", [0, 1, 2])
def test_two_level_matches_golden(ctx, uploaded, case, pred, precision, narrow_mode):
    print(f"[golden] {case} {pred} precision={precision} narrow={narrow_mode}", flush=True)
    _, _, da, db = uploaded[case]
    out = pz.join(ctx, da, db, predicate=PREDICATES[pred], precision=precision, narrow_mode=narrow_mode)
    np.testing.assert_array_equal(out.pairs, expected(case, pred))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("pred", list(PREDICATES))
@pytest.mark.parametrize("dedup,sizing", [(pz.DEDUP_SORT, pz.SIZING_TWO_PASS), (pz.DEDUP_HASH, pz.SIZING_FIXED)])
def test_global_matches_golden(ctx, uploaded, case, pred, dedup, sizing):
    """The set-wide (reference-shaped) path: edges traced through a BVH over all
    triangles, and for intersection / containment one +Z ray per object through
    the other set's BVH."""
    _, _, da, db = uploaded[case]
    out = pz.join(ctx, da, db, predicate=PREDICATES[pred], strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64,
                  dedup=dedup, sizing=sizing, capacity=1 << 10)
    np.testing.assert_array_equal(out.pairs, expected(case, pred))


def test_batch_of_all_golden_queries(ctx):
    queries = [(case, pred) for case in CASES for pred in PREDICATES]
    parts = [meshes(case) for case, _ in queries]
    da = ctx.upload_batch([p[0] for p in parts])
    db = ctx.upload_batch([p[1] for p in parts])
    outs, summary = pz.join_batch(ctx, da, db, [dict(predicate=PREDICATES[p]) for _, p in queries])
    for (case, pred), o in zip(queries, outs):
        np.testing.assert_array_equal(o.pairs, expected(case, pred), err_msg=f"{case} {pred}")


def _scaled(m, f):
    """The same mesh with every coordinate multiplied by a power of two (exact in fp32)."""
    return pz.Mesh(m.vertices * f, m.indices, m.triangle_to_object, m.edge_starts * f, m.edge_ends * f,
                   m.edge_objects)


@pytest.mark.parametrize("case", CASES)
def test_golden_scaled_beyond_fp32_filter_range(ctx, case):
    """Coordinates above 2^39 switch the fp32 filter off (differences could
    exceed its error analysis) and the walks to their fp64-filtered form.
    Scaling by 2^40 is exact and keeps every orientation sign, so the golden
    pair sets must come back unchanged on every path."""
    f = np.float32(2.0 ** 40)
    a, b = meshes(case)
    A, B = _scaled(a, f), _scaled(b, f)
    assert max(np.abs(A.vertices).max(), np.abs(B.vertices).max()) > 2.0 ** 39
    da, db = ctx.upload(A), ctx.upload(B)
    for pred in PREDICATES:
        for kw in (dict(), dict(narrow_mode=1), dict(narrow_mode=2), dict(precision=pz.PREC_FP64), dict(broad_mode=1)):
            print(f"[golden x2^40] {case} {pred} {kw}", flush=True)
            out = pz.join(ctx, da, db, predicate=PREDICATES[pred], **kw)
            np.testing.assert_array_equal(out.pairs, expected(case, pred), err_msg=f"{case} {pred} {kw}")
    out = pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL)
    np.testing.assert_array_equal(out.pairs, expected(case, "overlap"))
    outs, _ = pz.join_batch(ctx, ctx.upload_batch([A, A]), ctx.upload_batch([B, B]),
                            [dict(predicate=PREDICATES["overlap"]), dict(predicate=PREDICATES["containment"])])
    np.testing.assert_array_equal(outs[0].pairs, expected(case, "overlap"))
    np.testing.assert_array_equal(outs[1].pairs, expected(case, "containment"))
