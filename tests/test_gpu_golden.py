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
@pytest.mark.parametrize("narrow_mode", [0, 1])
def test_two_level_matches_golden(ctx, uploaded, case, pred, precision, narrow_mode):
    print(f"[golden] {case} {pred} precision={precision} narrow={narrow_mode}", flush=True)
    _, _, da, db = uploaded[case]
    out = pz.join(ctx, da, db, predicate=PREDICATES[pred], precision=precision, narrow_mode=narrow_mode)
    np.testing.assert_array_equal(out.pairs, expected(case, pred))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dedup,sizing", [(pz.DEDUP_SORT, pz.SIZING_TWO_PASS), (pz.DEDUP_HASH, pz.SIZING_FIXED)])
def test_global_matches_golden(ctx, uploaded, case, dedup, sizing):
    _, _, da, db = uploaded[case]
    out = pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64, dedup=dedup, sizing=sizing,
                  capacity=1 << 10)
    np.testing.assert_array_equal(out.pairs, expected(case, "overlap"))


def test_batch_of_all_golden_queries(ctx):
    queries = [(case, pred) for case in CASES for pred in PREDICATES]
    parts = [meshes(case) for case, _ in queries]
    da = ctx.upload_batch([p[0] for p in parts])
    db = ctx.upload_batch([p[1] for p in parts])
    outs, summary = pz.join_batch(ctx, da, db, [dict(predicate=PREDICATES[p]) for _, p in queries])
    for (case, pred), o in zip(queries, outs):
        np.testing.assert_array_equal(o.pairs, expected(case, pred), err_msg=f"{case} {pred}")
