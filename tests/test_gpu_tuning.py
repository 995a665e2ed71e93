"""Size-class launches of the LDS index builder and the autotuner.

Every workgroup size must give the oracle's pair set (the BVH shape never
changes results).  The autotuner must time each size class present in the
workload and keep a valid choice per class.
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed(ctx):
    # levels 1-4: 80, 320, 1280 and 5120 triangles -> all three LDS classes and the global pipeline
    kw = dict(box=60.0, radius_min=1.0, radius_max=4.0, level_min=1, level_max=4)
    a = pz.generate(pz.GEN_SPHERES_MIXED, count=300, seed=7, **kw)
    b = pz.generate(pz.GEN_SPHERES_MIXED, count=300, seed=8, **kw)
    return a, b, ctx.upload(a), ctx.upload(b), orc.face_overlap(a, b)


@pytest.mark.parametrize("threads", [128, 256, 512, 1024])
def test_every_workgroup_size_gives_the_oracle_pairs(ctx, mixed, threads):
    a, b, da, db, exp = mixed
    saved = pz.tune_get(ctx)
    try:
        q = {128: 256, 256: 512, 512: 1024, 1024: 1536}[threads]
        pz.tune_set(ctx, [threads] * 3, narrow_queue=(q, q))
        out = pz.join(ctx, da, db)
        np.testing.assert_array_equal(out.pairs, exp)
        inter = pz.join(ctx, da, db, predicate=pz.PRED_INTERSECTION)
        np.testing.assert_array_equal(inter.pairs, orc.predicate_join(a, b, pz.PRED_INTERSECTION)[0])
    finally:
        pz.tune_set(ctx, saved["build_threads"], saved["narrow_queue"], saved["narrow_chunk"])


def test_autotune_times_each_class(ctx, mixed):
    a, b, da, db, exp = mixed
    saved = pz.tune_get(ctx)
    try:
        rep = pz.autotune(ctx, da, db)
        print(f"[tune] {rep}", flush=True)
        assert rep["joins_run"] == 13
        assert rep["narrow_chunk"][rep["queue_class"]] in pz.CHUNK_CANDIDATES
        q = rep["narrow_queue"][rep["queue_class"]]
        assert rep["queue_class"] == 1  # level-4 spheres (5,120 triangles) exceed the LDS builder
        assert q in pz.QUEUE_CANDIDATES
        # report times are rounded to 1e-4 ms: the chosen entry is a minimum up to that rounding
        assert rep["narrow_ms"][str(q)] <= min(rep["narrow_ms"].values()) + 1e-4
        timed = 0
        for k in range(3):
            ms = rep["build_ms"][f"class{k}"]
            if all(v is None for v in ms.values()):
                continue  # no needed object of this size class
            timed += 1
            assert all(v is not None and v > 0 for v in ms.values()), (k, ms)
            assert ms[str(rep["build_threads"][k])] <= min(ms.values()) + 1e-4, (k, ms)
        assert timed >= 2
        assert pz.tune_get(ctx)["build_threads"] == rep["build_threads"]
        assert pz.tune_get(ctx)["narrow_queue"] == rep["narrow_queue"]
        np.testing.assert_array_equal(pz.join(ctx, da, db).pairs, exp)
    finally:
        pz.tune_set(ctx, saved["build_threads"], saved["narrow_queue"], saved["narrow_chunk"])


def test_tune_set_rejects_bad_sizes(ctx):
    with pytest.raises(pz.PierceError):
        pz.tune_set(ctx, [100, 256, 512])
    with pytest.raises(pz.PierceError):
        pz.tune_set(ctx, [128, 256, 512], narrow_queue=(300, 512))
