"""The LDS builder at the sizes where its code paths change.

Objects of 1 .. 2,049 triangles: every power of two and its neighbours (the
bitonic sort's block sizes and register chunks, the 128 / 512 / 2,048 size
classes, 2,049 = the first object built by the global pipeline), built with
every workgroup size (with 128 or 256 threads the larger objects hold more
triangles per thread than the registers keep, so those are re-gathered).
The objects are triangle soups, so their triangles cross each other and
share centroids; one object repeats a single triangle, one is degenerate
(collinear corners).  Checked against the oracle's Face semantics.
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 4, 5, 7, 8, 9, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1023,
         1024, 1025, 1279, 1535, 1536, 1537, 2047, 2048, 2049]


def soup_set(seed, sizes, centres, extent):
    """One object per size: n random triangles around its centre."""
    rng = np.random.default_rng(seed)
    verts, tris, t2o = [], [], []
    nv = 0
    for o, (n, c) in enumerate(zip(sizes, centres)):
        p = c + rng.uniform(-extent, extent, (n, 3, 3)) * rng.uniform(0.2, 1.0, (n, 1, 1))
        p += rng.uniform(-extent, extent, (n, 1, 3)) * 0.5
        if o == len(sizes) - 2:  # one triangle repeated n times: equal keys up to the index bits
            p[:] = p[0]
        if o == len(sizes) - 1:  # collinear corners: zero-area triangles
            p[:, 2] = 0.5 * (p[:, 0] + p[:, 1])
        verts.append(p.reshape(-1, 3))
        tris.append(np.arange(nv, nv + 3 * n, dtype=np.uint32).reshape(-1, 3))
        t2o.append(np.full(n, o, np.int32))
        nv += 3 * n
    m = pz.Mesh(np.concatenate(verts).astype(np.float32), np.concatenate(tris), np.concatenate(t2o))
    return pz.extract_edges(m)


@pytest.fixture(scope="module")
def soups():
    sizes = SIZES + [700, 700]
    k = len(sizes)
    grid = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij"), -1).reshape(-1, 3)[:k]
    centres = grid.astype(np.float64) * 6.0
    a = soup_set(11, sizes, centres, 2.0)
    b = soup_set(12, sizes[::-1], centres + 1.5, 2.0)  # each B object near an A object of another size
    exp = orc.face_overlap(a, b)
    assert len(exp) > 10
    return a, b, exp


@pytest.mark.parametrize("threads", [128, 256, 512, 1024])
def test_every_size_and_workgroup(ctx, soups, threads):
    a, b, exp = soups
    saved = pz.tune_get(ctx)
    try:
        pz.tune_set(ctx, [threads] * 3)
        da, db = ctx.upload(a), ctx.upload(b)
        print(f"[sizes] threads={threads}", flush=True)
        out = pz.join(ctx, da, db)
        np.testing.assert_array_equal(out.pairs, exp)
        for pred in (pz.PRED_INTERSECTION,):  # soups are open: the parity side only needs to run
            pz.join(ctx, da, db, predicate=pred)
    finally:
        pz.tune_set(ctx, saved["build_threads"], saved["narrow_queue"], saved["narrow_chunk"])


def test_self_join_of_every_size(ctx, soups):
    a, _, _ = soups
    da = ctx.upload(a)
    out = pz.join(ctx, da, da)
    pairs = set(map(tuple, out.pairs.tolist()))
    # every object with a non-degenerate triangle meets itself (a triangle meets itself)
    for o in range(a.num_objects - 1):
        assert (o, o) in pairs, o
    np.testing.assert_array_equal(out.pairs, orc.face_overlap(a, a))
