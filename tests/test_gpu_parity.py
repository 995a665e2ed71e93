"""GPU parity: the HIP join against the exact CPU oracle on the same seeded inputs.

The bar is bit-exact equality of the sorted unique pair sets (integer output).
Cases follow the reference's predicates
(pierce/query/src/applications/pierce_overlap_main.cpp,
pierce_overlap_two_pass_main.cpp, pierce_intersection_main.cpp,
pierce_containment_main.cpp).
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from tests.meshes import containment_sets, cube_mesh, equal_cube_offsets, nested_cubes

pytestmark = pytest.mark.gpu


def c1_sets(count=1000, box=100.0):
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=count, seed=42, box=box, jitter=0.02)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=count, seed=43, box=box, jitter=0.02)
    return a, b


@pytest.fixture(scope="module")
def c1(ctx):
    a, b = c1_sets()
    print(f"[c1] {a.indices.shape[0]} + {b.indices.shape[0]} triangles", flush=True)
    return a, b, ctx.upload(a), ctx.upload(b), orc.face_overlap(a, b)


@pytest.mark.parametrize("narrow_mode", [0, 1, 2])
@pytest.mark.parametrize("precision", [pz.PREC_FP32_FILTERED, pz.PREC_FP64])
@pytest.mark.parametrize("sizing", [pz.SIZING_TWO_PASS, pz.SIZING_FIXED])
def test_two_level_overlap_c1(ctx, c1, precision, sizing, narrow_mode):
    a, b, da, db, expected = c1
    print(f"[two-level c1] precision={precision} sizing={sizing} narrow={narrow_mode}", flush=True)
    out = pz.join(ctx, da, db, precision=precision, sizing=sizing, capacity=16, narrow_mode=narrow_mode)
    assert out.overflowed == (sizing == pz.SIZING_FIXED and out.num_candidates > 16)
    np.testing.assert_array_equal(out.pairs, expected)


@pytest.mark.parametrize("dedup,sizing", [(pz.DEDUP_SORT, pz.SIZING_TWO_PASS), (pz.DEDUP_HASH, pz.SIZING_FIXED)])
def test_global_overlap_c1(ctx, c1, dedup, sizing):
    a, b, da, db, expected = c1
    print(f"[global c1] dedup={dedup}", flush=True)
    out = pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64, dedup=dedup, sizing=sizing,
                  capacity=1 << 12)
    np.testing.assert_array_equal(out.pairs, expected)


def test_hash_overflow_falls_back(ctx, c1):
    a, b, da, db, expected = c1
    out = pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL, dedup=pz.DEDUP_HASH, sizing=pz.SIZING_FIXED, capacity=4)
    assert out.overflowed
    np.testing.assert_array_equal(out.pairs, expected)


@pytest.mark.parametrize("narrow_mode", [0, 1, 2])
@pytest.mark.parametrize("direction", [pz.DIR_1TO2, pz.DIR_2TO1])
def test_direction_split_c1(ctx, c1, direction, narrow_mode):
    a, b, da, db, expected = c1
    out = pz.join(ctx, da, db, direction=direction, narrow_mode=narrow_mode)
    np.testing.assert_array_equal(out.pairs, orc.edge_overlap(a, b, direction))


@pytest.mark.parametrize("narrow_mode", [0, 1, 2])
@pytest.mark.parametrize("direction", [pz.DIR_1TO2, pz.DIR_2TO1])
def test_intersection_direction(ctx, direction, narrow_mode):
    """--query-direction of pierce_intersection (pierce_intersection_main.cpp:103-155):
    one edge direction and the parity rays from that side's first vertices."""
    for name, (a, b) in (("ellipsoids", containment_sets()), ("nested cubes", nested_cubes())):
        exp, _ = orc.predicate_join(a, b, pz.PRED_INTERSECTION, direction=direction)
        both, _ = orc.predicate_join(a, b, pz.PRED_INTERSECTION)
        assert 0 < len(exp) <= len(both)
        da, db = ctx.upload(a), ctx.upload(b)
        for kw in (dict(), dict(precision=pz.PREC_FP64), dict(broad_mode=1)):
            print(f"[intersection dir {direction}] {name} narrow={narrow_mode} {kw}", flush=True)
            out = pz.join(ctx, da, db, predicate=pz.PRED_INTERSECTION, direction=direction, narrow_mode=narrow_mode,
                          **kw)
            np.testing.assert_array_equal(out.pairs, exp)
        # batched: each query carries its own direction
        outs, _ = pz.join_batch(ctx, ctx.upload_batch([a, a]), ctx.upload_batch([b, b]),
                                [dict(predicate=pz.PRED_INTERSECTION, direction=direction),
                                 dict(predicate=pz.PRED_INTERSECTION)], narrow_mode=narrow_mode)
        np.testing.assert_array_equal(outs[0].pairs, exp)
        np.testing.assert_array_equal(outs[1].pairs, both)


@pytest.mark.parametrize("direction", [pz.DIR_1TO2, pz.DIR_2TO1])
def test_containment_uses_both_edge_directions(ctx, direction):
    """pierce_containment has no direction: its edge phase always runs both ways
    (pierce_containment_main.cpp:385-550), so a direction in the descriptor
    leaves the result unchanged."""
    a, b = containment_sets()
    exp, _ = orc.predicate_join(a, b, pz.PRED_CONTAINMENT)
    da, db = ctx.upload(a), ctx.upload(b)
    out = pz.join(ctx, da, db, predicate=pz.PRED_CONTAINMENT, direction=direction)
    np.testing.assert_array_equal(out.pairs, exp)


def test_candidate_count_c1(ctx, c1):
    a, b, da, db, _ = c1
    assert pz.count_candidate_tests(ctx, da, db) == orc.count_candidate_tests(a, b)


@pytest.mark.parametrize("narrow_mode", [0, 1, 2])
@pytest.mark.parametrize("predicate", [pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT])
def test_predicates_vs_oracle(ctx, predicate, narrow_mode):
    a, b = containment_sets()
    print(f"[predicate {predicate} narrow {narrow_mode}]", flush=True)
    da, db = ctx.upload(a), ctx.upload(b)
    exp, _ = orc.predicate_join(a, b, predicate)
    out = pz.join(ctx, da, db, predicate=predicate, narrow_mode=narrow_mode)
    np.testing.assert_array_equal(out.pairs, exp)
    if predicate == pz.PRED_CONTAINMENT:
        assert len(exp) > 0


@pytest.mark.parametrize("strategy", [pz.STRATEGY_TWO_LEVEL, pz.STRATEGY_GLOBAL])
def test_containment_with_overlap_pairs(ctx, strategy):
    a, b = containment_sets()
    da, db = ctx.upload(a), ctx.upload(b)
    exp, types = orc.predicate_join(a, b, pz.PRED_CONTAINMENT, include_overlap=True)
    out = pz.join(ctx, da, db, predicate=pz.PRED_CONTAINMENT, include_overlap_pairs=1, strategy=strategy)
    np.testing.assert_array_equal(out.pairs, exp)
    np.testing.assert_array_equal(out.types, types)


@pytest.mark.parametrize("dedup,sizing", [(pz.DEDUP_SORT, pz.SIZING_TWO_PASS), (pz.DEDUP_HASH, pz.SIZING_FIXED)])
@pytest.mark.parametrize("direction", [pz.DIR_BOTH, pz.DIR_1TO2, pz.DIR_2TO1])
@pytest.mark.parametrize("predicate", [pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT])
def test_global_predicates_vs_oracle(ctx, predicate, direction, dedup, sizing):
    """Set-wide intersection and containment: edges against a BVH over all
    triangles, plus one +Z ray per source object through the other set's BVH
    (pierce_intersection_main.cpp:103-155, pierce_containment_main.cpp:385-550)."""
    for name, (a, b) in (("ellipsoids", containment_sets()), ("nested cubes", nested_cubes())):
        exp, _ = orc.predicate_join(a, b, predicate, direction=direction)
        da, db = ctx.upload(a), ctx.upload(b)
        print(f"[global predicate {predicate} dir {direction}] {name}", flush=True)
        out = pz.join(ctx, da, db, predicate=predicate, direction=direction, strategy=pz.STRATEGY_GLOBAL,
                      precision=pz.PREC_FP64, dedup=dedup, sizing=sizing, capacity=64)
        np.testing.assert_array_equal(out.pairs, exp)
        if predicate == pz.PRED_CONTAINMENT:
            assert out.parity_rays == b.num_objects


@pytest.mark.parametrize("precision", [pz.PREC_FP32_FILTERED, pz.PREC_FP64])
@pytest.mark.parametrize("strategy,narrow_mode", [(pz.STRATEGY_TWO_LEVEL, 0), (pz.STRATEGY_TWO_LEVEL, 1),
                                                  (pz.STRATEGY_TWO_LEVEL, 2),
                                                  (pz.STRATEGY_GLOBAL, 0)])
def test_touching_cubes_exact(ctx, precision, strategy, narrow_mode):
    """Exact face/edge/vertex contact and coplanar faces: the escalation paths decide these."""
    offs = equal_cube_offsets()
    a = cube_mesh([np.zeros(3, np.float32)] * len(offs), [2.0] * len(offs))
    b = cube_mesh(offs, [2.0] * len(offs))
    exp = orc.face_overlap(a, b)
    da, db = ctx.upload(a), ctx.upload(b)
    out = pz.join(ctx, da, db, precision=precision, strategy=strategy, narrow_mode=narrow_mode)
    np.testing.assert_array_equal(out.pairs, exp)
    assert out.tests_exact > 0  # coplanar contact needs exact signs


@pytest.mark.parametrize("predicate", [pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT])
def test_nested_cubes(ctx, predicate):
    a, b = nested_cubes()
    da, db = ctx.upload(a), ctx.upload(b)
    exp = orc.face_overlap(a, b) if predicate == pz.PRED_OVERLAP else orc.predicate_join(a, b, predicate)[0]
    out = pz.join(ctx, da, db, predicate=predicate)
    np.testing.assert_array_equal(out.pairs, exp)


def test_self_join_and_swap(ctx):
    a, b = containment_sets()
    da, db = ctx.upload(a), ctx.upload(b)
    self_pairs = pz.join(ctx, da, da).pairs
    assert set((i, i) for i in range(a.num_objects)) <= set(map(tuple, self_pairs.tolist()))
    ab = pz.join(ctx, da, db).pairs
    ba = pz.join(ctx, db, da).pairs
    assert set(map(tuple, ab.tolist())) == set(map(tuple, ba[:, ::-1].tolist()))


@pytest.mark.parametrize("broad_mode", [1, 2, 3])  # object LBVH walk, box scan, hashed grid
def test_sizing_modes_agree(ctx, c1, broad_mode):
    a, b, da, db, expected = c1
    for sizing, cap in ((pz.SIZING_TWO_PASS, 0), (pz.SIZING_FIXED, 1), (pz.SIZING_FIXED, 1 << 20),
                        (pz.SIZING_ESTIMATE, 0)):
        out = pz.join(ctx, da, db, sizing=sizing, capacity=cap, broad_mode=broad_mode)
        np.testing.assert_array_equal(out.pairs, expected)


def test_broad_modes_give_the_same_candidates(ctx, c1):
    a, b, da, db, expected = c1
    walk = pz.join(ctx, da, db, broad_mode=1)
    scan = pz.join(ctx, da, db, broad_mode=2)
    grid = pz.join(ctx, da, db, broad_mode=3)
    assert walk.num_candidates == scan.num_candidates == grid.num_candidates > 0
    np.testing.assert_array_equal(walk.pairs, scan.pairs)
    np.testing.assert_array_equal(grid.pairs, scan.pairs)
    with pytest.raises(pz.PierceError):
        pz.join(ctx, da, db, broad_mode=4)


def _scaled(m, factors):
    """Copy of m with object o scaled by factors[o] about its vertex mean
    (objects own contiguous vertex runs, as the generators write them)."""
    v = m.vertices.copy()
    t2o = m.triangle_to_object
    vobj = np.full(v.shape[0], -1, np.int64)
    vobj[m.indices.reshape(-1)] = np.repeat(t2o, 3)
    for o in np.nonzero(factors != 1.0)[0]:
        sel = vobj == o
        c = v[sel].mean(0)
        v[sel] = (c + (v[sel] - c) * np.float32(factors[o])).astype(np.float32)
    return pz.Mesh(v, m.indices, t2o)


def test_grid_broad_phase_mixed_sizes(ctx):
    """Hashed-grid broad phase against the box scan: a few mesh2 boxes over
    many cells (kept apart and tested by every query), mesh1 boxes over more
    cells than a query visits, flat and empty objects, negative coordinates."""
    rng = np.random.default_rng(7)
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=600, seed=11, box=60.0)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=600, seed=12, box=60.0)
    fa = np.ones(a.num_objects)
    fb = np.ones(b.num_objects)
    fb[rng.choice(b.num_objects, 6, replace=False)] = 12.0   # boxes over many cells
    fa[rng.choice(a.num_objects, 3, replace=False)] = 40.0   # queries over very many cells
    fb[rng.choice(b.num_objects, 20, replace=False)] = 0.05  # tiny boxes
    a, b = _scaled(a, fa), _scaled(b, fb)
    a.vertices -= np.float32(30.0)  # negative coordinates
    b.vertices -= np.float32(30.0)
    # flatten one mesh2 object onto a plane (zero extent on one axis)
    vb = b.vertices
    sel = np.unique(b.indices[b.triangle_to_object == 3])
    vb[sel, 2] = vb[sel[0], 2]
    # an object id without triangles in mesh2
    keep = b.triangle_to_object != 5
    a = pz.extract_edges(a)
    b = pz.extract_edges(pz.Mesh(b.vertices, b.indices[keep], b.triangle_to_object[keep]))
    da, db = ctx.upload(a), ctx.upload(b)
    scan = pz.join(ctx, da, db, broad_mode=2)
    grid = pz.join(ctx, da, db, broad_mode=3)
    assert grid.num_candidates == scan.num_candidates > 0
    np.testing.assert_array_equal(grid.pairs, scan.pairs)
    # the object-LBVH walk (a wave per mesh1 object, the 40x boxes meet most of
    # mesh2), exact and capacity-sized; capacity 64 overflows into the recount
    lbvh = pz.join(ctx, da, db, broad_mode=1)
    assert lbvh.num_candidates == scan.num_candidates
    np.testing.assert_array_equal(lbvh.pairs, scan.pairs)
    for kw in (dict(sizing=pz.SIZING_ESTIMATE), dict(sizing=pz.SIZING_FIXED, capacity=64)):
        o = pz.join(ctx, da, db, broad_mode=1, **kw)
        print(f"[mixed sizes] lbvh {kw}: candidates {o.num_candidates} overflowed {o.overflowed}", flush=True)
        assert o.num_candidates == scan.num_candidates
        np.testing.assert_array_equal(o.pairs, scan.pairs)
        if kw.get("capacity") == 64:
            assert o.overflowed
    for pred in (pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT):
        np.testing.assert_array_equal(pz.join(ctx, da, db, predicate=pred, broad_mode=3).pairs,
                                      pz.join(ctx, da, db, predicate=pred, broad_mode=2).pairs)


def test_empty_result(ctx):
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=5, seed=1, box=10.0)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=5, seed=2, box=10.0)
    b.vertices += np.float32(1000.0)
    b.edge_starts += np.float32(1000.0)
    b.edge_ends += np.float32(1000.0)
    da, db = ctx.upload(a), ctx.upload(b)
    for pred in (pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT):
        assert len(pz.join(ctx, da, db, predicate=pred).pairs) == 0


def _ragged(seed, copies=False):
    """Ellipsoids where some object ids have no triangles and some objects are
    single triangles (cut from a neighbouring mesh); with `copies`, scaled
    copies of the seed-31 set's meshes sit inside them."""
    kw = dict(copy_fraction=0.4, copy_scale_min=0.3, copy_scale_max=0.8, copy_source_count=40,
              copy_source_seed=31) if copies else {}
    m = pz.generate(pz.GEN_ELLIPSOIDS, count=40, seed=seed, level_min=2, level_max=2, box=25.0, **kw)
    t2o = m.triangle_to_object.copy()
    keep = (t2o % 7) != 3  # ids 3, 10, 17, ... lose every triangle
    idx, t2o = m.indices[keep], t2o[keep]
    # the first triangle of every fifth object becomes an object of its own
    n = int(m.triangle_to_object.max()) + 1
    first = {}
    for i, o in enumerate(t2o.tolist()):
        first.setdefault(o, i)
    for k, o in enumerate(sorted(o for o in first if o % 5 == 0)):
        t2o[first[o]] = n + 2 * k + 1  # odd ids above n: the even ones stay empty
    return pz.extract_edges(pz.Mesh(m.vertices, idx, t2o.astype(np.int32)))


@pytest.mark.parametrize("predicate", [pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT])
def test_ragged_object_ids_and_single_triangles(ctx, predicate):
    a, b = _ragged(31), _ragged(32, copies=True)
    exp = orc.face_overlap(a, b) if predicate == pz.PRED_OVERLAP else orc.predicate_join(a, b, predicate)[0]
    assert len(exp) > 0
    da, db = ctx.upload(a), ctx.upload(b)
    for kw in (dict(), dict(broad_mode=1), dict(narrow_mode=1), dict(narrow_mode=2), dict(precision=pz.PREC_FP64)):
        print(f"[ragged] predicate={predicate} {kw}", flush=True)
        np.testing.assert_array_equal(pz.join(ctx, da, db, predicate=predicate, **kw).pairs, exp)
    if predicate == pz.PRED_OVERLAP:
        np.testing.assert_array_equal(pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL).pairs, exp)


def test_shuffled_triangle_order(ctx):
    # Every triangle in random order: each 64-triangle step of the object
    # bounds mixes objects (segmented wave reduction), and no object is stored
    # contiguously, so all of them go through the sort-and-chunk builder (the
    # 5,120- and 20,480-triangle spheres in several chunks under a top tree).
    rng = np.random.default_rng(5)
    sets = []
    for seed in (51, 52):
        m = pz.generate(pz.GEN_SPHERES_MIXED, count=60, seed=seed, level_min=1, level_max=5, radius_min=1.0,
                        radius_max=4.0, box=30.0)
        p = rng.permutation(m.indices.shape[0])
        sets.append(pz.extract_edges(pz.Mesh(m.vertices, m.indices[p], m.triangle_to_object[p])))
    a, b = sets
    assert np.bincount(a.triangle_to_object).max() > 4096
    da, db = ctx.upload(a), ctx.upload(b)
    for predicate in (pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT):
        exp = orc.face_overlap(a, b) if predicate == pz.PRED_OVERLAP else orc.predicate_join(a, b, predicate)[0]
        print(f"[shuffled] predicate={predicate} expected={len(exp)}", flush=True)
        assert len(exp) > 0 or predicate == pz.PRED_CONTAINMENT
        out = pz.join(ctx, da, db, predicate=predicate)
        np.testing.assert_array_equal(out.pairs, exp)
        assert out.chunk_sorted_sets == 2  # objects not contiguous: chunks of the sorted order


def _spheres_up_to_level5(seed):
    return pz.generate(pz.GEN_SPHERES_MIXED, count=60, seed=seed, level_min=1, level_max=5, radius_min=1.0,
                       radius_max=4.0, box=30.0)


def test_object_chunks_in_own_order_and_incoherent_fallback(ctx):
    """Objects over 2,048 triangles stored contiguously are cut into chunks
    of their own triangle order (the spheres' subdivision order is coherent
    enough).  With each object's triangles shuffled in place they stay
    contiguous but incoherent: the 20,480-triangle spheres' ten chunks each
    span the whole sphere, the check after the chunk builds counts them, and
    the set is built again from the sorted order.  Both give the oracle's
    pairs."""
    rng = np.random.default_rng(7)
    plain, local = [], []
    for seed in (51, 52):
        m = _spheres_up_to_level5(seed)
        t2o = m.triangle_to_object
        starts = np.flatnonzero(np.r_[True, t2o[1:] != t2o[:-1]])
        ends = np.r_[starts[1:], len(t2o)]
        assert len(np.unique(t2o)) == len(starts)  # contiguous objects
        p = np.concatenate([s + rng.permutation(e - s) for s, e in zip(starts, ends)])
        plain.append(m)
        local.append(pz.extract_edges(pz.Mesh(m.vertices, m.indices[p], t2o[p])))
    assert np.bincount(plain[0].triangle_to_object).max() == 20480
    for (a, b), name in ((plain, "own order"), (local, "shuffled in place")):
        da, db = ctx.upload(a), ctx.upload(b)
        for predicate in (pz.PRED_OVERLAP, pz.PRED_CONTAINMENT):
            exp = orc.face_overlap(a, b) if predicate == pz.PRED_OVERLAP else orc.predicate_join(a, b, predicate)[0]
            out = pz.join(ctx, da, db, predicate=predicate)
            print(f"[own-order chunks] {name} predicate={predicate} expected={len(exp)} "
                  f"sorted sets={out.chunk_sorted_sets}", flush=True)
            np.testing.assert_array_equal(out.pairs, exp)
            if name == "own order":
                assert out.chunk_sorted_sets == 0
            else:
                assert out.chunk_sorted_sets >= 1
            np.testing.assert_array_equal(pz.join(ctx, da, db, predicate=predicate, narrow_mode=2).pairs, exp)


@pytest.mark.parametrize("count", [6000, 70000])  # two sort tiles; beyond the tiles (radix sort)
def test_many_result_pairs_sorted_and_agreeing(ctx, count):
    """A set joined with itself gives every (i, i) plus its neighbours: more
    result pairs than one sort tile (6,000 cubes) and than the tile merge
    (70,000 cubes).  The sieve's speculative download, the tree path's
    download and the set-wide path agree and are sorted."""
    m = pz.generate(pz.GEN_CUBES, count=count, seed=31, radius_min=1.0, radius_max=4.0,
                    box=6.0 * count ** (1.0 / 3.0))
    d = ctx.upload(m)
    outs = [pz.join(ctx, d, d, **kw) for kw in (dict(), dict(narrow_mode=2),
                                                dict(strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64))]
    p = outs[0].pairs
    assert len(p) > count
    keys = p[:, 0].astype(np.int64) << 32 | p[:, 1].astype(np.int64)
    assert np.all(np.diff(keys) > 0)  # sorted, unique
    assert set((i, i) for i in range(count)) <= set(map(tuple, p.tolist()))
    for o in outs[1:]:
        np.testing.assert_array_equal(o.pairs, p)
    for pred in (pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT):
        np.testing.assert_array_equal(pz.join(ctx, d, d, predicate=pred).pairs,
                                      pz.join(ctx, d, d, predicate=pred, narrow_mode=2).pairs)
