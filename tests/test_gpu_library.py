"""Sets instanced from a resident mesh library (pz_library_*): the device
copies match a numpy evaluation of the same maps bit for bit, and joins of
the instanced sets give the oracle's pairs for every predicate."""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from pierce_amd.stream import instantiate_host, random_rotations, stream_catalogue, stream_query

pytestmark = pytest.mark.gpu

PREDS = [pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT]


def oracle_pairs(a, b, pred):
    return orc.face_overlap(a, b) if pred == pz.PRED_OVERLAP else orc.predicate_join(a, b, pred)[0]


@pytest.fixture(scope="module")
def small_library(ctx):
    m = pz.generate(pz.GEN_QUERY_MIX, count=40, seed=77, box=0.0, tri_min=500, tri_median=1500, tri_max=4000)
    return m, ctx.library(m)


def random_instances(rng, n, nlib, side):
    inst = np.zeros(n, pz.INSTANCE_DTYPE)
    inst["mesh"] = rng.integers(0, nlib, size=n)
    A = random_rotations(rng, n) * rng.uniform(0.5, 2.0, size=(n, 1, 1))
    inst["xform"][:, :9] = A.reshape(n, 9).astype(np.float32)
    inst["xform"][:, 9:] = rng.uniform(0, side, size=(n, 3)).astype(np.float32)
    return inst


def test_instances_match_host_map(ctx, small_library):
    m, lib = small_library
    rng = np.random.default_rng(3)
    counts = [5, 0, 12, 7]
    inst = random_instances(rng, sum(counts), m.num_objects, 30.0)
    ds = lib.instantiate(counts, inst)
    got = ds.download()
    parts = instantiate_host(m, counts, inst)
    # the batched dataset is the parts back to back (object ids and vertex indices shifted)
    exp = pz.stream.concat_meshes([p for p in parts if p.indices.shape[0]])
    np.testing.assert_array_equal(got.vertices.view(np.uint32), exp.vertices.view(np.uint32))
    np.testing.assert_array_equal(got.indices, exp.indices)
    np.testing.assert_array_equal(got.triangle_to_object, exp.triangle_to_object)
    np.testing.assert_array_equal(got.edge_starts.view(np.uint32), exp.edge_starts.view(np.uint32))
    np.testing.assert_array_equal(got.edge_ends.view(np.uint32), exp.edge_ends.view(np.uint32))
    np.testing.assert_array_equal(got.edge_objects, exp.edge_objects)
    assert ds.info()["objects"] == sum(counts)


@pytest.mark.parametrize("sizing", [pz.SIZING_TWO_PASS, pz.SIZING_FIXED])
def test_instanced_batch_joins_match_oracle(ctx, small_library, sizing):
    m, lib = small_library
    rng = np.random.default_rng(4)
    ca, cb = [9, 14, 1, 6], [11, 3, 8, 6]
    ia = random_instances(rng, sum(ca), m.num_objects, 25.0)
    ib = random_instances(rng, sum(cb), m.num_objects, 25.0)
    da, db = lib.instantiate(ca, ia), lib.instantiate(cb, ib)
    assert da.vertex_runs and db.vertex_runs  # topology from the instance records
    pa, pb = instantiate_host(m, ca, ia), instantiate_host(m, cb, ib)
    preds = [PREDS[q % 3] for q in range(len(ca))]
    outs, _ = pz.join_batch(ctx, da, db, [dict(predicate=p) for p in preds], sizing=sizing, capacity=8)
    total = 0
    for q, (o, p) in enumerate(zip(outs, preds)):
        exp = oracle_pairs(pa[q], pb[q], p)
        total += len(exp)
        np.testing.assert_array_equal(o.pairs, exp, err_msg=f"query {q} predicate {p}")
    assert total > 0
    # the same parts uploaded from the host give the same pairs
    ha, hb = ctx.upload_batch(pa), ctx.upload_batch(pb)
    outs2, _ = pz.join_batch(ctx, ha, hb, [dict(predicate=p) for p in preds])
    for o, o2 in zip(outs, outs2):
        np.testing.assert_array_equal(o.pairs, o2.pairs)


def test_library_with_unreferenced_vertex_in_a_run(ctx, small_library):
    """A vertex inside an object's run that no triangle references: instanced
    sets then carry no vertex runs (bounds and sieve from the triangles), and
    joins still give the oracle's pairs."""
    m, _ = small_library
    k = int(m.indices[m.triangle_to_object == 0].min()) + 1
    V = np.insert(m.vertices, k, m.vertices[k], axis=0)  # the new vertex k is unreferenced
    I = m.indices.copy()
    I[I >= k] += 1
    mm = pz.extract_edges(pz.Mesh(V, I, m.triangle_to_object))
    lib = ctx.library(mm)
    rng = np.random.default_rng(8)
    ca, cb = [10, 7], [9, 12]
    ia = random_instances(rng, sum(ca), mm.num_objects, 25.0)
    ia["mesh"][0] = 0  # the object with the unreferenced vertex is instanced
    ib = random_instances(rng, sum(cb), mm.num_objects, 25.0)
    da, db = lib.instantiate(ca, ia), lib.instantiate(cb, ib)
    assert not da.vertex_runs and not db.vertex_runs
    pa, pb = instantiate_host(mm, ca, ia), instantiate_host(mm, cb, ib)
    outs, _ = pz.join_batch(ctx, da, db, [dict(predicate=PREDS[q]) for q in range(len(ca))])
    for q, o in enumerate(outs):
        np.testing.assert_array_equal(o.pairs, oracle_pairs(pa[q], pb[q], PREDS[q]), err_msg=f"query {q}")


def test_library_with_empty_object_ids(ctx, small_library):
    """Library object ids with no triangles (a gap in the ids) instanced
    between real objects: the instanced topology gives them no triangles and
    no vertex run, and joins still give the oracle's pairs."""
    m, _ = small_library
    t2o = m.triangle_to_object.copy()
    t2o[t2o >= 3] += 2  # ids 3 and 4 hold no triangles
    mm = pz.extract_edges(pz.Mesh(m.vertices, m.indices, t2o))
    lib = ctx.library(mm)
    rng = np.random.default_rng(9)
    ca, cb = [8, 9], [7, 10]
    ia = random_instances(rng, sum(ca), mm.num_objects, 25.0)
    ib = random_instances(rng, sum(cb), mm.num_objects, 25.0)
    ia["mesh"][[1, 4, 9]] = [3, 4, 3]  # empty objects among the instances of both parts
    ib["mesh"][[0, 12]] = [4, 3]
    da, db = lib.instantiate(ca, ia), lib.instantiate(cb, ib)
    assert da.vertex_runs and db.vertex_runs
    pa, pb = instantiate_host(mm, ca, ia), instantiate_host(mm, cb, ib)
    outs, _ = pz.join_batch(ctx, da, db, [dict(predicate=PREDS[q]) for q in range(len(ca))])
    for q, o in enumerate(outs):
        np.testing.assert_array_equal(o.pairs, oracle_pairs(pa[q], pb[q], PREDS[q]), err_msg=f"query {q}")


def test_stream_catalogue_queries(ctx):
    """A few queries of the library stream (small sets) against the oracle."""
    m, ids = stream_catalogue(ellipsoids_per_level=8, tubes=6, tube_tris=(5000, 8000, 12000))
    lib = ctx.library(m)
    qs = [stream_query(q, ids, seed0=11, median=12, max_meshes=40, spacing=(8.0, 20.0)) for q in range(6)]
    da = lib.instantiate([len(q[0]) for q in qs], np.concatenate([q[0] for q in qs]))
    db = lib.instantiate([len(q[1]) for q in qs], np.concatenate([q[1] for q in qs]))
    outs, _ = pz.join_batch(ctx, da, db, [dict(predicate=q[2]) for q in qs])
    for i, (q, o) in enumerate(zip(qs, outs)):
        a = instantiate_host(m, [len(q[0])], q[0])[0]
        b = instantiate_host(m, [len(q[1])], q[1])[0]
        np.testing.assert_array_equal(o.pairs, oracle_pairs(a, b, q[2]), err_msg=f"query {i}")


def test_library_rejects_bad_input(ctx, small_library):
    m, lib = small_library
    bad = np.zeros(1, pz.INSTANCE_DTYPE)
    bad["mesh"] = m.num_objects  # beyond the catalogue
    with pytest.raises(pz.PierceError):
        lib.instantiate([1], bad)
    # objects whose triangles are not contiguous cannot be instanced
    p = np.random.default_rng(1).permutation(m.indices.shape[0])
    shuffled = pz.Mesh(m.vertices, m.indices[p], m.triangle_to_object[p], m.edge_starts, m.edge_ends, m.edge_objects)
    with pytest.raises(pz.PierceError):
        ctx.library(shuffled)


def test_instancing_slots_reuse_storage(ctx, small_library):
    """A dataset built into a context slot is valid until the next
    instantiation into that slot; joins reject it afterwards."""
    m, lib = small_library
    rng = np.random.default_rng(9)
    ca, cb = [6, 9], [8, 4]
    ia = random_instances(rng, sum(ca), m.num_objects, 20.0)
    ib = random_instances(rng, sum(cb), m.num_objects, 20.0)
    owned = pz.join_batch(ctx, lib.instantiate(ca, ia), lib.instantiate(cb, ib), [dict(), dict()])[0]
    first_a = lib.instantiate(ca, ia, slot=0)
    db = lib.instantiate(cb, ib, slot=1)
    outs, _ = pz.join_batch(ctx, first_a, db, [dict(), dict()])
    for o, e in zip(outs, owned):
        np.testing.assert_array_equal(o.pairs, e.pairs)
    second_a = lib.instantiate(ca[::-1], np.concatenate([ia[ca[0]:], ia[:ca[0]]]), slot=0)
    with pytest.raises(pz.PierceError):
        pz.join_batch(ctx, first_a, db, [dict(), dict()])
    outs2, _ = pz.join_batch(ctx, second_a, db, [dict(), dict()])
    np.testing.assert_array_equal(outs2[0].pairs, pz.join_batch(
        ctx, lib.instantiate(ca[::-1], np.concatenate([ia[ca[0]:], ia[:ca[0]]])), lib.instantiate(cb, ib),
        [dict(), dict()])[0][0].pairs)
    ctx.trim()  # slot storage returned: slot datasets are no longer valid
    with pytest.raises(pz.PierceError):
        pz.join_batch(ctx, second_a, db, [dict(), dict()])
