"""Containment-ray tracking (pierce_intersection --enable-tracking;
pierce/query/src/cuda/mesh_intersection_anyhit.cu:46-240 and
pierce/query/src/applications/pierce_intersection_main.cpp:425-539, 883-922)
against the oracle's brute-force ray hits, and the two tracking CSVs.

Within one source the reference lists targets in the order its any-hit calls
arrive; here they come in ascending id order, so rows are compared as sets
per source, and their order is not pinned.
"""
import csv
import os
import subprocess

import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from tests.meshes import containment_sets, cube_mesh

pytestmark = pytest.mark.gpu


def expect(hits_row, max_targets):
    nz = np.flatnonzero(hits_row)
    odd = [int(t) for t in nz if hits_row[t] % 2 == 1]
    return nz, odd


def check(out, hits, max_targets, with_triangles):
    n = hits.shape[0]
    assert out["target_ids"].shape == (n, max_targets)
    for s in range(n):
        row = hits[s]
        nz, odd = expect(row, max_targets)
        ids = out["target_ids"][s]
        got = [int(x) for x in ids if x >= 0]
        assert (ids[len(got):] == -1).all() and (out["target_hits"][s][len(got):] == 0).all(), s
        assert got == sorted(got), s
        for t, h in zip(got, out["target_hits"][s][:len(got)]):
            assert row[t] == h and h % 2 == 1, (s, t)  # a listed object has all of its hits
        assert out["iterations"][s] == (1 if with_triangles[s] else 0), s
        assert out["candidate_count"][s] == min(len(nz), max_targets), s
        if len(nz) <= max_targets:
            assert got == odd and out["candidate_overflow"][s] == 0, s
        else:
            # the objects beyond the first max_targets met are not listed
            assert out["candidate_overflow"][s] >= 1, s


def has_triangles(m):
    flags = np.zeros(m.num_objects, bool)
    flags[m.triangle_to_object] = True
    return flags


@pytest.fixture(scope="module")
def sets():
    return containment_sets(n=150, box=40.0)


@pytest.mark.parametrize("max_targets", [256, 1])
def test_both_directions_against_the_oracle(ctx, sets, max_targets):
    # mesh2 holds shrunken copies of mesh1 objects: mesh2 -> mesh1 rays start
    # inside (odd counts); some rays cross two objects (overflow at 1 target)
    a, b = sets
    da, db = ctx.upload(a), ctx.upload(b)
    odd = overflowing = 0
    for src, tgt, ds, dt in ((a, b, da, db), (b, a, db, da)):
        print(f"[tracking] max_targets={max_targets}", flush=True)
        out = pz.track_containment_rays(ctx, ds, dt, max_targets)
        hits = orc.ray_target_hits(src, tgt)
        odd += int((hits % 2 == 1).sum())
        overflowing += int(((hits > 0).sum(1) > max_targets).sum())
        check(out, hits, max_targets, has_triangles(src))
    assert odd > 10
    assert max_targets > 1 or overflowing > 5


def test_cubes_gaps_and_single_triangle_target(ctx):
    # source cubes around a set of targets, one object id without triangles;
    # target sets of many cubes and of one triangle (the BVH is a single leaf)
    rng = np.random.default_rng(4)
    tc = [rng.choice([0.0, 1.0, 2.0, 3.0], 3).astype(np.float32) for _ in range(20)]
    te = [float(rng.choice([1.0, 2.0, 4.0])) for _ in range(20)]
    T = cube_mesh(tc, te)
    sc = [(rng.choice(np.arange(-1.0, 4.5, 0.5), 3) + 0.25).astype(np.float32) for _ in range(30)]
    S = cube_mesh(sc, [0.1] * 30)
    keep = S.triangle_to_object != 7  # object 7 has no triangles
    S = pz.extract_edges(pz.Mesh(S.vertices, S.indices[keep], S.triangle_to_object[keep]))
    one = pz.extract_edges(pz.Mesh(np.array([[-5, -5, 9], [9, -5, 9], [-5, 9, 9]], np.float32),
                                   np.array([[0, 1, 2]], np.uint32), np.array([0], np.int32)))
    dS = ctx.upload(S)
    for tgt in (T, one):
        out = pz.track_containment_rays(ctx, dS, ctx.upload(tgt), 256)
        hits = orc.ray_target_hits(S, tgt)
        assert hits.sum() > 0
        check(out, hits, 256, has_triangles(S))
    assert out["iterations"][7] == 0


def test_bad_max_targets(ctx, sets):
    a, b = sets
    da, db = ctx.upload(a), ctx.upload(b)
    for bad in (0, 4097):
        with pytest.raises(pz.PierceError):
            pz.track_containment_rays(ctx, da, db, bad)


def test_intersection_tool_tracking_csvs(sets, tmp_path):
    a, b = sets
    pz.write_pre(str(tmp_path / "a.pre"), a)
    pz.write_pre(str(tmp_path / "b.pre"), b)
    tool = os.path.join(pz.BIN_DIR, "pierce_intersection")
    print("[tracking] pierce_intersection --enable-tracking", flush=True)
    r = subprocess.run([tool, "--mesh1", "a.pre", "--mesh2", "b.pre", "--enable-tracking", "--pair-hits-output",
                        "hits.csv"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Pair hit tracking CSV: hits.csv" in r.stdout
    assert "Containment tracking CSV: intersection_containment_tracking.csv" in r.stdout
    with open(tmp_path / "intersection_pairs.csv") as f:
        pairs = {(int(x), int(y)) for x, y in list(csv.reader(f))[1:]}
    with open(tmp_path / "hits.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["direction", "source_object_id", "target_object_id", "target_ray_hits", "final_pair"]
    h12, h21 = orc.ray_target_hits(a, b), orc.ray_target_hits(b, a)
    want = {("mesh1_to_mesh2", s, t, int(h12[s, t])) for s, t in zip(*np.nonzero(h12 % 2))}
    want |= {("mesh2_to_mesh1", s, t, int(h21[s, t])) for s, t in zip(*np.nonzero(h21 % 2))}
    got = set()
    for d, s, t, h, fin in rows[1:]:
        s, t = int(s), int(t)
        key = (s, t) if d == "mesh1_to_mesh2" else (t, s)
        # an odd crossing count puts the source's point inside the target: an intersection pair
        assert fin == "1" and key in pairs, (d, s, t)
        got.add((d, s, t, int(h)))
    assert got == want
    with open(tmp_path / "intersection_containment_tracking.csv") as f:
        summ = list(csv.reader(f))
    assert summ[0] == ["direction", "source_object_id", "iterations", "candidate_count", "tracked_target_count",
                       "tracked_hit_total", "candidate_overflow_events", "final_pair_count",
                       "tracked_final_pair_count"]
    assert len(summ) == 1 + a.num_objects + b.num_objects
    for row in summ[1:]:
        d, s = row[0], int(row[1])
        it, cand, tracked, total, ovf, fin, tfin = map(int, row[2:])
        h = (h12 if d == "mesh1_to_mesh2" else h21)[s]
        assert (it, cand, ovf) == (1, int((h > 0).sum()), 0), row
        assert tracked == int((h % 2 == 1).sum()) and total == int(h[h % 2 == 1].sum()) and fin == tfin == tracked


def test_tracking_follows_the_query_direction(sets, tmp_path):
    # only mesh1 -> mesh2 queried: mesh2's sources keep empty tracking (no pair
    # rows, zero summary rows), as the reference's untouched buffers do
    a, b = sets
    pz.write_pre(str(tmp_path / "a.pre"), a)
    pz.write_pre(str(tmp_path / "b.pre"), b)
    tool = os.path.join(pz.BIN_DIR, "pierce_intersection")
    print("[tracking] --query-direction mesh1_to_mesh2", flush=True)
    r = subprocess.run([tool, "--mesh1", "a.pre", "--mesh2", "b.pre", "--enable-tracking", "--query-direction",
                        "mesh1_to_mesh2", "--containment-tracking-output", "summary.csv"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(tmp_path / "intersection_pair_hits.csv") as f:
        rows = list(csv.reader(f))[1:]
    h12 = orc.ray_target_hits(a, b)
    assert {(d, int(s), int(t)) for d, s, t, _, _ in rows} == \
        {("mesh1_to_mesh2", s, t) for s, t in zip(*np.nonzero(h12 % 2))}
    with open(tmp_path / "summary.csv") as f:
        summ = list(csv.reader(f))[1:]
    m2 = [row for row in summ if row[0] == "mesh2_to_mesh1"]
    assert len(m2) == b.num_objects and all(row[2:] == ["0"] * 7 for row in m2)
    assert all(row[2] == "1" for row in summ if row[0] == "mesh1_to_mesh2")
