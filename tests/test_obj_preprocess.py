"""OBJ input and the pierce_preprocess front-end (CPU only).

The reader follows the object semantics of
pierce/preprocess/src/ObjMeshDatasetLoader.cpp:14-168 (one object per shape
with faces, triangulated polygons, out-of-range vertices dropped with their
triangles); the CLI follows pierce/preprocess/src/pierce_preprocess.cpp:171-352.
tinyobjloader, which the reference parses with, is not available here: the
expected values below are written out by hand from those rules, and the edge
and grid payloads are compared with the already pinned a2/a3 functions.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import pierce_amd as pz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pierce_amd", "bin", "pierce_preprocess")

HAND_OBJ = (
    "# two objects, an empty group, a dropped triangle\n"
    "mtllib none.mtl\n"
    "v 0 0 0\n"
    "v 1 0 0\n"
    "v 0 1 0\n"
    "v 0 0 1\r\n"
    "vt 0 0\n"
    "vn 0 0 1\n"
    "g empty_group\n"
    "o first\n"
    "usemtl a\n"
    "f 1/1/1 2/1/1 3/1/1\n"
    "usemtl b\n"
    "s off\n"
    "f 1//1 2//1 4//1\n"
    "o second\n"
    "v 10 10 10\n"
    "v 11 10 10\n"
    "v 11 11 10\n"
    "v 10 11 10\n"
    "f -4 -3 -2 -1\n"
    "g third\n"
    "f 1 2 9\n"
    "o fourth\n"
    "  f 2 3 4\n"
)


def triangles(m: pz.Mesh):
    """Per triangle: (object, corner coordinates in corner order)."""
    v = m.vertices
    return [(int(o), tuple(tuple(float(c) for c in v[i]) for i in tri))
            for tri, o in zip(m.indices.tolist(), m.triangle_to_object.tolist())]


def test_hand_written_obj(tmp_path):
    p = tmp_path / "hand.obj"
    p.write_bytes(HAND_OBJ.encode())
    m = pz.read_obj(str(p))
    O, X, Y, Z = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
    a, b, c, d = (10.0, 10.0, 10.0), (11.0, 10.0, 10.0), (11.0, 11.0, 10.0), (10.0, 11.0, 10.0)
    assert triangles(m) == [
        (0, (O, X, Y)), (0, (O, X, Z)),  # `usemtl` does not start a new object
        (1, (a, b, c)), (1, (a, c, d)),  # quad as a fan; negative indices
        # object 2 ("third") keeps its id but loses its triangle (vertex 9 does not exist)
        (3, (X, Y, Z)),
    ]
    # per object: the distinct vertices its faces use (2 of "third"'s 3)
    assert m.vertices.shape == (4 + 4 + 2 + 3, 3)
    assert m.edge_starts.shape[0] == 0 and m.grid.shape[0] == 0


def test_round_trip_keeps_every_objects_triangles(tmp_path):
    m = pz.generate(pz.GEN_QUERY_MIX, count=12, seed=7, box=30.0, tri_min=40, tri_median=60, tri_max=120,
                    level_min=1, level_max=2)
    p = tmp_path / "mix.obj"
    pz.write_obj(str(p), m)
    r = pz.read_obj(str(p))
    assert r.num_objects == m.num_objects
    assert sorted(triangles(r)) == sorted(triangles(m))


@pytest.mark.parametrize("text, what", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "zero index"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\n", "no faces"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 x 3\n", "bad index"),
])
def test_bad_obj_is_an_error(tmp_path, text, what):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(pz.PierceError):
        pz.read_obj(str(p))


def test_only_obj_paths_are_read(tmp_path):
    p = tmp_path / "mesh.txt"
    p.write_text(HAND_OBJ)
    with pytest.raises(pz.PierceError):
        pz.read_obj(str(p))


def test_shuffle_moves_each_object_rigidly():
    m = pz.generate(pz.GEN_CUBES, count=20, seed=3, box=50.0)
    s1, s2 = pz.shuffle(m, seed=11), pz.shuffle(m, seed=11)
    np.testing.assert_array_equal(s1.vertices, s2.vertices)
    offsets = []
    for o in range(m.num_objects):
        vs = np.unique(m.indices[m.triangle_to_object == o])
        d = s1.vertices[vs].astype(np.float64) - m.vertices[vs].astype(np.float64)
        assert np.all(np.abs(d - d[0]) < 1e-4), o          # one offset per object
        assert np.all(np.abs(d[0]) <= 15.0 + 1e-3)
        offsets.append(d[0])
    assert len({tuple(np.round(x, 3)) for x in offsets}) == m.num_objects
    assert not np.array_equal(pz.shuffle(m, seed=12).vertices, s1.vertices)


def test_preprocess_cli_writes_edges_and_grid(tmp_path):
    m = pz.generate(pz.GEN_ELLIPSOIDS, count=25, seed=5, level_min=1, level_max=1, box=40.0)
    src = tmp_path / "set.obj"
    pz.write_obj(str(src), m)
    out, timing = tmp_path / "set_g5_0.pre", tmp_path / "set_timing.json"
    res = subprocess.run([BIN, "--mode", "mesh", "--dataset", str(src), "--output-geometry", str(out),
                          "--output-timing", str(timing), "--generate-grid", "--grid-cell-size", "5.0",
                          "--gamma", "0.5", "--bogus-flag"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "Warning: Unknown argument '--bogus-flag'" in res.stdout
    pre = pz.read_pre(str(out))
    want = pz.compute_grid(pz.extract_edges(pz.read_obj(str(src))), 5.0)
    np.testing.assert_array_equal(pre.vertices, want.vertices)
    np.testing.assert_array_equal(pre.indices, want.indices)
    np.testing.assert_array_equal(pre.triangle_to_object, want.triangle_to_object)
    np.testing.assert_array_equal(pre.edge_starts, want.edge_starts)
    np.testing.assert_array_equal(pre.edge_ends, want.edge_ends)
    np.testing.assert_array_equal(pre.edge_objects, want.edge_objects)
    assert pre.grid_cell_size == 5.0
    np.testing.assert_array_equal(pre.grid, want.grid)
    assert sorted(triangles(pre)) == sorted(triangles(m))
    # every object's edges are the unique edges of its triangles
    assert pre.edge_objects.shape[0] == sum(
        len({tuple(sorted(e)) for t in m.indices[m.triangle_to_object == o].tolist()
             for e in ((t[0], t[1]), (t[1], t[2]), (t[0], t[2]))}) for o in range(m.num_objects))
    phases = json.loads(timing.read_text())
    text = json.dumps(phases).lower()
    for name in ("loading dataset", "generating grid statistics", "extracting edges", "writing geometry data"):
        assert name.replace(" ", "_") in text or name in text, name


def test_preprocess_cli_rejects_dt_and_missing_dataset(tmp_path):
    res = subprocess.run([BIN, "--mode", "dt", "--dataset", str(tmp_path / "x.dt")], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 1 and "not supported" in res.stderr
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1 and "Dataset path is required" in res.stderr
