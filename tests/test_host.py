"""CPU-side tests: the C ABI loads and exports its header, .pre I/O, generators,
grid statistics and estimator arithmetic, CLI argument handling."""
import json
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import pierce_amd as pz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pierce_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;]*\);", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_library_exports_every_declared_symbol():
    lib = pz.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(pz.exported_symbols())


def test_no_device_is_an_error_not_a_fallback():
    if pz.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(pz.PierceError) as e:
        pz.Context(0)
    assert e.value.code == pz.PZ_E_NODEVICE


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_icosphere_closed_forms(level):
    m = pz.generate(pz.GEN_ELLIPSOIDS, count=3, seed=5, level_min=level, level_max=level)
    assert m.vertices.shape[0] == 3 * (10 * 4 ** level + 2)
    assert m.indices.shape[0] == 3 * 20 * 4 ** level
    assert m.edge_objects.shape[0] == 3 * 30 * 4 ** level
    # closed, consistently oriented: every directed edge appears once, its reverse once
    for o in range(3):
        tri = m.indices[m.triangle_to_object == o]
        directed = set()
        for a, b, c in tri.tolist():
            for e in ((a, b), (b, c), (c, a)):
                assert e not in directed
                directed.add(e)
        assert all((b, a) in directed for a, b in directed)
        # positive signed volume (outward orientation)
        v = m.vertices.astype(np.float64)
        vol = sum(np.dot(v[a], np.cross(v[b], v[c])) for a, b, c in tri.tolist())
        assert vol > 0


def test_generator_is_deterministic_and_seeded():
    a1 = pz.generate(pz.GEN_ELLIPSOIDS, count=50, seed=42, jitter=0.02)
    a2 = pz.generate(pz.GEN_ELLIPSOIDS, count=50, seed=42, jitter=0.02)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=50, seed=43, jitter=0.02)
    np.testing.assert_array_equal(a1.vertices, a2.vertices)
    assert not np.array_equal(a1.vertices, b.vertices)


def test_copies_are_scaled_inside_their_source():
    n = 40
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=7, level_min=2, level_max=2, radius_min=0.5, radius_max=2.0)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=8, level_min=2, level_max=2, radius_min=0.5, radius_max=2.0,
                    copy_fraction=1.0, copy_scale_min=0.5, copy_scale_max=0.5, copy_source_count=n,
                    copy_source_seed=7)
    # every B mesh is a copy: its box lies inside some A box
    def boxes(m):
        out = []
        for o in range(n):
            v = m.vertices[m.indices[m.triangle_to_object == o].reshape(-1)]
            out.append((v.min(0), v.max(0)))
        return out
    ba, bb = boxes(a), boxes(b)
    for lo, hi in bb:
        assert any(np.all(lo > la) and np.all(hi < ha) for la, ha in ba)


def test_pre_roundtrip_and_byte_layout(tmp_path):
    m = pz.compute_grid(pz.extract_edges(pz.generate(pz.GEN_ELLIPSOIDS, count=20, seed=3, level_min=1,
                                                     level_max=1, box=30.0)), 5.0)
    path = str(tmp_path / "a.pre")
    pz.write_pre(path, m)
    raw = open(path, "rb").read()
    V, T, E, G = m.vertices.shape[0], m.indices.shape[0], m.edge_objects.shape[0], m.grid.shape[0]
    assert len(raw) == 56 + 12 * V + 12 * T + 4 * T + 28 * E + 8 + 28 * G
    magic, ver, nv, ni, nm, ne, tt, hg, he = struct.unpack("<IIQQQQQBB", raw[:50])
    assert (magic, ver, nv, ni, nm, ne, tt, hg, he) == (0x52334442, 4, V, T, T, E, T, 1, 1)
    r = pz.read_pre(path)
    for f in ("vertices", "indices", "triangle_to_object", "edge_starts", "edge_ends", "edge_objects"):
        np.testing.assert_array_equal(getattr(r, f), getattr(m, f))
    np.testing.assert_array_equal(r.grid, m.grid)
    assert r.grid_cell_size == 5.0


def test_pre_rejects_wrong_version(tmp_path):
    p = tmp_path / "bad.pre"
    p.write_bytes(struct.pack("<II", 0x52334442, 3) + b"\0" * 48)
    with pytest.raises(pz.PierceError) as e:
        pz.read_pre(str(p))
    assert e.value.code == pz.PZ_E_IO and "version" in str(e.value)


def test_edge_extraction_matches_triangle_edges():
    m = pz.generate(pz.GEN_ELLIPSOIDS, count=30, seed=4, level_min=1, level_max=1, jitter=0.02)
    e = pz.extract_edges(m)
    # same undirected edge set per object as the triangles' own edges, first-seen orientation kept
    key = lambda s, t: tuple(sorted((tuple(s), tuple(t))))
    got = set((o, key(s, t)) for s, t, o in zip(e.edge_starts.tolist(), e.edge_ends.tolist(),
                                                e.edge_objects.tolist()))
    exp = set()
    for t, o in zip(m.indices.tolist(), m.triangle_to_object.tolist()):
        for k in range(3):
            exp.add((o, key(m.vertices[t[k]].tolist(), m.vertices[t[(k + 1) % 3]].tolist())))
    assert got == exp and len(e.edge_objects) == len(exp)
    # key order: lexicographic on the quantised (min endpoint, max endpoint, object)
    q = lambda p: tuple(int(round(float(x) * 1e5)) for x in p)
    keys = [(min(q(s), q(t)), max(q(s), q(t)), o) for s, t, o in zip(e.edge_starts, e.edge_ends, e.edge_objects)]
    assert keys == sorted(keys)


def test_grid_statistics_closed_form():
    # one axis-aligned cube of edge 4 centred at (10, 10, 10), cells of 5
    m = pz.generate(pz.GEN_CUBES, count=1, seed=0, radius_min=4.0, radius_max=4.0, box=0.0)
    m.vertices += np.float32(10.0)
    g = pz.compute_grid(m, 5.0).grid
    # box [8, 12]^3 touches cells 1..2 on each axis -> 8 cells; centre cell (2, 2, 2)
    assert len(g) == 8
    assert all(c["touch_count"] == 1 for c in g)
    assert sum(c["center_count"] for c in g) == 1
    centre = [c for c in g if c["center_count"] == 1][0]
    assert (centre["ix"], centre["iy"], centre["iz"]) == (2, 2, 2)
    assert np.allclose(g["avg_size_mean"], 4.0) and np.allclose(g["vol_ratio"], 1.0)


def test_estimator_formula():
    a = np.zeros(2, pz.GRID_DTYPE)
    b = np.zeros(2, pz.GRID_DTYPE)
    a["touch_count"] = [3, 0]
    b["touch_count"] = [4, 7]
    a["avg_size_mean"] = [1.0, 2.0]
    b["avg_size_mean"] = [0.5, 2.0]
    a["vol_ratio"] = [0.25, 1.0]
    b["vol_ratio"] = [1.0, 1.0]
    cv = 125.0
    got = pz.estimate_overlap_selectivity(a, b, cv, epsilon=0.001, gamma=0.8)
    s = 1.0 + 0.5 + 0.001
    exp = 3 * 4 * min(1.0, s ** 3 / cv * (0.25 ** 0.5) ** 0.8)
    assert abs(got - exp) < 1e-4 * exp


TOOLS = ["pierce_overlap", "pierce_overlap_two_pass", "pierce_intersection", "pierce_containment",
         "pierce_generate"]


@pytest.mark.parametrize("tool", TOOLS)
def test_cli_help_exits_zero(tool):
    r = subprocess.run([os.path.join(pz.BIN_DIR, tool), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage" in r.stdout


@pytest.mark.parametrize("tool", TOOLS[:4])
def test_cli_missing_inputs_exit_one(tool):
    r = subprocess.run([os.path.join(pz.BIN_DIR, tool), "--runs", "1", "--some-unknown-flag"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1


@pytest.mark.parametrize("tool", ["pierce_overlap", "pierce_intersection"])
def test_cli_bad_direction_exit_one(tool):
    r = subprocess.run([os.path.join(pz.BIN_DIR, tool), "--mesh1", "a", "--mesh2", "b", "--query-direction",
                        "sideways"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1


def test_cli_free_mem_fraction_range():
    r = subprocess.run([os.path.join(pz.BIN_DIR, "pierce_overlap"), "--mesh1", "a", "--mesh2", "b",
                        "--hash-table-free-mem-fraction", "1.5"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1


def test_generate_tool_writes_pre(tmp_path):
    out = str(tmp_path / "g.pre")
    r = subprocess.run([os.path.join(pz.BIN_DIR, "pierce_generate"), "--output", out, "--count", "12", "--level", "1",
                        "--generate-grid"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    m = pz.read_pre(out)
    assert m.indices.shape[0] == 12 * 80 and m.edge_objects.shape[0] == 12 * 120 and m.grid.shape[0] > 0


def test_grid_estimate_closed_form_and_oracle():
    from oracle import oracle as orc
    from tests.test_oracle import GRID_A, GRID_B, GRID_CASES
    for (cs, eps, gamma, alpha), (raw, a, final) in GRID_CASES:
        got = pz.estimate_grid(GRID_A, GRID_B, cs, epsilon=eps, gamma=gamma, alpha_correction=alpha)
        assert got["has_grid"] and got["matched_cells"] == 2
        assert (got["raw_pairs"], got["alpha"], got["estimated_pairs"]) == (raw, a, final)
        assert got["effective_size"] == (1.0, 1.0)
    assert not pz.estimate_grid(GRID_A, GRID_B[:0], 4.0)["has_grid"]
    assert not pz.estimate_grid(GRID_A, GRID_B, 0.0)["has_grid"]
    # random grids with repeated indices and untouched cells against the oracle
    rng = np.random.default_rng(5)
    for trial in range(20):
        g = []
        for n in (int(rng.integers(1, 300)), int(rng.integers(1, 300))):
            c = np.zeros(n, pz.GRID_DTYPE)
            c["ix"], c["iy"], c["iz"] = (rng.integers(0, 6, n) for _ in range(3))
            c["touch_count"] = rng.integers(0, 50, n) * (rng.random(n) > 0.1)
            c["avg_size_mean"] = rng.uniform(0.1, 6.0, n)
            c["vol_ratio"] = rng.uniform(0.05, 1.0, n)
            g.append(c)
        cs = float(rng.choice([0.5, 2.0, 5.0]))
        got = pz.estimate_grid(g[0], g[1], cs, alpha_correction=trial % 3 != 0)
        exp = orc.grid_estimate(g[0], g[1], cs, alpha_correction=trial % 3 != 0)
        assert got["matched_cells"] == exp["matched_cells"]
        assert abs(got["raw_pairs"] - exp["raw_pairs"]) <= 1e-6 * max(1.0, exp["raw_pairs"])
        assert got["alpha"] == pytest.approx(exp["alpha"], rel=1e-6)
        assert got["avg_size"] == pytest.approx(exp["avg_size"], rel=1e-6)
        assert abs(got["estimated_pairs"] - exp["estimated_pairs"]) <= 1
