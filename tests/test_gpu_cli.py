"""The four query tools end to end on .pre files: CSV contents against the
oracle, stdout lines and timing-JSON keys the reference harness parses
(benchmarks/overlap/adapters/pierce_adapter.py:228-376,
benchmarks/predicates/adapters/containment.py:134-172)."""
import csv
import json
import os
import re
import subprocess

import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def run(tool, args, cwd):
    print(f"[cli] {tool} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(pz.BIN_DIR, tool)] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read_pairs(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], [tuple(int(x) for x in r[:2]) for r in rows[1:]], [r[2] for r in rows[1:] if len(r) > 2]


@pytest.fixture(scope="module")
def pre_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("pre")
    gen = os.path.join(pz.BIN_DIR, "pierce_generate")
    common = ["--count", "300", "--level", "2", "--radius-min", "0.5", "--radius-max", "2.0", "--box", "60",
              "--generate-grid"]
    subprocess.run([gen, "--output", str(d / "a.pre"), "--seed", "7"] + common, check=True, capture_output=True)
    subprocess.run([gen, "--output", str(d / "b.pre"), "--seed", "8", "--copy-fraction", "0.2", "--copy-scale-min",
                    "0.3", "--copy-scale-max", "0.8", "--copy-source-count", "300", "--copy-source-seed", "7"] + common,
                   check=True, capture_output=True)
    a, b = pz.read_pre(str(d / "a.pre")), pz.read_pre(str(d / "b.pre"))
    return d, a, b


def test_two_pass_tool(pre_files, tmp_path):
    d, a, b = pre_files
    out = run("pierce_overlap_two_pass", ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre"), "--output",
                                          "t.json", "--warmup-runs", "1", "--ptx", "ignored"], tmp_path)
    hdr, pairs, _ = read_pairs(tmp_path / "mesh_overlap_results.csv")
    assert hdr == ["object_id_mesh1", "object_id_mesh2"]
    exp = orc.face_overlap(a, b)
    assert pairs == [tuple(p) for p in exp.tolist()]
    assert re.search(r"^Unique object pairs: %d$" % len(exp), out, re.M)
    assert re.search(r"^Mesh1 objects: 300$", out, re.M)
    phases = json.load(open(tmp_path / "t.json"))["phases"]
    assert any(k.startswith("raytrace_") for k in phases) and "download results_1" in phases


def test_overlap_tool_estimate_and_pairs_output(pre_files, tmp_path):
    d, a, b = pre_files
    out = run("pierce_overlap", ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre"), "--pairs-output",
                                 "p.csv", "--no-export", "--query-direction", "both"], tmp_path)
    hdr, pairs, _ = read_pairs(tmp_path / "p.csv")
    assert pairs == [tuple(p) for p in orc.face_overlap(a, b).tolist()]
    assert re.search(r"^Final Estimated Pairs:\s+\d+$", out, re.M)
    assert re.search(r"^Hash Table Query found \d+ unique pairs\.$", out, re.M)
    phases = json.load(open(tmp_path / "estimated_overlap_timing.json"))["phases"]
    assert "selectivity estimation_1" in phases and "raytrace_hash_mesh1tomesh2_1" in phases


def _line(out, label):
    m = re.findall(r"^%s:\s+(-?\d+)$" % re.escape(label), out, re.M)
    assert m, (label, out)
    return [int(x) for x in m]


@pytest.mark.parametrize("tool,title,full", [
    ("pierce_overlap", "=== Selectivity Estimation (Overlap - Direct) ===", True),
    ("pierce_intersection", "=== Selectivity Estimation ===", True),
    ("pierce_containment", "=== Containment Selectivity Estimation ===", False),
])
def test_verbose_estimate_block(pre_files, tmp_path, tool, title, full):
    # the estimation block of pierce_overlap_main.cpp:450-466 (and its
    # intersection / containment forms) against the oracle's whole-set estimate
    d, a, b = pre_files
    args = ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre"), "--verbose-estimate", "--warmup-runs", "0",
            "--no-export"]
    if tool != "pierce_containment":
        args.append("--estimate-only")
    out = run(tool, args, tmp_path)
    exp = orc.grid_estimate(a.grid, b.grid, a.grid_cell_size)
    assert exp["has_grid"] and exp["matched_cells"] > 0 and exp["estimated_pairs"] > 0
    assert title in out
    assert _line(out, "Matched Sparse Cells") == [exp["matched_cells"]]
    assert abs(_line(out, "Raw Potential Pairs")[0] - int(exp["raw_pairs"])) <= 1
    final = _line(out, "Final Estimated Pairs")
    assert abs(final[0] - exp["estimated_pairs"]) <= 1
    assert all(f == final[0] for f in final)  # the join sizes with the same estimate
    assert ("Replication Factor (alpha):" in out) == full
    assert ("Alpha Correction Enabled:" in out) == (tool == "pierce_overlap")


def test_intersection_tool(pre_files, tmp_path):
    d, a, b = pre_files
    out = run("pierce_intersection", ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre")], tmp_path)
    hdr, pairs, _ = read_pairs(tmp_path / "intersection_pairs.csv")
    assert hdr == ["a_object_id", "b_object_id"]
    exp, _ = orc.predicate_join(a, b, 1)
    assert pairs == [tuple(p) for p in exp.tolist()]
    assert re.search(r"^Actual Intersection Pairs: %d$" % len(exp), out, re.M)


def test_containment_tool(pre_files, tmp_path):
    d, a, b = pre_files
    out = run("pierce_containment", ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre"),
                                     "--include-overlap-pairs"], tmp_path)
    hdr, pairs, types = read_pairs(tmp_path / "mesh_containment_results.csv")
    assert hdr == ["a_object_id", "b_object_id", "pair_type"]
    exp, et = orc.predicate_join(a, b, 2, include_overlap=True)
    cont = [tuple(p) for p, t in zip(exp.tolist(), et) if t == 0]
    over = [tuple(p) for p, t in zip(exp.tolist(), et) if t == 1]
    assert pairs == cont + over and types == ["containment"] * len(cont) + ["overlap"] * len(over)
    assert re.search(r"^Containment pairs \(B in A\): %d$" % len(cont), out, re.M)
    assert re.search(r"^Reported pairs: %d$" % len(exp), out, re.M)


def test_tool_rejects_pre_without_edges(tmp_path):
    m = pz.generate(pz.GEN_ELLIPSOIDS, count=4, seed=1, edges_mode=pz.EDGES_NONE)
    pz.write_pre(str(tmp_path / "n.pre"), m)
    r = subprocess.run([os.path.join(pz.BIN_DIR, "pierce_overlap_two_pass"), "--mesh1", str(tmp_path / "n.pre"),
                        "--mesh2", str(tmp_path / "n.pre")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "precomputed edges" in r.stderr


def test_obj_through_preprocess_to_two_pass(tmp_path):
    """OBJ datasets -> pierce_preprocess (.pre v4 with edges and grid) -> the
    two-pass tool, against the oracle on the meshes as read from the OBJ."""
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=60, seed=21, level_min=1, level_max=2, box=25.0)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=60, seed=22, level_min=1, level_max=2, box=25.0)
    for name, m in (("a", a), ("b", b)):
        pz.write_obj(str(tmp_path / f"{name}.obj"), m)
        subprocess.run([os.path.join(pz.BIN_DIR, "pierce_preprocess"), "--mode", "mesh", "--dataset",
                        str(tmp_path / f"{name}.obj"), "--output-geometry", str(tmp_path / f"{name}.pre"),
                        "--output-timing", str(tmp_path / f"{name}_timing.json"), "--generate-grid",
                        "--grid-cell-size", "5.0"], check=True, capture_output=True, timeout=120)
    run("pierce_overlap_two_pass", ["--mesh1", "a.pre", "--mesh2", "b.pre", "--output", "t.json"], tmp_path)
    _, pairs, _ = read_pairs(tmp_path / "mesh_overlap_results.csv")
    exp = orc.face_overlap(pz.read_obj(str(tmp_path / "a.obj")), pz.read_obj(str(tmp_path / "b.obj")))
    assert len(exp) > 0
    assert pairs == [tuple(p) for p in exp.tolist()]


def test_hash_contention_lines(pre_files, tmp_path):
    """--track-hash-contention prints the lines pierce_adapter.py:290-304 parses;
    probes minus contended probes = successful inserts >= unique pairs."""
    d, a, b = pre_files
    out = run("pierce_overlap", ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre"), "--strategy", "global",
                                 "--track-hash-contention", "--no-export", "--runs", "2"], tmp_path)
    pat = r"^Hash contention \((run \d+|last run)\): (\d+)/(\d+) accesses \((\d+\.\d\d)%\)$"
    lines = re.findall(pat, out, re.M)
    assert [l[0] for l in lines] == ["run 1", "run 2", "last run"], out
    n = len(orc.face_overlap(a, b))
    for _, cont, acc, pct in lines:
        cont, acc = int(cont), int(acc)
        assert 0 <= cont <= acc and acc - cont >= n > 0
        assert abs(float(pct) - 100.0 * cont / acc) < 0.006
    assert re.search(r"^Hash contention tracking: enabled$", out, re.M)


def test_hash_counters_in_the_result(ctx, pre_files):
    _, a, b = pre_files
    da, db = ctx.upload(a), ctx.upload(b)
    g = pz.join(ctx, da, db, strategy=pz.STRATEGY_GLOBAL, sizing=pz.SIZING_ESTIMATE)
    assert 0 <= g.hash_contentions <= g.hash_accesses and g.hash_accesses - g.hash_contentions >= len(g.pairs) > 0
    t = pz.join(ctx, da, db)  # two-level: no hash set
    assert t.hash_accesses == 0 and t.hash_contentions == 0
    np.testing.assert_array_equal(g.pairs, t.pairs)


# Query-time sums of the reference adapters, per tool: the phase keys each one
# adds up (pierce_adapter.py:338-376 for the exact and direct_estimation modes,
# predicates/adapters/intersection.py:207-221, containment.py:163-172).
ADAPTER_KEYS = {
    "pierce_overlap": ["selectivity estimation", "raytrace_hash_mesh1tomesh2", "raytrace_hash_mesh2tomesh1",
                       "raytrace_overlap_hash_mesh1tomesh2", "raytrace_overlap_hash_mesh2tomesh1",
                       "compact_hash_table_pairs", "download results"],
    "pierce_overlap_two_pass": None,  # every raytrace_* key + gpu deduplication + download results
    "pierce_intersection": ["selectivity estimation", "raytrace_hash_mesh1tomesh2", "raytrace_hash_mesh2tomesh1",
                            "raytrace_overlap_hash_mesh1tomesh2", "raytrace_overlap_hash_mesh2tomesh1",
                            "raytrace_containment_hash_mesh1tomesh2", "raytrace_containment_hash_mesh2tomesh1",
                            "compact_hash_table_pairs", "download results"],
    "pierce_containment": ["selectivity estimation", "raytrace_overlap_hash_mesh1tomesh2",
                           "raytrace_overlap_hash_mesh2tomesh1", "raytrace_containment_hash_mesh1tomesh2",
                           "raytrace_containment_hash_mesh2tomesh1", "compact_hash_table_pairs (containment)",
                           "compact_hash_table_pairs (overlap)", "download results"],
}


def adapter_query_ms(tool, phases):
    """The adapter's query time from a timing JSON (keys lower-cased with a
    _<n> suffix, summed per name), using duration_us for precision."""
    vals = {}
    for k, v in phases.items():
        name = re.sub(r"_\d+$", "", k.lower())
        vals[name] = vals.get(name, 0.0) + v["duration_us"] / 1000.0
    keys = ADAPTER_KEYS[tool]
    if keys is None:
        return sum(v for k, v in vals.items() if k.startswith("raytrace_")) + vals.get("gpu deduplication", 0.0) + \
            vals.get("download results", 0.0)
    return sum(vals.get(k, 0.0) for k in keys)


@pytest.fixture(scope="module")
def larger_pre(tmp_path_factory):
    d = tmp_path_factory.mktemp("pre_large")
    gen = os.path.join(pz.BIN_DIR, "pierce_generate")
    common = ["--count", "4000", "--level", "2", "--radius-min", "0.5", "--radius-max", "2.0", "--box", "150",
              "--generate-grid"]
    subprocess.run([gen, "--output", str(d / "a.pre"), "--seed", "17"] + common, check=True, capture_output=True)
    subprocess.run([gen, "--output", str(d / "b.pre"), "--seed", "18", "--copy-fraction", "0.2", "--copy-scale-min",
                    "0.3", "--copy-scale-max", "0.8", "--copy-source-count", "4000", "--copy-source-seed", "17"]
                   + common, check=True, capture_output=True)
    return d


@pytest.mark.parametrize("tool", list(ADAPTER_KEYS))
@pytest.mark.parametrize("strategy", ["two-level", "global"])
def test_adapter_query_time_covers_the_join(larger_pre, tmp_path, tool, strategy):
    """Every device phase of the join sits under a key the tool's adapter sums,
    so the adapter's query time is at least 95% of the join's own time."""
    d = larger_pre
    out = run(tool, ["--mesh1", str(d / "a.pre"), "--mesh2", str(d / "b.pre"), "--output", "t.json",
                     "--warmup-runs", "2", "--no-export", "--strategy", strategy], tmp_path)
    join_ms = float(re.search(r"^Join time \(device-resident geometry to sorted host pairs\): ([\d.e+-]+) ms$",
                              out, re.M).group(1))
    q = adapter_query_ms(tool, json.load(open(tmp_path / "t.json"))["phases"])
    print(f"[adapter] {tool} {strategy}: adapter query {q:.3f} ms, join {join_ms:.3f} ms", flush=True)
    assert q >= 0.95 * join_ms
    assert q <= 1.05 * join_ms + 0.05  # nothing counted twice
