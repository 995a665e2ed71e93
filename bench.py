#!/usr/bin/env python3
"""Join throughput of the MI355X spatial-join engine on one GPU.

Headline workload: A = 100,000 ellipsoid icospheres at subdivision 3 (1,280
triangles each), radii U[0.5, 2], centres U[0, 1000)^3; B = 100,000 meshes of
which 20% are copies of random A meshes scaled by U[0.3, 0.8] about their
centroid and 80% fresh ellipsoids; fp32 vertices, seeded (A 42, B 43).

One step = one overlap join, device-resident geometry -> sorted unique host
pair set, including every index build.  Two paths run on the same GPU:
  chosen : two-level join (object broad phase, pair sieve, per-object BVHs for
           pairs beyond the sieve, early exit), fp32-filtered exact predicates,
           two-pass output sizing (learned and estimated sizing timed beside it);
  plain  : one BVH per set over all triangles, every edge traced, fp64-filtered
           exact predicates throughout, exact two-pass count/scan/write/sort/unique.
Both are checked against each other pair for pair.  The CPU leg times the
oracle's Face and TOUCH restatements on a spatial sub-box of the workload and,
after every GPU measurement, uses the oracle as the checker of recorded GPU
results (the sub-box sample, C1, a sample of the stream's queries).

The same line carries the other workloads at their stated sizes: the same
sets under intersection and containment, 1,000 vs 1,000 ellipsoids, the
10,000 vs 10,000 complexity sweep (three predicates as one batch), 2,000 vs
2,000 branching tubes, and a stream of 10,000 queries sized by the learned
estimator.

Prints one JSON line (see the repository's measurement contract).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "join ms + Gtri-tests/s, chosen vs plain GPU path, 0 mismatched pairs"
HBM_PEAK_GBS = 8000.0
COUNTER_TABLE = os.path.join(ROOT, "profiles", "kernel_counters_c2_overlap.md")


def alg_bytes(kernel, k):
    """Algorithmic bytes of one profiled kernel over one join, in the payload
    units of the .pre format: 12 B per vertex, 12 + 4 B per triangle (indices
    and object id), 28 B per edge, 8 B per result pair.  Index kernels count
    the payload of the triangles they index (IT) and of those triangles'
    vertices (IV); what they write is a design choice, reported separately
    (design_bytes).  None for kernels without a payload unit."""
    if kernel == "object_stats":    # both sets' triangles and vertices, once
        return 16 * k["T"] + 12 * k["V"]
    if kernel == "object_bounds_runs":  # both sets' vertices, once (objects own vertex runs)
        return 12 * k["V"]
    if kernel in ("build_small", "build_chunks", "tri_keys", "narrow_dual", "gather_leaves", "refit"):
        return 16 * k["IT"] + 12 * k["IV"]
    if kernel == "pair_sieve":      # the triangles and vertices of each sieved pair's two objects
        return 16 * k["ST"] + 12 * k["SV"]
    if kernel == "narrow":          # the indexed edges
        return 28 * k["IE"]
    if kernel == "edge_count":
        return 4 * k["E"]
    return None


def design_bytes(kernel, k):
    """What the kernel reads and writes by its own design, per join."""
    if kernel == "build_small":     # indices + 3 corners read, one 32-B node and one 48-B leaf written
        return (12 + 36 + 32 + 48) * k["IT"]
    if kernel == "narrow_dual":     # each indexed leaf (48 B) and node (32 B) at most once per side
        return (48 + 32) * k["IT"]
    if kernel == "object_stats":
        return 16 * k["T"] + 12 * k["V"]
    if kernel == "object_bounds_runs":
        return 12 * k["V"]
    if kernel == "pair_sieve":      # index triples and vertex runs read once per pair, nothing written per triangle
        return 12 * k["ST"] + 12 * k["SV"]
    return None


def counter_table(path=COUNTER_TABLE):
    """Per-kernel rows of the committed rocprofv3 counter summary (markdown
    table with a `kernel` column); {} when absent."""
    rows = {}
    if not os.path.exists(path):
        return rows
    head = None
    for line in open(path):
        if not line.startswith("|"):
            head = None if not line.strip() else head
            continue
        cells = [c.strip().strip("*`") for c in line.strip().strip("|").split("|")]
        if head is None:
            head = cells
            continue
        if set(cells[0]) <= set("-: "):
            continue
        if head and head[0] == "kernel" and len(cells) == len(head):
            rows[cells[0]] = dict(zip(head, cells))
    return rows


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """CPU threads for the host-side baselines: this process's CPU share
    (affinity), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def gen_sets(pz, n):
    common = dict(level_min=3, level_max=3, radius_min=0.5, radius_max=2.0, box=1000.0)
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=42, **common)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=43, copy_fraction=0.2, copy_scale_min=0.3,
                    copy_scale_max=0.8, copy_source_count=n, copy_source_seed=42, **common)
    return a, b


def sub_box(pz, m, side):
    """Objects whose first vertex lies in [0, side)^3, re-numbered compactly."""
    t2o = m.triangle_to_object
    first = np.full(t2o.max() + 1, -1, np.int64)
    order = np.arange(len(t2o))
    first[t2o[::-1]] = order[::-1]
    p0 = m.vertices[m.indices[first, 0]]
    keep = np.all(p0 < side, axis=1)
    new_id = np.cumsum(keep) - 1
    tmask = keep[t2o]
    idx = m.indices[tmask]
    used = np.unique(idx)
    remap = np.zeros(m.vertices.shape[0], np.uint32)
    remap[used] = np.arange(len(used), dtype=np.uint32)
    emask = keep[m.edge_objects]
    return pz.Mesh(m.vertices[used], remap[idx], new_id[t2o[tmask]].astype(np.int32),
                   m.edge_starts[emask], m.edge_ends[emask], new_id[m.edge_objects[emask]].astype(np.int32))


def time_join(pz, ctx, da, db, steps, **desc):
    outs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        outs.append(pz.join(ctx, da, db, **desc))
    dt = (time.perf_counter() - t0) / steps * 1000.0
    return dt, outs[-1]


def pair_mismatch(p, q):
    a = set(map(tuple, np.asarray(p).tolist()))
    b = set(map(tuple, np.asarray(q).tolist()))
    return len(a ^ b)


def counters_of(o):
    return {"tests": o.tests_total, "escalated_fp64": o.tests_fp64, "escalated_exact": o.tests_exact,
            "fp32_only_misclassified": o.tests_fp32_misclassified, "narrow_rounds": o.narrow_rounds,
            "narrow_fallback_pairs": o.narrow_fallback_pairs, "index_height": o.index_height}


def tune_of(t):
    return {k: t[k] for k in ("build_threads", "narrow_queue", "narrow_chunk")}


def c2_predicates(pz, ctx, da, db, sa, sb, chosen, plain, steps, plain_steps, checks):
    """Intersection and containment on the headline sets.  The plain path is
    the set-wide one (every edge through a BVH over all triangles, one +Z ray
    per object through the other set's BVH, fp64 tests, two-pass sizing); the
    two-level walk with fp64 tests is reported beside it.  The chosen path's
    pairs on the spatial sample go to the CPU leg's oracle check."""
    plain_tl = dict(strategy=pz.STRATEGY_TWO_LEVEL, precision=pz.PREC_FP64, sizing=pz.SIZING_TWO_PASS)
    res = {}
    dsa, dsb = ctx.upload(sa), ctx.upload(sb)
    for name, pred in (("c2_intersection", pz.PRED_INTERSECTION), ("c2_containment", pz.PRED_CONTAINMENT)):
        try:
            kw = dict(chosen, predicate=pred)
            pz.join(ctx, da, db, **kw)
            ms_c, out = time_join(pz, ctx, da, db, steps, **kw)
            pz.join(ctx, da, db, predicate=pred, **plain_tl)
            ms_tl, tout = time_join(pz, ctx, da, db, plain_steps, predicate=pred, **plain_tl)
            ctx.trim()
            pz.join(ctx, da, db, predicate=pred, **plain)
            ms_p, pout = time_join(pz, ctx, da, db, plain_steps, predicate=pred, **plain)
            ctx.trim()
            got = pz.join(ctx, dsa, dsb, **kw)
            checks.append((name, sa, sb, pred, got.pairs))
            res[name] = {
                "workload": f"{'intersection' if pred == pz.PRED_INTERSECTION else 'containment'} join on the "
                            "headline sets (100000 vs 100000 ellipsoid icospheres s=3, 20% of B scaled copies of A)",
                "plain": "set-wide BVH per set, every edge traced and one +Z ray per object, fp64-filtered exact "
                         "tests, two-pass sizing",
                "join_ms_chosen": round(ms_c, 3), "join_ms_plain": round(ms_p, 3),
                "speedup_vs_plain": round(ms_p / ms_c, 2), "mismatched_pairs": pair_mismatch(out.pairs, pout.pairs),
                "join_ms_two_level_fp64": round(ms_tl, 3),
                "two_level_fp64_mismatched_pairs": pair_mismatch(out.pairs, tout.pairs),
                "phases_plain": [(n, round(m, 3)) for n, m in pout.phases],
                "pairs": int(len(out.pairs)), "candidates": out.num_candidates,
                "pairs_per_s": round(len(out.pairs) / (ms_c * 1e-3), 1),
                "parity_rays": out.parity_rays, "parity_inside": out.parity_inside,
                "sizing_overflowed": out.overflowed, **counters_of(out),
                "oracle_sample": f"the objects of both sets whose first vertex lies in the CPU sample box: "
                                 f"{sa.num_objects} + {sb.num_objects}",
                "phases_chosen": [(n, round(m, 3)) for n, m in out.phases]}
            log(f"[bench] {name}: chosen {ms_c:.3f} ms, plain {ms_p:.3f} ms, pairs {len(out.pairs)}, "
                f"mismatches {res[name]['mismatched_pairs']}")
        except Exception as e:
            res[name] = {"error": str(e)}
            log(f"[bench] {name}: error {e}")
    dsa.close()
    dsb.close()
    return res


def single_config(pz, ctx, name, a, b, chosen, steps, plain, plain_steps, checks=None):
    """One configuration: autotuned chosen path vs the plain path, pair for pair."""
    da, db = ctx.upload(a), ctx.upload(b)
    tune = pz.autotune(ctx, da, db, **chosen)  # on this workload, untimed
    pz.join(ctx, da, db, **chosen)
    ms_c, out = time_join(pz, ctx, da, db, steps, **chosen)
    ctx.trim()  # the chosen path's work buffers need not stay beside the plain path's
    pz.join(ctx, da, db, **plain)
    ms_p, pout = time_join(pz, ctx, da, db, plain_steps, **plain)
    r = {"objects": [a.num_objects, b.num_objects], "triangles": [int(a.indices.shape[0]), int(b.indices.shape[0])],
         "join_ms_chosen": round(ms_c, 3), "join_ms_plain": round(ms_p, 3), "speedup_vs_plain": round(ms_p / ms_c, 2),
         "mismatched_pairs": pair_mismatch(out.pairs, pout.pairs), "pairs": int(len(out.pairs)),
         "candidates": out.num_candidates, "pairs_per_s": round(len(out.pairs) / (ms_c * 1e-3), 1),
         "steps": [steps, plain_steps], "sizing_overflowed": out.overflowed, **counters_of(out),
         "plain_index_height": pout.index_height, "autotune": tune_of(tune),
         "phases_chosen": [(n, round(m, 3)) for n, m in out.phases]}
    if checks is not None:
        checks.append((name.lower(), a, b, pz.PRED_OVERLAP, out.pairs))
    da.close()
    db.close()
    ctx.trim()
    log(f"[bench] {name}: chosen {ms_c:.3f} ms, plain {ms_p:.3f} ms, mismatches {r['mismatched_pairs']}")
    return r


def c1_config(pz, ctx, chosen, plain, steps, checks):
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=1000, seed=42, box=100.0, jitter=0.02)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=1000, seed=43, box=100.0, jitter=0.02)
    return dict(workload="overlap, 1000 vs 1000 ellipsoid icospheres s=2 (320 tris), radii U[1,3], centres in "
                         "100^3, jitter 0.02 r",
                **single_config(pz, ctx, "C1", a, b, chosen, steps, plain, steps, checks=checks))


def c3_config(pz, ctx, chosen, plain, steps):
    kw = dict(level_min=1, level_max=6, radius_min=1.0, radius_max=4.0, box=500.0)
    a = pz.generate(pz.GEN_SPHERES_MIXED, count=10000, seed=42, **kw)
    b = pz.generate(pz.GEN_SPHERES_MIXED, count=10000, seed=43, **kw)
    preds = [pz.PRED_OVERLAP, pz.PRED_INTERSECTION, pz.PRED_CONTAINMENT]
    da3, db3 = ctx.upload_batch([a] * 3), ctx.upload_batch([b] * 3)
    q = [dict(predicate=p) for p in preds]
    tune = pz.autotune(ctx, da3, db3, **chosen)  # on this workload, untimed
    pz.join_batch(ctx, da3, db3, q, **chosen)
    t0 = time.perf_counter()
    for _ in range(steps):
        outs, summ = pz.join_batch(ctx, da3, db3, q, **chosen)
    ms_c = (time.perf_counter() - t0) / steps * 1e3
    da3.close()
    db3.close()
    ctx.trim()
    da, db = ctx.upload(a), ctx.upload(b)
    pz.join(ctx, da, db, **plain)  # warm
    t0 = time.perf_counter()
    pouts = [pz.join(ctx, da, db, predicate=p, **plain) for p in preds]
    ms_p = (time.perf_counter() - t0) * 1e3
    mism = sum(pair_mismatch(o.pairs, po.pairs) for o, po in zip(outs, pouts))
    da.close()
    db.close()
    ctx.trim()
    log(f"[bench] C3: batch {ms_c:.3f} ms, plain {ms_p:.3f} ms, mismatches {mism}")
    return {"workload": "overlap + intersection + containment, 10000 vs 10000 spheres, subdivision 1-6 with "
                        "P ~ 2^-l (80-81920 tris), radius U[1,4], centres in 500^3",
            "chosen": "the three predicates as one batched join (fp32-filtered, two-pass sizing)",
            "plain": "one join per launch and predicate on the set-wide path (edges through a BVH over all "
                     "triangles, +Z rays per object), fp64 filter, two-pass sizing",
            "objects": [a.num_objects, b.num_objects],
            "triangles": [int(a.indices.shape[0]), int(b.indices.shape[0])],
            "join_ms_chosen": round(ms_c, 3), "join_ms_plain": round(ms_p, 3),
            "speedup_vs_plain": round(ms_p / ms_c, 2), "mismatched_pairs": int(mism),
            "pairs": [int(len(o.pairs)) for o in outs], **counters_of(summ), "autotune": tune_of(tune)}


def c4_config(pz, ctx, chosen, plain, count, steps, plain_steps):
    """count vs count branching tubes; 2,000 is the stated size (box 1000).
    Smaller counts keep the tube density (box scaled by the cube root)."""
    box = 1000.0 * (count / 2000.0) ** (1.0 / 3.0)
    kw = dict(box=box, radius_min=0.5, radius_max=3.0)
    t = time.perf_counter()
    a = pz.generate(pz.GEN_TUBES, count=count, seed=42, **kw)
    b = pz.generate(pz.GEN_TUBES, count=count, seed=43, degenerate_fraction=0.01, copy_source_count=count,
                    copy_source_seed=42, **kw)
    log(f"[bench] C4: generated {count} + {count} tubes in {time.perf_counter() - t:.1f} s: "
        f"T={a.indices.shape[0]}+{b.indices.shape[0]}")
    free0, total = ctx.mem_info()
    r = single_config(pz, ctx, "C4", a, b, chosen, steps, plain, plain_steps)
    return dict(workload=f"overlap, {count} vs {count} branching tubes (lognormal median 200k tris, sigma 0.5, "
                         f"clipped [50k, 2M]; tube radius U[0.5, 3]; 16 segments a ring), centres in {box:.0f}^3, "
                         "1% of B near-degenerate copies of A (1e-6 offsets or coplanar rotations)",
                hbm_free_before_gb=round(free0 / 1e9, 1), hbm_total_gb=round(total / 1e9, 1),
                payload_gb=round(sum(16 * m.indices.shape[0] + 12 * m.vertices.shape[0] +
                                     28 * m.edge_objects.shape[0] for m in (a, b)) / 1e9, 2), **r)


def c5_config(pz, ctx, nqueries, ntrain, batch_tris, checks):
    """The library stream: nqueries queries of (set A, set B, predicate) with
    learned output sizing against two-pass sizing, both as batched joins."""
    from pierce_amd.stream import instantiate_host, run_library_stream, stream_catalogue, stream_query

    t = time.perf_counter()
    cat, ids = stream_catalogue()
    lib = ctx.library(cat)
    tpo = np.bincount(cat.triangle_to_object, minlength=cat.num_objects)
    train = [stream_query(q, ids, seed0=1) for q in range(ntrain)]
    held = [stream_query(q, ids, seed0=2) for q in range(nqueries)]
    t_gen = time.perf_counter() - t
    # training targets from the exact two-pass join itself
    t = time.perf_counter()
    touts, _, X = run_library_stream(ctx, lib, train, tpo, with_features=True, max_tris=batch_tris,
                                     sizing=pz.SIZING_TWO_PASS)
    T = np.stack([pz.log_counts([o.num_candidates for o in touts]), pz.log_counts([len(o.pairs) for o in touts])], 1)
    t_train_joins = time.perf_counter() - t
    model = pz.Model(ctx, seed=5)
    t = time.perf_counter()
    losses = model.train(X, T)
    t_fit = time.perf_counter() - t
    log(f"[bench] C5: catalogue + queries {t_gen:.1f} s, {ntrain} training joins {t_train_joins:.1f} s, "
        f"fit {t_fit:.2f} s, loss {losses[0]:.3f} -> {losses[-1]:.4f}")
    run_library_stream(ctx, lib, held[:64], tpo, max_tris=batch_tris, sizing=pz.SIZING_LEARNED, model=model)  # warm
    t = time.perf_counter()
    exact, s_exact, _ = run_library_stream(ctx, lib, held, tpo, max_tris=batch_tris, sizing=pz.SIZING_TWO_PASS)
    t_exact = time.perf_counter() - t
    t = time.perf_counter()
    got, s_learned, _ = run_library_stream(ctx, lib, held, tpo, max_tris=batch_tris, sizing=pz.SIZING_LEARNED,
                                           model=model)
    t_learned = time.perf_counter() - t
    mism = sum(pair_mismatch(e.pairs, g.pairs) for e, g in zip(exact, got))
    cand = np.array([e.num_candidates for e in exact], np.float64)
    pairs = np.array([len(e.pairs) for e in exact], np.float64)
    pc = np.array([g.predicted for g in got], np.float64)
    pp = np.array([g.predicted_pairs for g in got], np.float64)
    ms_exact = sum(s.ms_total for s in s_exact)
    ms_learned = sum(s.ms_total for s in s_learned)
    # a sample of the smaller queries for the CPU leg's oracle check, from host-built copies of the instances
    small = sorted(range(nqueries), key=lambda q: len(held[q][0]) + len(held[q][1]))[:4]
    for q in small:
        a = instantiate_host(cat, [len(held[q][0])], held[q][0])[0]
        b = instantiate_host(cat, [len(held[q][1])], held[q][1])[0]
        checks.append(("c5", a, b, held[q][2], got[q].pairs))
    tri_q = np.array([tpo[a["mesh"]].sum() + tpo[b["mesh"]].sum() for a, b, _ in held], np.float64)
    res = {"workload": f"{nqueries} queries (set A, set B, predicate uniform); set sizes lognormal(median 500, "
                       "sigma 0.9) clipped to [1, 50000]; per mesh 50% ellipsoids (subdivision 1-4, radii U[1,3], "
                       "random orientation), 30% axis-aligned cubes (edge U[1,10], 12 tris), 20% branching tubes "
                       "(5k-50k tris, lognormal median 15k, sigma 0.6, random orientation); centres U[0,L)^3 with "
                       "L = s (nA + nB)^(1/3), s ~ U[30, 120] per query; sets instanced on the GPU from a catalogue "
                       f"of {cat.num_objects} meshes ({ids['ellipsoids_per_level']} ellipsoids per level, 1 cube, "
                       f"{ids['tubes']} tubes)",
           "queries": nqueries, "train_queries": ntrain, "batches": len(s_exact),
           "median_meshes_per_set": float(np.median([len(q[0]) for q in held])),
           "triangles_per_query_mean": round(float(tri_q.mean()), 1),
           "triangles_total": int(tri_q.sum()),
           "loss_first_epoch": round(float(losses[0]), 4), "loss_last_epoch": round(float(losses[-1]), 4),
           "log2_margins": [round(float(x), 3) for x in model.get()[2]],
           "mismatched_pairs": int(mism), "oracle_sample_queries": len(small),
           "pairs_total": int(pairs.sum()), "candidates_total": int(cand.sum()),
           "median_abs_log2_err_candidates": round(float(np.median(np.abs(np.log2(1 + pc) - np.log2(1 + cand)))), 3),
           "median_abs_log2_err_pairs": round(float(np.median(np.abs(np.log2(1 + pp) - np.log2(1 + pairs)))), 3),
           "query_capacity_covers_exact": round(float(np.mean([g.capacity >= e.num_candidates
                                                               for e, g in zip(exact, got)])), 4),
           "batch_overflow_rate": round(float(np.mean([s.overflowed for s in s_learned])), 4),
           "device_ms_two_pass": round(ms_exact, 3), "device_ms_learned": round(ms_learned, 3),
           "learned_speedup": round(ms_exact / ms_learned, 4),
           "queries_per_s_two_pass": round(nqueries / (ms_exact * 1e-3), 1),
           "queries_per_s_learned": round(nqueries / (ms_learned * 1e-3), 1),
           "gtri_per_s_learned": round(float(tri_q.sum()) / (ms_learned * 1e-3) / 1e9, 2),
           "wall_s_two_pass": round(t_exact, 1), "wall_s_learned": round(t_learned, 1)}
    log(f"[bench] C5: {nqueries} queries in {len(s_exact)} batches: learned {ms_learned:.1f} ms vs two-pass "
        f"{ms_exact:.1f} ms, mismatches {mism}")
    lib.close()
    model.close()
    ctx.trim()
    return res


def cpu_baseline(pz, ctx, A, B, side, chosen):
    """The oracle's Face and TOUCH restatements on the objects of both sets
    whose first vertex lies in [0, side)^3, with the reference's query-time
    windows (overlap_query.cpp:159-179, touch_query.cpp:403-503)."""
    from oracle import oracle as orc

    threads = host_threads()
    orc.set_threads(threads)
    sa, sb = sub_box(pz, A, side), sub_box(pz, B, side)
    log(f"[bench] CPU sample: {sa.num_objects} + {sb.num_objects} objects, "
        f"{sa.indices.shape[0] + sb.indices.shape[0]} triangles, {threads} threads")
    pairs_cpu, us = orc.face_overlap(sa, sb, with_time=True)
    n_touch, us_touch = orc.touch_overlap_count(sa, sb)
    n_cand_s = orc.count_candidate_tests(sa, sb)
    dsa, dsb = ctx.upload(sa), ctx.upload(sb)
    gs = pz.join(ctx, dsa, dsb, **chosen)
    dsa.close()
    dsb.close()
    mism = pair_mismatch(gs.pairs, pairs_cpu)
    cpu = {"value": round(n_cand_s / (us * 1e-6) / 1e9, 6), "unit": "Gtri-tests/s", "cores": threads,
           "kind": "port",
           "sample": (f"Face restatement (oracle/oracle.cpp) on the objects of both sets whose first vertex lies in "
                      f"[0,{side:g})^3: {sa.num_objects}+{sb.num_objects} objects, N_cand={n_cand_s}, "
                      f"{len(pairs_cpu)} pairs, query {us / 1e3:.1f} ms on {threads} threads; TOUCH restatement "
                      f"{us_touch / 1e3:.1f} ms, {n_touch} pairs; GPU chosen path on the same sample: "
                      f"{mism} mismatched pairs"),
           "query_ms": round(us / 1e3, 3), "pairs": int(len(pairs_cpu)),
           "touch_query_ms": round(us_touch / 1e3, 3), "touch_pairs": int(n_touch),
           "touch_value": round(n_cand_s / (us_touch * 1e-6) / 1e9, 6),
           "nproc": os.cpu_count(), "affinity_cpus": threads, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
           "gpu_mismatched_pairs": mism}
    log(f"[bench] CPU sample: Face {us / 1e3:.1f} ms, TOUCH {us_touch / 1e3:.1f} ms ({n_touch} pairs), "
        f"mismatches vs GPU {mism}")
    return cpu, sa, sb


def oracle_checks(pz, checks, other):
    """The CPU leg's checker: the oracle's pairs for every recorded GPU result
    (after all GPU timing), as oracle_mismatched_pairs per configuration."""
    from oracle import oracle as orc

    for name, a, b, pred, got in checks:
        exp = orc.face_overlap(a, b) if pred == pz.PRED_OVERLAP else orc.predicate_join(a, b, pred)[0]
        r = other.setdefault(name, {})
        r["oracle_mismatched_pairs"] = r.get("oracle_mismatched_pairs", 0) + pair_mismatch(got, exp)
        r["oracle_pairs"] = r.get("oracle_pairs", 0) + int(len(exp))
        log(f"[bench] oracle check {name}: {len(exp)} pairs, mismatches {pair_mismatch(got, exp)}")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--objects", type=int, default=100000, help="meshes per set")
    ap.add_argument("--plain-steps", type=int, default=None, help="timed steps of the plain path (default: --steps)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-side", type=float, default=500.0, help="side of the CPU sample sub-box")
    ap.add_argument("--no-learned", action="store_true", help="two-pass sizing only (no estimator training)")
    ap.add_argument("--no-extra", action="store_true", help="skip the other configurations")
    ap.add_argument("--configs", default="c2p,c1,c3,c4,c5",
                    help="other configurations to run (c2p = intersection and containment on the headline sets)")
    ap.add_argument("--train-queries", type=int, default=768, help="queries of the estimator's training stream")
    ap.add_argument("--c4-tubes", type=int, default=2000, help="tubes per set of the C4 workload (2000 as stated)")
    ap.add_argument("--c5-queries", type=int, default=10000, help="queries of the C5 library stream")
    ap.add_argument("--c5-train", type=int, default=1000, help="training queries of the C5 estimator")
    ap.add_argument("--c5-batch-tris", type=int, default=120_000_000, help="triangles per side of a C5 batch")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (scale of the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the join runs on a single device (sharding over several GPUs is "
              "designed, not run); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch

        n = 1 << 28
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return

    import pierce_amd as pz

    configs = set() if args.no_extra else set(c.strip() for c in args.configs.split(",") if c.strip())
    log(f"[bench] generating {args.objects} + {args.objects} meshes")
    t = time.perf_counter()
    A, B = gen_sets(pz, args.objects)
    log(f"[bench] generated in {time.perf_counter() - t:.1f} s: T={A.indices.shape[0]}+{B.indices.shape[0]}")
    counts = dict(V=A.vertices.shape[0] + B.vertices.shape[0], T=A.indices.shape[0] + B.indices.shape[0],
                  E=A.edge_objects.shape[0] + B.edge_objects.shape[0])
    psteps = args.plain_steps or args.steps
    other = {}

    with pz.Context(0) as ctx:
        model, learned = None, None
        if not args.no_learned:
            # Learned sizing: the MLP is trained on the GPU on (features, exact two-pass counts)
            # of a mixed-kind query stream, before and outside every timed region.
            from pierce_amd.stream import make_stream, training_data

            t = time.perf_counter()
            X, T, _ = training_data(ctx, make_stream(args.train_queries, seed0=1000), batch=48)
            t1 = time.perf_counter()
            model = pz.Model(ctx, seed=4)
            losses = model.train(X, T)
            margins = model.get()[2]
            learned = {"train_queries": args.train_queries, "stream_and_exact_joins_s": round(t1 - t, 2),
                       "train_s": round(time.perf_counter() - t1, 3), "epochs": int(len(losses)),
                       "loss_first_epoch": round(float(losses[0]), 4), "loss_last_epoch": round(float(losses[-1]), 4),
                       "log2_margins": [round(float(x), 3) for x in margins]}
            log(f"[bench] estimator trained on {args.train_queries} queries: loss {losses[0]:.3f} -> {losses[-1]:.4f}")
        da, db = ctx.upload(A), ctx.upload(B)
        log("[bench] uploaded; counting candidate tests")
        n_cand = pz.count_candidate_tests(ctx, da, db)
        ctx.trim()
        log(f"[bench] N_cand = {n_cand}")

        # Two-pass sizing: on this workload its count pass costs less than the learned
        # estimate it would replace (both are timed below when a model is trained).
        chosen = dict(strategy=pz.STRATEGY_TWO_LEVEL, precision=pz.PREC_FP32_FILTERED, sizing=pz.SIZING_TWO_PASS)
        plain = dict(strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64, sizing=pz.SIZING_TWO_PASS,
                     dedup=pz.DEDUP_SORT)
        for _ in range(args.warmup):
            pz.join(ctx, da, db, **chosen)
        # autotuner: workgroup size of the LDS index builder per size class, on this workload (untimed)
        tune = pz.autotune(ctx, da, db, **chosen)
        log(f"[bench] autotune: {tune['build_threads']} ({tune['joins_run']} joins)")
        log("[bench] chosen path warm; timing")
        ms_chosen, out = time_join(pz, ctx, da, db, args.steps, **chosen)
        # per-kernel device times from HIP events, in separate steps (timers disable graph replay)
        pz.profile_reset()
        pz.profile_enable(True)
        time_join(pz, ctx, da, db, args.steps, **chosen)
        pz.profile_enable(False)
        kernels = {}
        for k in ("object_stats", "object_bounds_runs", "edge_count", "tri_keys", "edge_scatter", "karras", "refit", "gather_leaves",
                  "radix_scatter", "build_small", "build_chunks", "narrow_dual", "narrow", "parity", "pair_sieve",
                  "escalate_tris", "sieve_parity", "features", "mlp_infer"):
            tot, n = pz.profile_query(k)
            if n:
                kernels[k] = (tot, n)
        log(f"[bench] chosen {ms_chosen:.3f} ms/join, pairs={len(out.pairs)}; kernels "
            + ", ".join(f"{k}={v[0] / args.steps:.2f}ms" for k, v in kernels.items()))

        ms_plain, mism, plain_phases = None, None, None
        if not args.no_plain:
            ctx.trim()
            for _ in range(min(args.warmup, 1)):
                pz.join(ctx, da, db, **plain)
            ms_plain, pout = time_join(pz, ctx, da, db, psteps, **plain)
            mism = pair_mismatch(out.pairs, pout.pairs)
            plain_phases = [(n, round(m, 3)) for n, m in pout.phases]
            log(f"[bench] plain {ms_plain:.3f} ms/join, pairs={len(pout.pairs)}, mismatches={mism}")

        if model:
            # the same chosen path with the other two sizings of its candidate buffer
            lout = None
            for name, alt in (("learned", dict(chosen, sizing=pz.SIZING_LEARNED, model=model)),
                              ("estimate", dict(chosen, sizing=pz.SIZING_ESTIMATE))):
                pz.join(ctx, da, db, **alt)
                ms_alt, aout = time_join(pz, ctx, da, db, args.steps, **alt)
                learned[f"c2_chosen_{name}_ms"] = round(ms_alt, 3)
                learned[f"c2_chosen_{name}_mismatched_pairs"] = pair_mismatch(out.pairs, aout.pairs)
                learned[f"c2_chosen_{name}_overflowed"] = bool(aout.overflowed)
                lout = aout if name == "learned" else lout
            learned["c2_chosen_two_pass_ms"] = round(ms_chosen, 3)
            log(f"[bench] chosen path sizing: two-pass {ms_chosen:.3f} ms, learned {learned['c2_chosen_learned_ms']} ms, "
                f"estimate {learned['c2_chosen_estimate_ms']} ms")
            learned["c2_candidates_predicted"] = lout.predicted
            learned["c2_candidates_exact"] = lout.num_candidates
            learned["c2_capacity"] = lout.capacity
            learned["c2_overflowed"] = lout.overflowed
            learned["c2_pairs_predicted"] = lout.predicted_pairs
            if not args.no_plain:
                # the reference-shaped path with the learned hash-set sizing instead of two passes
                glearn = dict(strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64, sizing=pz.SIZING_LEARNED,
                              dedup=pz.DEDUP_HASH, model=model)
                pz.join(ctx, da, db, **glearn)
                ms_gl, gout = time_join(pz, ctx, da, db, psteps, **glearn)
                learned["global_fp64_learned_hash_ms"] = round(ms_gl, 3)
                learned["global_fp64_two_pass_ms"] = round(ms_plain, 3)
                learned["global_learned_speedup"] = round(ms_plain / ms_gl, 3)
                learned["global_learned_mismatches"] = pair_mismatch(out.pairs, gout.pairs)
                learned["global_learned_overflowed"] = gout.overflowed
                log(f"[bench] global fp64 with learned hash sizing {ms_gl:.3f} ms/join (two-pass {ms_plain:.3f})")
            ctx.trim()

        # CPU baseline: the oracle's Face and TOUCH restatements on a spatial sub-box of the same workload.
        cpu, sa, sb = None, None, None
        if not args.no_cpu:
            cpu, sa, sb = cpu_baseline(pz, ctx, A, B, args.cpu_side, chosen)
        checks = []  # (config, mesh1, mesh2, predicate, GPU pairs) for the CPU leg's oracle
        if "c2p" in configs:
            if sa is None:
                sa, sb = sub_box(pz, A, args.cpu_side), sub_box(pz, B, args.cpu_side)
            other.update(c2_predicates(pz, ctx, da, db, sa, sb, chosen, plain, args.steps, psteps, checks))
        da.close()
        db.close()
        del A, B, sa, sb
        ctx.trim()

        for name, fn in (("c1", lambda: c1_config(pz, ctx, chosen, plain, args.steps, checks)),
                         ("c3", lambda: c3_config(pz, ctx, chosen, plain, args.steps)),
                         ("c4", lambda: c4_config(pz, ctx, chosen, plain, args.c4_tubes, min(args.steps, 5),
                                                  min(psteps, 2))),
                         ("c5", lambda: c5_config(pz, ctx, args.c5_queries, args.c5_train, args.c5_batch_tris,
                                                  checks))):
            if name not in configs:
                continue
            try:
                other[name] = fn()
            except Exception as e:  # keep the headline line even if a side configuration fails
                other[name] = {"error": str(e)}
                log(f"[bench] {name}: error {e}")
                ctx.trim()

        if not args.no_cpu:
            oracle_checks(pz, checks, other)

        # Roofline of the dominant profiled kernel (largest device time in the timed steps).
        iv = int(round(out.index_triangles * counts["V"] / counts["T"]))  # indexed objects' vertices
        sv = int(round(out.sieve_triangles * counts["V"] / counts["T"]))  # sieved pairs' vertices
        counts.update(IT=out.index_triangles, IV=iv, IE=out.index_edges, ST=out.sieve_triangles, SV=sv)
        dom = max(kernels, key=lambda k: kernels[k][0]) if kernels else None
        roofline = None
        if dom:
            tot, n = kernels[dom]
            avg_ms = tot / n
            b = alg_bytes(dom, counts)
            db_ = design_bytes(dom, counts)
            table = counter_table()
            row = table.get(dom, {})
            traffic = float(row["traffic_bytes"]) if row.get("traffic_bytes") not in (None, "", "-") else None
            if b is not None:
                b = b * args.steps / n  # per launch
                ach = b / (avg_ms * 1e-3) / 1e9
                roofline = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic,
                            "alg_bytes_per_launch": int(b),
                            "alg_units": "16 B per triangle + 12 B per vertex of the .pre payload the launch covers",
                            "design_bytes_per_launch": None if db_ is None else int(db_ * args.steps / n),
                            "avg_launch_ms": round(avg_ms, 4), "launches": n,
                            "counters": {k: v for k, v in row.items() if k != "kernel"} or None}
    # the whole join against the HBM roofline: the .pre payload of both sets read once plus the result keys
    join_alg = 16 * counts["T"] + 12 * counts["V"] + 28 * counts["E"] + 8 * len(out.pairs)
    value = n_cand / (ms_chosen * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gtri-tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_chosen, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"overlap join, {args.objects} vs {args.objects} ellipsoid icospheres s=3 "
                               f"(1280 tris), 20% of B scaled copies of A, centres in 1000^3",
                   "objects_per_set": args.objects, "triangles_per_set": counts["T"] // 2,
                   "edges_per_set": counts["E"] // 2, "predicate": "overlap",
                   "chosen": "two-level, fp32-filtered exact, two-pass sizing",
                   "plain": "set-wide BVH, fp64-filtered exact, two-pass count/scan/write/sort/unique"},
        "join_ms_chosen": round(ms_chosen, 4),
        "join_ms_plain": None if ms_plain is None else round(ms_plain, 4),
        "speedup_vs_plain": None if ms_plain is None else round(ms_plain / ms_chosen, 3),
        "mismatched_pairs": mism,
        "pairs": int(len(out.pairs)),
        "n_cand": int(n_cand),
        "pairs_per_s": round(len(out.pairs) / (ms_chosen * 1e-3), 1),
        "join_alg_bytes": int(join_alg),
        "join_roofline_frac": round(join_alg / (ms_chosen * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
        "tests_executed": out.tests_total,
        "escalated_fp64": out.tests_fp64,
        "escalated_exact": out.tests_exact,
        "fp32_only_misclassified": out.tests_fp32_misclassified,
        "graph_launches_per_join": out.graph_launches,
        "narrow_rounds": out.narrow_rounds,
        "narrow_tasks_spilled": out.tasks_spilled,
        "narrow_fallback_pairs": out.narrow_fallback_pairs,
        "index_height": out.index_height,
        "phases_chosen": [(n, round(m, 3)) for n, m in out.phases],
        "phases_plain": plain_phases,
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kernels.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_sizing": learned,
        "autotune": tune,
        "other_configs": other,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
