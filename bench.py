#!/usr/bin/env python3
"""Join throughput of the MI355X spatial-join engine on one GPU.

Workload (default): A = 100,000 ellipsoid icospheres at subdivision 3 (1,280
triangles each), radii U[0.5, 2], centres U[0, 1000)^3; B = 100,000 meshes of
which 20% are copies of random A meshes scaled by U[0.3, 0.8] about their
centroid and 80% fresh ellipsoids; fp32 vertices, seeded (A 42, B 43).

One step = one overlap join, device-resident geometry -> sorted unique host
pair set, including every index build.  Two paths run on the same GPU:
  chosen : two-level join (object broad phase + per-object BVHs, early exit),
           fp32-filtered exact predicates;
  plain  : one BVH per set over all triangles, every edge traced, fp64-filtered
           exact predicates throughout, exact two-pass count/scan/write/sort/unique.
Both are checked against each other pair for pair; the CPU oracle (Face
restatement) is timed and checked on a spatial sub-box of the same workload.

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

# Algorithmic bytes of each profiled kernel in one join (all its launches: one
# per set, or one per size class and set), from the .pre payload units (12 B
# per vertex, 12 + 4 B per triangle, 28 B per edge) and the index sizes the
# join reports.  Keys: kernel name -> f(counts) -> bytes per join.
def alg_bytes(kernel, k):
    if kernel == "object_stats":    # indices + map per triangle, each vertex once
        return 16 * k["T"] + 12 * k["V"]
    if kernel == "edge_count":      # source-object ids
        return 4 * k["E"]
    if kernel == "tri_keys":        # payload read + (key, triangle) written for indexed triangles
        return 16 * k["T"] + 12 * k["V"] + 12 * k["IT"]
    if kernel == "edge_scatter":    # ids of all edges + endpoints read and 32-B records written for indexed edges
        return 4 * k["E"] + 24 * k["IE"] + 32 * k["IE"]
    if kernel == "refit":           # per indexed triangle: order, key, indices, 3 vertices read; one node written
        return (4 + 8 + 12 + 36 + 64) * k["IT"]
    if kernel == "gather_leaves":   # per indexed triangle: order, indices, 3 vertices, map read; 48 B written
        return (4 + 12 + 36 + 4 + 48) * k["IT"]
    if kernel == "karras":          # keys read (~neighbour loads), node children + parents written
        return (8 + 16 + 8) * k["IT"]
    if kernel == "build_small":     # per indexed triangle: indices + 3 vertices read, one node + one leaf written
        return (12 + 36 + 64 + 48) * k["IT"]
    if kernel == "narrow_dual":     # every indexed triangle (48 B) and internal node (64 B) of both objects read once
        return (48 + 64) * k["IT"]
    if kernel == "narrow":          # every indexed edge record read once
        return 32 * k["IE"]
    return None


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
# workload, scaled by a 1 GiB copy (fetch x 2.0, write x 1.0); kernel -> bytes.
MEASURED_TRAFFIC = {
    "build_small": 1.709e9 + 2.877e9,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gen_sets(pz, n, with_grid=False):
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


def stream_sizing(pz, ctx, model, n, batch=48):
    """Held-out mixed-kind query stream: learned vs two-pass sizing of the batched join."""
    from pierce_amd.stream import make_stream, run_stream

    held = make_stream(n, seed0=50000)
    run_stream(ctx, held[:batch], batch=batch, sizing=pz.SIZING_LEARNED, model=model)  # warm-up
    exact, s_exact, _ = run_stream(ctx, held, batch=batch, sizing=pz.SIZING_TWO_PASS)
    got, s_learned, _ = run_stream(ctx, held, batch=batch, sizing=pz.SIZING_LEARNED, model=model)
    mism = sum(pair_mismatch(e.pairs, g.pairs) for e, g in zip(exact, got))
    cand = np.array([e.num_candidates for e in exact], np.float64)
    pairs = np.array([len(e.pairs) for e in exact], np.float64)
    pc = np.array([g.predicted for g in got], np.float64)
    pp = np.array([g.predicted_pairs for g in got], np.float64)
    ms_exact = sum(s.ms_total for s in s_exact)
    ms_learned = sum(s.ms_total for s in s_learned)
    res = {"queries": n, "batch": batch,
           "workload": "per query: set sizes lognormal(median 40, sigma 0.9) clipped to [2, 400]; 50% ellipsoids "
                       "(subdivision 1-4), 30% cubes, 20% tubes (500-4000 tris); predicate uniform; density and "
                       "B offset drawn per query",
           "mismatched_pairs": int(mism),
           "median_abs_log2_err_candidates": round(float(np.median(np.abs(np.log2(1 + pc) - np.log2(1 + cand)))), 3),
           "median_abs_log2_err_pairs": round(float(np.median(np.abs(np.log2(1 + pp) - np.log2(1 + pairs)))), 3),
           "query_capacity_covers_exact": round(float(np.mean([g.capacity >= e.num_candidates
                                                               for e, g in zip(exact, got)])), 4),
           "batch_overflow_rate": round(float(np.mean([s.overflowed for s in s_learned])), 4),
           "device_ms_two_pass": round(ms_exact, 3), "device_ms_learned": round(ms_learned, 3),
           "queries_per_s_two_pass": round(n / (ms_exact * 1e-3), 1),
           "queries_per_s_learned": round(n / (ms_learned * 1e-3), 1)}
    log(f"[bench] stream of {n}: learned {ms_learned:.1f} ms vs two-pass {ms_exact:.1f} ms, mismatches {mism}")
    return res


def extra_configs(pz, ctx, model, steps):
    """The other configurations: chosen path vs the plain GPU path on the same GPU.

    C1 (1000 vs 1000 ellipsoids, s=2) is also checked against the CPU oracle in full.
    C3 (10k vs 10k spheres, s in 1..6) runs its three predicates as one batch.
    C4 is reduced to 200 vs 200 tubes (same triangle-count law, density kept).
    Each configuration is autotuned on its own workload before it is timed."""
    from oracle import oracle as orc

    learned = dict(sizing=pz.SIZING_LEARNED, model=model) if model else dict(sizing=pz.SIZING_TWO_PASS)
    chosen = dict(strategy=pz.STRATEGY_TWO_LEVEL, precision=pz.PREC_FP32_FILTERED, **learned)
    plain = dict(strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64, sizing=pz.SIZING_TWO_PASS, dedup=pz.DEDUP_SORT)
    plain_tl = dict(strategy=pz.STRATEGY_TWO_LEVEL, precision=pz.PREC_FP64, sizing=pz.SIZING_TWO_PASS)
    res = {}

    def single(name, a, b, oracle=False, plain_steps=1):
        da, db = ctx.upload(a), ctx.upload(b)
        tune = pz.autotune(ctx, da, db, **chosen)  # on this workload, untimed
        pz.join(ctx, da, db, **chosen)
        ms_c, out = time_join(pz, ctx, da, db, steps, **chosen)
        pz.join(ctx, da, db, **plain)
        ms_p, pout = time_join(pz, ctx, da, db, plain_steps, **plain)
        r = {"objects": [a.num_objects, b.num_objects], "triangles": [int(a.indices.shape[0]), int(b.indices.shape[0])],
             "join_ms_chosen": round(ms_c, 3), "join_ms_plain": round(ms_p, 3), "speedup_vs_plain": round(ms_p / ms_c, 2),
             "mismatched_pairs": pair_mismatch(out.pairs, pout.pairs), "pairs": int(len(out.pairs)),
             "tests": out.tests_total, "escalated_fp64": out.tests_fp64, "escalated_exact": out.tests_exact,
             "fp32_only_misclassified": out.tests_fp32_misclassified, "sizing_overflowed": out.overflowed,
             "narrow_rounds": out.narrow_rounds, "autotune": {k: tune[k] for k in ("build_threads", "narrow_queue",
                                                                                  "narrow_chunk")}}
        if oracle:
            exp = orc.face_overlap(a, b)
            r["oracle_mismatched_pairs"] = pair_mismatch(out.pairs, exp)
        da.close()
        db.close()
        log(f"[bench] {name}: chosen {ms_c:.3f} ms, plain {ms_p:.3f} ms, mismatches {r['mismatched_pairs']}")
        return r

    try:
        a = pz.generate(pz.GEN_ELLIPSOIDS, count=1000, seed=42, box=100.0, jitter=0.02)
        b = pz.generate(pz.GEN_ELLIPSOIDS, count=1000, seed=43, box=100.0, jitter=0.02)
        res["c1"] = dict(workload="overlap, 1000 vs 1000 ellipsoid icospheres s=2 (320 tris), radii U[1,3], "
                                  "centres in 100^3, jitter 0.02 r", **single("C1", a, b, oracle=True, plain_steps=steps))
    except Exception as e:  # keep the headline line even if a side configuration fails
        res["c1"] = {"error": str(e)}
    try:
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
        da, db = ctx.upload(a), ctx.upload(b)
        for d in (plain, plain_tl):
            pz.join(ctx, da, db, **d)  # warm
        t0 = time.perf_counter()
        pouts = [pz.join(ctx, da, db, **plain)]
        pouts += [pz.join(ctx, da, db, predicate=p, **plain_tl) for p in preds[1:]]
        ms_p = (time.perf_counter() - t0) * 1e3
        mism = sum(pair_mismatch(o.pairs, po.pairs) for o, po in zip(outs, pouts))
        res["c3"] = {"workload": "overlap + intersection + containment, 10000 vs 10000 spheres, subdivision 1-6 with "
                                 "P ~ 2^-l (80-81920 tris), radius U[1,4], centres in 500^3",
                     "chosen": "the three predicates as one batched join (fp32-filtered, learned sizing)",
                     "plain": "one join per launch: overlap on the set-wide path, intersection and containment on "
                              "the two-level path, fp64 filter, two-pass sizing",
                     "objects": [a.num_objects, b.num_objects],
                     "triangles": [int(a.indices.shape[0]), int(b.indices.shape[0])],
                     "join_ms_chosen": round(ms_c, 3), "join_ms_plain": round(ms_p, 3),
                     "speedup_vs_plain": round(ms_p / ms_c, 2), "mismatched_pairs": int(mism),
                     "pairs": [int(len(o.pairs)) for o in outs], "tests": summ.tests_total,
                     "escalated_fp64": summ.tests_fp64, "escalated_exact": summ.tests_exact,
                     "fp32_only_misclassified": summ.tests_fp32_misclassified, "narrow_rounds": summ.narrow_rounds,
                     "autotune": {k: tune[k] for k in ("build_threads", "narrow_queue", "narrow_chunk")}}
        log(f"[bench] C3: batch {ms_c:.3f} ms, plain {ms_p:.3f} ms, mismatches {mism}")
        for d in (da3, db3, da, db):
            d.close()
    except Exception as e:
        res["c3"] = {"error": str(e)}
    try:
        kw = dict(box=464.0)
        a = pz.generate(pz.GEN_TUBES, count=200, seed=42, **kw)
        b = pz.generate(pz.GEN_TUBES, count=200, seed=43, degenerate_fraction=0.01, copy_source_count=200,
                        copy_source_seed=42, **kw)
        res["c4_reduced"] = dict(workload="overlap, 200 vs 200 branching tubes (lognormal median 200k tris, "
                                          "clipped [50k, 2M]), centres in 464^3 (the 2000-mesh density), 1% "
                                          "near-degenerate copies", **single("C4 (reduced)", a, b))
    except Exception as e:
        res["c4_reduced"] = {"error": str(e)}
    return res


def pair_mismatch(p, q):
    a = set(map(tuple, p.tolist()))
    b = set(map(tuple, q.tolist()))
    return len(a ^ b)


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
    ap.add_argument("--no-extra", action="store_true", help="skip the C1 / C3 / reduced C4 comparison lines")
    ap.add_argument("--train-queries", type=int, default=768, help="queries of the estimator's training stream")
    ap.add_argument("--stream-queries", type=int, default=480, help="held-out queries of the sizing stream (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (scale of the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the join runs on a single device (multi-GPU sharding is described "
              "in DESIGN.md, not run); use --gpus 1", file=sys.stderr)
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

    log(f"[bench] generating {args.objects} + {args.objects} meshes")
    t = time.perf_counter()
    A, B = gen_sets(pz, args.objects)
    log(f"[bench] generated in {time.perf_counter() - t:.1f} s: T={A.indices.shape[0]}+{B.indices.shape[0]}")
    counts = dict(V=A.vertices.shape[0] + B.vertices.shape[0], T=A.indices.shape[0] + B.indices.shape[0],
                  E=A.edge_objects.shape[0] + B.edge_objects.shape[0])

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
        log(f"[bench] N_cand = {n_cand}")

        chosen = dict(strategy=pz.STRATEGY_TWO_LEVEL, precision=pz.PREC_FP32_FILTERED,
                      sizing=pz.SIZING_LEARNED if model else pz.SIZING_TWO_PASS)
        if model:
            chosen["model"] = model
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
        for k in ("object_stats", "edge_count", "tri_keys", "edge_scatter", "karras", "refit", "gather_leaves",
                  "radix_scatter", "build_small", "narrow_dual", "narrow", "parity", "features", "mlp_infer"):
            tot, n = pz.profile_query(k)
            if n:
                kernels[k] = (tot, n)
        log(f"[bench] chosen {ms_chosen:.3f} ms/join, pairs={len(out.pairs)}; kernels "
            + ", ".join(f"{k}={v[0] / args.steps:.2f}ms" for k, v in kernels.items()))

        ms_plain, mism, plain_phases = None, None, None
        if not args.no_plain:
            psteps = args.plain_steps or args.steps
            for _ in range(min(args.warmup, 1)):
                pz.join(ctx, da, db, **plain)
            ms_plain, pout = time_join(pz, ctx, da, db, psteps, **plain)
            mism = pair_mismatch(out.pairs, pout.pairs)
            plain_phases = [(n, round(m, 3)) for n, m in pout.phases]
            log(f"[bench] plain {ms_plain:.3f} ms/join, pairs={len(pout.pairs)}, mismatches={mism}")

        if model:
            learned["c2_candidates_predicted"] = out.predicted
            learned["c2_candidates_exact"] = out.num_candidates
            learned["c2_capacity"] = out.capacity
            learned["c2_overflowed"] = out.overflowed
            learned["c2_pairs_predicted"] = out.predicted_pairs
            if not args.no_plain:
                # the reference-shaped path with the learned hash-set sizing instead of two passes
                glearn = dict(strategy=pz.STRATEGY_GLOBAL, precision=pz.PREC_FP64, sizing=pz.SIZING_LEARNED,
                              dedup=pz.DEDUP_HASH, model=model)
                pz.join(ctx, da, db, **glearn)
                ms_gl, gout = time_join(pz, ctx, da, db, args.plain_steps or args.steps, **glearn)
                learned["global_fp64_learned_hash_ms"] = round(ms_gl, 3)
                learned["global_fp64_two_pass_ms"] = round(ms_plain, 3)
                learned["global_learned_speedup"] = round(ms_plain / ms_gl, 3)
                learned["global_learned_mismatches"] = pair_mismatch(out.pairs, gout.pairs)
                learned["global_learned_overflowed"] = gout.overflowed
                log(f"[bench] global fp64 with learned hash sizing {ms_gl:.3f} ms/join (two-pass {ms_plain:.3f})")
            if args.stream_queries > 0:
                learned["stream"] = stream_sizing(pz, ctx, model, args.stream_queries)
        other = None if args.no_extra else extra_configs(pz, ctx, model, args.steps)

        # Roofline of the dominant profiled kernel (largest device time in the timed steps).
        counts.update(IT=out.index_triangles, IE=out.index_edges)
        dom = max(kernels, key=lambda k: kernels[k][0]) if kernels else None
        roofline = None
        if dom:
            tot, n = kernels[dom]
            avg_ms = tot / n
            b = alg_bytes(dom, counts)
            if b is not None:
                b = b * args.steps / n  # per launch
                ach = b / (avg_ms * 1e-3) / 1e9
                roofline = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                            "traffic": MEASURED_TRAFFIC.get(dom),
                            "alg_bytes_per_launch": int(b), "avg_launch_ms": round(avg_ms, 4), "launches": n}

        # CPU baseline: the oracle's Face restatement on a spatial sub-box of the same workload.
        cpu = None
        sample_mism = None
        if not args.no_cpu:
            from oracle import oracle as orc

            threads = min(16, os.cpu_count() or 1)
            orc.set_threads(threads)
            sa, sb = sub_box(pz, A, args.cpu_side), sub_box(pz, B, args.cpu_side)
            log(f"[bench] CPU sample: {sa.num_objects} + {sb.num_objects} objects, "
                f"{sa.indices.shape[0] + sb.indices.shape[0]} triangles, {threads} threads")
            pairs_cpu, us = orc.face_overlap(sa, sb, with_time=True)
            n_cand_s = orc.count_candidate_tests(sa, sb)
            dsa, dsb = ctx.upload(sa), ctx.upload(sb)
            gs = pz.join(ctx, dsa, dsb, **chosen)
            sample_mism = pair_mismatch(gs.pairs, pairs_cpu)
            cpu = {"value": round(n_cand_s / (us * 1e-6) / 1e9, 6), "unit": "Gtri-tests/s", "cores": threads,
                   "kind": "port",
                   "sample": (f"Face restatement (oracle/oracle.cpp) on the objects of both sets whose first vertex "
                              f"lies in [0,{args.cpu_side:g})^3: {sa.num_objects}+{sb.num_objects} objects, "
                              f"N_cand={n_cand_s}, {len(pairs_cpu)} pairs, query {us / 1e3:.1f} ms; "
                              f"GPU chosen path on the same sample: {sample_mism} mismatched pairs"),
                   "query_ms": round(us / 1e3, 3), "pairs": int(len(pairs_cpu))}
            log(f"[bench] CPU sample query {us / 1e3:.1f} ms, mismatches vs GPU {sample_mism}")

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
                   "chosen": "two-level, fp32-filtered exact, " + (
                       "learned sizing (bf16 MLP) with two-pass fallback" if model else "two-pass sizing"),
                   "plain": "set-wide BVH, fp64-filtered exact, two-pass count/scan/write/sort/unique"},
        "join_ms_chosen": round(ms_chosen, 4),
        "join_ms_plain": None if ms_plain is None else round(ms_plain, 4),
        "speedup_vs_plain": None if ms_plain is None else round(ms_plain / ms_chosen, 3),
        "mismatched_pairs": mism,
        "pairs": int(len(out.pairs)),
        "n_cand": int(n_cand),
        "pairs_per_s": round(len(out.pairs) / (ms_chosen * 1e-3), 1),
        "tests_executed": out.tests_total,
        "escalated_fp64": out.tests_fp64,
        "escalated_exact": out.tests_exact,
        "fp32_only_misclassified": out.tests_fp32_misclassified,
        "graph_launches_per_join": out.graph_launches,
        "narrow_rounds": out.narrow_rounds,
        "narrow_tasks_spilled": out.tasks_spilled,
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
