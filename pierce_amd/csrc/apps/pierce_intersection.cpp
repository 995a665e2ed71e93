// pierce_intersection: surface overlap or one solid containing the other
// (pierce/query/src/applications/pierce_intersection_main.cpp:541-978).
// The pairs CSV is always written (default intersection_pairs.csv).  With
// --enable-tracking the containment rays' per-source tracking is written too,
// in the two CSV layouts of pierce_intersection_main.cpp:425-539.
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <vector>

#include "app_common.hpp"

namespace {

constexpr int kTrackedTargets = 256;  // PairHitTracking.h:31

// One direction's tracking over its source objects.
struct Tracking {
    std::vector<int32_t> ids;
    std::vector<uint32_t> hits, candidates, overflow, iterations;
    void init(uint64_t n) {
        ids.assign(n * kTrackedTargets, -1);
        hits.assign(n * kTrackedTargets, 0u);
        candidates.assign(n, 0u);
        overflow.assign(n, 0u);
        iterations.assign(n, 0u);
    }
    bool run(pz_ctx* ctx, const pz_dataset* src, const pz_dataset* tgt) {
        return pz_track_containment_rays(ctx, src, tgt, kTrackedTargets, ids.data(), hits.data(), candidates.data(),
                                         overflow.data(), iterations.data()) == PZ_OK;
    }
};

// Is (mesh1 a, mesh2 b) among the result pairs (sorted by (mesh1, mesh2))?
bool final_pair(const pz_join_result& r, int32_t a, int32_t b) {
    const pz_pair* b0 = r.pairs;
    const pz_pair* e = b0 + r.num_pairs;
    const pz_pair* it = std::lower_bound(b0, e, std::make_pair(a, b), [](const pz_pair& x, std::pair<int32_t, int32_t> k) {
        return x.object_id_mesh1 != k.first ? x.object_id_mesh1 < k.first : x.object_id_mesh2 < k.second;
    });
    return it != e && it->object_id_mesh1 == a && it->object_id_mesh2 == b;
}

// direction,source_object_id,target_object_id,target_ray_hits,final_pair
bool write_pair_hits(const std::string& path, const Tracking& t12, const Tracking& t21, const pz_join_result& r) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::cerr << "Error: Failed to open intersection tracking output: " << path << std::endl;
        return false;
    }
    std::fprintf(f, "direction,source_object_id,target_object_id,target_ray_hits,final_pair\n");
    for (int d = 0; d < 2; ++d) {
        const Tracking& t = d == 0 ? t12 : t21;
        const uint64_t n = t.candidates.size();
        for (uint64_t s = 0; s < n; ++s)
            for (int i = 0; i < kTrackedTargets; ++i) {
                const int32_t tgt = t.ids[s * kTrackedTargets + i];
                if (tgt < 0) continue;
                const bool fin = d == 0 ? final_pair(r, (int32_t)s, tgt) : final_pair(r, tgt, (int32_t)s);
                std::fprintf(f, "%s,%llu,%d,%u,%d\n", d == 0 ? "mesh1_to_mesh2" : "mesh2_to_mesh1",
                             (unsigned long long)s, tgt, t.hits[s * kTrackedTargets + i], fin ? 1 : 0);
            }
    }
    std::fclose(f);
    return true;
}

// direction,source_object_id,iterations,candidate_count,tracked_target_count,
// tracked_hit_total,candidate_overflow_events,final_pair_count,tracked_final_pair_count
bool write_summary(const std::string& path, const Tracking& t12, const Tracking& t21, const pz_join_result& r) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::cerr << "Error: Failed to open containment tracking output: " << path << std::endl;
        return false;
    }
    std::fprintf(f, "direction,source_object_id,iterations,candidate_count,tracked_target_count,tracked_hit_total,"
                    "candidate_overflow_events,final_pair_count,tracked_final_pair_count\n");
    for (int d = 0; d < 2; ++d) {
        const Tracking& t = d == 0 ? t12 : t21;
        const uint64_t n = t.candidates.size();
        for (uint64_t s = 0; s < n; ++s) {
            int tracked = 0, fin = 0;
            unsigned long long total = 0;
            for (int i = 0; i < kTrackedTargets; ++i) {
                const int32_t tgt = t.ids[s * kTrackedTargets + i];
                if (tgt < 0) continue;
                ++tracked;
                total += t.hits[s * kTrackedTargets + i];
                fin += (d == 0 ? final_pair(r, (int32_t)s, tgt) : final_pair(r, tgt, (int32_t)s)) ? 1 : 0;
            }
            std::fprintf(f, "%s,%llu,%u,%u,%d,%llu,%u,%d,%d\n", d == 0 ? "mesh1_to_mesh2" : "mesh2_to_mesh1",
                         (unsigned long long)s, t.iterations[s], t.candidates[s], tracked, total, t.overflow[s], fin,
                         fin);
        }
    }
    std::fclose(f);
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    app::Timer timer;
    timer.start("Data Reading");
    app::Common c;
    c.output = "estimated_intersection_timing.json";
    float gamma = 0.8f, epsilon = 0.001f, load_factor = 0.5f;
    std::string direction = "both", pairs_out = "intersection_pairs.csv";
    std::string hits_out = "intersection_pair_hits.csv", tracking_out = "intersection_containment_tracking.csv";
    bool estimate_only = false, tracking = false, profiling = false;
    int max_iter = 100;
    app::parse(argc, argv, c, false, [&](const std::string& a, int& i, int argc_, char** argv_) {
        auto has = [&]() { return i + 1 < argc_; };
        if (a == "--gamma" && has()) gamma = std::stof(argv_[++i]);
        else if (a == "--epsilon" && has()) epsilon = std::stof(argv_[++i]);
        else if (a == "--estimate-only") estimate_only = true;
        else if (a == "--query-direction" && has()) direction = argv_[++i];
        else if (a == "--overlap-max-iterations" && has()) max_iter = std::atoi(argv_[++i]);
        else if (a == "--hash-load-factor" && has()) load_factor = std::stof(argv_[++i]);
        else if (a == "--enable-tracking") tracking = true;
        else if (a == "--enable-profiling-stats") profiling = true;
        else if (a == "--pairs-output" && has()) pairs_out = argv_[++i];
        else if (a == "--pair-hits-output" && has()) hits_out = argv_[++i];
        else if (a == "--containment-tracking-output" && has()) tracking_out = argv_[++i];
        else return false;
        return true;
    });
    if (c.help) {
        auto o = app::common_help(false);
        o.insert(o.end() - 1,
                 {{"--gamma <float>", "Estimation gamma (default: 0.8)"},
                  {"--epsilon <float>", "Estimation epsilon (default: 0.001)"},
                  {"--estimate-only", "Run only selectivity estimation"},
                  {"--query-direction <both|mesh1_to_mesh2|mesh2_to_mesh1>", "Control query direction (default: both)"},
                  {"--overlap-max-iterations <int>", "Accepted (> 0); every hit is enumerated exactly"},
                  {"--hash-load-factor <float>", "Load factor in (0,1] for sized buffers (default: 0.5)"},
                  {"--enable-tracking", "Enable optional containment candidate tracking and CSV exports"},
                  {"--enable-profiling-stats", "Write device test counters into the timing JSON"},
                  {"--pairs-output <path>", "Intersection pairs CSV path (default: intersection_pairs.csv)"},
                  {"--pair-hits-output <path>", "Pair hit tracking CSV path (default: intersection_pair_hits.csv)"},
                  {"--containment-tracking-output <path>",
                   "Containment tracking CSV path (default: intersection_containment_tracking.csv)"}});
        app::print_help(argv[0], "--mesh1 <path> --mesh2 <path> [options]",
                        "Intersection query: overlapping surfaces or one object inside the other.", o);
        return 0;
    }
    int dir = direction == "both" ? PZ_DIR_BOTH
              : direction == "mesh1_to_mesh2" ? PZ_DIR_1TO2
              : direction == "mesh2_to_mesh1" ? PZ_DIR_2TO1
                                              : -1;
    if (dir < 0) {
        std::cerr << "Invalid query direction: " << direction << std::endl;
        std::cerr << "Valid values for --query-direction are: both, mesh1_to_mesh2, mesh2_to_mesh1" << std::endl;
        return 1;
    }
    if (!(load_factor > 0.f && load_factor <= 1.f)) {
        std::cerr << "Invalid --hash-load-factor. Expected value in (0, 1]." << std::endl;
        return 1;
    }
    if (max_iter <= 0) {
        std::cerr << "--overlap-max-iterations must be > 0." << std::endl;
        return 1;
    }
    if (c.mesh1.empty() || c.mesh2.empty()) {
        std::cerr << "Usage: " << argv[0] << " --mesh1 <path> --mesh2 <path> [options]" << std::endl;
        return 1;
    }
    timer.next("Application Creation");
    pz_ctx* ctx = nullptr;
    if (pz_ctx_create(c.device, &ctx) != PZ_OK) return app::fail_join("device initialisation");
    uint64_t nt1 = 0, nt2 = 0, no1 = 0, no2 = 0;
    pz_dataset* d1 = app::load(ctx, c.mesh1, "Mesh1", timer, &nt1, &no1);
    if (!d1) {
        std::cerr << "Error loading mesh1." << std::endl;
        return 1;
    }
    pz_dataset* d2 = app::load(ctx, c.mesh2, "Mesh2", timer, &nt2, &no2);
    if (!d2) {
        std::cerr << "Error loading mesh2." << std::endl;
        return 1;
    }
    pz_join_desc d;
    pz_join_desc_defaults(&d);
    d.predicate = PZ_PRED_INTERSECTION;
    d.direction = dir;
    d.precision = c.precision;
    d.gamma = gamma;
    d.epsilon = epsilon;
    d.load_factor = load_factor;
    d.sizing = c.sizing_override >= 0 && c.sizing_override != PZ_SIZING_LEARNED ? c.sizing_override
                                                                                 : PZ_SIZING_ESTIMATE;
    if (c.verbose_estimate)
        app::print_estimate("=== Selectivity Estimation ===", "==============================", d1, d2, epsilon, gamma,
                            true, true, false);
    pz_join_result r;
    if (estimate_only) {
        d.sizing = PZ_SIZING_ESTIMATE;
        if (pz_join(ctx, d1, d2, &d, &r) != PZ_OK) return app::fail_join("estimate");
        std::cout << "Final Estimated Pairs:     " << r.predicted_pairs << std::endl;
        pz_join_result_release(&r);
        timer.finish(c.output);
        return 0;
    }
    std::cout << "\n=== Query Configuration ===" << std::endl;
    std::cout << "Query Direction:    " << direction << std::endl;
    std::cout << "Tracking:           " << (tracking ? "enabled" : "disabled") << std::endl;
    std::cout << "Profiling Stats:    " << (profiling ? "enabled" : "disabled") << std::endl;
    std::cout << "===========================\n" << std::endl;
    timer.next("Warmup");
    if (c.warmup > 0) std::cout << "Running " << c.warmup << " warmup iterations (hash query)..." << std::endl;
    for (int w = 0; w < c.warmup; ++w) {
        if (pz_join(ctx, d1, d2, &d, &r) != PZ_OK) return app::fail_join("warmup query");
        pz_join_result_release(&r);
    }
    timer.next("Query");
    std::cout << "Running Intersection Query..." << std::endl;
    for (int run = 0; run < c.runs; ++run) {
        if (run) pz_join_result_release(&r);
        if (pz_join(ctx, d1, d2, &d, &r) != PZ_OK) return app::fail_join("query");
    }
    timer.next("Output");
    app::add_phases(timer, r);
    // query-time keys of predicates/adapters/intersection.py:207-221
    timer.add("Raytrace_Overlap_Hash_Mesh1ToMesh2", app::uncovered_ms(r, false, {"Parity"}));
    timer.add("Raytrace_Containment_Hash_Mesh1ToMesh2", app::phase_ms(r, {"Parity"}));
    std::cout << "Mesh1 objects: " << no1 << std::endl;
    std::cout << "Mesh2 objects: " << no2 << std::endl;
    std::cout << "Final Estimated Pairs:     " << r.predicted_pairs << std::endl;
    std::cout << "Actual Intersection Pairs: " << r.num_pairs << std::endl;
    app::print_join_time(r);
    if (profiling) {
        // The counters of pierce_intersection_main.cpp:924-948
        // (MeshIntersectionProfilingStats, mesh_intersection.h:7-22), from this
        // engine's walk counters.  The reference's iterations are repeated
        // closest-hit traces per edge; here a walk tests segment-triangle pairs
        // until the first proven hit, and a parity ray is one traversal.  Caps
        // and same-hit suppression do not exist here (0).
        const unsigned long long ov_tests = r.tests_total - r.parity_tests;
        const unsigned long long walks = r.walks_started, rays = r.parity_rays;
        timer.counter("Profile_Overlap_Trace_Calls", walks);
        timer.counter("Profile_Overlap_Iterations_Total", ov_tests);
        timer.counter("Profile_Overlap_Hits_Total", r.pairs_touching);
        timer.counter("Profile_Overlap_Max_Iterations_Per_Trace", 0);
        timer.counter("Profile_Containment_Rays_Total", rays);
        timer.counter("Profile_Containment_Iterations_Total", rays);
        timer.counter("Profile_Containment_Hits_Total", r.parity_crossings);
        timer.counter("Profile_Containment_Max_Iterations_Per_Ray", rays ? 1 : 0);
        timer.counter("Profile_Containment_Same_Hit_Suppressed", 0);
        timer.counter("Profile_Containment_Candidate_Additions", r.parity_rays_crossing);
        timer.counter("Profile_Containment_Candidate_Toggles", r.parity_crossings - r.parity_rays_crossing);
        timer.counter("Profile_Containment_Candidate_Overflow", 0);
        timer.counter("Profile_Containment_Targets_Total", r.parity_inside);
        timer.counter("Profile_Overlap_Avg_Iterations_x1000", walks ? ov_tests * 1000ull / walks : 0ull);
        timer.counter("Profile_Containment_Avg_Iterations_x1000", rays ? 1000ull : 0ull);
        timer.counter("tests_total", r.tests_total);
        timer.counter("tests_fp64", r.tests_fp64);
        timer.counter("tests_exact", r.tests_exact);
        timer.counter("tests_fp32_misclassified", r.tests_fp32_misclassified);
        timer.counter("edges_traced", r.edges_traced);
        timer.counter("candidate_pairs", r.num_candidates);
    }
    app::write_pairs_csv(pairs_out, "a_object_id,b_object_id", r);
    if (tracking) {
        // directions that were not queried keep empty tracking (no rows in the
        // pair-hit CSV, zero rows in the summary), as the reference's buffers do
        Tracking t12, t21;
        t12.init(no1);
        t21.init(no2);
        if (dir != PZ_DIR_2TO1 && !t12.run(ctx, d1, d2)) return app::fail_join("tracking mesh1_to_mesh2");
        if (dir != PZ_DIR_1TO2 && !t21.run(ctx, d2, d1)) return app::fail_join("tracking mesh2_to_mesh1");
        if (!write_pair_hits(hits_out, t12, t21, r) || !write_summary(tracking_out, t12, t21, r)) return 1;
    }
    std::cout << "Intersection pairs CSV: " << pairs_out << std::endl;
    if (tracking) {
        std::cout << "Pair hit tracking CSV: " << hits_out << std::endl;
        std::cout << "Containment tracking CSV: " << tracking_out << std::endl;
    }
    pz_join_result_release(&r);
    pz_dataset_release(d1);
    pz_dataset_release(d2);
    pz_ctx_destroy(ctx);
    timer.finish(c.output);
    return 0;
}
