// pierce_containment: mesh2 (B) objects strictly inside mesh1 (A) objects
// with no surface contact (pierce/query/src/applications/pierce_containment_main.cpp:236-644).
#include <cstdio>
#include <iostream>

#include "app_common.hpp"

int main(int argc, char** argv) {
    app::Timer timer;
    timer.start("Data Reading");
    app::Common c;
    c.output = "containment_timing.json";
    bool include_overlap = false;
    float gamma = 0.8f, epsilon = 0.001f, load_factor = 0.5f;
    int max_iter = 256;
    app::parse(argc, argv, c, true, [&](const std::string& a, int& i, int argc_, char** argv_) {
        auto has = [&]() { return i + 1 < argc_; };
        if (a == "--include-overlap-pairs") include_overlap = true;
        else if (a == "--overlap-max-iterations" && has()) max_iter = std::atoi(argv_[++i]);
        else if (a == "--gamma" && has()) gamma = std::stof(argv_[++i]);
        else if (a == "--epsilon" && has()) epsilon = std::stof(argv_[++i]);
        else if (a == "--hash-load-factor" && has()) load_factor = std::stof(argv_[++i]);
        else return false;
        return true;
    });
    (void)max_iter;
    if (c.help) {
        auto o = app::common_help(true);
        o.insert(o.end() - 1,
                 {{"--include-overlap-pairs", "Include overlap/touch pairs in output (union of overlap + strict containment)"},
                  {"--overlap-max-iterations <int>", "Accepted; every hit is enumerated exactly"},
                  {"--gamma <float>", "Estimation gamma (default: 0.8)"},
                  {"--epsilon <float>", "Estimation epsilon (default: 0.001)"},
                  {"--hash-load-factor <float>", "Load factor for sized buffers (default: 0.5)"}});
        app::print_help(argv[0], "--mesh1 <A_dataset> --mesh2 <B_dataset> [options]",
                        "Containment query: finds pairs where B objects are fully contained in A objects.", o);
        return 0;
    }
    if (c.mesh1.empty() || c.mesh2.empty()) {
        std::cerr << "Error: --mesh1 and --mesh2 are required" << std::endl;
        return 1;
    }
    timer.next("Application Creation");
    pz_ctx* ctx = nullptr;
    if (pz_ctx_create(c.device, &ctx) != PZ_OK) return app::fail_join("device initialisation");
    uint64_t nt1 = 0, nt2 = 0, no1 = 0, no2 = 0;
    pz_dataset* d1 = app::load(ctx, c.mesh1, "Mesh1", timer, &nt1, &no1);
    if (!d1) return 1;
    pz_dataset* d2 = app::load(ctx, c.mesh2, "Mesh2", timer, &nt2, &no2);
    if (!d2) return 1;
    pz_join_desc d;
    pz_join_desc_defaults(&d);
    d.predicate = PZ_PRED_CONTAINMENT;
    d.include_overlap_pairs = include_overlap ? 1 : 0;
    d.precision = c.precision;
    d.gamma = gamma;
    d.epsilon = epsilon;
    d.load_factor = load_factor;
    d.sizing = c.sizing_override >= 0 && c.sizing_override != PZ_SIZING_LEARNED ? c.sizing_override
                                                                                 : PZ_SIZING_ESTIMATE;
    if (c.verbose_estimate)
        app::print_estimate("=== Containment Selectivity Estimation ===", "==========================================",
                            d1, d2, epsilon, gamma, true, false, false);
    pz_join_result r;
    timer.next("Warmup");
    for (int w = 0; w < c.warmup; ++w) {
        if (pz_join(ctx, d1, d2, &d, &r) != PZ_OK) return app::fail_join("warmup query");
        pz_join_result_release(&r);
    }
    timer.next("Query");
    std::cout << "\n=== Executing containment query ===" << std::endl;
    for (int run = 0; run < c.runs; ++run) {
        if (run) pz_join_result_release(&r);
        if (pz_join(ctx, d1, d2, &d, &r) != PZ_OK) return app::fail_join("query");
    }
    timer.next("Output");
    app::add_phases(timer, r);
    // query-time keys of predicates/adapters/containment.py:163-172
    timer.add("Raytrace_Overlap_Hash_Mesh2ToMesh1", app::uncovered_ms(r, false, {"Parity"}));
    timer.add("Raytrace_Containment_Hash_Mesh2ToMesh1", app::phase_ms(r, {"Parity"}));
    uint64_t ncont = 0, nover = 0;
    for (uint64_t i = 0; i < r.num_pairs; ++i) {
        if (r.pair_types && r.pair_types[i]) ++nover;
        else ++ncont;
    }
    std::cout << "\n=== Containment Query Summary ===" << std::endl;
    std::cout << "A triangles: " << nt1 << "  objects: " << no1 << std::endl;
    std::cout << "B triangles: " << nt2 << "  objects: " << no2 << std::endl;
    std::cout << "A objects: " << no1 << std::endl;
    std::cout << "B objects: " << no2 << std::endl;
    std::cout << "Containment pairs (B in A): " << ncont << std::endl;
    app::print_join_time(r);
    if (include_overlap) {
        std::cout << "Overlap/Touch pairs (A vs B): " << nover << std::endl;
        std::cout << "Reported pairs: " << r.num_pairs << std::endl;
    }
    if (c.export_results) {
        const char* path = "mesh_containment_results.csv";
        std::cout << "Exporting results to " << path << std::endl;
        FILE* f = std::fopen(path, "w");
        if (f) {
            // containment pairs first, then overlap pairs (pierce_containment_main.cpp:606-624)
            std::fprintf(f, include_overlap ? "a_object_id,b_object_id,pair_type\n" : "a_object_id,b_object_id\n");
            for (int pass = 0; pass < 2; ++pass)
                for (uint64_t i = 0; i < r.num_pairs; ++i) {
                    int type = r.pair_types ? r.pair_types[i] : 0;
                    if (type != pass) continue;
                    if (include_overlap)
                        std::fprintf(f, "%d,%d,%s\n", r.pairs[i].object_id_mesh1, r.pairs[i].object_id_mesh2,
                                     type ? "overlap" : "containment");
                    else
                        std::fprintf(f, "%d,%d\n", r.pairs[i].object_id_mesh1, r.pairs[i].object_id_mesh2);
                }
            std::fclose(f);
        }
    }
    pz_join_result_release(&r);
    pz_dataset_release(d1);
    pz_dataset_release(d2);
    pz_ctx_destroy(ctx);
    timer.finish(c.output);
    std::cout << "\nContainment query completed successfully." << std::endl;
    return 0;
}
