// pierce_generate: writes the synthetic workloads as .pre v4 files (edges in
// EdgePreprocessor key order, optional grid statistics), the form the query
// tools and the CPU baselines read.
#include <cstring>
#include <iostream>
#include <string>

#include "app_common.hpp"

int main(int argc, char** argv) {
    std::string out, kind = "ellipsoids";
    pz_gen_params p;
    pz_gen_defaults(PZ_GEN_ELLIPSOIDS, &p);
    float grid = 0.f;
    bool help = false;
    pz_gen_params over = p;
    std::memset(&over, 0, sizeof(over));
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto num = [&]() { return i + 1 < argc ? std::stod(argv[++i]) : 0.0; };
        if (a == "--help" || a == "-h") help = true;
        else if (a == "--output" && i + 1 < argc) out = argv[++i];
        else if (a == "--kind" && i + 1 < argc) kind = argv[++i];
        else if (a == "--count") over.count = (uint64_t)num();
        else if (a == "--seed") over.seed = (uint64_t)num();
        else if (a == "--level") over.level_min = over.level_max = (int)num();
        else if (a == "--level-min") over.level_min = (int)num();
        else if (a == "--level-max") over.level_max = (int)num();
        else if (a == "--radius-min") over.radius_min = (float)num();
        else if (a == "--radius-max") over.radius_max = (float)num();
        else if (a == "--box") over.box = (float)num();
        else if (a == "--jitter") over.jitter = (float)num();
        else if (a == "--copy-fraction") over.copy_fraction = (float)num();
        else if (a == "--copy-scale-min") over.copy_scale_min = (float)num();
        else if (a == "--copy-scale-max") over.copy_scale_max = (float)num();
        else if (a == "--copy-source-count") over.copy_source_count = (uint64_t)num();
        else if (a == "--copy-source-seed") over.copy_source_seed = (uint64_t)num();
        else if (a == "--generate-grid") grid = grid > 0 ? grid : 5.0f;
        else if (a == "--grid-cell-size") grid = (float)num();
        else std::cerr << "Ignoring unknown argument: " << a << std::endl;
    }
    if (help || out.empty()) {
        std::cout << "Usage: " << argv[0]
                  << " --output <file.pre> [--kind ellipsoids|spheres|cubes|tubes] [--count N] [--seed S]\n"
                     "  [--level L | --level-min L --level-max L] [--radius-min r] [--radius-max r] [--box B]\n"
                     "  [--jitter f] [--copy-fraction f --copy-scale-min s --copy-scale-max s\n"
                     "   --copy-source-count N --copy-source-seed S] [--generate-grid] [--grid-cell-size c]"
                  << std::endl;
        return help ? 0 : 1;
    }
    int k = kind == "spheres" ? PZ_GEN_SPHERES_MIXED
            : kind == "cubes" ? PZ_GEN_CUBES
            : kind == "tubes" ? PZ_GEN_TUBES
                              : PZ_GEN_ELLIPSOIDS;
    pz_gen_defaults(k, &p);
    p.edges_mode = PZ_EDGES_PRE;
    if (over.count) p.count = over.count;
    if (over.seed) p.seed = over.seed;
    if (over.level_max) {
        p.level_min = over.level_min ? over.level_min : over.level_max;
        p.level_max = over.level_max;
    }
    if (over.radius_min > 0) p.radius_min = over.radius_min;
    if (over.radius_max > 0) p.radius_max = over.radius_max;
    if (over.box > 0) p.box = over.box;
    p.jitter = over.jitter;
    p.copy_fraction = over.copy_fraction;
    p.copy_scale_min = over.copy_scale_min;
    p.copy_scale_max = over.copy_scale_max;
    p.copy_source_count = over.copy_source_count;
    p.copy_source_seed = over.copy_source_seed;
    pz_host_mesh* m = nullptr;
    if (pz_generate(&p, &m) != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        return 1;
    }
    if (grid > 0 && pz_host_mesh_compute_grid(m, grid) != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        return 1;
    }
    pz_mesh_view v;
    pz_host_mesh_view(m, &v);
    if (pz_pre_write(out.c_str(), &v) != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        return 1;
    }
    std::cout << "Geometry data written to: " << out << " (Binary): " << v.num_vertices << " vertices, "
              << v.num_triangles << " triangles, " << v.num_edges << " edges, " << v.num_grid_cells
              << " grid cells" << std::endl;
    pz_host_mesh_free(m);
    return 0;
}
