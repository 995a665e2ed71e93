// pierce_preprocess: turns a Wavefront .obj dataset into a .pre v4 file with
// the per-object edge payload and, on request, the grid statistics the
// query tools' estimate reads.  Command line, phases and output follow
// pierce/preprocess/src/pierce_preprocess.cpp:171-352:
//   --mode <mesh|dt> --dataset <path> --output-geometry <file>
//   --output-timing <file> [--shuffle] [--generate-grid] [--grid-cell-size S]
// (`--gamma <v>` is consumed and ignored; unknown arguments draw a warning).
// `--seed <n>` fixes the --shuffle offsets, which the reference draws from an
// unseeded generator.  TDBase .dt input (--mode dt) is not supported.
#include <cstring>
#include <iostream>
#include <random>
#include <string>

#include "app_common.hpp"

int main(int argc, char** argv) {
    std::string dataset, out_geometry = "geometry_data.txt", out_timing = "preprocessing_timing.json";
    std::string mode = "mesh";
    bool shuffle = false, grid = false, seeded = false;
    float cell = 1.0f;
    uint64_t seed = 0;
    std::cout << "Arguments received:" << std::endl;
    for (int i = 0; i < argc; ++i) std::cout << "  [" << i << "] " << argv[i] << std::endl;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--dataset" && i + 1 < argc) dataset = argv[++i];
        else if (a == "--mode" && i + 1 < argc) mode = argv[++i];
        else if (a == "--output-geometry" && i + 1 < argc) out_geometry = argv[++i];
        else if (a == "--output-timing" && i + 1 < argc) out_timing = argv[++i];
        else if (a == "--shuffle" || a == "shuffle") shuffle = true;
        else if (a == "--generate-grid") grid = true;
        else if (a == "--grid-cell-size" && i + 1 < argc) cell = std::stof(argv[++i]);
        else if (a == "--gamma" && i + 1 < argc) ++i;
        else if (a == "--seed" && i + 1 < argc) {
            seed = std::stoull(argv[++i]);
            seeded = true;
        } else if (a == "--help" || a == "-h") {
            app::print_help(argv[0], "[--mode <mesh|dt>] [--dataset <path>] [--output-geometry <file>] "
                                     "[--output-timing <file>]", "",
                            {{"--mode <mesh|dt>", "Dataset loading mode (default: mesh); mesh: .obj file"},
                             {"--dataset <path>", "Path to dataset"},
                             {"--output-geometry <path>", "Path of the .pre v4 output"},
                             {"--output-timing <path>", "Path to JSON file for preprocessing timing output"},
                             {"--shuffle", "Randomly translate each loaded object (offsets in [-15, 15])"},
                             {"--seed <n>", "Seed of the --shuffle offsets (default: random)"},
                             {"--generate-grid", "Generate grid statistics for selectivity estimation"},
                             {"--grid-cell-size <S>", "Grid cell size (default: 1.0)"}});
            return 0;
        } else {
            std::cout << "Warning: Unknown argument '" << a << "'" << std::endl;
        }
    }
    if (mode != "mesh" && mode != "dt") {
        std::cerr << "Error: Unknown mode '" << mode << "'. Valid modes: mesh, dt" << std::endl;
        return 1;
    }
    if (dataset.empty()) {
        std::cerr << "Error: Dataset path is required. Use --dataset <path>" << std::endl;
        return 1;
    }
    std::cout << "Dataset Preprocessing Tool" << std::endl;
    std::cout << "Mode: " << mode << std::endl;
    std::cout << "Input dataset: " << dataset << std::endl;
    std::cout << "Output geometry: " << out_geometry << std::endl;
    std::cout << "Output timing: " << out_timing << std::endl;

    app::Timer timer;
    timer.start("Loading Dataset");
    std::string ext = dataset.size() >= 4 ? dataset.substr(dataset.size() - 4) : "";
    for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".dt" || (dataset.size() >= 3 && dataset.substr(dataset.size() - 3) == ".dt")) {
        std::cerr << "Error: TDBase .dt datasets are not supported by this build: " << dataset << std::endl;
        return 1;
    }
    if (ext != ".obj") {
        std::cerr << "Error: Unsupported dataset format: " << dataset << std::endl;
        return 1;
    }
    pz_host_mesh* m = nullptr;
    if (pz_obj_read(dataset.c_str(), &m) != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        std::cerr << "Error: Failed to load dataset or dataset is empty." << std::endl;
        return 1;
    }
    pz_mesh_view v;
    pz_host_mesh_view(m, &v);
    int32_t nobj = 0;
    for (uint64_t t = 0; t < v.num_triangles; ++t) nobj = std::max(nobj, v.triangle_to_object[t] + 1);
    std::cout << "Successfully loaded " << nobj << " object(s)" << std::endl;
    std::cout << "Total vertices: " << v.num_vertices << std::endl;
    std::cout << "Total triangles: " << v.num_triangles << std::endl;

    if (shuffle) {
        timer.next("Shuffling Objects (Random Translation)");
        if (!seeded) seed = std::random_device{}();
        std::cout << "Applying random translations to " << nobj << " objects (seed " << seed << ")..." << std::endl;
        if (pz_host_mesh_shuffle(m, seed, -15.0f, 15.0f) != PZ_OK) {
            std::cerr << "Error: " << pz_last_error() << std::endl;
            return 1;
        }
    }
    if (grid) {
        timer.next("Generating Grid Statistics");
        if (pz_host_mesh_compute_grid(m, cell) != PZ_OK) {
            std::cerr << "Error: " << pz_last_error() << std::endl;
            return 1;
        }
        pz_host_mesh_view(m, &v);
        std::cout << "Grid statistics generated. Cell Size: " << cell
                  << ". Populated sparse cells: " << v.num_grid_cells << std::endl;
    }
    timer.next("Extracting Edges");
    if (pz_host_mesh_extract_edges(m) != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        return 1;
    }
    pz_host_mesh_view(m, &v);
    if (v.num_edges == 0) {
        std::cerr << "Error: Edge extraction failed. No valid edge payload generated." << std::endl;
        return 1;
    }
    std::cout << "Extracted " << v.num_edges << " unique edges (per object)" << std::endl;
    timer.next("Writing Geometry Data");
    if (pz_pre_write(out_geometry.c_str(), &v) != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        return 1;
    }
    std::cout << "Geometry data written to: " << out_geometry << " (Binary)" << std::endl;
    timer.finish(out_timing);

    std::cout << "\n=== Preprocessing Complete ===" << std::endl;
    std::cout << "Mode: " << mode << std::endl;
    std::cout << "Processed dataset: " << dataset << std::endl;
    std::cout << "Total objects (triangles): " << v.num_triangles << std::endl;
    std::cout << "Total vertices: " << v.num_vertices << std::endl;
    std::cout << "Total triangles: " << v.num_triangles << std::endl;
    std::cout << "Total extracted edges: " << v.num_edges << std::endl;
    if (grid) {
        std::cout << "Grid cell size: " << cell << std::endl;
        std::cout << "Populated sparse grid cells: " << v.num_grid_cells << std::endl;
    }
    std::cout << "Geometry data saved to: " << out_geometry << std::endl;
    std::cout << "Timing data saved to: " << out_timing << std::endl;
    pz_host_mesh_free(m);
    return 0;
}
