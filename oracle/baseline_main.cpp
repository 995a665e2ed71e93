// ORACLE - test infrastructure / CPU baseline only.
// Command-line fronts of the Face and TOUCH restatements, with the
// reference baselines' argument surface and output lines
// (baselines/face/src/overlap_query.cpp:51-188, touch_query.cpp:312-511):
//   face_overlap  <A.pre> <B.pre> [threads] [--output-csv path]
//   touch_overlap <A.pre> <B.pre> [threads]
// The .pre reader here is independent of the engine's (BinaryIO.h:84-151 layout).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
struct or_mesh_view {
    const float* vertices;
    uint64_t num_vertices;
    const uint32_t* indices;
    uint64_t num_triangles;
    const int32_t* triangle_to_object;
    const float* edge_starts;
    const float* edge_ends;
    const int32_t* edge_objects;
    uint64_t num_edges;
};
struct or_pairs;
or_pairs* or_face_overlap(const or_mesh_view* a, const or_mesh_view* b);
int64_t or_pairs_count(const or_pairs* r);
const int32_t* or_pairs_data(const or_pairs* r);
double or_pairs_query_us(const or_pairs* r);
void or_pairs_free(or_pairs* r);
int64_t or_touch_overlap_count(const or_mesh_view* a, const or_mesh_view* b, double* query_us);
}

namespace {
struct Pre {
    std::vector<float> v;
    std::vector<uint32_t> idx;
    std::vector<int32_t> t2o;
    or_mesh_view view() const {
        return {v.data(), v.size() / 3, idx.data(), idx.size() / 3, t2o.data(), nullptr, nullptr, nullptr, 0};
    }
};

bool read_pre(const std::string& path, Pre& p) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    unsigned char h[56];
    in.read(reinterpret_cast<char*>(h), 56);
    if (in.gcount() != 56) return false;
    uint32_t magic, version;
    uint64_t nv, ni, nm;
    std::memcpy(&magic, h, 4);
    std::memcpy(&version, h + 4, 4);
    std::memcpy(&nv, h + 8, 8);
    std::memcpy(&ni, h + 16, 8);
    std::memcpy(&nm, h + 24, 8);
    if (magic != 0x52334442u || version != 4 || nm != ni) return false;
    p.v.resize(3 * nv);
    p.idx.resize(3 * ni);
    p.t2o.resize(nm);
    in.read(reinterpret_cast<char*>(p.v.data()), 12 * nv);
    in.read(reinterpret_cast<char*>(p.idx.data()), 12 * ni);
    in.read(reinterpret_cast<char*>(p.t2o.data()), 4 * nm);
    return static_cast<bool>(in);
}
}  // namespace

int main(int argc, char** argv) {
#ifdef PZ_BASELINE_FACE
    const char* usage = " <datasetA.bin> <datasetB.bin> [threads] [--output-csv path]";
#else
    const char* usage = " <datasetA.bin> <datasetB.bin> [threads]";
#endif
    if (argc > 1 && (std::string(argv[1]) == "--help" || std::string(argv[1]) == "-h")) {
        std::cout << "Usage: " << argv[0] << usage << std::endl;
        return 0;
    }
    if (argc < 3) {
        std::cout << "Usage: " << argv[0] << usage << std::endl;
        return 1;
    }
    int threads = -1;
    std::string csv = "mesh_overlap_results.csv";
    for (int i = 3; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--output-csv" && i + 1 < argc) {
            csv = argv[++i];
            continue;
        }
        if (!a.empty() && a[0] != '-' && threads < 0) {
            threads = std::stoi(a);
            continue;
        }
        std::cerr << "Unknown argument: " << a << std::endl;
        return 1;
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
    std::cout << "OpenMP threads: " << omp_get_max_threads() << std::endl;
#endif
    Pre A, B;
    std::cout << "Loading datasets..." << std::endl;
    if (!read_pre(argv[1], A) || !read_pre(argv[2], B)) {
        std::cerr << "Error: could not read the .pre inputs" << std::endl;
        return 1;
    }
    or_mesh_view va = A.view(), vb = B.view();
#ifdef PZ_BASELINE_FACE
    or_pairs* r = or_face_overlap(&va, &vb);
    uint64_t us = static_cast<uint64_t>(or_pairs_query_us(r));
    int64_t n = or_pairs_count(r);
    std::cout << "Query Time: " << us << " us (" << us / 1000.0 << " ms)" << std::endl;
    std::cout << "Total Overlaps: " << n << std::endl;
    FILE* f = std::fopen(csv.c_str(), "w");
    if (f) {
        std::fprintf(f, "a_object_id,b_object_id\n");
        const int32_t* d = or_pairs_data(r);
        for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%d,%d\n", d[2 * i], d[2 * i + 1]);
        std::fclose(f);
        std::cout << "Wrote overlap pairs CSV: " << csv << std::endl;
    }
    or_pairs_free(r);
#else
    double qus = 0;
    int64_t n = or_touch_overlap_count(&va, &vb, &qus);
    uint64_t us = static_cast<uint64_t>(qus);
    std::cout << "Query Time: " << us << " us (" << us / 1000.0 << " ms)" << std::endl;
    std::cout << "Total Overlaps: " << n << std::endl;
#endif
    return 0;
}
