// Host-side dataset container and helpers shared by the C ABI, the
// generators and the command-line front-ends.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pierce_hip.h"

struct pz_host_mesh {
    std::vector<float> vertices;         // 3 * nV
    std::vector<uint32_t> indices;       // 3 * nT
    std::vector<int32_t> tri_to_object;  // nT
    std::vector<float> edge_starts;      // 3 * nE
    std::vector<float> edge_ends;        // 3 * nE
    std::vector<int32_t> edge_objects;   // nE
    std::vector<pz_grid_cell> grid;
    float grid_cell_size = 0.0f;
    bool has_grid = false;

    uint64_t num_vertices() const { return vertices.size() / 3; }
    uint64_t num_triangles() const { return indices.size() / 3; }
    uint64_t num_edges() const { return edge_objects.size(); }
    pz_mesh_view view() const;
};

namespace pz {

// Thread-local error message plumbing for the C ABI.
void set_error(const std::string& msg);
pz_status fail(pz_status code, const std::string& msg);

// Run fn(i) for i in [0, n) on up to hardware_concurrency threads, in
// contiguous chunks.  fn must be safe to call concurrently for distinct i.
template <class F>
void parallel_for(uint64_t n, F fn);

// Edge extraction with EdgePreprocessor semantics
// (pierce/preprocess/src/EdgePreprocessor.cpp:44-111).
void extract_edges_pre_order(const pz_mesh_view& v, std::vector<float>& starts,
                             std::vector<float>& ends, std::vector<int32_t>& objs);

// Grid statistics (pierce/preprocess/src/pierce_preprocess.cpp:49-167).
void compute_grid(const pz_mesh_view& v, float cell_size, std::vector<pz_grid_cell>& cells);

}  // namespace pz

#include <thread>

#include <atomic>

// Threads take runs of indices from a shared counter (about 16 runs per
// thread), so items of very different cost still balance.
template <class F>
void pz::parallel_for(uint64_t n, F fn) {
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    if (hw > 16) hw = 16;
    uint64_t nt = n < hw ? n : hw;
    if (nt <= 1) {
        for (uint64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    const uint64_t run = n / (nt * 16) > 0 ? n / (nt * 16) : 1;
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> ts;
    ts.reserve(nt);
    for (uint64_t t = 0; t < nt; ++t) {
        ts.emplace_back([n, run, &next, &fn]() {
            for (uint64_t lo = next.fetch_add(run); lo < n; lo = next.fetch_add(run)) {
                const uint64_t hi = lo + run < n ? lo + run : n;
                for (uint64_t i = lo; i < hi; ++i) fn(i);
            }
        });
    }
    for (auto& t : ts) t.join();
}
