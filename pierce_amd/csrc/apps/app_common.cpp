#include "app_common.hpp"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <unordered_map>

namespace app {

void Timer::start(const std::string& name) {
    phases_.clear();
    manual_.clear();
    counters_.clear();
    t0_ = clock::now();
    running_ = true;
    cur_ = name;
    cur_start_ = t0_;
}

void Timer::close_current() {
    auto now = clock::now();
    phases_.push_back({cur_, std::chrono::duration<double, std::micro>(now - cur_start_).count()});
    cur_start_ = now;
}

void Timer::next(const std::string& name) {
    if (!running_) return;
    close_current();
    cur_ = name;
}

void Timer::add(const std::string& name, double ms) { manual_.push_back({name, ms * 1000.0}); }

void Timer::counter(const std::string& name, unsigned long long v) { counters_.push_back({name, v}); }

double Timer::total_ms() const { return std::chrono::duration<double, std::milli>(t1_ - t0_).count(); }

void Timer::finish(const std::string& path) {
    if (!running_) return;
    close_current();
    t1_ = clock::now();
    running_ = false;
    std::cout << "\n=== Detailed Performance Summary ===" << std::endl;
    auto line = [](const Phase& p) {
        long long us = static_cast<long long>(p.us);
        std::cout << std::left << std::setw(30) << (p.name + ":") << std::right << std::setw(20) << us
                  << " microseconds (" << std::fixed << std::setprecision(2) << us / 1000.0 << " ms)" << std::endl;
        std::cout.unsetf(std::ios::floatfield);
    };
    for (auto& p : phases_) line(p);
    for (auto& p : manual_) line(p);
    if (!counters_.empty()) {
        std::cout << "\n=== Profiling Counters ===" << std::endl;
        for (auto& c : counters_)
            std::cout << std::left << std::setw(30) << (c.first + ":") << std::right << std::setw(20) << c.second
                      << std::endl;
    }
    long long total_us = static_cast<long long>(total_ms() * 1000.0);
    std::cout << std::left << std::setw(25) << "Total Execution Time:" << std::right << std::setw(10) << total_us
              << " microseconds (" << std::fixed << std::setprecision(2) << total_us / 1000.0 << " ms)" << std::endl;
    std::cout.unsetf(std::ios::floatfield);

    std::ofstream f(path);
    if (!f) {
        std::cerr << "Warning: Could not write timing results to file." << std::endl;
        return;
    }
    // keys: lower-cased name + "_<occurrence>" (pierce/query/src/timer.cpp:142-159)
    std::unordered_map<std::string, int> seen;
    std::vector<Phase> all(phases_);
    all.insert(all.end(), manual_.begin(), manual_.end());
    f << "{\n  \"phases\": {\n";
    for (size_t i = 0; i < all.size(); ++i) {
        std::string k;
        for (char ch : all[i].name) k.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(ch))));
        k += "_" + std::to_string(++seen[k]);
        long long us = static_cast<long long>(all[i].us);
        f << "    \"" << k << "\": {\n      \"duration_us\": " << us << ",\n      \"duration_ms\": " << std::fixed
          << std::setprecision(2) << us / 1000.0 << "\n    }" << (i + 1 < all.size() ? "," : "") << "\n";
    }
    f << "  },\n  \"counters\": {\n";
    for (size_t i = 0; i < counters_.size(); ++i) {
        std::string k;
        for (char ch : counters_[i].first) k.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(ch))));
        f << "    \"" << k << "\": " << counters_[i].second << (i + 1 < counters_.size() ? "," : "") << "\n";
    }
    f << "  },\n  \"total\": {\n    \"duration_us\": " << total_us << ",\n    \"duration_ms\": " << std::fixed
      << std::setprecision(2) << total_us / 1000.0 << "\n  }\n}\n";
}

void print_help(const char* exe, const std::string& usage, const std::string& desc,
                const std::vector<std::pair<std::string, std::string>>& opts) {
    std::cout << "Usage: " << exe << " " << usage << std::endl;
    if (!desc.empty()) std::cout << desc << std::endl;
    std::cout << "Options:" << std::endl;
    for (auto& o : opts) std::cout << "  " << std::left << std::setw(34) << o.first << o.second << std::endl;
}

std::vector<std::pair<std::string, std::string>> common_help(bool no_export) {
    std::vector<std::pair<std::string, std::string>> o = {
        {"--mesh1 <path>", "Path to first mesh dataset (.pre v4)"},
        {"--mesh2 <path>", "Path to second mesh dataset (.pre v4)"},
        {"--output <path>", "Path to JSON file for performance timing output"},
        {"--ptx <path>", "Accepted for compatibility; no PTX is used on gfx950"},
        {"--runs <number>", "Number of measured query runs (default: 1)"},
        {"--warmup-runs <number>", "Number of warmup iterations (default: 2)"},
        {"--device <index>", "HIP device (default: 0)"},
        {"--strategy <two-level|global>", "Join strategy (default: two-level)"},
        {"--precision <fp32|fp64>", "fp32-filtered exact tests or fp64-filtered exact tests (default: fp32)"},
        {"--sizing <two-pass|estimate|learned|fixed>", "Output sizing (default depends on the tool)"},
        {"--model <path>", "Selectivity model weights for --sizing learned"},
        {"--verbose-estimate", "Print the estimate's matched cells, raw pairs, sizes and alpha"},
    };
    if (no_export) o.push_back({"--no-export", "Do not export results to CSV"});
    o.push_back({"--help, -h", "Show this help message"});
    return o;
}

pz_dataset* load(pz_ctx* ctx, const std::string& path, const char* label, Timer& t, uint64_t* ntri, uint64_t* nobj) {
    t.next(std::string("Load ") + label);
    std::cout << "Loading " << label << " from: " << path << std::endl;
    pz_host_mesh* m = nullptr;
    pz_status st = pz_pre_read(path.c_str(), &m);
    if (st != PZ_OK) {
        std::cerr << "Error: " << pz_last_error() << std::endl;
        std::cerr << "Error: Failed to load " << label << " from " << path << std::endl;
        return nullptr;
    }
    pz_mesh_view v;
    pz_host_mesh_view(m, &v);
    if (v.num_vertices == 0) {
        std::cerr << "Error: Failed to load " << label << " from " << path << std::endl;
        pz_host_mesh_free(m);
        return nullptr;
    }
    std::cout << label << " loaded: " << v.num_vertices << " vertices, " << v.num_triangles << " triangles"
              << std::endl;
    t.next(std::string("Upload ") + label);
    pz_dataset* d = nullptr;
    st = pz_dataset_upload(ctx, &v, &d);
    pz_host_mesh_free(m);
    if (st != PZ_OK) {
        if (st == PZ_E_NO_EDGES)
            std::cerr << "Error: " << label << " file '" << path
                      << "' does not contain precomputed edges. Re-run preprocessing with edge extraction."
                      << std::endl;
        else
            std::cerr << "Error: " << pz_last_error() << std::endl;
        return nullptr;
    }
    uint64_t info[5];
    pz_dataset_info(d, info);
    *ntri = info[1];
    *nobj = info[3];
    std::cout << label << " uploaded to GPU: " << info[2] << " unique edges" << std::endl;
    return d;
}

void add_phases(Timer& t, const pz_join_result& r) {
    for (int i = 0; i < r.num_phases; ++i) t.add(r.phases[i].name, r.phases[i].ms);
}

double phase_ms(const pz_join_result& r, std::initializer_list<const char*> keys) {
    double s = 0;
    for (int i = 0; i < r.num_phases; ++i)
        for (const char* k : keys)
            if (std::strstr(r.phases[i].name, k)) {
                s += r.phases[i].ms;
                break;
            }
    return s;
}

double uncovered_ms(const pz_join_result& r, bool exact, std::initializer_list<const char*> except) {
    static const char* const kEstimateKeys[] = {
        "selectivity estimation", "raytrace_hash_mesh1tomesh2", "raytrace_hash_mesh2tomesh1",
        "raytrace_overlap_hash_mesh1tomesh2", "raytrace_overlap_hash_mesh2tomesh1",
        "raytrace_containment_hash_mesh1tomesh2", "raytrace_containment_hash_mesh2tomesh1",
        "compact_hash_table_pairs", "compact_hash_table_pairs (containment)", "compact_hash_table_pairs (overlap)",
        "download results"};
    double s = 0;
    for (int i = 0; i < r.num_phases; ++i) {
        std::string k;
        for (const char* ch = r.phases[i].name; *ch; ++ch)
            k.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(*ch))));
        bool covered = false;
        if (exact) {
            covered = k.rfind("raytrace_", 0) == 0 || k == "download results" || k == "gpu deduplication";
        } else {
            for (const char* e : kEstimateKeys) covered = covered || k == e;
        }
        bool skip = covered;
        for (const char* e : except)
            if (std::strstr(r.phases[i].name, e)) skip = true;
        if (!skip) s += r.phases[i].ms;
    }
    return s;
}

void print_join_time(const pz_join_result& r) {
    std::cout << "Join time (device-resident geometry to sorted host pairs): " << r.ms_total << " ms" << std::endl;
}

bool write_pairs_csv(const std::string& path, const char* header, const pz_join_result& r) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::cerr << "Error: Could not open pairs output file: " << path << std::endl;
        return false;
    }
    std::fprintf(f, "%s\n", header);
    for (uint64_t i = 0; i < r.num_pairs; ++i)
        std::fprintf(f, "%d,%d\n", r.pairs[i].object_id_mesh1, r.pairs[i].object_id_mesh2);
    std::fclose(f);
    return true;
}

void print_estimate(const char* title, const char* rule, const pz_dataset* a, const pz_dataset* b, float epsilon,
                    float gamma, bool alpha_correction, bool full, bool alpha_line) {
    pz_grid_estimate e;
    if (pz_dataset_estimate_grid(a, b, epsilon, gamma, alpha_correction ? 1 : 0, &e) != PZ_OK) {
        std::cerr << "Error: estimate: " << pz_last_error() << std::endl;
        return;
    }
    if (!e.has_grid) {
        std::cout << "Skipping estimation: Grid data not found in one or both datasets." << std::endl;
        std::cout << "Run pierce_preprocess with --generate-grid to enable estimation." << std::endl;
        return;
    }
    std::cout << "\n" << title << std::endl;
    std::cout << "Matched Sparse Cells:      " << e.matched_cells << std::endl;
    std::cout << "Raw Potential Pairs:       " << (long long)e.raw_pairs << std::endl;
    if (full) {
        std::cout << "Avg Object Size (Mesh1):   " << e.avg_size[0] << std::endl;
        std::cout << "Avg Object Size (Mesh2):   " << e.avg_size[1] << std::endl;
        std::cout << "Avg VolRatio (Mesh1):      " << e.avg_vol_ratio[0] << std::endl;
        std::cout << "Avg VolRatio (Mesh2):      " << e.avg_vol_ratio[1] << std::endl;
        std::cout << "Effective Size (Mesh1):    " << e.effective_size[0] << std::endl;
        std::cout << "Effective Size (Mesh2):    " << e.effective_size[1] << std::endl;
        if (alpha_line)
            std::cout << "Alpha Correction Enabled:  " << (alpha_correction ? "yes" : "no") << std::endl;
        std::cout << "Replication Factor (alpha):" << e.alpha << std::endl;
    }
    std::cout << "Final Estimated Pairs:     " << e.estimated_pairs << std::endl;
    std::cout << rule << "\n" << std::endl;
}

int fail_join(const char* what) {
    std::cerr << "Error: " << what << ": " << pz_last_error() << std::endl;
    return 1;
}

}  // namespace app
