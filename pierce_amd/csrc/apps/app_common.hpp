// Shared pieces of the four query front-ends: argument scanning with the
// reference's rules (unknown flags ignored, pierce/query/src/applications/app_cli_options.h:18-30),
// the phase timer with the reference's JSON layout
// (pierce/query/src/timer.cpp:124-192), dataset loading and CSV output.
#pragma once

#include <chrono>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "pierce_hip.h"

namespace app {

// Phase timer: wall-clock phases plus measurements added from device events.
class Timer {
   public:
    void start(const std::string& name);
    void next(const std::string& name);
    void add(const std::string& name, double ms);  // manual measurement
    void counter(const std::string& name, unsigned long long v);
    void finish(const std::string& json_path);  // prints summary, writes JSON
    double total_ms() const;

   private:
    struct Phase {
        std::string name;
        double us;
    };
    using clock = std::chrono::steady_clock;
    void close_current();
    std::vector<Phase> phases_, manual_;
    std::vector<std::pair<std::string, unsigned long long>> counters_;
    std::string cur_;
    clock::time_point cur_start_, t0_, t1_;
    bool running_ = false;
};

struct Common {
    std::string mesh1, mesh2, output, ptx;
    int runs = 1;
    int warmup = 2;
    bool help = false;
    bool export_results = true;
    int device = 0;
    int strategy = PZ_STRATEGY_TWO_LEVEL;
    int precision = PZ_PREC_FP32_FILTERED;
    int sizing_override = -1;  // --sizing
    std::string model_path;
    bool verbose_estimate = false;  // --verbose-estimate
};

// Scan argv; `app` handles application flags and returns true when it consumed
// one (advancing i past values); everything unknown is ignored.
template <class F>
void parse(int argc, char** argv, Common& c, bool allow_no_export, F app) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](std::string& dst) {
            if (i + 1 < argc) {
                dst = argv[++i];
                return true;
            }
            return false;
        };
        if (a == "--help" || a == "-h") {
            c.help = true;
            continue;
        }
        if (a == "--mesh1" && val(c.mesh1)) continue;
        if (a == "--mesh2" && val(c.mesh2)) continue;
        if (a == "--output" && val(c.output)) continue;
        if (a == "--ptx" && val(c.ptx)) continue;  // accepted: there is no PTX on gfx950
        if (a == "--runs" && i + 1 < argc) {
            c.runs = std::atoi(argv[++i]);
            continue;
        }
        if (a == "--warmup-runs" && i + 1 < argc) {
            c.warmup = std::atoi(argv[++i]);
            continue;
        }
        if (allow_no_export && a == "--no-export") {
            c.export_results = false;
            continue;
        }
        if (a == "--device" && i + 1 < argc) {
            c.device = std::atoi(argv[++i]);
            continue;
        }
        if (a == "--strategy" && i + 1 < argc) {
            std::string v = argv[++i];
            c.strategy = v == "global" ? PZ_STRATEGY_GLOBAL : PZ_STRATEGY_TWO_LEVEL;
            continue;
        }
        if (a == "--precision" && i + 1 < argc) {
            std::string v = argv[++i];
            c.precision = v == "fp64" ? PZ_PREC_FP64 : PZ_PREC_FP32_FILTERED;
            continue;
        }
        if (a == "--sizing" && i + 1 < argc) {
            std::string v = argv[++i];
            c.sizing_override = v == "two-pass"   ? PZ_SIZING_TWO_PASS
                                : v == "estimate" ? PZ_SIZING_ESTIMATE
                                : v == "learned"  ? PZ_SIZING_LEARNED
                                                  : PZ_SIZING_FIXED;
            continue;
        }
        if (a == "--model" && val(c.model_path)) continue;
        if (a == "--verbose-estimate") {
            c.verbose_estimate = true;
            continue;
        }
        app(a, i, argc, argv);
    }
    if (c.runs < 1) c.runs = 1;
    if (c.warmup < 0) c.warmup = 0;
}

void print_help(const char* exe, const std::string& usage, const std::string& desc,
                const std::vector<std::pair<std::string, std::string>>& opts);
std::vector<std::pair<std::string, std::string>> common_help(bool no_export);

// Load a .pre file and upload it; prints the reference's progress lines.
// Returns nullptr (after printing the error) on failure.
pz_dataset* load(pz_ctx* ctx, const std::string& path, const char* label, Timer& t, uint64_t* ntri,
                 uint64_t* nobj);

// The reference's verbose selectivity-estimation block (the estimate functions
// with verbose set: pierce/query/src/applications/pierce_overlap_main.cpp:450-466,
// pierce_intersection_main.cpp:291-311, pierce_containment_main.cpp:127-133).  The
// reference's tools pass verbose = false at every call, so this prints only
// under --verbose-estimate.  `full` selects the overlap/intersection layout
// (sizes and alpha); without it only cells, raw and final are printed.
void print_estimate(const char* title, const char* rule, const pz_dataset* a, const pz_dataset* b, float epsilon,
                    float gamma, bool alpha_correction, bool full, bool alpha_line);

// Library phases into the timer (as manual measurements).
void add_phases(Timer& t, const pz_join_result& r);

// Sum of library phases whose names contain any of `keys`.
double phase_ms(const pz_join_result& r, std::initializer_list<const char*> keys);
// The join's phases that no query-time key of the tool's benchmark adapter
// covers, minus those whose name contains one of `except`.  exact == true: the
// adapter's exact mode, which sums every "raytrace_*" key, "gpu deduplication"
// and "download results" (pierce_adapter.py:338-342); otherwise the estimate
// modes' fixed key lists (pierce_adapter.py:343-357,
// predicates/adapters/intersection.py:207-221, containment.py:163-172).  A tool
// reports this sum under one of its adapter's raytrace keys, so the adapter's
// query time covers every device phase of the join.
double uncovered_ms(const pz_join_result& r, bool exact, std::initializer_list<const char*> except);
// "Join time (device-resident geometry to sorted host pairs): <ms> ms"
void print_join_time(const pz_join_result& r);

bool write_pairs_csv(const std::string& path, const char* header, const pz_join_result& r);

int fail_join(const char* what);

}  // namespace app
