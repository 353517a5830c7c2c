// search_result.cpp -- CSV emitter of the CLI.  Column set, number format
// (17 significant digits), escaping, header-on-new-file and the
// FSMELD_RESULTS_CSV destination follow src/search_result.cpp:39-101.
#include "search_result.h"

#include <chrono>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace {

std::string escape(const std::string& v) {
    if (v.find_first_of(",\"\r\n") == std::string::npos) return v;
    std::string out = "\"";
    for (char c : v) {
        if (c == '"') out += '"';
        out += c;
    }
    return out + "\"";
}

std::string now_iso() {
    const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    std::tm tm{};
    localtime_r(&t, &tm);
    std::ostringstream s;
    s << std::put_time(&tm, "%Y-%m-%dT%H:%M:%S");
    return s.str();
}

const char* header() {
    return "dataset,normalization,metric,m,w,xi,method,lb_time,total_time,valid_candidates,"
           "invalid_candidates,pruned_candidates,pruning_rate,survivors,matches,"
           "lbip_time,lbip_pruning_rate,timestamp";
}

std::string row_text(const ExperimentRow& r) {
    const double rate = r.valid_candidates == 0
                            ? 0.0
                            : (double)r.pruned_candidates / (double)r.valid_candidates;
    std::ostringstream o;
    o << std::setprecision(17) << escape(r.dataset) << ',' << escape(r.normalization) << ','
      << escape(r.metric) << ',' << r.m << ',' << r.w << ',' << r.xi << ',' << escape(r.method)
      << ',' << r.lb_time_seconds << ',' << r.total_time_seconds << ',' << r.valid_candidates
      << ',' << r.invalid_candidates << ',' << r.pruned_candidates << ',' << rate << ','
      << r.survivors << ',' << r.matches << ',';
    if (r.has_lbip) o << r.lbip_time_seconds;
    o << ',';
    if (r.has_lbip) o << r.lbip_pruning_rate;
    o << ',' << now_iso();
    return o.str();
}

}  // namespace

std::string dataset_name_of(const char* target_path) {
    if (!target_path || !*target_path) return "unknown";
    std::filesystem::path p(target_path);
    const auto parent = p.parent_path();
    return parent.empty() ? p.stem().string() : parent.filename().string();
}

void emit_experiment_row(const ExperimentRow& r) {
    const std::string row = row_text(r);
    std::cout << "EXPERIMENT_CSV_HEADER=" << header() << '\n';
    std::cout << "EXPERIMENT_CSV=" << row << std::endl;
    const char* env = std::getenv("FSMELD_RESULTS_CSV");
    const std::filesystem::path path = (env && *env) ? std::filesystem::path(env)
                                                     : std::filesystem::path("experiment_results.csv");
    std::error_code ec;
    const bool need_header =
        !std::filesystem::exists(path, ec) || std::filesystem::file_size(path, ec) == 0;
    std::ofstream out(path, std::ios::app);
    if (!out) {
        std::cerr << "[Warning] Cannot append experiment CSV: " << path.string() << std::endl;
        return;
    }
    if (need_header) out << header() << '\n';
    out << row << '\n';
    std::cout << "Experiment CSV saved to: " << path.string() << std::endl;
}
