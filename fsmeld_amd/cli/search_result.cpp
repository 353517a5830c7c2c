// search_result.cpp -- the experiment CSV record (format in search_result.h).
#include "search_result.h"

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <utility>
#include <vector>

namespace {

using Cells = std::vector<std::pair<const char*, std::string>>;

std::string real(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

std::string count(std::uint64_t v) { return std::to_string(v); }

// RFC 4180 style: only cells that need it are quoted
std::string text(const std::string& v) {
    bool plain = true;
    for (char c : v)
        if (c == ',' || c == '"' || c == '\r' || c == '\n') plain = false;
    if (plain) return v;
    std::string q;
    q.reserve(v.size() + 2);
    q.push_back('"');
    for (char c : v) {
        q.push_back(c);
        if (c == '"') q.push_back('"');
    }
    q.push_back('"');
    return q;
}

std::string local_time_now() {
    const std::time_t now = std::time(nullptr);
    std::tm tm{};
    localtime_r(&now, &tm);
    char buf[32];
    std::strftime(buf, sizeof(buf), "%Y-%m-%dT%H:%M:%S", &tm);
    return buf;
}

// The record as (column, cell) pairs in column order.
Cells cells_of(const ExperimentRow& r) {
    const double rate = r.valid_candidates
                            ? (double)r.pruned_candidates / (double)r.valid_candidates
                            : 0.0;
    Cells c;
    c.emplace_back("dataset", text(r.dataset));
    c.emplace_back("normalization", text(r.normalization));
    c.emplace_back("metric", text(r.metric));
    c.emplace_back("m", std::to_string(r.m));
    c.emplace_back("w", std::to_string(r.w));
    c.emplace_back("xi", real(r.xi));
    c.emplace_back("method", text(r.method));
    c.emplace_back("lb_time", real(r.lb_time_seconds));
    c.emplace_back("total_time", r.total_time_seconds ? real(*r.total_time_seconds) : "");
    c.emplace_back("valid_candidates", count(r.valid_candidates));
    c.emplace_back("invalid_candidates", count(r.invalid_candidates));
    c.emplace_back("pruned_candidates", count(r.pruned_candidates));
    c.emplace_back("pruning_rate", real(rate));
    c.emplace_back("survivors", count(r.survivors));
    c.emplace_back("matches", r.matches ? count(*r.matches) : "");
    c.emplace_back("lbip_time", r.lbip_time_seconds ? real(*r.lbip_time_seconds) : "");
    c.emplace_back("lbip_pruning_rate", r.lbip_time_seconds ? real(r.lbip_pruning_rate) : "");
    c.emplace_back("timestamp", local_time_now());
    return c;
}

std::string join(const Cells& c, bool names) {
    std::string s;
    for (size_t i = 0; i < c.size(); i++) {
        if (i) s.push_back(',');
        s += names ? std::string(c[i].first) : c[i].second;
    }
    return s;
}

}  // namespace

std::string dataset_name_of(const char* target_path) {
    if (!target_path || !*target_path) return "unknown";
    const std::filesystem::path p(target_path);
    if (p.has_parent_path()) return p.parent_path().filename().string();
    return p.stem().string();
}

std::string experiment_csv_header() { return join(cells_of(ExperimentRow{}), true); }

std::string experiment_csv_row(const ExperimentRow& r) { return join(cells_of(r), false); }

void emit_experiment_row(const ExperimentRow& r) {
    const Cells cells = cells_of(r);
    const std::string header = join(cells, true), row = join(cells, false);
    std::cout << "EXPERIMENT_CSV_HEADER=" << header << '\n'
              << "EXPERIMENT_CSV=" << row << std::endl;
    const char* env = std::getenv("FSMELD_RESULTS_CSV");
    const std::filesystem::path path = (env && *env) ? env : "experiment_results.csv";
    std::error_code ec;
    const auto size = std::filesystem::file_size(path, ec);
    const bool fresh = ec || size == 0;  // missing, unreadable or empty
    std::ofstream out(path, std::ios::app);
    if (!out) {
        std::cerr << "[Warning] Cannot append experiment CSV: " << path.string() << std::endl;
        return;
    }
    if (fresh) out << header << '\n';
    out << row << '\n';
    std::cout << "Experiment CSV saved to: " << path.string() << std::endl;
}
