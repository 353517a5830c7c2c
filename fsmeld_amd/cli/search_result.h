// search_result.h -- the experiment CSV record a search / exact run appends.
//
// The record is an interface shared with the reference (include/search_result.h,
// src/search_result.cpp of lzzt0123/FSMELD), so the columns, their order and
// the cell formats are the same:
//
//   dataset, normalization, metric, m, w, xi, method, lb_time, total_time,
//   valid_candidates, invalid_candidates, pruned_candidates, pruning_rate,
//   survivors, matches, lbip_time, lbip_pruning_rate, timestamp
//
//   - reals print as %.17g, counts as decimal integers;
//   - pruning_rate = pruned_candidates / valid_candidates (0 when none valid);
//   - total_time, matches and the two lbip cells are empty when absent;
//   - a text cell holding ',', '"', CR or LF is quoted, inner quotes doubled;
//   - timestamp is local time as YYYY-MM-DDTHH:MM:SS.
#pragma once

#include <cstdint>
#include <optional>
#include <string>

struct ExperimentRow {
    std::string dataset;
    std::string normalization = "zscore";
    std::string metric;
    int m = 0;
    int w = 0;
    double xi = 0.0;
    std::string method;
    double lb_time_seconds = 0.0;
    std::optional<double> total_time_seconds;
    std::uint64_t valid_candidates = 0;
    std::uint64_t invalid_candidates = 0;
    std::uint64_t pruned_candidates = 0;
    std::uint64_t survivors = 0;
    std::optional<std::uint64_t> matches;
    std::optional<double> lbip_time_seconds;  // both lbip cells are present together
    double lbip_pruning_rate = 0.0;
};

// Dataset label of a target path: the name of the folder holding the file,
// or the file's stem when the path names no folder ("unknown" for none).
std::string dataset_name_of(const char* target_path);

// Header and row text (one table gives both, so they cannot disagree).
std::string experiment_csv_header();
std::string experiment_csv_row(const ExperimentRow& r);

// Prints EXPERIMENT_CSV_HEADER= / EXPERIMENT_CSV= lines and appends the row
// to $FSMELD_RESULTS_CSV (default experiment_results.csv), writing the header
// first when that file is missing or empty.
void emit_experiment_row(const ExperimentRow& r);
