// search_result.h -- experiment CSV row of a search / exact run
// (same columns and formatting as include/search_result.h and
// src/search_result.cpp of the reference).
#pragma once

#include <cstdint>
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
    double total_time_seconds = 0.0;
    std::uint64_t valid_candidates = 0;
    std::uint64_t invalid_candidates = 0;
    std::uint64_t pruned_candidates = 0;
    std::uint64_t survivors = 0;
    std::uint64_t matches = 0;
    bool has_lbip = false;
    double lbip_time_seconds = 0.0;
    double lbip_pruning_rate = 0.0;
};

std::string dataset_name_of(const char* target_path);
void emit_experiment_row(const ExperimentRow& r);
