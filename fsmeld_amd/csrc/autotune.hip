// autotune.hip -- the engine's autotuner (host code): picks the launch shapes
// of fsm_tuning by timing sample searches on the device they will run on.
//
// One knob at a time, over the searches that knob affects:
//   - LB_IP workgroup size per segment class (m_k <= 1024, 2048, 4096, 8192),
//     over the workgroup sizes whose LDS fits a CU, judged by ms_lbip;
//   - the register-band ELD's refill threshold per metric, judged by ms_eld;
//   - GLB's persistent grid, judged by ms_glb over every sample.
// Each candidate runs every affected search `repeats` times and keeps each
// search's smallest stage time (device time from HIP events, so host noise
// does not enter); the candidate with the smallest sum wins, the current value
// on ties.  Every candidate's matches must equal those of the shapes the
// tuner started from: a shape changes speed, never results.
#include <algorithm>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/fsmeld.h"
#include "engine_internal.h"

using namespace fsm;

namespace {

constexpr int kNtCand[] = {128, 256, 512, 1024};
constexpr int kRefillCand[] = {4, 8, 12, 16, 24};
constexpr int kGlbCand[] = {1024, 2048, 4096, 8192};

struct Sample {
    const fsm_query_desc* d;
    const double* q;
    int64_t qlen;
    std::vector<int64_t> want;  // matches under the starting shapes
};

enum Stage { kLbip, kGlb, kEld };

// Sum over the samples of each one's smallest stage time in `repeats` runs;
// -1 on a failed query, -2 when a run's matches differ from the sample's.
double time_samples(fsm_target* t, std::vector<Sample*>& ss, Stage stage, int repeats,
                    int* rc_out) {
    double total = 0.0;
    for (Sample* s : ss) {
        double best = std::numeric_limits<double>::infinity();
        for (int r = 0; r < repeats; r++) {
            fsm_result* res = nullptr;
            const int rc = fsm_query(t, s->d, s->q, s->qlen, &res);
            if (rc) {
                *rc_out = rc;
                return -1.0;
            }
            fsm_stats st;
            fsm_result_stats(res, &st);
            const int64_t n = fsm_result_count(res);
            const int64_t* idx = fsm_result_indices(res);
            const bool same = (int64_t)s->want.size() == n &&
                              std::equal(s->want.begin(), s->want.end(), idx);
            fsm_result_free(res);
            if (!same) return -2.0;
            best = std::min(best, stage == kLbip ? st.ms_lbip : (stage == kGlb ? st.ms_glb
                                                                              : st.ms_eld));
        }
        total += best;
    }
    return total;
}

}  // namespace

extern "C" int fsm_autotune(fsm_target* t, int64_t n, const fsm_query_desc* descs,
                            const double* const* queries, const int64_t* query_lens,
                            int repeats, fsm_tuning* out) {
    if (!t || n < 1 || !descs || !queries || !query_lens || !out)
        return report_error(FSM_ERR_ARG, "null argument or no sample searches");
    if (repeats < 1) return report_error(FSM_ERR_ARG, "repeats must be at least 1");
    for (int64_t i = 0; i < n; i++)
        if (descs[i].mode != FSM_MODE_SEARCH)
            return report_error(FSM_ERR_ARG, "the autotuner times search queries only");
    fsm_tuning tune;
    int rc = fsm_tuning_get(t, &tune);
    if (rc) return rc;
    const int64_t N = fsm_target_length(t);
    // the starting shapes' matches (this also computes each length's
    // statistics once before anything is timed)
    std::vector<Sample> all((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        Sample& s = all[(size_t)i];
        s.d = &descs[i];
        s.q = queries[i];
        s.qlen = query_lens[i];
        fsm_result* res = nullptr;
        rc = fsm_query(t, s.d, s.q, s.qlen, &res);
        if (rc) return rc;
        s.want.assign(fsm_result_indices(res), fsm_result_indices(res) + fsm_result_count(res));
        fsm_result_free(res);
    }
    // Tries the candidate values of one field over samples ss; the field
    // ends at the fastest (the current value on ties).
    auto tune_field = [&](int* field, const int* cand, int ncand, std::vector<Sample*>& ss,
                          Stage stage, auto valid) -> int {
        if (ss.empty()) return FSM_OK;
        const int start = *field;
        int err = FSM_OK;
        double best = time_samples(t, ss, stage, repeats, &err);
        if (best == -1.0) return err;
        int best_v = start;
        for (int c = 0; c < ncand; c++) {
            const int v = cand[c];
            if (v == start || !valid(v)) continue;
            *field = v;
            int r2 = fsm_tuning_set(t, &tune);
            if (r2) return r2;
            const double ms = time_samples(t, ss, stage, repeats, &err);
            if (ms == -1.0) return err;
            if (ms == -2.0)
                return report_error(FSM_ERR_HIP, "a launch shape changed a search's matches");
            if (ms < best) {
                best = ms;
                best_v = v;
            }
        }
        *field = best_v;
        return fsm_tuning_set(t, &tune);
    };
    // LB_IP: per segment class, the sizes that fit every sample's geometry
    for (int cls = 0; cls < 4; cls++) {
        std::vector<Sample*> ss;
        std::vector<Geometry> gs;
        for (Sample& s : all) {
            Geometry G;
            if (make_geometry(N, s.d->m, &G) && lbip_tune_class(G) == cls) {
                ss.push_back(&s);
                gs.push_back(G);
            }
        }
        rc = tune_field(&tune.lbip_threads[cls], kNtCand, 4, ss, kLbip, [&](int v) {
            for (const Geometry& G : gs)
                if (!lbip_threads_ok(G, v)) return false;
            return true;
        });
        if (rc) return rc;
    }
    // ELD refill threshold, per metric
    for (int metric = FSM_METRIC_DTW; metric <= FSM_METRIC_SWALE; metric++) {
        std::vector<Sample*> ss;
        for (Sample& s : all)
            if (s.d->metric == metric) ss.push_back(&s);
        rc = tune_field(&tune.eld_refill[metric], kRefillCand, 5, ss, kEld,
                        [](int) { return true; });
        if (rc) return rc;
    }
    // GLB grid, over every sample
    std::vector<Sample*> every;
    for (Sample& s : all) every.push_back(&s);
    rc = tune_field(&tune.glb_blocks, kGlbCand, 4, every, kGlb, [](int) { return true; });
    if (rc) return rc;
    *out = tune;
    return FSM_OK;
}
