// k_eld.hip -- ELD entry points: dispatch by metric to the per-metric
// launchers (k_eld_<metric>.hip, templates in k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {

#define FSM_EXTERN(MM)                                                                         \
    extern template hipError_t launch_metric<MM>(                                              \
        hipStream_t, const VerifyArgs&, const int64_t*, const double*,                         \
        const unsigned long long*, int64_t, const double*, bool, int64_t*, double*,            \
        unsigned long long*, unsigned long long*, unsigned long long*, double*, int64_t);
FSM_EXTERN(DTW)
FSM_EXTERN(ERP)
FSM_EXTERN(EDR)
FSM_EXTERN(LCSS)
FSM_EXTERN(MSM)
FSM_EXTERN(SWALE)
#undef FSM_EXTERN

int64_t eld_scratch_threads() { return (int64_t)512 * 256; }

bool eld_uses_scratch(int metric, int m, int w) {
    return !uses_register_band(metric, m, w) && !uses_wave_band(metric, m, w);
}

static hipError_t dispatch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                               const double* in_total, const unsigned long long* in_count,
                               int64_t n_direct, const double* rows, bool keep_all,
                               int64_t* out_idx, double* out_dist, unsigned long long* out_count,
                               unsigned long long* queue, unsigned long long* work,
                               double* scratch, int64_t scratch_threads) {
#define FSM_M(MM) \
    launch_metric<MM>(st, a, in, in_total, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, queue, work, scratch, scratch_threads)
    switch (a.p.metric) {
        case DTW: return FSM_M(DTW);
        case ERP: return FSM_M(ERP);
        case EDR: return FSM_M(EDR);
        case LCSS: return FSM_M(LCSS);
        case MSM: return FSM_M(MSM);
        case SWALE: return FSM_M(SWALE);
    }
#undef FSM_M
    return hipErrorInvalidValue;
}

hipError_t launch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const double* in_total, const unsigned long long* in_count, bool keep_all,
                      int64_t* out_idx, double* out_dist, unsigned long long* out_count,
                      unsigned long long* queue, unsigned long long* work, double* scratch,
                      int64_t scratch_threads) {
    return dispatch_eld(st, a, in, in_total, in_count, 0, nullptr, keep_all, out_idx, out_dist,
                        out_count, queue, work, scratch, scratch_threads);
}

hipError_t launch_eld_pairs(hipStream_t st, const VerifyArgs& a, const double* cands, int64_t n,
                            double* out, unsigned long long* queue, double* scratch,
                            int64_t scratch_threads) {
    return dispatch_eld(st, a, nullptr, nullptr, nullptr, n, cands, true, nullptr, out, nullptr,
                        queue, nullptr, scratch, scratch_threads);
}

}  // namespace fsm
