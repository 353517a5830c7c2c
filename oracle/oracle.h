/*
 * oracle.h -- C ABI of the CPU oracle for FSMELD's LB_IP -> GLB -> ELD cascade.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a CPU restatement of the
 * reference's algorithm, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the product
 * (fsmeld_amd/) links, loads or calls it.
 *
 * Metric codes follow include/funcs.h:10-18 of the reference
 * (DTW=1, EDR=2, ERP=3, LCSS=4, MSM=5).
 */
#ifndef FSMELD_ORACLE_H
#define FSMELD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stage functions (one macro-segment or one candidate) ---- */
void orc_mvmean(const double* a, int len_a, int l, double* miu, double* si);
void orc_invalid_mask(const double* t, const double* si, int segment_size, int m,
                      uint8_t* invalid_mask);
void orc_min_max_window(const double* a, int n, int m, double* min_out, double* max_out);
void orc_lower_upper_lemire(const double* a, int n, int r, double* l, double* u);
void orc_normalized_arrays(const double* U, const double* L, double* U_n, double* L_n,
                           int length, const double* max_mean, const double* min_mean,
                           const double* max_std, const double* min_std);
/* thresholds t_u / t_l for one segment (heuristic of find_t_lp.cpp) */
void orc_thresholds(int metric, const double* U_n, const double* L_n, int l,
                    const double* q_normal, const double* q_sorted, int m, double arg,
                    double* t_u, double* t_l);
void orc_q_to_h(int metric, const double* q, int m, double t_l, double t_u, double arg,
                double* h_l, double* h_u);
void orc_s_to_g(const double* L_n, const double* U_n, int n, double t_l, double t_u,
                double* g_l, double* g_u);
/* direct sliding inner product: z[j+m-1] = sum_{s<m} g[j+s] h[s], j in [0, n-m] */
void orc_sip_direct(const double* g, int n, const double* h, int m, double* z);
/* 0: z-normalised matching (default); 1: raw matching (paper App. D: query
 * and windows as they are, a window invalid only when it holds a non-finite
 * sample).  Process-wide, like the reference's set_normalization_mode. */
void orc_set_normalization_mode(int raw);
/* LB_IP value of one window from its two sliding sums (trans, src/fsmeld.cpp:84-106) */
double orc_trans(int metric, double z_l, double z_u, int m);
int orc_lbip_exceeds_cutoff(double bound, double cutoff, int metric);
double orc_glb(int metric, const double* Q, const double* T, int m, double arg,
               const double* L_Q, const double* U_Q, int w, double bsf);
double orc_eld(int metric, const double* q, const double* t, int m, int w, double arg);

/* ---- query-level drivers over a whole FP64 target held in memory ---- */
typedef struct {
    uint64_t valid;        /* kk1 */
    uint64_t invalid;
    uint64_t after_lbip;   /* kk2 (search) / LB survivors (cutoff) */
    uint64_t after_glb;    /* kk3 (search) */
    uint64_t exact_evals;
    uint64_t matches;
    double threshold;      /* cutoff: final top-k threshold */
    int64_t ts_n, k_m_l, forsize, tail;
    int m_k;
} orc_stats;

/* Returns 0 on success, negative on bad arguments, -4 when fewer than k
 * valid candidates exist (cutoff).  Match / rank arrays are caller owned
 * with capacity `cap`; *n_out receives the full count even when larger. */
int orc_search(const double* T, int64_t N, const double* q_raw, int m, double cutoff,
               int metric, int w, double arg, int64_t excl_b, int64_t excl_e,
               int64_t seg_begin, int64_t seg_end, int threads,
               orc_stats* st, int64_t* matches, int64_t cap, int64_t* n_out);
int orc_exact(const double* T, int64_t N, const double* q_raw, int m, double cutoff,
              int metric, int w, double arg, int64_t excl_b, int64_t excl_e,
              int threads, orc_stats* st, int64_t* matches, double* dists,
              int64_t cap, int64_t* n_out);
/* threads > 1 scans contiguous segment ranges in parallel and merges their
 * heaps: the same ranks and threshold, counters of the split schedule. */
int orc_cutoff(const double* T, int64_t N, const double* q_raw, int m, int k,
               int metric, int w, double arg, int64_t excl_b, int64_t excl_e,
               int threads, orc_stats* st, int64_t* ranks, double* rank_dists);

/* FSMELD_exact with top_k = k: every window whose exact distance is at or
 * below the k-th smallest finite one (*kth), ascending; -4 below k. */
int orc_exact_topk(const double* T, int64_t N, const double* q_raw, int m, int k, int metric,
                   int w, double arg, int64_t excl_b, int64_t excl_e, int threads,
                   orc_stats* st, int64_t* matches, double* dists, int64_t cap,
                   int64_t* n_out, double* kth);

/* Per-window stage outputs in global window order (ts_n entries), for
 * stage-level parity: mu/sigma/invalid of the owning segment, the LB_IP
 * value (NaN for invalid windows) and the per-segment thresholds. */
int orc_window_stats(const double* T, int64_t N, int m, int64_t excl_b, int64_t excl_e,
                     double* mu, double* sigma, uint8_t* invalid);
int orc_lbip_all(const double* T, int64_t N, const double* q_raw, int m, int metric,
                 int w, double arg, double* lb, double* t_u, double* t_l);

/* z-normalisation of the query exactly as the drivers do it. */
void orc_normalize_query(const double* q_raw, int m, double* q_normal);

#ifdef __cplusplus
}
#endif
#endif
