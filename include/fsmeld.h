/*
 * fsmeld.h -- C ABI of the MI355X (gfx950) FSMELD engine: exact z-normalised
 * subsequence range search under elastic distances with the
 * LB_IP -> GLB -> ELD cascade, computed by HIP kernels on one GPU.
 *
 * The reference is one executable with a C++ free-function API
 * (include/funcs.h, include/fsmeld_internal.h of lzzt0123/FSMELD); each entry
 * point below names the reference function it replaces.  Conventions:
 *   - plain pointers and sizes, host memory on both sides of the boundary;
 *   - int status codes (FSM_OK = 0, negative on failure); fsm_last_error()
 *     returns a message for the calling thread's last failure;
 *   - result objects are library owned and released with fsm_result_free();
 *   - a target handle is reentrant for sequential callers; concurrent calls on
 *     one handle must be serialised by the caller.
 * Window indices are 0-based candidate start positions in the target.
 */
#ifndef FSMELD_H
#define FSMELD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FSM_OK 0
#define FSM_ERR_ARG (-1)          /* invalid argument (message in fsm_last_error) */
#define FSM_ERR_OOM (-2)          /* device or host allocation failed */
#define FSM_ERR_HIP (-3)          /* HIP runtime error */
#define FSM_ERR_FEWER_THAN_K (-4) /* cutoff: fewer valid candidates than k */
#define FSM_ERR_NODEV (-5)        /* no usable GPU */
#define FSM_ERR_IO (-6)           /* target file unreadable */

/* metric codes: include/funcs.h:10-18 of the reference */
#define FSM_METRIC_DTW 1
#define FSM_METRIC_EDR 2
#define FSM_METRIC_ERP 3
#define FSM_METRIC_LCSS 4
#define FSM_METRIC_MSM 5
/* SWALE (include/funcs.h:17, src/ElasticMeasures.cpp:200-231): match cost 1
 * within eps, mismatch and gap cost 5; implemented by the reference but
 * rejected by its command line (src/main.cpp:57-59), so only this ABI takes it */
#define FSM_METRIC_SWALE 6

/* query modes: the reference's three commands (src/main.cpp:90-104) */
#define FSM_MODE_SEARCH 0 /* FSMELD()                 include/funcs.h:67-68 */
#define FSM_MODE_CUTOFF 1 /* FSMELD_find_threshold()  include/funcs.h:64-65 */
#define FSM_MODE_EXACT 2  /* FSMELD_exact(top_k = 0)  include/funcs.h:70-72 */
/* FSMELD_exact(top_k = k), src/fsmeld.cpp:1097-1107, :1273-1290: the exact
 * distance of every valid window (no filter), the k-th smallest of them as
 * the threshold (stats.threshold), and every window at or below it (ties
 * included), ascending by index */
#define FSM_MODE_EXACT_TOPK 3

typedef struct fsm_target fsm_target;
typedef struct fsm_result fsm_result;

typedef struct {
    int mode;              /* FSM_MODE_* */
    int metric;            /* FSM_METRIC_* */
    int m;                 /* query / window length, m >= 2 */
    int w;                 /* Sakoe-Chiba radius, 0 <= w < m */
    double arg;            /* ERP gap g, EDR/LCSS/SWALE epsilon, MSM cost c; ignored by DTW */
    double cutoff;         /* search / exact: distance cutoff xi >= 0 */
    int k;                 /* cutoff / exact top-k: rank k >= 1 */
    int64_t exclude_begin; /* candidate exclusion [begin, end); -1 = none */
    int64_t exclude_end;
    int raw;               /* 0: z-normalised matching (the reference's only mode,
                              NormalizationMode::ZScore, include/funcs.h:20-28);
                              1: raw matching (NormalizationMode::Raw, paper App. D):
                              query and windows as they are, a window is invalid only
                              when it holds a non-finite sample */
    int lbip_f32;          /* LB_IP's sliding inner products (src/sip.cpp:16-52):
                              0: exact 64-bit fixed-point sums of the FP64 penalties
                              rounded down;
                              1: an FP32 contraction on the matrix cores with a
                              certified rounding slack subtracted (a looser but still
                              valid lower bound), for m_k <= 8192 (m <= 1024); longer
                              segments keep the fixed-point sums (stats.lbip_f32) */
} fsm_query_desc;

typedef struct {
    uint64_t valid;       /* kk1: valid, non-excluded windows */
    uint64_t invalid;     /* windows skipped as invalid or excluded */
    uint64_t after_lbip;  /* kk2: survivors of LB_IP */
    uint64_t after_glb;   /* kk3: survivors of GLB */
    uint64_t exact_evals; /* exact elastic distances computed */
    uint64_t matches;     /* result count */
    double threshold;     /* cutoff mode: k-th smallest distance */
    int64_t ts_n, k_m_l, forsize, tail;
    int m_k;
    /* device time per stage from HIP events on the engine's stream, ms */
    double ms_stats, ms_lbip, ms_glb, ms_eld;
    double ms_total; /* host wall time of the call, ms */
    /* band cells the ELD kernel evaluated for live candidates (iterations of
     * the register-band kernel x (2w + 1)); 0 when the row-band kernel ran */
    uint64_t eld_cells;
    /* of ms_stats: the second statistics kernel (mean / std / invalid flag
     * from the running sums); the target envelope, when this query computed
     * it (ms_lbip no longer includes it) */
    double ms_stats_finish, ms_env;
    /* windows whose device distance lay within 1e-9 (relative) of the
     * decision boundary (the cutoff; for cutoff / top-k the k-th distance)
     * and were recomputed on the host in FP64, whose value then decided;
     * how many of those host values differed from the device's */
    uint64_t near_rechecked, near_changed;
    /* 1 when LB_IP's sums came from the FP32 matrix-core contraction */
    int lbip_f32;
} fsm_stats;

/* ---- device and errors ---- */
int fsm_device_count(void);
const char* fsm_last_error(void);

/* ---- device-resident target (replaces the reference's per-segment ifstream
 *      reads, src/fsmeld.cpp:740-746 and read_tail_segment :159-168) ---- */
int fsm_target_open(const double* host_values, int64_t n, fsm_target** out);
int fsm_target_open_file(const char* path, fsm_target** out);
int64_t fsm_target_length(const fsm_target* t);
void fsm_target_close(fsm_target* t);

/* Another handle to the same device-resident samples, with its own HIP
 * stream and working buffers: handles made this way may run queries
 * concurrently from different host threads (the reference runs one query per
 * process, src/main.cpp:40-110).  The samples are freed with the last handle. */
int fsm_target_share(fsm_target* base, fsm_target** out);

/* ---- query: search / cutoff / exact over the whole target ----
 * query_raw holds at least m raw (un-normalised) query values; the engine
 * z-normalises them as src/fsmeld.cpp:622-639 does.
 * search/exact: indices ascending, distances are the exact ELD values.
 * cutoff: the k (distance, index)-smallest windows in rank order, as printed
 *         by src/fsmeld.cpp:544-553; stats.threshold is the k-th distance. */
int fsm_query(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
              int64_t query_len, fsm_result** out);
int fsm_result_stats(const fsm_result* r, fsm_stats* out);

/* The same query over a target FILE of any size with at most device_budget
 * bytes of device memory (the reference's per-segment file streaming,
 * src/fsmeld.cpp:740-746, read_tail_segment :159-168): the file is read in
 * chunks of whole macro-segments (the last one EOF-aligned, as the
 * reference's tail segment), each uploaded while the previous chunk's
 * cascade runs; results, thresholds and counters equal fsm_query's on the
 * whole target.  FSM_ERR_OOM when the budget holds no one-segment chunk. */
int fsm_query_file(const char* path, const fsm_query_desc* desc, const double* query_raw,
                   int64_t query_len, uint64_t device_budget, fsm_result** out);
int64_t fsm_result_count(const fsm_result* r);
const int64_t* fsm_result_indices(const fsm_result* r);
const double* fsm_result_distances(const fsm_result* r);
void fsm_result_free(fsm_result* r);

/* ---- many queries over one target (no reference counterpart: the
 *      reference answers one query per process) ----
 * Runs queries[0 .. n) (query i: descs[i], query_lens[i] raw values) with up
 * to `concurrency` in flight, each on its own handle (kept in t between
 * calls) driven by its own host thread; results[i] and status[i] receive what
 * fsm_query would return for query i alone.  Returns FSM_OK once every query
 * has run, or an error if the handles could not be set up. */
int fsm_query_batch(fsm_target* t, int64_t n, const fsm_query_desc* descs,
                    const double* const* queries, const int64_t* query_lens, int concurrency,
                    fsm_result** results, int* status);

/* ---- launch shapes and the autotuner (no reference counterpart) ----
 * Every field 0 keeps the engine's built-in rule; values outside a field's
 * valid set are ignored where they would not fit (an LB_IP workgroup whose
 * LDS exceeds a CU's).  Shapes change speed only, never results. */
typedef struct {
    int lbip_threads[4]; /* LB_IP workgroup size (128, 256, 512, 1024) for segments of
                            m_k <= 1024, 2048, 4096, 8192 (longer: fixed 1024) */
    int eld_refill[7];   /* register-band ELD: idle lanes of a wave that trigger its
                            refill from the candidate queue, by FSM_METRIC_* (index 0
                            unused) */
    int glb_blocks;      /* GLB: persistent workgroups of 256 threads */
} fsm_tuning;

/* The target's launch shapes (shared by all its handles); set them while no
 * query runs on any handle of the target. */
int fsm_tuning_get(const fsm_target* t, fsm_tuning* out);
int fsm_tuning_set(fsm_target* t, const fsm_tuning* in);

/* Project-written autotuner: times the sample searches (descs[i].mode must be
 * FSM_MODE_SEARCH) under each candidate launch shape -- per LB_IP segment
 * class the workgroup sizes that fit, per metric the ELD refill thresholds
 * {4, 8, 12, 16, 24}, the GLB grid {1024, 2048, 4096, 8192} -- by the HIP-
 * event stage times of the stage each shape drives (the smallest of
 * `repeats` runs per search), keeps the fastest, checks that every candidate
 * returns the same matches, applies the result to the target and returns it
 * in *out.  Classes and metrics without a sample search keep the built-in
 * rule. */
int fsm_autotune(fsm_target* t, int64_t n, const fsm_query_desc* descs,
                 const double* const* queries, const int64_t* query_lens, int repeats,
                 fsm_tuning* out);

/* Device bytes a handle holds while running a query of (metric, m, w) over n
 * samples (working buffers only; the samples themselves are shared). */
uint64_t fsm_query_bytes(int64_t n, int metric, int m, int w);

/* Frees a handle's working buffers (they are re-created on its next query),
 * and the target's shared window statistics / envelopes that no handle is
 * using.  Those are kept per window length m and radius w for all handles of
 * a target (fsm_target_share) once a length or radius comes back within 64
 * requests, up to half of device memory, least recently used first out. */
void fsm_target_trim(fsm_target* t);

/* enable = 0: every query on this handle recomputes its window statistics
 * and target envelope into the handle's own buffers, as the reference
 * computes them for every query (src/fsmeld.cpp:750-773); enable = 1 (the
 * default): they are reused from an earlier query of the same m / w. */
int fsm_target_reuse_stats(fsm_target* t, int enable);

/* Caps the device bytes of that shared cache (default: half of device
 * memory); a new length or radius that would exceed the cap stays in the
 * requesting handle's own buffers.  Idle entries beyond a lowered cap are
 * freed at once; entries still in use are freed by later requests that need
 * the room, or by fsm_target_trim, once no handle uses them. */
int fsm_target_cache_limit(fsm_target* t, uint64_t bytes);

/* ---- serving loop (no reference counterpart) ----
 * A pool of worker threads, each driving its own handle that shares t's
 * device samples, serves submitted queries first come, first served.  A
 * quarter of mem_budget is the target's shared statistics / envelope cache
 * (fsm_target_cache_limit is set to it); a free worker starts the oldest
 * queued query as soon as the working memory of the running and cached
 * queries plus its own estimate (fsm_query_bytes) fits the other three
 * quarters; idle workers release cached buffers while the oldest query does
 * not fit.  Queries that arrive while others run are admitted as
 * workers and memory free up, and their kernels overlap on the device. */
#define FSM_PENDING 1             /* fsm_server_poll: not finished yet */
#define FSM_ERR_CANCELLED (-7)    /* the query was cancelled */
typedef struct fsm_server fsm_server;
int fsm_server_create(fsm_target* t, uint64_t mem_budget, int max_workers, fsm_server** out);
/* Copies the query; *ticket identifies it.  FSM_ERR_OOM if its estimate alone
 * exceeds the workers' three quarters of the budget.  fsm_query's own errors (argument validation included)
 * arrive as the *status of fsm_server_poll. */
int fsm_server_submit(fsm_server* s, const fsm_query_desc* desc, const double* query_raw,
                      int64_t query_len, uint64_t* ticket);
/* Waits up to wait_ms (0: no wait, < 0: until done).  FSM_OK: *out and *status
 * receive fsm_query's result and status, *latency_ms the time from submit to
 * completion, and the ticket is retired; FSM_PENDING: still queued / running;
 * FSM_ERR_CANCELLED: cancelled (ticket retired); FSM_ERR_ARG: unknown ticket. */
int fsm_server_poll(fsm_server* s, uint64_t ticket, int wait_ms, fsm_result** out, int* status,
                    double* latency_ms);
/* A queued query is dropped at once; a running one finishes on the device and
 * its result is discarded. */
int fsm_server_cancel(fsm_server* s, uint64_t ticket);
int fsm_server_info(fsm_server* s, int* workers, int64_t* queued, int64_t* running);
/* Cancels what is queued, waits for running queries, frees everything. */
void fsm_server_destroy(fsm_server* s);

/* ---- stage-level entry points (parity tests; outputs in host memory) ---- */

/* Per-window mean, std and invalid flag in global window order (ts_n each),
 * each from the recurrence of the macro-segment that evaluates the window.
 * Replaces mvmean() include/funcs.h:39 + precompute_invalid_mask_fast()
 * include/fsmeld_internal.h:22-25 (exclusion [xb, xe) applied as there). */
int fsm_stage_window_stats(fsm_target* t, int m, int64_t xb, int64_t xe, double* mu,
                           double* sigma, uint8_t* invalid);

/* LB_IP value per window (ts_n; NaN for invalid windows) and the per-segment
 * thresholds t_u, t_l (forsize + [tail > 0] each; may be NULL).  Replaces the
 * per-segment LB_IP stage of src/fsmeld.cpp:756-808 (min_max_window,
 * lower_upper_lemire, compute_normalized_arrays, find_t*_lp*, q_to_h_*,
 * s_to_g_*, inv_and_padding_y / FFT_x / cov_IFFT, trans). */
int fsm_stage_lbip(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                   double* lb, double* t_u, double* t_l);

/* GLB value for each given window (glb_elastic, include/funcs.h:84-85, called
 * as src/fsmeld.cpp:827-829 with bsf = desc->cutoff). */
int fsm_stage_glb(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                  const int64_t* windows, int64_t n, double* out);

/* GLB keep decision for each given window (keep[i] = 1 when
 * glb_elastic(...) <= desc->cutoff, src/fsmeld.cpp:827-831), computed by the
 * production filter kernel that search and cutoff use. */
int fsm_stage_glb_filter(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                         const int64_t* windows, int64_t n, uint8_t* keep);

/* Exact elastic distance for each given window
 * (exact_elastic_distance, include/fsmeld_internal.h:27). */
int fsm_stage_eld(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                  const int64_t* windows, int64_t n, double* out);

/* ---- stage-level entry points on device memory ----
 * The same stages with the window list and the outputs in DEVICE memory
 * (caller-owned), enqueued on the caller's `stream` (a hipStream_t; NULL: the
 * null stream): the target's own stream first waits for what `stream` has
 * enqueued, and `stream` then waits for the stage (events, no host
 * synchronisation), so the outputs are ready when `stream` reaches that
 * point.  The query stays a host array of m samples, normalised on the host
 * as the reference does (src/fsmeld.cpp:622-643).  The window-list entries
 * check their indices on the device first and read one word back
 * (FSM_ERR_ARG when one lies outside [0, ts_n)).  Outputs as in the host
 * entries above, element for element. */
int fsm_stage_window_stats_dev(fsm_target* t, int m, int64_t xb, int64_t xe, double* mu,
                               double* sigma, uint8_t* invalid, void* stream);
int fsm_stage_lbip_dev(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                       double* lb, double* t_u, double* t_l, void* stream);
int fsm_stage_glb_dev(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                      const int64_t* windows, int64_t n, double* out, void* stream);
int fsm_stage_glb_filter_dev(fsm_target* t, const fsm_query_desc* desc,
                             const double* query_raw, const int64_t* windows, int64_t n,
                             uint8_t* keep, void* stream);
int fsm_stage_eld_dev(fsm_target* t, const fsm_query_desc* desc, const double* query_raw,
                      const int64_t* windows, int64_t n, double* out, void* stream);

/* Exact elastic distance of n already-normalised rows cands[n][m] against q[m]
 * (dtw/erp/edr/lcss/msm, include/funcs.h:76-81, argument order of
 * src/fsmeld.cpp:170-183).  No target needed; all values must be finite
 * (FSM_ERR_ARG otherwise). */
int fsm_eld_pairs(int metric, int m, int w, double arg, const double* q, const double* cands,
                  int64_t n, double* out);

/* ---- the same without device allocation (no reference counterpart) ----
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc): fsm_eld_pairs with q[m],
 * cands[n][m] and out[n] in DEVICE memory, the kernels enqueued on `stream`
 * (a hipStream_t; NULL: the null stream) and no synchronisation, so out is
 * ready when the stream reaches that point.  No hipMalloc, hipFree or other
 * allocation: the gap costs and the work counter live in a static device
 * array of 256 + 8 * FSM_NOALLOC_MAX_M bytes per GPU, shared by every call on
 * that GPU, so calls that overlap on one GPU race.  Limits: m <=
 * FSM_NOALLOC_MAX_M and (m, w) on the register / wave band kernels (w <= 127,
 * m <= 4094); FSM_ERR_ARG beyond them.  Values must be finite (they are on
 * the device and not checked; non-finite ones give unspecified distances). */
#define FSM_NOALLOC_MAX_M 4096
int fsm_eld_pairs_noalloc(int metric, int m, int w, double arg, const double* q,
                          const double* cands, int64_t n, double* out, void* stream);

/* ---- seeded synthetic workloads (host only; no GPU needed) ----
 * SplitMix64-seeded xoshiro256** with Box-Muller normals.
 * Target: random walk of N(0,1) increments; spike_frac of positions get an
 * added N(0, spike_sd) outlier; n_const_runs runs t[a..a+const_len) = t[a];
 * n_nans single NaN samples.  Query: a copy of m target values from a seeded
 * offset whose window is finite and free of repeated samples, plus
 * N(0, noise_sd) noise; *offset receives the source position. */
int fsm_synth_target(uint64_t seed, int64_t n, double spike_frac, double spike_sd,
                     int n_const_runs, int const_len, int n_nans, double* out);
int fsm_synth_query(uint64_t seed, const double* target, int64_t n, int m, double noise_sd,
                    double* out_query, int64_t* offset);

#ifdef __cplusplus
}
#endif
#endif
