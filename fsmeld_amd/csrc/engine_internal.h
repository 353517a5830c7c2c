// engine_internal.h -- types shared by the gfx950 kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsm {

enum Metric : int { DTW = 1, EDR = 2, ERP = 3, LCSS = 4, MSM = 5, SWALE = 6 };
// include/funcs.h:30-31: SWALE's cost of a match within eps (the reward r)
// and of a mismatch or a gap (p)
constexpr double kSwaleR = 1.0;
constexpr double kSwaleP = 5.0;

// src/ElasticMeasures.cpp:7 -- value of out-of-band DP cells.
constexpr double kEldSentinel = 1e18;
// include/def.h:7 -- value glb_elastic returns when it abandons early.
constexpr double kGlbAbandon = 1008610086.0;

// Macro-segmentation of one query length over a target of N samples
// (src/fsmeld.cpp:28-37, :647-661).  Segment s < forsize starts at s*kml;
// the tail segment (index forsize, present when tail > 0) starts at N - m_k
// and evaluates only its windows [tail_first, kml).
struct Geometry {
    int64_t N;
    int m;
    int m_k;
    int64_t kml;
    int64_t ts_n;
    int64_t forsize;
    int64_t tail;
    int tail_first;
    int64_t S;  // forsize + (tail > 0)
};

// The macro-segmentation of a target of N samples for length m; false when
// the target is too short (N < 2m - 1) or m_k exceeds 2^30.
bool make_geometry(int64_t N, int m, Geometry* G);

// Device-side query tables, uploaded once per query.
struct QueryTables {
    const double* qn;   // z-normalised query [m]
    const double* LQ;   // query envelope, radius w [m]
    const double* UQ;
    const double* qs;   // sorted qn [m]
    const double* S1;   // prefix sums over qs of q, q^2, (q - arg)^2 [m + 1]
    const double* S2;
    const double* SD;
    const double* TU;   // fl(qs - arg), fl(qs + arg) [m]
    const double* TL;
    const double* XG;   // (qn - arg)^2 [m], the ERP gap cost of each query sample
    const double* MQ;   // fl(LQ + UQ) / 2, the envelope midpoints [m]
    const double* HQ;   // h = fl(UQ - LQ) / 2: fl(h + arg) (EDR, LCSS), fl(h glb_msm_rho(arg))
                        // (MSM), h (others) [m]
    double qabs;        // max |qn|, which also bounds |LQ| and |UQ|
};

// 1 / c for GLB's MSM terms, which it measures in units of c.  Outside this
// range 1 / c would leave the normal range, and the fast pass is not used;
// c = 0 gives 0, which makes every term 0 (as min(d, 0) is).
__host__ __device__ inline double glb_msm_rho(double c) {
    return (c >= 1e-100 && c <= 1e100) ? 1.0 / c : 0.0;
}

struct QueryParams {
    int metric;
    int m;
    int w;
    double arg;
    double cutoff;
    int64_t xb, xe;  // exclusion [xb, xe), xb < 0: none
    // LB_IP's sliding sums: 0 exact fixed point, 1 the FP32 matrix-core
    // contraction with certified slack (segments of m_k <= 8192; longer ones
    // keep the fixed-point sums)
    int lbip_f32;
};

// Window statistics live segment-major: entry s*kml + j is window j of
// segment s.  Global window g maps to that layout by stat_index().
__host__ __device__ inline int64_t stat_index(const Geometry& G, int64_t g) {
    const int64_t full = G.forsize * G.kml;
    return g < full ? g : g + G.tail_first;
}

__host__ __device__ inline bool is_excluded(const QueryParams& P, int64_t g) {
    return P.xb >= 0 && P.xe > P.xb && g >= P.xb && g < P.xe;
}

// NaN-propagating min / max (any NaN operand gives NaN).
__host__ __device__ inline double nmin(double a, double b) { return (a < b || a != a) ? a : b; }
__host__ __device__ inline double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

// Block-aggregated append: every thread of the block calls it (uniform
// control flow); lanes with keep = true write value (and dval) to
// consecutive slots of out / outd claimed with one atomicAdd per block.
// sh must hold blockDim.x / 64 + 1 entries.
__device__ inline void block_append(bool keep, int64_t value, double dval, int64_t* out,
                                    double* outd, unsigned long long* count,
                                    unsigned long long* sh) {
    const uint64_t mask = __ballot(keep);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (int)(blockDim.x >> 6);
    if (lane == 0) sh[wid] = (unsigned long long)__builtin_popcountll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int k = 0; k < nw; k++) {
            const unsigned long long c = sh[k];
            sh[k] = tot;
            tot += c;
        }
        sh[nw] = tot ? atomicAdd(count, tot) : 0ull;
    }
    __syncthreads();
    if (keep) {
        const uint64_t below = lane ? (mask & ((1ull << lane) - 1)) : 0ull;
        const unsigned long long pos = sh[nw] + sh[wid] + (unsigned long long)__builtin_popcountll(below);
        out[pos] = value;
        if (outd) outd[pos] = dval;
    }
    __syncthreads();
}

// Exact elastic distance of one window on the host in FP64 (x the query, y
// the normalised candidate), for re-verification near a decision boundary.
double host_eld(int metric, const double* x, const double* y, int m, int w, double arg);

// Records msg as the calling thread's fsm_last_error() and returns code.
int report_error(int code, const char* msg);

// ------------------------------------------------------------ launchers ---
// k_window_sums then k_window_finish (raw: mean 0, std 1 and only the
// non-finite test); `between` (or null) is recorded between the two
hipError_t launch_window_stats(hipStream_t st, const double* T, const Geometry& G, double* MU,
                               double* SG, uint8_t* INV, hipEvent_t between = nullptr,
                               bool raw = false);
hipError_t launch_envelope(hipStream_t st, const double* T, int64_t N, int w, double* Lraw,
                           double* Uraw);

struct LbipArgs {
    const double* T;
    const double* MU;
    const double* SG;
    const uint8_t* INV;
    const double* Lraw;
    const double* Uraw;
    QueryTables q;
    QueryParams p;
    Geometry G;
    double* LB;        // [ts_n]
    double* TUo;       // [S] or null
    double* TLo;       // [S] or null
    double* rep_lb;    // [S] or null: smallest LB of the segment's valid windows
    int64_t* rep_idx;  // [S] or null
    void* slab;        // per-block global scratch
    size_t slab_bytes_per_block;
    int blocks;
    int nt;            // workgroup size for the kDual layout (0: the built-in rule)
};
size_t lbip_slab_bytes(const Geometry& G);
// nt: a kDual workgroup size (lbip_threads_ok), 0 for the built-in rule
size_t lbip_lds_bytes(const Geometry& G, int nt = 0);
// the autotuner's class of G (0: m_k <= 1024, 1: 2048, 2: 4096, 3: 8192) or
// -1 when G's layout has a fixed workgroup size
int lbip_tune_class(const Geometry& G);
// true when LB_IP may run G's segments in workgroups of nt threads
bool lbip_threads_ok(const Geometry& G, int nt);
// true when the FP32 contraction serves queries of this geometry
bool lbip_f32_supported(const Geometry& G);
hipError_t launch_lbip(hipStream_t st, const LbipArgs& a);

// Compaction: windows g in [0, ts_n) that are valid, not excluded and (for
// use_lb) not pruned by the LB_IP test at P.cutoff.  counters[0] += kept,
// counters[1] += valid.
hipError_t launch_compact(hipStream_t st, const Geometry& G, const QueryParams& P,
                          const uint8_t* INV, const double* LB, bool use_lb, int64_t* out,
                          unsigned long long* counters);

// list[i] for i < *count with sign * keys[i] <= tau, appended to out.
hipError_t launch_select_le(hipStream_t st, const int64_t* list, const double* keys,
                            const unsigned long long* count, double sign, double tau,
                            int64_t* out, unsigned long long* out_count);

// The k-th smallest finite value of keys[0 .. *count) (radix select on the
// device, one host round trip per digit; hist: 2048 device counters).
// *n_finite receives the number of finite keys; *kth is left alone when
// that is below k.
hipError_t kth_smallest(hipStream_t st, const double* keys, const unsigned long long* count,
                        int64_t k, unsigned long long* hist, double* kth, int64_t* n_finite);
// (list[i], keys[i]) for i < *count with keys[i] <= tau, appended to out /
// out_keys (unordered).
hipError_t launch_select_le_kv(hipStream_t st, const int64_t* list, const double* keys,
                               const unsigned long long* count, double tau, int64_t* out,
                               double* out_keys, unsigned long long* out_count);

struct VerifyArgs {
    const double* T;
    const double* MU;
    const double* SG;
    const double* Lraw;
    const double* Uraw;
    QueryTables q;
    QueryParams p;
    Geometry G;
    // launch shapes chosen by the autotuner (0: the built-in values): GLB's
    // persistent workgroups; the register-band ELD's idle lanes per refill
    int glb_blocks;
    int eld_refill;
};
// GLB filter over in[0 .. *in_count): survivors appended to out (out_count),
// with, when out_total != null, each survivor's certified consumption bound
// for the ELD abandon test beside it (out_total[k] belongs to out[k]); values
// written to vals[i] when vals != null (stage-level entry point).  queue: one
// device counter the launch resets and uses to hand out candidates.
hipError_t launch_glb(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const unsigned long long* in_count, int64_t* out,
                      unsigned long long* out_count, double* out_total, double* vals,
                      unsigned long long* queue);
// Exact distances over in[0 .. *in_count) (in_total: the GLB bounds of in[],
// or null).  keep_all: write dist[i] for every input; else append (index,
// distance) for distances <= P.cutoff.  queue: one device counter the launch
// resets and uses to hand out candidates.
// work (or null) += lane-iterations spent on live candidates by the
// register-band kernel.
hipError_t launch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const double* in_total, const unsigned long long* in_count, bool keep_all,
                      int64_t* out_idx, double* out_dist, unsigned long long* out_count,
                      unsigned long long* queue, unsigned long long* work, double* scratch,
                      int64_t scratch_threads);
// Exact distances of normalised rows cands[n][m] (no target); a.q.qn / a.q.XG
// and a.p.{metric, m, w, arg} describe the query.
hipError_t launch_eld_pairs(hipStream_t st, const VerifyArgs& a, const double* cands,
                            int64_t n, double* out, unsigned long long* queue, double* scratch,
                            int64_t scratch_threads);
int64_t eld_scratch_threads();
// true when (metric, m, w) runs the row-band path, which needs the scratch
// buffer of 2 (2w + 2) doubles per thread of eld_scratch_threads()
bool eld_uses_scratch(int metric, int m, int w);

}  // namespace fsm
