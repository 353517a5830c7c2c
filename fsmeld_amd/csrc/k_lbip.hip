// k_lbip.hip -- LB_IP stage on gfx950: one workgroup per macro-segment.
//
// Per segment (src/fsmeld.cpp:756-808):
//   1. min/max of window mean and std over the windows containing each
//      position (min_max_window, src/lower_upper_lemire.cpp:94-126): per tile
//      of 64 positions, wave scans of the window runs at the tile's ends and
//      the extremes of the whole 64-window chunks between (segments up to
//      16384 positions, everything in LDS); van Herk prefix/suffix scans over
//      blocks of m windows in global scratch for longer segments;
//   2. centred envelope L/U of the raw segment (lower_upper_lemire, :8-44),
//      taken from the target-wide envelope except within w of the edges;
//   3. normalised envelopes (src/lb_normalize.cpp:9-32);
//   4. thresholds t_u, t_l: the argmax over the segment's envelope values of
//      h(t) * G(t), the objective of src/find_t_lp.cpp (h(t) in closed form
//      over the sorted query; the ranks G of the few values whose bucket can
//      hold the maximum, from a sort of those values only: select_threshold);
//   5. penalties h (src/q_to_h.cpp) in exact 64-bit fixed point, rounded
//      down, and their prefix sums CH;
//   6. masks g (src/lb.cpp) as bitmaps built with wavefront ballots, and
//      the ascending lists of their transitions;
//   7. per window j: z = sum over runs [a, b) of set mask bits inside
//      [j, j+m) of CH[b-j] - CH[a-j] (walked as the mask's transitions, a
//      wave's 64 windows together: list_sum) -- the sliding inner product of
//      src/sip.cpp evaluated exactly, so the bound is a certified lower bound
//      of the FP64 penalties; then trans (src/fsmeld.cpp:84-106).
// Min / max / compare steps are exact, so for NaN-free segments steps 1-3
// reproduce the reference's bits.  Near NaN samples the min / max propagate
// NaN, which only clears mask bits (admissible).
#include <algorithm>

#include "engine_internal.h"

namespace fsm {

__host__ __device__ inline int npad_of(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Workgroup size per segment: 128 threads up to 1024 keys when there are
// enough segments to fill the device with two-wave workgroups (short segments
// are then dominated by the workgroup's barriers), 256 up to 2048 (the
// autotuner's choice for 2048 keys on every workload it was run on), 512 up
// to 4096 (two workgroups and 16 waves per CU with both envelopes in LDS),
// 1024 beyond, so that long segments spread their per-position work over more
// waves.  (npad, the keys' power of two, equals m_k.)
inline int lbip_threads(int npad, int64_t segments) {
    if (npad <= 1024) return segments >= 4096 ? 128 : 256;
    if (npad <= 2048) return 256;
    return npad <= 4096 ? 512 : 1024;
}


// Threshold selection by buckets (select_threshold): thr_buckets buckets of
// value-ordered keys and a candidate list of up to thr_cap keys, sized per
// workgroup shape to what its share of the CU's LDS leaves: 128 threads
// (eight workgroups per CU): 64 buckets, npad / 8 candidates; 256 (m_k <=
// 1024): 256 buckets, npad / 2; 512 (m_k 2048 - 4096, two workgroups per CU
// beside up to 64 KiB of envelopes) and 1024: 256 buckets, npad / 8.  (At
// most one bucket per thread: the bucket scans run a thread per bucket.)
// For 512 threads, 256 buckets instead of 128 halve the keys that reach the
// candidate sort (more of them take the one-wave sort) for the same bucket
// scoring time: LB_IP 4.86 -> 4.73 ms per 2^26 search at m = 512.
__host__ __device__ inline int thr_buckets(int npad, int nt) {
    (void)npad;
    return nt == 128 ? 64 : 256;
}
__host__ __device__ inline int thr_cap(int npad, int nt) {
    const int c = nt == 256 ? npad / 2 : npad / 8;
    return c > 64 ? c : 64;
}
__host__ __device__ inline size_t thr_lds_bytes(int npad, int nt) {
    return sizeof(double) * (size_t)thr_cap(npad, nt) +
           (size_t)thr_buckets(npad, nt) * (2 * sizeof(uint64_t) + 4 * sizeof(int));
}

// Layouts by segment length.  kDual (m_k <= 8192): U_n and -L_n both in LDS
// (two m_k arrays), from one pass over the window statistics; with m_k = 4096
// and 512 threads, 71 KiB, two workgroups per CU.  kSingle (m_k = 16384): one
// m_k array, U_n for the upper threshold and mask, then -L_n (from one pass
// over the statistics that parks -L_n in the block's global scratch);
// 156 KiB with 1024 threads.  Beside the arrays: a zone for the chunk
// extremes and then the threshold scratch, and the two mask bitmaps.
// kSlab (longer segments): van Herk arrays and both envelopes in the block's
// global scratch.
enum LbipLayout { kDual = 0, kSingle = 1, kSlab = 2 };
__host__ __device__ inline int lbip_layout(int mk) {
    return mk <= 8192 ? kDual : (mk <= 16384 ? kSingle : kSlab);
}
// Sparse-table levels above the 64-window chunk extremes: a tile's middle
// spans at most m / 64 - 1 whole chunks, so levels 1 .. floor(log2(m / 64))
// (ranges of 2^k chunks) answer it with two reads per quantity (kDual).
__host__ __device__ inline int chunk_levels(int m) {
    int k = 0;
    while ((2 << k) <= m / 64) k++;
    return k;
}
// kSingle has no LDS to spare: as many levels as fit the threshold scratch
// that shares the zone (a tile's middle then takes about len / 2^levels
// reads per quantity).
__host__ __device__ inline int chunk_levels_in(int m, int kml, size_t budget) {
    const size_t c = 4 * sizeof(double) * (size_t)((kml + 63) / 64);
    int k = chunk_levels(m);
    while (k > 0 && c * (size_t)(1 + k) > budget) k--;
    return k;
}
__host__ __device__ inline size_t zone_bytes(int npad, int nt, int kml, int levels, int layout) {
    // kDual: the levels and the chunks' suffixes from one offset (chunk_extremes)
    const size_t c = 4 * sizeof(double) * (size_t)((kml + 63) / 64) *
                     (size_t)(1 + levels + (layout == kDual ? 1 : 0));
    const size_t t = thr_lds_bytes(npad, nt);
    return c > t ? c : t;
}
// The levels each layout keeps (kSlab computes its ranges otherwise).
__host__ __device__ inline int lbip_levels(int layout, int m, int kml, int npad, int nt) {
    if (layout == kDual) return chunk_levels(m);
    if (layout == kSingle) return chunk_levels_in(m, kml, thr_lds_bytes(npad, nt));
    return 0;
}

// FP32 matrix-core bounds (kDual only): words read past the end of the
// upper mask bitmap (at most 7, see f32_bounds)
constexpr int kMaskPadWords = 8;

size_t lbip_lds_bytes(const Geometry& G, int nt_want) {
    const int nthr = G.m_k - 2 * G.m + 2;
    const int npad = npad_of(nthr);
    const int nw = (G.m_k + 63) / 64;
    const int nt = nt_want ? nt_want : lbip_threads(npad, G.S);
    const int lay = lbip_layout(G.m_k);
    // (+ kMaskPadWords: the FP32 contraction reads a few words past the
    // upper mask's end, whose bits only ever meet zero penalties)
    if (lay != kSlab)
        return sizeof(double) * (size_t)G.m_k * (lay == kDual ? 2 : 1) +
               zone_bytes(npad, nt, (int)G.kml, lbip_levels(lay, G.m, (int)G.kml, npad, nt), lay) +
               sizeof(uint64_t) * (2 * (size_t)nw + kMaskPadWords);
    // kSlab: the keys in global scratch; the threshold scratch, then the
    // penalty prefix sums and the mask bitmaps, in LDS
    const size_t late = sizeof(int64_t) * 2 * (size_t)(G.m + 1) + sizeof(uint64_t) * 4 * (size_t)nw;
    return std::max(late, thr_lds_bytes(npad, nt));
}

// Per-block global scratch: the threshold's full-sort buffer (npad keys);
// with kSlab also the van Herk arrays and both envelopes.
bool lbip_f32_supported(const Geometry& G) { return lbip_layout(G.m_k) == kDual; }

int lbip_tune_class(const Geometry& G) {
    if (lbip_layout(G.m_k) != kDual) return -1;
    return G.m_k <= 1024 ? 0 : (G.m_k <= 2048 ? 1 : (G.m_k <= 4096 ? 2 : 3));
}

// kDual runs in any of the four workgroup sizes whose LDS fits a CU; the
// mask transition lists take one bitmap word per thread (nw <= nt holds for
// m_k <= 8192 and nt >= 128)
bool lbip_threads_ok(const Geometry& G, int nt) {
    if (lbip_layout(G.m_k) != kDual) return false;
    if (nt != 128 && nt != 256 && nt != 512 && nt != 1024) return false;
    return (G.m_k + 63) / 64 <= nt && lbip_lds_bytes(G, nt) <= 160 * 1024;
}

size_t lbip_slab_bytes(const Geometry& G) {
    const int nthr = G.m_k - 2 * G.m + 2;
    const int npad = npad_of(nthr);
    size_t b = sizeof(double) * (size_t)npad;
    if (lbip_layout(G.m_k) == kSingle)  // -L_n while U_n's threshold is chosen
        b += sizeof(double) * (size_t)G.m_k;
    if (lbip_layout(G.m_k) == kSlab)
        b += sizeof(double) * (8 * (size_t)G.kml + 2 * (size_t)G.m_k);
    return (b + 255) & ~(size_t)255;
}

// ---------------------------------------------------------------- helpers ---
__device__ inline int upper_bound(const double* a, int n, double x) {  // first a[i] > x
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo;
}
__device__ inline int lower_bound(const double* a, int n, double x) {  // first a[i] >= x
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] >= x) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// sum over qs[a..b) of (q - t)^2 from prefix sums
__device__ inline double quad_range(const QueryTables& Q, int a, int b, double t) {
    const double s1 = Q.S1[b] - Q.S1[a];
    const double s2 = Q.S2[b] - Q.S2[a];
    const double c = (double)(b - a);
    return s2 - s1 * t * 2 + c * (t * t);
}

// Objective h(t) for the upper threshold (src/find_t_lp.cpp:583-825):
// DTW sum_{q>t}(q-t)^2; ERP sum_{q>t} min((q-t)^2,(q-g)^2);
// MSM sum_{q>t} min(q-t, c); EDR/LCSS #{q - eps > t}.
__device__ double h_upper(const QueryTables& Q, int metric, int m, double C, double t) {
    switch (metric) {
        case DTW: {
            const int i = upper_bound(Q.qs, m, t);
            return quad_range(Q, i, m, t);
        }
        case ERP: {
            const int i = upper_bound(Q.qs, m, t);
            if (t < C) {
                const double x = (t + C) * 0.5;
                const int f = upper_bound(Q.qs, m, x);
                return (Q.SD[m] - Q.SD[f]) + quad_range(Q, i, f > i ? f : i, t);
            }
            return quad_range(Q, i, m, t);
        }
        case MSM: {
            const int i1 = upper_bound(Q.qs, m, t);
            const int i2 = upper_bound(Q.TU, m, t);
            const int lo = i1 < i2 ? i1 : i2;
            const double mid = (Q.S1[i2] - Q.S1[lo]) - (double)(i2 - lo) * t;
            return C * (double)(m - i2) + mid;
        }
        case SWALE: {  // src/find_t_lp.cpp:827-927: p per q - eps > t, r per q >= t beyond
            const int big = m - upper_bound(Q.TU, m, t);   // fl(q - eps) > t
            const int ge = m - lower_bound(Q.qs, m, t);    // q >= t (includes big)
            return kSwaleP * (double)big + kSwaleR * (double)(ge - big);
        }
        default:  // EDR, LCSS
            return (double)(m - upper_bound(Q.TU, m, t));
    }
}

// Objective h(t) for the lower threshold (src/find_t_lp.cpp:218-582).
__device__ double h_lower(const QueryTables& Q, int metric, int m, double C, double t) {
    switch (metric) {
        case DTW: {
            const int i = lower_bound(Q.qs, m, t);
            return fabs(quad_range(Q, 0, i, t));
        }
        case ERP: {
            const int i = lower_bound(Q.qs, m, t);
            if (t > C) {
                const double x = (t + C) * 0.5;
                const int f = lower_bound(Q.qs, m, x);
                const int ff = f < i ? f : i;
                return Q.SD[ff] + fabs(quad_range(Q, ff, i, t));
            }
            return fabs(quad_range(Q, 0, i, t));
        }
        case MSM: {
            const int is = lower_bound(Q.TL, m, t);
            const int iq = lower_bound(Q.qs, m, t);
            const int hi = iq > is ? iq : is;
            const double mid = (Q.S1[hi] - Q.S1[is]) - (double)(hi - is) * t;
            return C * (double)is + fabs(mid);
        }
        case SWALE: {  // src/find_t_lp.cpp:399-509: p per q + eps < t, r per q <= t beyond
            const int small = lower_bound(Q.TL, m, t);     // fl(q + eps) < t
            const int le = upper_bound(Q.qs, m, t);        // q <= t (includes small)
            return kSwaleP * (double)small + kSwaleR * (double)(le - small);
        }
        default:
            return (double)lower_bound(Q.TL, m, t);
    }
}

// src/q_to_h.cpp:11-134 for one query point.
__device__ inline void penalties(int metric, double q, double tl, double tu, double C,
                                 double* hl, double* hu) {
    double vl = 0.0, vu = 0.0;
    switch (metric) {
        case MSM:
            if (q < tl) vl = fmin(C, tl - q);
            if (q > tu) vu = fmin(C, q - tu);
            break;
        case DTW:
            if (q < tl) vl = (tl - q) * (tl - q);
            if (q > tu) vu = (q - tu) * (q - tu);
            break;
        case ERP:
            if (q < tl)
                vl = ((q - tl) * (q - tl) > (q - C) * (q - C)) ? (q - C) * (q - C)
                                                                : (q - tl) * (q - tl);
            if (q > tu)
                vu = ((q - tu) * (q - tu) > (q - C) * (q - C)) ? (q - C) * (q - C)
                                                                : (q - tu) * (q - tu);
            break;
        case SWALE:  // src/q_to_h.cpp:60-82, :128-150
            if (q < tl - C) vl = kSwaleP;
            else if (q < tl) vl = kSwaleR;
            if (q > tu + C) vu = kSwaleP;
            else if (q > tu) vu = kSwaleR;
            break;
        default:  // EDR, LCSS
            if (q < tl - C) vl = 1.0;
            if (q > tu + C) vu = 1.0;
            break;
    }
    *hl = vl;
    *hu = vu;
}

// std::min(C, x) of the reference returns C when equal; fmin returns the
// same value, and penalties are never NaN for finite t.

// min / max of a window statistic over windows [lo, hi] (hi - lo < m) from
// the van Herk prefix value at hi (ph) and suffix value at lo (sl)
__device__ inline double stat_range(double ph, double sl, int64_t lo, int64_t hi, int m,
                                    bool is_min) {
    const int64_t blo = lo / m, bhi = hi / m;
    if (blo != bhi) return is_min ? nmin(sl, ph) : nmax(sl, ph);
    if (lo == blo * m) return ph;
    return sl;
}

// Bitonic sort, ascending, of keys that are never NaN (the threshold choice
// maps NaN to +inf and counts it out), so the values compare totally and
// each comparator is one min and one max.  Every comparator is ascending: the
// first step of each merge pairs index i with its mirror i ^ (k - 1) in the
// k-block, the later steps with i ^ j.  Equal values (+0 / -0 included) may
// leave in either order; the thresholds read values and ranks only, as
// std::sort leaves equal keys unordered.

// min / max of two values that are never NaN (a compare and two selects,
// without fmin's NaN quieting)
__device__ __forceinline__ double vmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double vmax(double a, double b) { return (b < a) ? a : b; }

// In LDS or global scratch, any power-of-two n.
__device__ void bitonic_sort(double* keys, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const bool mirror = j == (k >> 1);
            for (int q = threadIdx.x; q < (n >> 1); q += blockDim.x) {
                // bit j of i clear (j a power of two: q / j, q % j as bit operations)
                const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const int p = mirror ? (i ^ (k - 1)) : (i + j);
                const double a = keys[i], b = keys[p];
                keys[i] = vmin(a, b);
                keys[p] = vmax(a, b);
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}

// Block argmax with "first index wins" among equal scores; NaN never wins.
struct ArgBest {
    double v;
    int64_t i;
};
__device__ inline ArgBest better_max(ArgBest a, ArgBest b) {
    if (b.v != b.v) return a;
    if (a.v != a.v) return b;
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}
__device__ inline ArgBest better_min(ArgBest a, ArgBest b) {
    if (b.v != b.v) return a;
    if (a.v != a.v) return b;
    if (b.v < a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}

template <bool MAX>
__device__ ArgBest block_best(ArgBest x, ArgBest* sh) {
    for (int o = 32; o > 0; o >>= 1) {
        ArgBest y;
        y.v = __shfl_xor(x.v, o, 64);
        y.i = __shfl_xor(x.i, o, 64);
        x = MAX ? better_max(x, y) : better_min(x, y);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wid] = x;
    __syncthreads();
    ArgBest r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++)
        r = MAX ? better_max(r, sh[k]) : better_min(r, sh[k]);
    __syncthreads();
    return r;
}

// Block-wide sum of one int per thread (ends with a barrier).
__device__ int block_count(int c, int* sh) {
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
    __syncthreads();
    return r;
}

__device__ double block_sum(double x, double* sh) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wid] = x;
    __syncthreads();
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
    __syncthreads();
    return r;
}

// Inclusive NaN-propagating min and max scans over the wave's 64 lanes in lane
// order, with gfx9 data-parallel primitives: row_shr 1, 2, 4, 8 within each
// 16-lane row, then row_bcast:15 (lane 15 of rows 0 / 2 into rows 1 / 3) and
// row_bcast:31 (lane 31 into rows 2 and 3).  A lane without a source, or in a
// row outside the step's row mask, keeps the identity (+inf for the min, -inf
// for the max).  min / max are associative, so the values equal a sequential
// scan's (up to the sign of a zero and NaN payloads).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64_or(double v, double id) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long o = (unsigned long long)__double_as_longlong(id);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, CTRL, ROWS, 0xf,
                                               false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32),
                                               CTRL, ROWS, 0xf, false);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void scan_step(double& vmin, double& vmax) {
    const double a = dpp_f64_or<CTRL, ROWS>(vmin, __builtin_inf());
    const double b = dpp_f64_or<CTRL, ROWS>(vmax, -__builtin_inf());
    vmin = nmin(a, vmin);
    vmax = nmax(b, vmax);
}
__device__ __forceinline__ void wave_scan_min_max(double& vmin, double& vmax) {
    scan_step<0x111, 0xf>(vmin, vmax);  // row_shr:1
    scan_step<0x112, 0xf>(vmin, vmax);  // row_shr:2
    scan_step<0x114, 0xf>(vmin, vmax);  // row_shr:4
    scan_step<0x118, 0xf>(vmin, vmax);  // row_shr:8
    scan_step<0x142, 0xa>(vmin, vmax);  // row_bcast:15
    scan_step<0x143, 0xc>(vmin, vmax);  // row_bcast:31
}
// lane 63's value, in every lane
__device__ __forceinline__ double lane63(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Inclusive min (MN) and / or max (MX) scans over the wave's lanes in lane
// order, for NaN-free values (a lane without a source keeps the identity).
template <int CTRL, int ROWS, bool MN, bool MX>
__device__ __forceinline__ void scan_step_nf(double& vmn, double& vmx) {
    if (MN) vmn = vmin(vmn, dpp_f64_or<CTRL, ROWS>(vmn, __builtin_inf()));
    if (MX) vmx = vmax(vmx, dpp_f64_or<CTRL, ROWS>(vmx, -__builtin_inf()));
}
template <bool MN, bool MX>
__device__ __forceinline__ void wave_scan_nf(double& vmn, double& vmx) {
    scan_step_nf<0x111, 0xf, MN, MX>(vmn, vmx);
    scan_step_nf<0x112, 0xf, MN, MX>(vmn, vmx);
    scan_step_nf<0x114, 0xf, MN, MX>(vmn, vmx);
    scan_step_nf<0x118, 0xf, MN, MX>(vmn, vmx);
    scan_step_nf<0x142, 0xa, MN, MX>(vmn, vmx);
    scan_step_nf<0x143, 0xc, MN, MX>(vmn, vmx);
}

// ------------------------------------------ window-statistic ranges, LDS ---
// Position i of a segment is covered by the windows [lo(i), hi(i)] =
// [max(0, i - m + 1), min(i, k_m_l - 1)] (src/lower_upper_lemire.cpp:94-126).
// A wave takes 64 consecutive positions p0 .. p0 + 63 (a tile; p0 a multiple
// of 64, m >= 64).  With A / B the tile's first / last position, each range
// splits into
//   [lo(i), lo(B))     a suffix of the left run [lo(A), lo(B)),
//   [lo(B), hi(A)]     the middle shared by the tile (not empty for m >= 64),
//   (hi(A), hi(i)]     a prefix of the right run (hi(A), hi(B)].
// The runs are at most 63 windows: one load per lane and a wave scan each
// (the left run loaded in reverse, so its scan gives suffixes).  The middle
// is a partial 64-window chunk at either end (one load per lane each) and
// the whole chunks between, whose extremes a first pass put in LDS.
// NaN statistics are tracked apart (ballots of the NaN lanes; chunk extremes
// are NaN when the chunk holds one) and make the result NaN, as the
// NaN-propagating min / max of a sequential scan do; the finite extremes are
// exact, so the values equal the sequential scan's up to the sign of a zero
// and NaN payloads.
// The tile's window geometry (wave-uniform but for lo / hi) and its load
// indices, each clamped into the run it reads so that every lane loads
// unconditionally (the in* flags say which loads count).
struct TileGeom {
    int loB, hiA, lo, hi, ca, cb;
    int il, ir, ih, it;
    bool inl, inr, inh, intl;
    bool interior;  // neither range end clamped: middle = [p0 + 64 - m, p0]
};
__device__ __forceinline__ TileGeom tile_geom(int p0, int m, int kml, int mk, int lane) {
    TileGeom g;
    const int iB = (p0 + 63 < mk - 1) ? p0 + 63 : mk - 1;
    const int i = (p0 + lane < iB) ? p0 + lane : iB;
    const int loA = (p0 - m + 1 > 0) ? p0 - m + 1 : 0;
    g.loB = (iB - m + 1 > 0) ? iB - m + 1 : 0;
    g.hiA = (p0 < kml - 1) ? p0 : kml - 1;
    const int hiB = (iB < kml - 1) ? iB : kml - 1;
    g.lo = (i - m + 1 > 0) ? i - m + 1 : 0;
    g.hi = (i < kml - 1) ? i : kml - 1;
    g.ca = g.loB >> 6;
    g.cb = g.hiA >> 6;
    g.inl = lane < g.loB - loA;
    g.inr = lane < hiB - g.hiA;
    g.inh = (g.ca != g.cb) ? (lane < 64 - (g.loB & 63)) : (g.loB + lane <= g.hiA);
    g.intl = (g.ca != g.cb) && lane <= (g.hiA & 63);
    g.il = (g.loB - 1 - lane > loA) ? g.loB - 1 - lane : loA;
    g.ir = (g.hiA + 1 + lane < hiB) ? g.hiA + 1 + lane : hiB;
    g.ih = (g.loB + lane < g.hiA) ? g.loB + lane : g.hiA;
    g.it = ((g.cb << 6) + lane < g.hiA) ? (g.cb << 6) + lane : g.hiA;
    g.interior = p0 + 64 - m >= 0 && p0 <= kml - 1;
    return g;
}
struct TileVals {
    double l, r, h, t;
};
__device__ __forceinline__ TileVals tile_load(const double* __restrict__ src, const TileGeom& g) {
    return TileVals{src[g.il], src[g.ir], src[g.ih], src[g.it]};
}

// Scan part of one statistic: the left / right runs' scans per lane and the
// middle's partial chunks reduced over the wave; NaN lanes as ballots.
template <bool MN, bool MX>
struct TileScan {
    double lmn, lmx, rmn, rmx, mmn, mmx;
    uint64_t nanl, nanr;
    bool nanm;
};
// MID: the middle's partial chunks from the wave (h, t); otherwise only the
// runs, the middle left at the identity for the caller to fill.
template <bool MN, bool MX, bool MID>
__device__ __forceinline__ TileScan<MN, MX> tile_scan(const TileVals& v, const TileGeom& g) {
    const double inf = __builtin_inf();
    TileScan<MN, MX> r;
    r.nanl = __ballot(g.inl && v.l != v.l);
    r.nanr = __ballot(g.inr && v.r != v.r);
    const bool okl = g.inl && v.l == v.l, okr = g.inr && v.r == v.r;
    r.lmn = okl ? v.l : inf;
    r.lmx = okl ? v.l : -inf;
    r.rmn = okr ? v.r : inf;
    r.rmx = okr ? v.r : -inf;
    wave_scan_nf<MN, MX>(r.lmn, r.lmx);  // lane j: [lo(B) - 1 - j, lo(B))
    wave_scan_nf<MN, MX>(r.rmn, r.rmx);  // lane j: (hi(A), hi(A) + 1 + j]
    r.mmn = inf;
    r.mmx = -inf;
    r.nanm = false;
    if constexpr (MID) {
        r.nanm = __ballot((g.inh && v.h != v.h) || (g.intl && v.t != v.t)) != 0ull;
        const bool okh = g.inh && v.h == v.h, okt = g.intl && v.t == v.t;
        const double tmn = okt ? v.t : inf, tmx = okt ? v.t : -inf;
        r.mmn = vmin(okh ? v.h : inf, tmn);
        r.mmx = vmax(okh ? v.h : -inf, tmx);
        wave_scan_nf<MN, MX>(r.mmn, r.mmx);
        if (MN) r.mmn = lane63(r.mmn);
        if (MX) r.mmx = lane63(r.mmx);
    }
    return r;
}
// The position's extremes from the scans, once the whole chunks are folded
// into mmn / mmx (NaN propagating).
template <bool MN, bool MX>
__device__ __forceinline__ void tile_finish(const TileScan<MN, MX>& r, const TileGeom& g,
                                            double* omn, double* omx) {
    double mmn = r.mmn, mmx = r.mmx;
    bool nan = r.nanm || (MN ? mmn != mmn : mmx != mmx);
    const int jl = (g.loB - 1 - g.lo) & 63, jr = (g.hi - g.hiA - 1) & 63;
    const double gln = MN ? __shfl(r.lmn, jl, 64) : 0.0, glx = MX ? __shfl(r.lmx, jl, 64) : 0.0;
    const double grn = MN ? __shfl(r.rmn, jr, 64) : 0.0, grx = MX ? __shfl(r.rmx, jr, 64) : 0.0;
    if (g.lo < g.loB) {
        nan = nan || (r.nanl & ((2ull << jl) - 1)) != 0ull;
        mmn = vmin(mmn, gln);
        mmx = vmax(mmx, glx);
    }
    if (g.hi > g.hiA) {
        nan = nan || (r.nanr & ((2ull << jr) - 1)) != 0ull;
        mmn = vmin(mmn, grn);
        mmx = vmax(mmx, grx);
    }
    if (MN) *omn = nan ? __builtin_nan("") : mmn;
    if (MX) *omx = nan ? __builtin_nan("") : mmx;
}

// The tile's extremes of both statistics: the mean's minimum (MUMN) and / or
// maximum (MUMX), the std's minimum and maximum.  C: the chunk extremes.
// INT (an interior tile of kDual): the middle [p0 + 64 - m, p0] is chunk
// ca's suffix from offset (64 - m) mod 64 (the block after the sparse
// table, chunk_extremes), the whole chunks, and window p0 (mp, sp), with no
// wave reduction.
template <bool MUMN, bool MUMX, bool FULL, bool INT>
__device__ __forceinline__ void tile_ext(const TileVals& vm, const TileVals& vs,
                                         const TileGeom& g, const double* C, int nch, int levels,
                                         double mp, double sp, double* mumin, double* mumax,
                                         double* sgmin, double* sgmax) {
    TileScan<MUMN, MUMX> a = tile_scan<MUMN, MUMX, !INT>(vm, g);
    TileScan<true, true> b = tile_scan<true, true, !INT>(vs, g);
    if constexpr (INT) {
        if (g.ca < g.cb) {
            const double* X = C + (size_t)(levels + 1) * 4 * nch;
            const int c = g.ca;
            if constexpr (MUMN) a.mmn = nmin(a.mmn, X[c]);
            if constexpr (MUMX) a.mmx = nmax(a.mmx, X[nch + c]);
            b.mmn = nmin(b.mmn, X[2 * nch + c]);
            b.mmx = nmax(b.mmx, X[3 * nch + c]);
        }
        if constexpr (MUMN) a.mmn = nmin(a.mmn, mp);
        if constexpr (MUMX) a.mmx = nmax(a.mmx, mp);
        b.mmn = nmin(b.mmn, sp);
        b.mmx = nmax(b.mmx, sp);
    }
    // whole chunks ca + 1 .. cb - 1 from the sparse table: level k holds
    // ranges of 2^k chunks; consecutive ones and a last one ending at cb - 1
    // (overlapping: the NaN-propagating min / max are idempotent) cover the
    // middle, two when the table reaches floor(log2(len)) (FULL: it always
    // does)
    const int len = g.cb - g.ca - 1;
    if (len > 0) {
        int k = 31 - __builtin_clz(len);
        if constexpr (!FULL) k = k < levels ? k : levels;
        const int step = 1 << k;
        const double* L = C + (size_t)k * 4 * nch;
        auto fold = [&](int c) {
            if constexpr (MUMN) a.mmn = nmin(a.mmn, L[c]);
            if constexpr (MUMX) a.mmx = nmax(a.mmx, L[nch + c]);
            b.mmn = nmin(b.mmn, L[2 * nch + c]);
            b.mmx = nmax(b.mmx, L[3 * nch + c]);
        };
        if constexpr (FULL) {
            fold(g.ca + 1);
        } else {
            for (int c = g.ca + 1; c <= g.cb - step; c += step) fold(c);
        }
        fold(g.cb - step);
    }
    tile_finish<MUMN, MUMX>(a, g, mumin, mumax);
    tile_finish<true, true>(b, g, sgmin, sgmax);
}

// Extremes of the 64-window chunks [64 c, 64 c + 64) of the segment's mean
// and std: C[c], C[nch + c] (mean min, max), C[2 nch + c], C[3 nch + c]
// (std min, max); NaN when the chunk holds a NaN statistic.  With s0 >= 0,
// also the extremes of each chunk's windows from offset s0 on, in the same
// layout after the sparse table's levels.
template <int NT>
__device__ void chunk_extremes(const double* __restrict__ MU, const double* __restrict__ SG,
                               int kml, int nch, double* C, int levels, int s0) {
    const double inf = __builtin_inf(), nan = __builtin_nan("");
    // eight threads per chunk, eight consecutive windows each (their loads
    // in flight together), then a three-step reduction within the eight
    const int part = threadIdx.x & 7;
    for (int c0 = 0; c0 < nch; c0 += NT / 8) {
        const int c = c0 + (int)(threadIdx.x >> 3);
        // [0]: the whole chunk, [1]: its windows from offset s0 on
        double a[2] = {inf, inf}, b[2] = {-inf, -inf}, e[2] = {inf, inf}, f[2] = {-inf, -inf};
        int nanm[2] = {0, 0}, nans[2] = {0, 0};
        if (c < nch) {
            const int j0 = c * 64 + part * 8;
            double vm[8], vs[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int j = (j0 + i < kml) ? j0 + i : kml - 1;
                vm[i] = MU[j];
                vs[i] = SG[j];
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (j0 + i < kml) {
#pragma unroll
                    for (int x = 0; x < 2; x++) {
                        if (x == 1 && (s0 < 0 || part * 8 + i < s0)) continue;
                        if (vm[i] == vm[i]) {
                            a[x] = vmin(a[x], vm[i]);
                            b[x] = vmax(b[x], vm[i]);
                        } else {
                            nanm[x] = 1;
                        }
                        if (vs[i] == vs[i]) {
                            e[x] = vmin(e[x], vs[i]);
                            f[x] = vmax(f[x], vs[i]);
                        } else {
                            nans[x] = 1;
                        }
                    }
                }
            }
        }
        const int nx = s0 >= 0 ? 2 : 1;
        for (int x = 0; x < nx; x++) {
#pragma unroll
            for (int o = 1; o < 8; o <<= 1) {
                a[x] = vmin(a[x], __shfl_xor(a[x], o, 64));
                b[x] = vmax(b[x], __shfl_xor(b[x], o, 64));
                e[x] = vmin(e[x], __shfl_xor(e[x], o, 64));
                f[x] = vmax(f[x], __shfl_xor(f[x], o, 64));
                nanm[x] |= __shfl_xor(nanm[x], o, 64);
                nans[x] |= __shfl_xor(nans[x], o, 64);
            }
            if (part == 0 && c < nch) {
                double* D = C + (x ? (size_t)(levels + 1) * 4 * nch : 0);
                D[c] = nanm[x] ? nan : a[x];
                D[nch + c] = nanm[x] ? nan : b[x];
                D[2 * nch + c] = nans[x] ? nan : e[x];
                D[3 * nch + c] = nans[x] ? nan : f[x];
            }
        }
    }
    // sparse-table level k: extremes of chunks c .. c + 2^k - 1 (entries
    // whose range passes the last chunk are never read)
    for (int k = 1; k <= levels; k++) {
        __syncthreads();
        const double* P = C + (size_t)(k - 1) * 4 * nch;
        double* D = C + (size_t)k * 4 * nch;
        const int h = 1 << (k - 1);
        for (int x = threadIdx.x; x < 4 * nch; x += NT) {
            const int q = x / nch, c = x - q * nch;
            const double u = P[x], v = (c + h < nch) ? P[x + h] : u;
            D[x] = (q & 1) ? nmax(u, v) : nmin(u, v);
        }
    }
}

// Keys of one threshold: how many are not NaN, whether one is -inf, and the
// finite range.
struct KeyStats {
    int cnt = 0;
    int ninf = 0;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    __device__ void add(double v) {
        cnt += (v == v);
        ninf |= (v == -__builtin_inf());
        if (__builtin_isfinite(v)) {
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
    }
    // the smallest key (rank 0) once reduced over the block
    __device__ double kmin() const { return ninf ? -__builtin_inf() : lo; }
};

// The normalised upper envelope U_n (PART & 1: into Xu) and / or the negated
// lower envelope -L_n (PART & 2: into Xl) of every position of the segment
// (src/lb_normalize.cpp:9-32), from the raw envelopes Eu / El (the
// target-wide ones, or for positions within w of the segment's edges the
// segment's own, src/lower_upper_lemire.cpp:8-44) and the statistics'
// extremes over the position's windows.  ku / kl: the keys (positions m - 1
// .. m_k - m).
template <int PART, int NT, bool FULL>
__device__ void norm_env_lds(const double* __restrict__ MU, const double* __restrict__ SG,
                             const double* __restrict__ t, const double* __restrict__ Eu,
                             const double* __restrict__ El, const double* C, int m, int mk,
                             int kml, int w, int levels, double* Xu, double* Xl, KeyStats* ku,
                             KeyStats* kl) {
    constexpr bool U = PART & 1, L = PART & 2;
    const int lane = threadIdx.x & 63;
    const int nch = (kml + 63) / 64;
    auto finish = [&](int i, double Uv, double Lv, double mumin, double mumax, double sgmin,
                      double sgmax) {
        if (i < w || i >= mk - w) {
            const int lo = i - w > 0 ? i - w : 0;
            const int hi = i + w < mk - 1 ? i + w : mk - 1;
            Uv = t[lo];
            Lv = t[lo];
            for (int x = lo + 1; x <= hi; x++) {
                if (U) Uv = nmax(Uv, t[x]);
                if (L) Lv = nmin(Lv, t[x]);
            }
        }
        const bool key = i >= m - 1 && i <= mk - m;
        if constexpr (U) {
            const double nu = Uv - mumin;
            const double v = (nu >= 0) ? nu / sgmin : nu / sgmax;
            Xu[i] = v;
            if (key) ku->add(v);
        }
        if constexpr (L) {
            const double nl = Lv - mumax;
            const double v = -((nl <= 0) ? nl / sgmin : nl / sgmax);
            Xl[i] = v;
            if (key) kl->add(v);
        }
    };
    if (m >= 64) {
        for (int p0 = (threadIdx.x >> 6) * 64; p0 < mk; p0 += NT) {
            const int i = p0 + lane;  // < mk: m_k is a multiple of 64 for m >= 64
            const TileGeom g = tile_geom(p0, m, kml, mk, lane);
            // (not used within w of the edges)
            const double Uv = U ? Eu[i] : 0.0, Lv = L ? El[i] : 0.0;
            double mumin = 0.0, mumax = 0.0, sgmin, sgmax;
            if (FULL && g.interior) {  // (wave-uniform)
                const TileVals vm{MU[g.il], MU[g.ir], 0.0, 0.0}, vs{SG[g.il], SG[g.ir], 0.0, 0.0};
                tile_ext<U, L, FULL, true>(vm, vs, g, C, nch, levels, MU[g.hiA], SG[g.hiA],
                                           &mumin, &mumax, &sgmin, &sgmax);
            } else {
                const TileVals vm = tile_load(MU, g), vs = tile_load(SG, g);
                tile_ext<U, L, FULL, false>(vm, vs, g, C, nch, levels, 0.0, 0.0, &mumin, &mumax,
                                            &sgmin, &sgmax);
            }
            finish(i, Uv, Lv, mumin, mumax, sgmin, sgmax);
        }
    } else {  // ranges of fewer than 64 windows: a direct loop per position
        for (int i = threadIdx.x; i < mk; i += NT) {
            const int lo = i - m + 1 > 0 ? i - m + 1 : 0;
            const int hi = i < kml - 1 ? i : kml - 1;
            double mumin = MU[lo], mumax = MU[lo], sgmin = SG[lo], sgmax = SG[lo];
            for (int x = lo + 1; x <= hi; x++) {
                mumin = nmin(mumin, MU[x]);
                mumax = nmax(mumax, MU[x]);
                sgmin = nmin(sgmin, SG[x]);
                sgmax = nmax(sgmax, SG[x]);
            }
            const bool edge = i < w || i >= mk - w;
            finish(i, (U && !edge) ? Eu[i] : 0.0, (L && !edge) ? El[i] : 0.0, mumin, mumax,
                   sgmin, sgmax);
        }
    }
}

// Exclusive scan of per-thread int64 totals; returns this thread's offset
// (and the block's total in *total when given).
__device__ int64_t block_exscan(int64_t x, int64_t* sh, int64_t* total = nullptr) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t inc = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    int64_t off = 0, tot = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
        if (k < wid) off += sh[k];
        tot += sh[k];
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - x;
}

// A KeyStats per thread reduced over the block, into every thread (one
// barrier pair).
__device__ void block_keystats(KeyStats* ks, int* shc, int64_t* shn, double* shm) {
    int c = ks->cnt, n = ks->ninf;
    double a = ks->lo, b = ks->hi;
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o, 64);
        n |= __shfl_xor(n, o, 64);
        a = fmin(a, __shfl_xor(a, o, 64));
        b = fmax(b, __shfl_xor(b, o, 64));
    }
    const int nwv = (int)(blockDim.x >> 6), wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        shc[wid] = c;
        shn[wid] = n;
        shm[wid] = a;
        shm[nwv + wid] = b;
    }
    __syncthreads();
    c = 0;
    n = 0;
    a = __builtin_inf();
    b = -__builtin_inf();
    for (int k = 0; k < nwv; k++) {
        c += shc[k];
        n |= (int)shn[k];
        a = fmin(a, shm[k]);
        b = fmax(b, shm[nwv + k]);
    }
    __syncthreads();
    ks->cnt = c;
    ks->ninf = n;
    ks->lo = a;
    ks->hi = b;
}

// ------------------------------------------------------- threshold choice ---
// The threshold is the key at the argmax, over the ascending ranks p of the
// nval non-NaN keys, of h(key_p) (p + 1) (better_max: the larger score, then
// the smaller rank; NaN scores never win), or the smallest key when h of it is
// NaN (src/find_t_lp.cpp:206-216).  h is non-increasing in exact arithmetic
// and herr(a) bounds |computed h(t) - exact h(t)| for |t| <= a.
//
// Buckets.  Keys map monotonically to value-ordered buckets (0: -inf, 1 ..
// nbk-2: equal widths of the finite range, last: +inf), each with its
// count and its min / max as order-preserving integers.  A bucket's keys take
// the ranks [start, end) of the full order, and its largest key sits at rank
// end - 1, so h(max) end is a score the full order attains.  Every score in
// the bucket is at most (h(min) + 2 herr) end, so a bucket whose bound is
// below the best attained score holds no maximum.  The kept buckets' keys are
// gathered, sorted, ranked start + their place in the bucket and scored as the
// full order scores them.  Beyond thr_cap candidates: the full sort.
__device__ inline unsigned long long ord_key(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double key_of(unsigned long long o) {
    return __longlong_as_double((long long)((o >> 63) ? (o & 0x7fffffffffffffffull) : ~o));
}
__device__ inline int bucket_of(double v, double lo, double scale, int nb) {
    if (v == -__builtin_inf()) return 0;
    if (v == __builtin_inf()) return nb - 1;
    const double x = (v - lo) * scale;  // non-decreasing in v (NaN only when scale is 0)
    return 1 + (int)fmin(fmax(x, 0.0), (double)(nb - 3));
}

// keys[0 .. nthr): the values, NaN for positions that do not count; ks: their
// statistics, reduced over the block.  tlds: thr_lds_bytes(npad, NT) of LDS.
// The full sort (fallback) orders the keys in sortbuf[0 .. npad), which may be
// keys itself.
template <int NT, class HF, class EF>
__device__ double select_threshold(const double* keys, int nthr, int npad, const KeyStats& ks,
                                   HF hf, EF herr, unsigned char* tlds, double* sortbuf,
                                   ArgBest* sh_best, int64_t* sh_scan, double* sh_key) {
    const int nval = ks.cnt;
    const double lo = ks.lo, hi = ks.hi;
    if (nval == 0) return __builtin_nan("");
    // the rank-0 key, and h there first: the bucket pass's atomics hide the
    // latency of its searches
    const double kmin = ks.kmin();
    const double h0 = hf(kmin);
    if (h0 != h0) return kmin;  // src/find_t_lp.cpp:206-216 keeps index 0
    const int nbk = thr_buckets(npad, NT);
    const int tid = threadIdx.x;
    const int cap = thr_cap(npad, NT);
    double* cand = (double*)tlds;
    unsigned long long* bmin = (unsigned long long*)(cand + cap);
    unsigned long long* bmax = bmin + nbk;
    int* cnt = (int*)(bmax + nbk);
    int* bst = cnt + nbk;  // first rank
    int* cst = bst + nbk;  // first candidate slot; -1: bucket dropped
    int* fil = cst + nbk;
    const double scale =
        (hi > lo && hi - lo < __builtin_inf()) ? (double)(nbk - 2) / (hi - lo) : 0.0;
    for (int b = tid; b < nbk; b += NT) {
        cnt[b] = 0;
        fil[b] = 0;
        bmin[b] = ~0ull;
        bmax[b] = 0ull;
    }
    __syncthreads();
    {
        // each thread a contiguous run of keys (positions): neighbouring
        // positions' envelope values mostly share a bucket, so a thread
        // folds its run per bucket and updates the shared counts once per
        // bucket change instead of once per key (lanes hitting one bucket
        // serialise on its LDS word)
        const int per = (nthr + NT - 1) / NT;
        const int i1 = (tid + 1) * per < nthr ? (tid + 1) * per : nthr;
        int rb = -1, rc = 0;
        unsigned long long rmn = ~0ull, rmx = 0ull;
        for (int i = tid * per; i < i1; i++) {
            const double v = keys[i];
            if (v != v) continue;
            const int b = bucket_of(v, lo, scale, nbk);
            const unsigned long long o = ord_key(v);
            if (b != rb) {
                if (rc) {
                    atomicAdd(&cnt[rb], rc);
                    atomicMin(&bmin[rb], rmn);
                    atomicMax(&bmax[rb], rmx);
                }
                rb = b;
                rc = 0;
                rmn = ~0ull;
                rmx = 0ull;
            }
            rc++;
            rmn = o < rmn ? o : rmn;
            rmx = o > rmx ? o : rmx;
        }
        if (rc) {
            atomicAdd(&cnt[rb], rc);
            atomicMin(&bmin[rb], rmn);
            atomicMax(&bmax[rb], rmx);
        }
    }
    __syncthreads();
    const int c = (tid < nbk) ? cnt[tid] : 0;
    const int st = (int)block_exscan(c, sh_scan);
    if (tid < nbk) bst[tid] = st;

    ArgBest att{__builtin_nan(""), 0};
    double ub = 0.0;
    if (tid < nbk && c > 0) {
        const double vmn = key_of(bmin[tid]), vmx = key_of(bmax[tid]);
        const int e = st + c;
        att = ArgBest{hf(vmx) * (double)e, (int64_t)(e - 1)};
        const double hb = hf(vmn) + 2.0 * herr(fmax(fabs(vmn), fabs(vmx)));
        ub = (hb >= 0.0) ? hb * (double)e : hb * (double)(st + 1);
        ub += fabs(ub) * 1e-12 + 1e-300;
    }
    const ArgBest lb = block_best<true>(att, sh_best);
    const bool keep = tid < nbk && c > 0 && !(ub < lb.v);
    const int kc = keep ? c : 0;
    int64_t ktot;
    const int cs = (int)block_exscan(kc, sh_scan, &ktot);
    const int K = (int)ktot;
    if (tid < nbk) cst[tid] = keep ? cs : -1;
    __syncthreads();
    if (K <= cap) {
        const int kpad = npad_of(K);
        for (int i = tid; i < nthr; i += NT) {
            const double v = keys[i];
            if (v != v) continue;
            const int b = bucket_of(v, lo, scale, nbk);
            const int o = cst[b];
            if (o >= 0) cand[o + atomicAdd(&fil[b], 1)] = v;
        }
        if (K <= 64) {
            // one wave sorts the candidates in its registers (a bitonic
            // network over lanes, no block barriers) and takes the argmax
            __threadfence_block();
            __syncthreads();
            const int lane = tid & 63;
            if (tid < 64) {
                double v = lane < K ? cand[lane] : __builtin_inf();
                for (int k = 2; k <= 64; k <<= 1) {
                    for (int j = k >> 1; j > 0; j >>= 1) {
                        const double o = __shfl_xor(v, j, 64);
                        // ascending within blocks of k whose bit k is clear
                        const bool take_min = ((lane & j) == 0) == ((lane & k) == 0);
                        v = take_min ? vmin(v, o) : vmax(v, o);
                    }
                }
                ArgBest x{__builtin_nan(""), 0};
                int64_t p = 0;
                if (lane < K) {
                    const int b = bucket_of(v, lo, scale, nbk);
                    p = (int64_t)bst[b] + (lane - cst[b]);
                    x = ArgBest{hf(v) * (double)(p + 1), p};
                }
                for (int o = 32; o > 0; o >>= 1) {
                    ArgBest y;
                    y.v = __shfl_xor(x.v, o, 64);
                    y.i = __shfl_xor(x.i, o, 64);
                    x = better_max(x, y);
                }
                const uint64_t win = __ballot(lane < K && x.v == x.v && p == x.i);
                const double key = __shfl(v, win ? __builtin_ctzll(win) : 0, 64);
                if (lane == 0) *sh_key = (x.v == x.v) ? key : kmin;
            }
            __syncthreads();
            const double t = *sh_key;
            __syncthreads();
            return t;
        }
        for (int j = K + tid; j < kpad; j += NT) cand[j] = __builtin_inf();
        __threadfence_block();
        __syncthreads();
        bitonic_sort(cand, kpad);
        ArgBest best{__builtin_nan(""), 0};
        double bkey = kmin;
        for (int j = tid; j < K; j += NT) {
            const double v = cand[j];
            const int b = bucket_of(v, lo, scale, nbk);
            const int64_t p = (int64_t)bst[b] + (j - cst[b]);
            const ArgBest x{hf(v) * (double)(p + 1), p};
            const ArgBest nb = better_max(best, x);
            if (nb.i == p && nb.v == nb.v) bkey = v;
            best = nb;
        }
        const ArgBest r = block_best<true>(best, sh_best);
        if (r.v == r.v && best.v == r.v && best.i == r.i) *sh_key = bkey;
        __syncthreads();
        const double t = (r.v == r.v) ? *sh_key : kmin;
        __syncthreads();
        return t;
    }
    // full order: NaN and padding sort last
    for (int i = tid; i < npad; i += NT) {
        const double v = (i < nthr) ? keys[i] : __builtin_nan("");
        sortbuf[i] = (v != v) ? __builtin_inf() : v;
    }
    __threadfence_block();
    __syncthreads();
    bitonic_sort(sortbuf, npad);
    ArgBest best{__builtin_nan(""), 0};
    for (int p = tid; p < nval; p += NT) {
        const double tv = sortbuf[p];
        best = better_max(best, ArgBest{hf(tv) * (double)(p + 1), p});
    }
    best = block_best<true>(best, sh_best);
    const double t = sortbuf[best.i];
    __syncthreads();
    return t;
}

// sum over runs [a, b) of set mask bits inside [j, j+m) of CH[b-j] - CH[a-j],
// by parts: over the transitions t of the mask inside (j, j+m) (bit t of tw:
// g[t] != g[t-1]) of +CH[t-j] where a run ends and -CH[t-j] where one starts
// (they alternate, the first an end when g[j] is set; CH[0] = 0), plus CH[m]
// when g[j+m-1] is set.  Exact int64: every partial sum lies in [-CH[m], CH[m]].
__device__ inline int64_t trans_sum(const uint64_t* gw, const uint64_t* tw, const int64_t* CH,
                                    int j, int m) {
    int64_t acc = 0;
    const int hi = j + m - 1;  // last position of the window
    if (m > 1) {
        bool end = (gw[j >> 6] >> (j & 63)) & 1;
        const int lo = j + 1;
        int wi = lo >> 6;
        const int wend = hi >> 6;
        uint64_t word = tw[wi] & (~0ull << (lo & 63));
        while (true) {
            if (wi == wend) word &= (2ull << (hi & 63)) - 1;  // bits <= hi
            while (word) {
                const int t = (wi << 6) + __builtin_ctzll(word);
                const int64_t c = CH[t - j];
                acc += end ? c : -c;
                end = !end;
                word &= word - 1;
            }
            if (wi == wend) break;
            word = tw[++wi];
        }
    }
    if ((gw[hi >> 6] >> (hi & 63)) & 1) acc += CH[m];
    return acc;
}

// The same sum for the 64 windows j0 .. j0 + 63 of a wave (lane: j), from
// the ascending list P[0 .. n) of the mask's transitions: the wave walks the
// transitions t in (j0, j0 + 63 + m) together (t uniform), each lane adding
// -CH[t - j] where a run starts (g[t] set) and +CH[t - j] where one ends, for
// j < t < j + m.  The caller adds CH[m] when g[j + m - 1] is set.
__device__ inline int64_t list_sum(const int* P, int n, const uint64_t* gw, const int64_t* CH,
                                   int j0, int j, int m) {
    int lo = 0, hi = n;  // first t > j0
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (P[mid] > j0) hi = mid; else lo = mid + 1;
    }
    const int lim = j0 + 63 + m;
    int64_t acc = 0;
    for (int k = lo; k < n; k++) {
        const int t = P[k];
        if (t >= lim) break;
        const int d = t - j;
        const bool in = d > 0 && d < m;
        const int64_t c = CH[in ? d : 0];
        const bool start = (gw[t >> 6] >> (t & 63)) & 1;
        acc += in ? (start ? -c : c) : 0;
    }
    return acc;
}

// ------------------------------------------ FP32 matrix-core bounds ---
// The sliding inner products of src/sip.cpp as a Hankel-mask x shifted-penalty
// contraction on the matrix cores (v_mfma_f32_16x16x4_f32).  For a tile of 256
// windows j0 + 16 r + n (r, n < 16):
//   z[j0 + 16 r + n] = sum_k A[r][k] B[k][n],  A[r][k] = g[j0 + 16 r + k],
//   B[k][n] = h[k - n] (zero outside [0, m)),  k < 4 nst, nst = ceil((m + 15) / 4),
// the lower side (g_l, h_l) and the upper side (g_u, h_u) in accumulators of
// their own.  Operand lanes: A[lane % 16][lane / 16], B[lane / 16][lane % 16];
// result D[4 (lane / 16) + i][lane % 16] in element i.  Each lane keeps 64 mask
// bits of its row in registers (16 steps of 4) and reads B from the padded
// penalty rows hf[15 + k - n].
//
// Certification: the penalties enter rounded down to FP32 and every product is
// g h with g in {0, 1}, so the terms are non-negative and exact.  Any order of
// at most n = 4 nst FP32 additions of them, each within one ulp (u = 2^-23,
// whatever the unit's rounding), leaves the computed s^ within gamma_n =
// n u / (1 - n u) of the terms' exact sum s, so s >= s^ / (1 + gamma_n) >=
// s^ (1 - gamma_n); subnormal results add at most 2^-149 per addition.  The
// bound is s^ (1 - gamma_n - 2^-48) - n 2^-149 per side, clamped at 0 (the
// 2^-48 covers the FP64 multiply and add that form it): a certified lower
// bound of the FP32 penalties' sum, itself at most the FP64 penalties' sum.
typedef float f32x4 __attribute__((ext_vector_type(4)));

// the 64 mask bits from bit 64 wi + sh on, words lo = [wi], hi = [wi + 1]
__device__ __forceinline__ uint64_t bits64(uint64_t lo, uint64_t hi, int sh) {
    return (lo >> sh) | ((hi << 1) << (63 - sh));
}

__host__ __device__ inline int f32_steps(int m) { return (m + 18) / 4; }
__host__ __device__ inline int f32_hlen(int m) { return 4 * f32_steps(m) + 16; }

// Both sides' sums of the tile at j0 for this lane's four windows.
__device__ __forceinline__ void f32_tile(const uint64_t* GLw, const uint64_t* GUw,
                                         const float* HLf, const float* HUf, int j0, int nst,
                                         int lane, f32x4& al, f32x4& au) {
    const int ra = lane & 15, kq = lane >> 4;
    const int p = j0 + 16 * ra + kq;  // A[ra][kq]'s bit at step 0
    const int sh = p & 63;
    int wi = p >> 6;
    uint64_t l0 = GLw[wi], u0 = GUw[wi];
    const float* hl = HLf + 15 + kq - ra;
    const float* hu = HUf + 15 + kq - ra;
    al = f32x4{0.f, 0.f, 0.f, 0.f};
    au = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t0 = 0; t0 < nst; t0 += 16) {
        const uint64_t l1 = GLw[wi + 1], u1 = GUw[wi + 1];
        const uint64_t RL = bits64(l0, l1, sh), RU = bits64(u0, u1, sh);
        const uint32_t rl[2] = {(uint32_t)RL, (uint32_t)(RL >> 32)};
        const uint32_t ru[2] = {(uint32_t)RU, (uint32_t)(RU >> 32)};
#pragma unroll
        for (int tt = 0; tt < 16; tt++) {
            if (t0 + tt >= nst) break;
            const float a_l = (float)((rl[tt >> 3] >> (4 * (tt & 7))) & 1u);
            const float a_u = (float)((ru[tt >> 3] >> (4 * (tt & 7))) & 1u);
            al = __builtin_amdgcn_mfma_f32_16x16x4f32(a_l, hl[4 * (t0 + tt)], al, 0, 0, 0);
            au = __builtin_amdgcn_mfma_f32_16x16x4f32(a_u, hu[4 * (t0 + tt)], au, 0, 0, 0);
        }
        l0 = l1;
        u0 = u1;
        wi++;
    }
}

// ------------------------------------------------------------------ kernel ---
// kDual / kSingle keep the normalised envelopes in LDS (X, and X2 = -L_n for
// kDual), then the penalty prefix sums and mask transitions in X's space;
// the statistics' ranges come from tiles of 64 positions (tile_ext) and the
// chunk extremes, so nothing but the bounds leaves the workgroup.  kSlab runs
// van Herk prefix / suffix scans over blocks of m windows into the block's
// global scratch, with U_n, L_n and the keys there too.
// One kernel per layout, so that none carries another's registers; every
// workgroup size fits four waves per SIMD.
// F32 (kDual only): the bounds from the FP32 matrix-core contraction
// (f32_tile) instead of the exact fixed-point run sums.
template <int NT, int LAYOUT, bool F32 = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
    k_lbip(LbipArgs a) {
    static_assert(!F32 || LAYOUT == kDual, "FP32 bounds need the kDual layout");
    extern __shared__ unsigned char lds_raw[];
    __shared__ ArgBest sh_best[NT / 64];
    __shared__ double sh_sum[NT / 64];
    __shared__ int64_t sh_scan[NT / 64];
    __shared__ int sh_cnt[NT / 64];
    __shared__ double sh_mm[2 * (NT / 64)];
    __shared__ double sh_key;

    const Geometry G = a.G;
    const QueryTables Q = a.q;
    const int m = G.m, mk = G.m_k, w = a.p.w, metric = a.p.metric;
    const double C = a.p.arg;
    const int kml = (int)G.kml;
    const int nthr = mk - 2 * m + 2;
    const int npad = npad_of(nthr);
    const int nw = (mk + 63) / 64;
    constexpr bool xl = LAYOUT != kSlab;
    const int levels = lbip_levels(LAYOUT, m, kml, npad, NT);  // sparse-table levels in the zone
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

    double* slab = (double*)((char*)a.slab + (size_t)blockIdx.x * a.slab_bytes_per_block);
    // LDS layout: X (and X2), the zone (chunk extremes, then threshold
    // scratch), masks
    double* X = (double*)lds_raw;
    double* X2 = X + mk;
    unsigned char* zone = lds_raw + sizeof(double) * (size_t)mk * (LAYOUT == kDual ? 2 : 1);
    // global-scratch layout: 8 van Herk arrays of kml (2 * stat + dir), U_n,
    // L_n, keys
    double* PQ = slab;
    double* Un = slab + 8 * (size_t)kml;
    double* Ln = Un + mk;
    int64_t *CHl, *CHu;
    uint64_t *GLw, *GUw, *TLw, *TUw;  // masks, transitions: bit p = g[p] xor g[p - 1]
    if (xl) {
        GLw = (uint64_t*)(zone + zone_bytes(npad, NT, kml, levels, LAYOUT));
        GUw = GLw + nw;
        CHl = (int64_t*)X;  // X is dead once both masks are built
        CHu = CHl + (m + 1);
        TLw = (uint64_t*)(CHu + (m + 1));
        TUw = TLw + nw;
    } else {
        CHl = (int64_t*)lds_raw;
        CHu = CHl + (m + 1);
        GLw = (uint64_t*)(CHu + (m + 1));
        GUw = GLw + nw;
        TLw = GUw + nw;
        TUw = TLw + nw;
    }
    double* keys = xl ? X + (m - 1) : Ln + mk;
    double* sortbuf = xl ? slab : keys;  // the threshold's full-sort fallback
    unsigned char* tlds = xl ? zone : lds_raw;

    // (|computed h - exact h| is far below 1e-9 of the terms' magnitude:
    // prefix sums of at most m terms of |q| + |t| + |C| or its square)
    const double qmax = fmax(fabs(Q.qs[0]), fabs(Q.qs[m - 1]));
    auto herr = [&](double at) {
        const double A = qmax + at + fabs(C);
        return 1e-9 * (double)m * (A * A + A);
    };
    auto hf_u = [&](double t) { return h_upper(Q, metric, m, C, t); };
    auto hf_l = [&](double t) { return h_lower(Q, metric, m, C, -t); };
    // threshold from keys k of stats ks: upper (ascending ranks of U_n, G =
    // rank + 1) or lower (ascending ranks of -L_n are L_n's from the top;
    // negation is exact, so t_l = -the key)
    auto threshold = [&](const double* k, KeyStats ks, bool upper) {
        block_keystats(&ks, sh_cnt, sh_scan, sh_mm);
        if (upper)
            return select_threshold<NT>(k, nthr, npad, ks, hf_u, herr, tlds, sortbuf, sh_best,
                                        sh_scan, &sh_key);
        const double tn = select_threshold<NT>(k, nthr, npad, ks, hf_l, herr, tlds, sortbuf,
                                               sh_best, sh_scan, &sh_key);
        return (tn == tn) ? -tn : tn;
    };
    // mask bitmaps: g_u = U_n <= t_u from X[] = U_n, g_l = L_n >= t_l from
    // X[] = -L_n
    auto mask = [&](const double* Xs, double thr, bool upper, uint64_t* out) {
        for (int wi = wid; wi < nw; wi += NT / 64) {
            const int pos = wi * 64 + lane;
            const bool in = pos < mk;
            const uint64_t b = __ballot(in && (upper ? Xs[pos] <= thr : -Xs[pos] >= thr));
            if (lane == 0) out[wi] = b;
        }
    };

    for (int64_t s = blockIdx.x; s < G.S; s += gridDim.x) {
        const bool tail = (s == G.forsize);
        const int64_t seg0 = tail ? G.N - mk : s * G.kml;
        const double* t = a.T + seg0;
        const double* MU = a.MU + s * G.kml;
        const double* SG = a.SG + s * G.kml;
        const uint8_t* INV = a.INV + s * G.kml;
        const int jfirst = tail ? G.tail_first : 0;
        const int64_t gbase = s * G.kml;
        double t_u, t_l;

        if constexpr (LAYOUT == kDual) {
            double* Cx = (double*)zone;
            if (m >= 64) {
                chunk_extremes<NT>(MU, SG, kml, (kml + 63) / 64, Cx, levels, (64 - m) & 63);
                __syncthreads();
            }
            KeyStats ku, kl;
            norm_env_lds<3, NT, true>(MU, SG, t, a.Uraw + seg0, a.Lraw + seg0, Cx, m, mk, kml, w,
                                      levels, X, X2, &ku, &kl);
            __syncthreads();  // (the zone's chunk extremes are dead from here)
            t_u = threshold(X + (m - 1), ku, true);
            mask(X, t_u, true, GUw);
            t_l = threshold(X2 + (m - 1), kl, false);
            mask(X2, t_l, false, GLw);
            __syncthreads();
        } else if constexpr (LAYOUT == kSingle) {
            // one pass over the statistics for both envelopes: U_n into X,
            // -L_n into the block's global scratch (after the full-sort
            // buffer), copied into X once U_n's mask is built
            double* Cx = (double*)zone;
            double* Lg = slab + npad;
            KeyStats ku, kl;
            if (m >= 64) {
                chunk_extremes<NT>(MU, SG, kml, (kml + 63) / 64, Cx, levels, -1);
                __syncthreads();
            }
            norm_env_lds<3, NT, false>(MU, SG, t, a.Uraw + seg0, a.Lraw + seg0, Cx, m, mk, kml, w,
                                       levels, X, Lg, &ku, &kl);
            __syncthreads();
            t_u = threshold(keys, ku, true);
            mask(X, t_u, true, GUw);
            __syncthreads();
            for (int i = threadIdx.x; i < mk; i += NT) X[i] = Lg[i];
            __syncthreads();
            t_l = threshold(keys, kl, false);
            mask(X, t_l, false, GLw);
            __syncthreads();
        } else {
            // 1. van Herk prefix (dir 0) / suffix (dir 1) scans over blocks of
            //    m windows: one wave per (statistic, direction, block), 64
            //    windows per step, min and max of the same statistic in one
            //    pass.  The NaN-propagating min / max are associative, so the
            //    values equal a sequential scan's.
            {
                const int nb = (kml + m - 1) / m;
                const double inf = __builtin_inf();
                for (int item = wid; item < 4 * nb; item += NT / 64) {
                    const int sg = item / (2 * nb);  // 0: mean, 1: std
                    const int rem = item - sg * 2 * nb;
                    const int dir = rem / nb;
                    const int b = rem - dir * nb;
                    const double* src = sg ? SG : MU;
                    double* omin = PQ + (size_t)(2 * (2 * sg) + dir) * kml;
                    double* omax = PQ + (size_t)(2 * (2 * sg + 1) + dir) * kml;
                    const int j0 = b * m;
                    const int j1 = (j0 + m < kml) ? j0 + m : kml;
                    double cmin = inf, cmax = -inf;
                    // the loads of up to kScanG chunks are issued before the
                    // chunk-to-chunk carry chain that consumes them
                    constexpr int kScanG = 4;
                    for (int c0 = 0; c0 < j1 - j0; c0 += 64 * kScanG) {
                        double v[kScanG];
#pragma unroll
                        for (int u = 0; u < kScanG; u++) {
                            const int c = c0 + 64 * u;
                            // prefix: ascending windows; suffix: descending
                            const int j = (dir == 0) ? j0 + c + lane : j1 - 1 - c - lane;
                            const bool in = (dir == 0) ? (j < j1) : (j >= j0);
                            v[u] = in ? src[j] : 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < kScanG; u++) {
                            const int c = c0 + 64 * u;
                            if (c >= j1 - j0) break;
                            const int j = (dir == 0) ? j0 + c + lane : j1 - 1 - c - lane;
                            const bool in = (dir == 0) ? (j < j1) : (j >= j0);
                            double vmn = in ? v[u] : inf, vmx = in ? v[u] : -inf;
                            wave_scan_min_max(vmn, vmx);
                            vmn = nmin(cmin, vmn);
                            vmx = nmax(cmax, vmx);
                            if (in) {
                                omin[j] = vmn;
                                omax[j] = vmx;
                            }
                            cmin = lane63(vmn);
                            cmax = lane63(vmx);
                        }
                    }
                }
            }
            __threadfence_block();
            __syncthreads();

            // 2-3. envelopes, stat ranges, normalised envelopes; upper keys.
            //     Each thread takes kEnvG positions at a time and issues their
            //     loads (envelope, the eight prefix / suffix values) first.
            constexpr int kEnvG = 2;
            KeyStats ku;
            for (int i0 = threadIdx.x; i0 < mk; i0 += kEnvG * NT) {
                double Lg[kEnvG], Ug[kEnvG], PS[kEnvG][8];
#pragma unroll
                for (int e = 0; e < kEnvG; e++) {
                    const int i = i0 + e * NT;
                    if (i >= mk) break;
                    const bool edge = i < w || i >= mk - w;
                    Lg[e] = edge ? 0.0 : a.Lraw[seg0 + i];
                    Ug[e] = edge ? 0.0 : a.Uraw[seg0 + i];
                    const int64_t lo = i - m + 1 > 0 ? i - m + 1 : 0;
                    const int64_t hi = i < kml - 1 ? i : kml - 1;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        PS[e][2 * q] = PQ[(size_t)(2 * q) * kml + hi];          // prefix at hi
                        PS[e][2 * q + 1] = PQ[(size_t)(2 * q + 1) * kml + lo];  // suffix at lo
                    }
                }
#pragma unroll
                for (int e = 0; e < kEnvG; e++) {
                    const int i = i0 + e * NT;
                    if (i >= mk) break;
                    double Lv = Lg[e], Uv = Ug[e];
                    if (i < w || i >= mk - w) {
                        const int lo = i - w > 0 ? i - w : 0;
                        const int hi = i + w < mk - 1 ? i + w : mk - 1;
                        Lv = t[lo];
                        Uv = t[lo];
                        for (int x = lo + 1; x <= hi; x++) {
                            Lv = nmin(Lv, t[x]);
                            Uv = nmax(Uv, t[x]);
                        }
                    }
                    const int64_t lo = i - m + 1 > 0 ? i - m + 1 : 0;
                    const int64_t hi = i < kml - 1 ? i : kml - 1;
                    const double mumin = stat_range(PS[e][0], PS[e][1], lo, hi, m, true);
                    const double mumax = stat_range(PS[e][2], PS[e][3], lo, hi, m, false);
                    const double sgmin = stat_range(PS[e][4], PS[e][5], lo, hi, m, true);
                    const double sgmax = stat_range(PS[e][6], PS[e][7], lo, hi, m, false);
                    const double nu = Uv - mumin;
                    const double un = (nu >= 0) ? nu / sgmin : nu / sgmax;
                    const double nl = Lv - mumax;
                    const double ln = (nl <= 0) ? nl / sgmin : nl / sgmax;
                    Un[i] = un;
                    Ln[i] = ln;
                    if (i >= m - 1 && i <= mk - m) {
                        keys[i - (m - 1)] = un;
                        ku.add(un);
                    }
                }
            }
            __threadfence_block();
            __syncthreads();

            // 4a. upper threshold: ascending ranks of U_n, G = rank + 1
            t_u = threshold(keys, ku, true);

            // 4b. lower threshold: ranks from the top of L_n, G = rank + 1: the
            //     ascending ranks of -L_n (negation is exact)
            KeyStats kl;
            for (int i = threadIdx.x; i < nthr; i += blockDim.x) {
                const double nl = -Ln[i + m - 1];
                keys[i] = nl;
                kl.add(nl);
            }
            __threadfence_block();
            __syncthreads();
            t_l = threshold(keys, kl, false);

            // mask bitmaps: g_l = L_n >= t_l, g_u = U_n <= t_u
            for (int wi = wid; wi < nw; wi += NT / 64) {
                const int pos = wi * 64 + lane;
                const bool in = pos < mk;
                const uint64_t bl = __ballot(in && (Ln[pos] >= t_l));
                const uint64_t bu = __ballot(in && (Un[pos] <= t_u));
                if (lane == 0) {
                    GLw[wi] = bl;
                    GUw[wi] = bu;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (a.TUo) a.TUo[s] = t_u;
            if (a.TLo) a.TLo[s] = t_l;
        }

        ArgBest rep{__builtin_nan(""), -1};
        // the bound of window j from a lower bound zd of its z_l + z_u
        auto bound = [&](int j, bool finite, double zd) {
            const int64_t g = gbase + (j - jfirst);
            double lb;
            if (INV[j] || !finite) {
                lb = __builtin_nan("");
            } else {
                if (metric == DTW || metric == ERP) lb = sqrt(fmax(0.0, zd));
                else if (metric == LCSS) lb = fmax(0.0, zd / (double)m);
                else lb = zd;
                if (!is_excluded(a.p, g)) rep = better_min(rep, ArgBest{lb, g});
            }
            a.LB[g] = lb;
        };
        if constexpr (F32) {
            // 5'. penalties rounded down to FP32 in zero-padded rows:
            //     HLf[15 + s] = h_l[s], HUf[15 + s] = h_u[s] (X is dead)
            const int nst = f32_steps(m), hlen = f32_hlen(m);
            float* HLf = (float*)X;
            float* HUf = HLf + hlen;
            double hs = 0.0;
            for (int i = threadIdx.x; i < hlen; i += NT) {
                const int sidx = i - 15;
                float fl = 0.f, fu = 0.f;
                if (sidx >= 0 && sidx < m) {
                    double hl, hu;
                    penalties(metric, Q.qn[sidx], t_l, t_u, C, &hl, &hu);
                    hs += hl + hu;
                    fl = __double2float_rd(hl);
                    fu = __double2float_rd(hu);
                }
                HLf[i] = fl;
                HUf[i] = fu;
            }
            hs = block_sum(hs, sh_sum);  // (its barriers publish the rows)
            const bool finite = __builtin_isfinite(hs);
            // 7'. tiles of 256 windows, one per wave at a time
            const double nr = 4.0 * (double)nst;
            const double gam = nr * 0x1p-23 / (1.0 - nr * 0x1p-23);
            const double keep = 1.0 - gam - 0x1p-48;
            const double sub = nr * 0x1p-149;
            constexpr float kFltMax = 3.4028234663852886e38f;  // an overflowed sum
            auto lo = [&](float v) { return (double)(v <= kFltMax ? v : kFltMax) * keep - sub; };
            const int ntiles = (kml - jfirst + 255) / 256;
            for (int tile = wid; tile < ntiles; tile += NT / 64) {
                const int j0 = jfirst + 256 * tile;
                f32x4 al, au;
                f32_tile(GLw, GUw, HLf, HUf, j0, nst, lane, al, au);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int j = j0 + 16 * (4 * (lane >> 4) + i) + (lane & 15);
                    if (j < kml) bound(j, finite, fmax(0.0, lo(al[i]) + lo(au[i])));
                }
            }
        } else {
            // 5. penalties in fixed point: hI = floor(h * 2^F), exact prefix sums
            double hs = 0.0;
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                double hl, hu;
                penalties(metric, Q.qn[i], t_l, t_u, C, &hl, &hu);
                hs += hl + hu;
            }
            hs = block_sum(hs, sh_sum);
            const bool finite = __builtin_isfinite(hs);
            int F = 0;
            if (finite && hs > 0.0) {
                F = 60 - ilogb(hs);
                if (F > 1000) F = 1000;
            }
            {
                const int per = (m + NT - 1) / NT;
                const int i0 = threadIdx.x * per;
                const int i1 = (i0 + per < m) ? i0 + per : m;
                int64_t tl_sum = 0, tu_sum = 0;
                for (int i = i0; i < i1; i++) {
                    double hl, hu;
                    penalties(metric, Q.qn[i], t_l, t_u, C, &hl, &hu);
                    if (finite) {
                        tl_sum += (int64_t)floor(ldexp(hl, F));
                        tu_sum += (int64_t)floor(ldexp(hu, F));
                    }
                }
                const int64_t ol = block_exscan(tl_sum, sh_scan);
                const int64_t ou = block_exscan(tu_sum, sh_scan);
                int64_t cl = ol, cu = ou;
                if (threadIdx.x == 0) {
                    CHl[0] = 0;
                    CHu[0] = 0;
                }
                for (int i = i0; i < i1; i++) {
                    double hl, hu;
                    penalties(metric, Q.qn[i], t_l, t_u, C, &hl, &hu);
                    if (finite) {
                        cl += (int64_t)floor(ldexp(hl, F));
                        cu += (int64_t)floor(ldexp(hu, F));
                    }
                    CHl[i + 1] = cl;
                    CHu[i + 1] = cu;
                }
            }

            // 6. mask transitions: bitmaps, and for kDual / kSingle their
            //    ascending lists (positions t with g[t] != g[t - 1]) after the
            //    bitmaps in X's space, one word per thread (nw <= NT), when they
            //    fit (lcap entries each)
            for (int wi = threadIdx.x; wi < nw; wi += NT) {
                const uint64_t pl = wi ? GLw[wi - 1] >> 63 : 0ull;
                const uint64_t pu = wi ? GUw[wi - 1] >> 63 : 0ull;
                TLw[wi] = GLw[wi] ^ ((GLw[wi] << 1) | pl);
                TUw[wi] = GUw[wi] ^ ((GUw[wi] << 1) | pu);
            }
            __syncthreads();
            int* PL = (int*)(TUw + nw);
            const int lcap = (LAYOUT == kDual) ? mk : mk / 2;
            int* PU = PL + lcap;
            int nl = 0, nu = 0;
            bool lists = false;
            if constexpr (xl) {
                const int wi = threadIdx.x;
                const int64_t c = (wi < nw) ? ((int64_t)__builtin_popcountll(TLw[wi]) |
                                               ((int64_t)__builtin_popcountll(TUw[wi]) << 32))
                                            : 0;
                int64_t tot;
                const int64_t off = block_exscan(c, sh_scan, &tot);  // low / high halves
                nl = (int)(tot & 0xffffffff);
                nu = (int)(tot >> 32);
                lists = nl <= lcap && nu <= lcap;
                if (lists && wi < nw) {
                    int ol = (int)(off & 0xffffffff), ou = (int)(off >> 32);
                    for (uint64_t b = TLw[wi]; b; b &= b - 1) PL[ol++] = wi * 64 + __builtin_ctzll(b);
                    for (uint64_t b = TUw[wi]; b; b &= b - 1) PU[ou++] = wi * 64 + __builtin_ctzll(b);
                }
                __syncthreads();
            }

            // 7. per-window bound (z rounded down to FP64; the scaling by 2^-F
            //    is exact)
            const double scale = ldexp(1.0, -F);
            auto zfix = [&](int64_t z) { return __ll2double_rd(z) * scale; };
            if (lists) {
                for (int j0 = jfirst + wid * 64; j0 < kml; j0 += NT) {
                    const int j = j0 + lane;
                    const int jc = j < kml ? j : kml - 1;
                    int64_t z = list_sum(PL, nl, GLw, CHl, j0, jc, m) +
                                list_sum(PU, nu, GUw, CHu, j0, jc, m);
                    const int hi = jc + m - 1;
                    if ((GLw[hi >> 6] >> (hi & 63)) & 1) z += CHl[m];
                    if ((GUw[hi >> 6] >> (hi & 63)) & 1) z += CHu[m];
                    if (j < kml) bound(j, finite, zfix(z));
                }
            } else {
                for (int j = jfirst + threadIdx.x; j < kml; j += blockDim.x)
                    bound(j, finite,
                          zfix(trans_sum(GLw, TLw, CHl, j, m) + trans_sum(GUw, TUw, CHu, j, m)));
            }
        }
        if (a.rep_lb) {
            rep = block_best<false>(rep, sh_best);
            if (threadIdx.x == 0) {
                a.rep_lb[s] = rep.v;
                a.rep_idx[s] = (rep.v == rep.v) ? rep.i : -1;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_lbip(hipStream_t st, const LbipArgs& a) {
    const int npad = npad_of(a.G.m_k - 2 * a.G.m + 2);
    const dim3 g((unsigned)a.blocks);
    const int nt = (a.nt && lbip_threads_ok(a.G, a.nt)) ? a.nt : lbip_threads(npad, a.G.S);
    const size_t lds = lbip_lds_bytes(a.G, lbip_layout(a.G.m_k) == kDual ? nt : 0);
    const bool f32 = a.p.lbip_f32 != 0;
    switch (lbip_layout(a.G.m_k)) {
        case kDual:
            if (f32) {
                if (nt == 128)
                    hipLaunchKernelGGL((k_lbip<128, kDual, true>), g, dim3(128), lds, st, a);
                else if (nt == 256)
                    hipLaunchKernelGGL((k_lbip<256, kDual, true>), g, dim3(256), lds, st, a);
                else if (nt == 512)
                    hipLaunchKernelGGL((k_lbip<512, kDual, true>), g, dim3(512), lds, st, a);
                else
                    hipLaunchKernelGGL((k_lbip<1024, kDual, true>), g, dim3(1024), lds, st, a);
            } else if (nt == 128) hipLaunchKernelGGL((k_lbip<128, kDual>), g, dim3(128), lds, st, a);
            else if (nt == 256) hipLaunchKernelGGL((k_lbip<256, kDual>), g, dim3(256), lds, st, a);
            else if (nt == 512) hipLaunchKernelGGL((k_lbip<512, kDual>), g, dim3(512), lds, st, a);
            else hipLaunchKernelGGL((k_lbip<1024, kDual>), g, dim3(1024), lds, st, a);
            break;
        case kSingle:  // m_k = 16384: 1024 threads
            hipLaunchKernelGGL((k_lbip<1024, kSingle>), g, dim3(1024), lds, st, a);
            break;
        default:  // m_k > 16384: 1024 threads
            hipLaunchKernelGGL((k_lbip<1024, kSlab>), g, dim3(1024), lds, st, a);
    }
    return hipGetLastError();
}

}  // namespace fsm
