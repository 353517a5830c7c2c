// k_eld.hip -- exact elastic distances (ELD) on gfx950, one candidate per lane.
//
// Every cell is computed with the reference's expression and operand order
// (src/ElasticMeasures.cpp: dtw :24-61, lcss :63-93, erp :95-130, edr
// :132-165, msm :167-198); a cell's value depends only on its three
// neighbours, so any evaluation order gives the reference's bits.
//
// Anti-diagonal schedule.  Band cells (i, j), |i - j| <= w, are held in
// registers by slot s = i - j + W.  Anti-diagonal d = i + j updates the slots
// of one parity from the other parity (anti-diagonal d-1: "up" (i-1, j) and
// "left" (i, j-1)) and from their own old value (d-2: "diag" (i-1, j-1)), so
// the cells of one anti-diagonal are independent: a lane has W+1 independent
// cells per step.  Row / column indices are wave-uniform, so boundary cases
// are scalar branches.  Query samples are wave-uniform; candidate samples are
// normalised once each ((t - mu) / sigma, src/fsmeld.cpp:822-826) into a
// register ring.
//
// Early abandon (search / exact / cutoff collection only): every path crosses
// one of any two consecutive anti-diagonals, and after that cell it still has
// to consume every candidate sample beyond the cell's column.  Consuming
// sample y_k costs at least its distance to the query envelope over the band
// (GLB's per-position term: d^2 for DTW, min(d^2, (y-g)^2) for ERP, [d > eps]
// for EDR, min(d, c) for MSM).  So the final cost is at least
//   min over the two anti-diagonals of the cell value + sum of those terms
//   over the samples beyond the largest column of the two anti-diagonals,
// and once that exceeds the cutoff (squared for DTW/ERP, relative margin
// 1e-12 plus the rounding terms below) the window cannot match and is
// dropped.  LCSS counts matches instead: at most one per remaining sample
// that lies within eps of the envelope.  The distances of surviving windows
// are unchanged.
#include "engine_internal.h"

namespace fsm {

// DP values are sums of non-negative costs (or LCSS counts): never NaN and
// never -0, so fmin / fmax (v_min_f64 / v_max_f64) return exactly what
// std::min({...}) / std::max({...}) return.  This file is compiled with
// -fno-honor-nans (all inputs are finite: valid windows, a validated query),
// which lets them lower to single v_min_f64 / v_max_f64 instructions.
__device__ inline double dmin3(double a, double b, double c) { return fmin(fmin(a, b), c); }
__device__ inline double dmax3(double a, double b, double c) { return fmax(fmax(a, b), c); }
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

// src/ElasticMeasures.cpp:173-176 -- c if np lies between x and y (inclusive),
// else c + the distance to the nearer one.  Written as c + max(np - hi,
// lo - np, 0): the two differences are the same rounded values the reference
// takes the min of, and c + 0 = c.
__device__ inline double msm_cost(double np, double x, double y, double c) {
    const double hi = fmax(x, y);
    const double lo = fmin(x, y);
    return c + fmax(fmax(np - hi, lo - np), 0.0);
}

// Candidate sample source: raw target samples normalised on the fly, or rows
// that are already normalised.
struct CandSrc {
    const double* t;
    double mu, sd;
    bool raw;
    int m;
    __device__ double at(int j) const {
        if (j < 0 || j >= m) return 0.0;
        return raw ? (t[j] - mu) / sd : t[j];
    }
};

__device__ inline double qat(const double* __restrict__ q, int i, int m) {
    return (i >= 0 && i < m) ? q[i] : 0.0;
}

// One cell.  DTW / LCSS / MSM index samples by (i, j); ERP / EDR use the
// (m+1)^2 grid with gap row / column and samples x[i-1], y[j-1] (passed as
// xp / yp, with xg = (x[i-1]-g)^2, yg = (y[j-1]-g)^2).  EDGE = false: the
// caller guarantees i > 0 and j > 0.
template <int METRIC, bool EDGE>
__device__ inline double cell(int i, int j, double up, double left, double diag, double xi,
                              double xp, double yj, double yp, double xg, double yg,
                              double arg) {
    const double S = kEldSentinel;
    if (METRIC == DTW) {
        const double d = (xi - yj) * (xi - yj);
        if (EDGE) {
            if (i == 0 && j == 0) return d;
            const double u = (i > 0) ? up : S;
            const double l = (j > 0) ? left : S;
            const double g = (i > 0 && j > 0) ? diag : S;
            return dmin3(u, l, g) + d;
        }
        return dmin3(up, left, diag) + d;
    } else if (METRIC == ERP) {
        if (EDGE) {
            if (i == 0 && j == 0) return 0.0;
            if (i == 0) return left + yg;
            if (j == 0) return up + xg;
        }
        const double match = diag + (xp - yp) * (xp - yp);
        return dmin3(match, left + yg, up + xg);
    } else if (METRIC == EDR) {
        if (EDGE) {
            if (i == 0 && j == 0) return 0.0;
            if (i == 0) return (double)j;
            if (j == 0) return (double)i;
        }
        const double sub = (fabs(xp - yp) <= arg) ? 0.0 : 1.0;
        return dmin3(up + 1.0, left + 1.0, diag + sub);
    } else if (METRIC == LCSS) {
        double d1 = diag, u = up, l = left;
        if (EDGE) {
            d1 = (i > 0 && j > 0) ? diag : 0.0;
            u = (i > 0) ? up : 0.0;
            l = (j > 0) ? left : 0.0;
        }
        return (fabs(xi - yj) <= arg) ? dmax3(d1 + 1, u, l) : dmax3(d1, u, l);
    } else {  // MSM
        if (EDGE) {
            if (i == 0 && j == 0) return fabs(xi - yj);
            if (i == 0) return left + msm_cost(yj, xi, yp, arg);
            if (j == 0) return up + msm_cost(xi, xp, yj, arg);
        }
        const double c1 = diag + fabs(xi - yj);
        const double c2 = up + msm_cost(xi, xp, yj, arg);
        const double c3 = left + msm_cost(yj, xi, yp, arg);
        return dmin3(c1, c2, c3);
    }
}

// Query staging in LDS: xs[kXPad + i] = x[i] for i in [0, m), 0 elsewhere in
// [-kXPad, m + kXPad); for ERP xg holds (x[i] - g)^2 likewise.
constexpr int kXPad = 40;
constexpr size_t kEldLdsCap = 64 * 1024;

// Anti-diagonal DP for one lane, compile-time band half-width W >= w
// (EXACT: w == W, so every slot is in band).  Iteration b handles
// anti-diagonals dA = 2b + W (even slots 2r, cells (b + r, b + W - r)) and
// dB = dA + 1 (odd slots 2r+1, cells (b + 1 + r, b + W - r)).  The candidate
// ring holds Y[k] = y[b - 1 + k], k in [0, W+1]; a cell reads y[j] =
// Y[j - b + 1] and y[j-1] = Y[j - b]; query samples come from LDS at
// wave-uniform addresses.
// One iteration (two anti-diagonals) of the band update; EDGE selects the
// boundary-checked form.
template <int METRIC, int W, bool EXACT, bool EDGE>
__device__ __forceinline__ void ad_step(int b, int M, int w, double arg, double* A, double* B,
                                        const double* X, const double* Y) {
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
#pragma unroll
    for (int r = 0; r <= W; r++) {
        const int i = b + r, j = b + W - r;
        const int delta = 2 * r - W;
        const bool inb = EXACT || (delta <= w && delta >= -w);
        const bool inr = !EDGE || (i >= 0 && j >= 0 && i <= M && j <= M);
        if (inb && inr) {
            const double up = (r >= 1) ? B[r - 1] : S;
            const double left = (r <= W - 1) ? B[r] : S;
            A[r] = cell<METRIC, EDGE>(i, j, up, left, A[r], X[r + 1], X[r], Y[W - r + 1],
                                      Y[W - r],
                                      METRIC == ERP ? (X[r] - arg) * (X[r] - arg) : 0.0,
                                      METRIC == ERP ? (Y[W - r] - arg) * (Y[W - r] - arg) : 0.0,
                                      arg);
        } else if (EDGE) {
            A[r] = S;
        }
    }
#pragma unroll
    for (int r = 0; r < W; r++) {
        const int i = b + 1 + r, j = b + W - r;
        const int delta = 2 * r + 1 - W;
        const bool inb = EXACT || (delta <= w && delta >= -w);
        const bool inr = !EDGE || (i >= 0 && j >= 0 && i <= M && j <= M);
        if (inb && inr) {
            B[r] = cell<METRIC, EDGE>(i, j, A[r], A[r + 1], B[r], X[r + 2], X[r + 1],
                                      Y[W - r + 1], Y[W - r],
                                      METRIC == ERP ? (X[r + 1] - arg) * (X[r + 1] - arg) : 0.0,
                                      METRIC == ERP ? (Y[W - r] - arg) * (Y[W - r] - arg) : 0.0,
                                      arg);
        } else if (EDGE) {
            B[r] = S;
        }
    }
}

template <int W>
__device__ __forceinline__ void ring_shift(double* X, double* Y, double xn, double yn) {
#pragma unroll
    for (int k = 0; k < W + 1; k++) {
        Y[k] = Y[k + 1];
        X[k] = X[k + 1];
    }
    Y[W + 1] = yn;
    X[W + 1] = xn;
}

// Lower bound on the cost of consuming candidate sample y whose band sees the
// query envelope [lo, hi]; for LCSS, 1 if y can match at all.  Rounding: the
// cells' costs are at least these values times (1 - 7u) (u = 2^-53), and the
// LCSS / EDR comparisons are widened by a relative 1e-12, far above the
// rounding of |x - y| against e.
template <int METRIC>
__device__ inline double consume_lb(double y, double lo, double hi, double arg) {
    const double e = fmax(fmax(lo - y, y - hi), 0.0);
    if (METRIC == DTW) return e * e;
    if (METRIC == ERP) return fmin(e * e, (y - arg) * (y - arg));
    if (METRIC == EDR) return (e > arg * (1.0 + 1e-12) + 1e-300) ? 1.0 : 0.0;
    if (METRIC == MSM) return fmin(e, arg);
    return (e <= arg * (1.0 + 1e-12) + 1e-300) ? 1.0 : 0.0;
}

// total: sum of consume_lb over all m samples, p: the same sum, in the same
// order, over samples 0 .. K (K = the frontier's column bound for this check).
// fl(total - p) is within (m + 1) u total of the exact suffix sum; the path
// beyond the frontier adds at most 2m + 2 rounded terms.
template <int METRIC, int W>
__device__ __forceinline__ bool ad_alive(const double* A, const double* B, int b, int M, int m,
                                         double thr, double total, double p) {
    if (METRIC == LCSS) {
        double mx = A[0];
#pragma unroll
        for (int k = 1; k <= W; k++) mx = fmax(mx, A[k]);
#pragma unroll
        for (int k = 0; k < W; k++) mx = fmax(mx, B[k]);
        // matches still possible after a cell of anti-diagonal >= 2b+W
        const double rem = fmin((double)((2 * M - (2 * b + W)) / 2 + 1), total - p);
        return !(1.0 - (mx + rem) / (double)m > thr);
    }
    double mn = A[0];
#pragma unroll
    for (int k = 1; k <= W; k++) mn = fmin(mn, A[k]);
#pragma unroll
    for (int k = 0; k < W; k++) mn = fmin(mn, B[k]);
    const double cb = fmax((total - p) - (double)(m + 8) * 1.2e-16 * total, 0.0);
    return !((mn + cb) * (1.0 - (double)(2 * m + 24) * 1.2e-16) > thr);
}

template <int METRIC, int W, bool EXACT>
__device__ double eld_ad(const double* xs, const double* xgs, const CandSrc& c, int m, int w,
                         double arg, double thr, const double* lqs, const double* uqs,
                         bool active, bool* dead_out) {
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR);
    // after iteration b the ring holds Y[k] = y[b + k]; the frontier's columns
    // are at most b + W (grid: samples up to b + W - 1 are consumed); LCSS
    // bounds matches by samples beyond the smallest column, b
    constexpr int KOFF = (METRIC == LCSS) ? 0 : (GRID1 ? W - 1 : W);
    const int M = GRID1 ? m : m - 1;
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
    double A[W + 1], B[W + 1];
    double X[W + 2], Y[W + 2];
    const int b0 = -((W + 1) / 2) - 1;
    const int bend = (2 * M - W) >= 0 ? (2 * M - W) / 2 : -((W - 2 * M + 1) / 2);
    // interior iterations [1, M - W]: every cell in range with i, j >= 1
    const int bi0 = 1;
    const int bi1 = M - W;
#pragma unroll
    for (int k = 0; k <= W; k++) {
        A[k] = S;
        B[k] = S;
    }
#pragma unroll
    for (int k = 0; k < W + 2; k++) {
        Y[k] = c.at(b0 - 1 + k);
        X[k] = xs[kXPad + b0 - 1 + k];
    }
    bool alive = active;
    const bool check = thr < 1e300;
    // p starts as the prefix of total through sample b0 + KOFF (the first
    // iteration's bound) and then follows the same summation order
    double total = 0.0, p = 0.0;
    if (check) {
        constexpr int U = 8;
        int k = 0;
        for (; k + U <= m; k += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = c.at(k + u);
#pragma unroll
            for (int u = 0; u < U; u++) {
                total += consume_lb<METRIC>(v[u], lqs[kXPad + k + u], uqs[kXPad + k + u], arg);
                p = (k + u == b0 + KOFF) ? total : p;
            }
        }
        for (; k < m; k++) {
            total += consume_lb<METRIC>(c.at(k), lqs[kXPad + k], uqs[kXPad + k], arg);
            p = (k == b0 + KOFF) ? total : p;
        }
    }
    int b = b0;
    // head
    for (; b <= bend && b < bi0; b++) {
        ad_step<METRIC, W, EXACT, true>(b, M, w, arg, A, B, X, Y);
        ring_shift<W>(X, Y, xs[kXPad + b + W + 1], c.at(b + W + 1));
        const int K = b + KOFF;
        if (check && b > b0 && K >= 0 && K < m)
            p += consume_lb<METRIC>(Y[KOFF], lqs[kXPad + K], uqs[kXPad + K], arg);
    }
    // interior, early abandon every 4 iterations
    int it = 0;
    for (; b <= bi1; b++, it++) {
        ad_step<METRIC, W, EXACT, false>(b, M, w, arg, A, B, X, Y);
        ring_shift<W>(X, Y, xs[kXPad + b + W + 1], c.at(b + W + 1));
        if (check) {
            const int K = b + KOFF;
            if (K < m) p += consume_lb<METRIC>(Y[KOFF], lqs[kXPad + K], uqs[kXPad + K], arg);
            if ((it & 3) == 3) {
                alive = alive && ad_alive<METRIC, W>(A, B, b, M, m, thr, total, p);
                if (!__any(alive)) break;
            }
        }
    }
    double result = S;
    if (__any(alive) || !check) {
        // tail (the final cell (M, M) sits in slot W of iteration bend)
        for (b = (b > bi1 ? b : bi1 + 1); b <= bend; b++) {
            ad_step<METRIC, W, EXACT, true>(b, M, w, arg, A, B, X, Y);
            ring_shift<W>(X, Y, xs[kXPad + b + W + 1], c.at(b + W + 1));
        }
        result = (W & 1) ? B[(W - 1) / 2] : A[W / 2];
    }
    *dead_out = !alive;
    if (METRIC == DTW || METRIC == ERP) {
        if (!__builtin_isfinite(result)) return S;
        return sqrt(smax(0.0, result));
    }
    if (METRIC == LCSS) return 1 - result / (double)m;
    return result;
}

// ---------------------------------------------------------------- generic ---
// Radii above the register variants: the same per-cell expressions evaluated
// row by row with the band in lane-interleaved global scratch.
struct MemArr {
    double* base;
    int64_t stride;
    __device__ double& operator[](int k) { return base[(int64_t)k * stride]; }
};

template <int METRIC>
__device__ double eld_rows(const double* __restrict__ q, const double* __restrict__ qg,
                           const CandSrc& c, int m, int w, double arg, MemArr b, MemArr y) {
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR);
    const int M = GRID1 ? m : m - 1;
    const int xoff = GRID1 ? 1 : 0;
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
    const int w2 = 2 * w;
    // band slot k = j - i + w; y[k] holds sample j - xoff of the current row
    for (int k = 0; k <= w2 + 1; k++) {
        b[k] = S;
        y[k] = c.at(k - w - xoff);
    }
    for (int i = 0; i <= M; i++) {
        const double xi = qat(q, i - xoff, m), xp = qat(q, i - 1 - xoff, m);
        const double xg = (METRIC == ERP) ? qat(qg, i - xoff, m) : 0.0;
        for (int k = 0; k <= w2; k++) {
            const int j = i - w + k;
            double v = S;
            if (j >= 0 && j <= M) {
                const double yj = y[k];
                const double yp = (k > 0) ? y[k - 1] : c.at(j - 1 - xoff);
                const double up = (k < w2) ? b[k + 1] : S;
                const double left = (k > 0) ? b[k - 1] : S;
                if (GRID1)
                    v = cell<METRIC, true>(i, j, up, left, b[k], 0.0, xi, 0.0, yj, xg,
                                     (yj - arg) * (yj - arg), arg);
                else
                    v = cell<METRIC, true>(i, j, up, left, b[k], xi, xp, yj, yp, 0.0, 0.0, arg);
            }
            b[k] = v;
        }
        const double ynew = c.at(i + 1 + w - xoff);
        for (int k = 0; k < w2; k++) y[k] = y[k + 1];
        y[w2] = ynew;
    }
    const double res = b[w];
    if (METRIC == DTW || METRIC == ERP) {
        if (!__builtin_isfinite(res)) return S;
        return sqrt(smax(0.0, res));
    }
    if (METRIC == LCSS) return 1 - res / (double)m;
    return res;
}

// ----------------------------------------------------------------- kernel ---
// in[] lists window indices (target mode) or, with rows != null, row numbers
// 0..n_direct-1 of normalised candidates.  W == 0: generic row-band path.
template <int METRIC, int W, bool EXACT>
__global__ void __launch_bounds__(256) k_eld(VerifyArgs a, const double* __restrict__ LQ,
                                             const double* __restrict__ UQ,
                                             const int64_t* __restrict__ in,
                                             const unsigned long long* __restrict__ in_count,
                                             int64_t n_direct, const double* __restrict__ rows,
                                             int keep_all, int64_t* __restrict__ out_idx,
                                             double* __restrict__ out_dist,
                                             unsigned long long* out_count,
                                             double* __restrict__ scratch) {
    extern __shared__ double xlds[];
    const int64_t n = in_count ? (int64_t)*in_count : n_direct;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t iters = (n + stride - 1) / stride;
    const int m = a.p.m, w = a.p.w;
    const double cut = a.p.cutoff;
    double thr = 1e308;
    // abandoning needs non-negative step costs (MSM: c >= 0)
    if (!keep_all && !(METRIC == MSM && a.p.arg < 0.0)) {
        if (METRIC == DTW || METRIC == ERP) thr = cut * cut * (1.0 + 1e-12) + 1e-300;
        else thr = cut * (1.0 + 1e-12) + 1e-300;
    }
    const int span = m + 2 * kXPad;
    // LDS: query, [ERP: gap costs], query envelope lower, upper (the last two
    // only when abandoning)
    constexpr int NX = (METRIC == ERP) ? 2 : 1;
    double* lqs = xlds + NX * span;
    double* uqs = lqs + span;
    if (W > 0) {
        for (int i = threadIdx.x; i < span; i += blockDim.x) {
            const int qi = i - kXPad;
            const bool in_q = qi >= 0 && qi < m;
            xlds[i] = in_q ? a.q.qn[qi] : 0.0;
            if (METRIC == ERP) xlds[span + i] = in_q ? a.q.XG[qi] : 0.0;
            if (!keep_all) {
                lqs[i] = in_q ? LQ[qi] : 0.0;
                uqs[i] = in_q ? UQ[qi] : 0.0;
            }
        }
        __syncthreads();
    }
    int64_t ci = tid;
    for (int64_t itx = 0; itx < iters; itx++, ci += stride) {
        const bool have = ci < n;
        CandSrc c;
        c.m = m;
        int64_t g = 0;
        if (rows) {
            g = have ? ci : 0;
            c.t = rows + g * (int64_t)m;
            c.raw = false;
            c.mu = 0.0;
            c.sd = 1.0;
        } else {
            g = have ? in[ci] : in[0];
            const int64_t si = stat_index(a.G, g);
            c.t = a.T + g;
            c.mu = a.MU[si];
            c.sd = a.SG[si];
            c.raw = true;
        }
        bool dead = false;
        double d;
        if constexpr (W > 0) {
            d = eld_ad<METRIC, W, EXACT>(xlds, xlds + span, c, m, w, a.p.arg, thr, lqs, uqs, have,
                                         &dead);
        } else {
            MemArr b{scratch + tid, stride * 2};
            MemArr y{scratch + tid + stride, stride * 2};
            d = eld_rows<METRIC>(a.q.qn, a.q.XG, c, m, w, a.p.arg, b, y);
        }
        if (keep_all) {
            if (have) out_dist[ci] = d;
            continue;
        }
        const bool keep = have && !dead && __builtin_isfinite(d) && d <= cut;  // src/fsmeld.cpp:846-847
        const uint64_t mask = __ballot(keep);
        if (mask) {
            const int lane = threadIdx.x & 63;
            const int leader = __builtin_ctzll(mask);
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(out_count, (unsigned long long)__builtin_popcountll(mask));
            base = __shfl(base, leader, 64);
            if (keep) {
                const uint64_t below = lane ? (mask & ((1ull << lane) - 1)) : 0ull;
                const unsigned long long pos = base + __builtin_popcountll(below);
                out_idx[pos] = g;
                out_dist[pos] = d;
            }
        }
    }
}

int64_t eld_scratch_threads() { return (int64_t)512 * 256; }

bool eld_uses_scratch(int metric, int m, int w) {
    const size_t lds = sizeof(double) * (size_t)(m + 2 * kXPad) * (metric == ERP ? 4 : 3);
    if (lds > kEldLdsCap) return true;
    return w > 32 || (w > 25 && metric == ERP);
}

template <int METRIC>
static hipError_t launch_metric(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                                const unsigned long long* in_count, int64_t n_direct,
                                const double* rows, bool keep_all, int64_t* out_idx,
                                double* out_dist, unsigned long long* out_count,
                                double* scratch, int64_t scratch_threads) {
    const int w = a.p.w;
    const dim3 blk(256), grid(2048);
    const int ka = keep_all ? 1 : 0;
    // register variants stage query, [ERP gap costs,] and envelope in LDS
    const size_t lds = sizeof(double) * (size_t)(a.p.m + 2 * kXPad) * (METRIC == ERP ? 4 : 3);
    const bool fits = lds <= kEldLdsCap;
#define FSM_ELD(WW, EX) \
    hipLaunchKernelGGL((k_eld<METRIC, WW, EX>), grid, blk, lds, st, a, a.q.LQ, a.q.UQ, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch)
    if (fits && w == 6) FSM_ELD(6, true);
    else if (fits && w == 25) FSM_ELD(25, true);
    else if (fits && w <= 8) FSM_ELD(8, false);
    else if (fits && w <= 16) FSM_ELD(16, false);
    else if (fits && w <= 25) FSM_ELD(25, false);
    else if (fits && w <= 32 && METRIC != ERP) FSM_ELD(32, false);
    else {
        if (!scratch) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_eld<METRIC, 0, false>), dim3((unsigned)(scratch_threads / 256)), blk, 0, st, a, a.q.LQ, a.q.UQ, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch);
    }
#undef FSM_ELD
    return hipGetLastError();
}

static hipError_t dispatch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                               const unsigned long long* in_count, int64_t n_direct,
                               const double* rows, bool keep_all, int64_t* out_idx,
                               double* out_dist, unsigned long long* out_count,
                               double* scratch, int64_t scratch_threads) {
    switch (a.p.metric) {
        case DTW: return launch_metric<DTW>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, scratch_threads);
        case ERP: return launch_metric<ERP>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, scratch_threads);
        case EDR: return launch_metric<EDR>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, scratch_threads);
        case LCSS: return launch_metric<LCSS>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, scratch_threads);
        case MSM: return launch_metric<MSM>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, scratch_threads);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const unsigned long long* in_count, bool keep_all, int64_t* out_idx,
                      double* out_dist, unsigned long long* out_count, double* scratch,
                      int64_t scratch_threads) {
    return dispatch_eld(st, a, in, in_count, 0, nullptr, keep_all, out_idx, out_dist,
                        out_count, scratch, scratch_threads);
}

hipError_t launch_eld_pairs(hipStream_t st, const VerifyArgs& a, const double* cands, int64_t n,
                            double* out, double* scratch, int64_t scratch_threads) {
    return dispatch_eld(st, a, nullptr, nullptr, n, cands, true, nullptr, out, nullptr, scratch,
                        scratch_threads);
}

}  // namespace fsm
