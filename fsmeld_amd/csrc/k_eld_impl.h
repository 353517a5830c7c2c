// k_eld_impl.h -- exact elastic distances (ELD) on gfx950: kernel templates and
// the per-metric launcher, instantiated once per metric (k_eld_<metric>.hip).
// One candidate per lane (register band), per wave (wide bands) or per thread
// with the band in global scratch (row band).
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
// cells per step.  The matrix boundary is padding (see ad_step), so every
// step runs the same branch-free code and each lane advances through its own
// candidate at its own iteration: a lane whose candidate ends takes the next
// one (lane refill), and a wave no longer waits for its slowest lane before
// starting new work.  Query samples come from LDS; candidate samples are
// normalised once each ((t - mu) / sigma, src/fsmeld.cpp:822-826) into a
// register ring.
//
// Early abandon (search / exact / cutoff collection only): every path crosses
// one of any two consecutive anti-diagonals, and after that cell it still has
// to consume every candidate sample beyond the cell's column.  Consuming
// sample y_k costs at least its distance to the query envelope over the band
// (GLB's per-position term: d^2 for DTW, min(d^2, (y-g)^2) for ERP, [d > eps]
// for EDR, min(d, c) for MSM), whose sum over the window the GLB pass hands
// over with each surviving candidate.  So the final cost is at least
//   min over the two anti-diagonals of the cell value + sum of those terms
//   over the samples beyond the largest column of the two anti-diagonals,
// and once that exceeds the cutoff (squared for DTW/ERP, relative margin
// 1e-12 plus the rounding terms below) the window cannot match and is
// dropped.  LCSS counts matches instead: at most one per remaining sample
// that lies within eps of the envelope.  The distances of surviving windows
// are unchanged.
#pragma once
#include <algorithm>

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

// MSM cell of the register band from e = x_i - y_j, dx = x_i - x_{i-1} and
// dy = y_j - y_{j-1}.  Rounded subtraction is monotone and odd, so
// fl(np - max(a, b)) = min(fl(np - a), fl(np - b)) and
// fl(min(a, b) - np) = -max(fl(np - a), fl(np - b)): msm_cost(x_i, x_{i-1}, y_j)
// = c + max(min(dx, e), -max(dx, e), 0) and msm_cost(y_j, x_i, y_{j-1})
// = c + max(min(-e, dy), -max(-e, dy), 0), bit for bit, sharing e with the
// match cost |x_i - y_j|.
//
// Both terms have the form max(min(d, f), -max(d, f), 0), which equals
// |clamp(f, min(0, d), max(0, d))|: for d > 0 both are clamp(f, 0, d); for
// d < 0 both are clamp(-f, 0, -d); for d = 0 both are 0 (a zero's sign is
// absorbed by c + s).  With f = -e and d = dy the clamp is
// -clamp(e, min(0, -dy), max(0, -dy)), and dyn = y_{j-1} - y_j is exactly -dy.
// The clamp bounds belong to a row (column) and serve the two cells of an
// iteration that share it; min, max and |.| are exact, so the values are the
// reference's bit for bit.
__device__ __forceinline__ double msm_split(double d, double e) {
    return fabs(fmax(fmin(e, fmax(d, 0.0)), fmin(d, 0.0)));
}

__device__ inline double msm_cell(double up, double left, double diag, double e, double dx,
                                  double dyn, double c) {
    const double c1 = diag + fabs(e);
    const double s2 = msm_split(dx, e);
    const double s3 = msm_split(dyn, e);
    const double c2 = up + (c + s2);
    const double c3 = left + (c + s3);
    return dmin3(c1, c2, c3);
}
// The same with the row's clamp bounds given: dxlo = min(0, dx), dxhi =
// max(0, dx) (the register kernel stages them per query row in LDS).
__device__ __forceinline__ double msm_cell_rb(double up, double left, double diag, double e,
                                              double dxlo, double dxhi, double dyn, double c) {
    const double c1 = diag + fabs(e);
    const double s2 = fabs(fmax(fmin(e, dxhi), dxlo));
    const double s3 = msm_split(dyn, e);
    const double c2 = up + (c + s2);
    const double c3 = left + (c + s3);
    return dmin3(c1, c2, c3);
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
    } else if (METRIC == SWALE) {  // src/ElasticMeasures.cpp:200-231, p = 5, r = 1
        if (EDGE) {
            if (i == 0 && j == 0) return 0.0;
            if (i == 0) return left + kSwaleP;
            if (j == 0) return up + kSwaleP;
        }
        const double mv = (fabs(xp - yp) <= arg) ? kSwaleR : kSwaleP;
        return dmin3(diag + mv, up + kSwaleP, left + kSwaleP);
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

// ------------------------------------------------------ register variants ---
// Query staging in LDS: xs[i] = x[i] for i in [0, m) and kPad for the kXPad
// entries on either side; with abandoning, the query envelope LQ / UQ follows.
constexpr int kXPad = 40;
constexpr size_t kEldLdsCap = 64 * 1024;
constexpr double kPad = 1e300;
// idle lanes that trigger a wave's refill (the integer EDR / LCSS kernels'
// iterations are cheaper, so their refills are spaced further apart)
template <int METRIC>
constexpr int kRefillMin = (METRIC == EDR || METRIC == LCSS || METRIC == SWALE) ? 12 : 8;
// queue positions a wave takes per atomic; refills draw on the wave's pool
// (short-lived candidates would otherwise make one atomic per few lanes on
// a single counter)
constexpr unsigned long long kEldChunk = 64;

// Padded boundary.  Query and candidate samples outside [0, m) read as kPad;
// the two anti-diagonals before the first iteration start as the
// out-of-range sentinel S (0 for LCSS) except the virtual cell (-1, -1) = 0,
// which seeds (0, 0) through its diagonal.  Then the interior recurrence
// gives every in-range cell the reference's value: a cell with a negative
// index costs a padded sample against a real one (>= 1e300: its square is
// infinite, an LCSS match impossible, an MSM split / merge between two pads
// costs c + ~1e300), so such cells stay at S or above and min / max always
// pick the reference's own operand; the row / column 0 formulas of
// src/ElasticMeasures.cpp fall out exactly (x + min(S, v, S) = x + v).  With
// odd W the first iteration recomputes anti-diagonal -1, whose cells see the
// virtual cell as up / left, each through a cost of at least 1e300 (EDR: its
// unit costs leave row / column -1 at j + 1 / i + 1, never below row /
// column 0's values).  For ERP / EDR cell (0, 0) matches pad against pad at
// cost 0 (EDR with eps < 0 costs 1: the virtual cell is -1 there).  Cells
// past row or column M never feed (M, M).

// One iteration (two anti-diagonals) of the band update.  Iteration b
// handles anti-diagonals dA = 2b + W (even slots 2r, cells (b + r, b + W - r))
// and dB = dA + 1 (odd slots 2r+1, cells (b + 1 + r, b + W - r)); the rings
// give X[k] = x[b - 1 + k], Y[k] = y[b - 1 + k].  A cell reads x[i] = X[r+1]
// (x[i-1] = X[r]) and y[j] = Y[j - b + 1] (y[j-1] = Y[j - b]).
// Register-band state.  EDR and LCSS values are edit and match counts:
// integers far below 2^24, so 32-bit integer arithmetic gives the reference's
// values exactly (its doubles hold the same integers), with half the registers
// and fewer instructions per cell.  Their sample comparisons stay in FP64.
// The out-of-band neighbour is a sentinel no reachable value exceeds (EDR) or
// 0 (LCSS, where cells take maxima).  EDR slots hold D(i, j) - (i + j), the
// same for a whole anti-diagonal, so D = min(up + 1, left + 1, diag + sub)
// becomes min(up', left', diag' + sub - 2): one three-way min and one add per
// cell on exactly the same integers (band_value adds i + j back).  The
// sentinel loses at most 2 per anti-diagonal, far from reachable values.
template <int METRIC> struct EldState {
    using type = double;
    static constexpr int waves = 2;  // waves per SIMD the register kernel is built for
};
template <> struct EldState<EDR> {
    using type = int;
    static constexpr int waves = 3;
};
template <> struct EldState<LCSS> {
    using type = int;
    static constexpr int waves = 3;
};
// SWALE's values are sums of its integer costs (1 and 5): int slots hold D
// itself, at most 10 m plus the sentinel's growth of 5 per anti-diagonal
template <> struct EldState<SWALE> {
    using type = int;
    static constexpr int waves = 3;
};
template <int METRIC> using eld_t = typename EldState<METRIC>::type;
constexpr int kEdrSentinel = 1 << 29;
__device__ __forceinline__ int band_min(int a, int b) { return min(a, b); }
__device__ __forceinline__ double band_min(double a, double b) { return fmin(a, b); }
// value of a slot on anti-diagonal d (= i + j)
template <int METRIC>
__device__ __forceinline__ double band_value(eld_t<METRIC> v, int d) {
    if constexpr (METRIC == EDR) return (double)(v + d);
    else return (double)v;
}
template <int METRIC>
__device__ __forceinline__ eld_t<METRIC> band_sentinel() {
    if constexpr (METRIC == LCSS) return 0;
    else if constexpr (METRIC == EDR || METRIC == SWALE) return kEdrSentinel;
    else return kEldSentinel;
}

// Slot update of the even anti-diagonal (A[r], cell (b + r, b + W - r)).
template <int METRIC, int W>
__device__ __forceinline__ void step_a(int r, double arg, eld_t<METRIC>* A,
                                       const eld_t<METRIC>* B, const double* X,
                                       const double* Y, const double* DL, const double* DH) {
    using VT = eld_t<METRIC>;
    const VT S = band_sentinel<METRIC>();  // out-of-band neighbour
    const VT up = (r >= 1) ? B[r - 1] : S;
    const VT left = (r <= W - 1) ? B[r] : S;
    if constexpr (METRIC == EDR) {  // x[i-1] = X[r], y[j-1] = Y[W - r]
        const int sub = (fabs(X[r] - Y[W - r]) <= arg) ? 0 : 1;
        A[r] = min(min(up, left), A[r] - 2 + sub);  // D - (i + j)
        return;
    } else if constexpr (METRIC == LCSS) {  // x[i] = X[r + 1], y[j] = Y[W - r + 1]
        const int hit = (fabs(X[r + 1] - Y[W - r + 1]) <= arg) ? 1 : 0;
        A[r] = max(A[r] + hit, max(up, left));
        return;
    } else if constexpr (METRIC == SWALE) {  // x[i-1] = X[r], y[j-1] = Y[W - r]
        const int mv = (fabs(X[r] - Y[W - r]) <= arg) ? 1 : 5;
        A[r] = min(min(up, left) + 5, A[r] + mv);
        return;
    }
    if (METRIC == MSM) {  // DL / DH[k]: the bounds of x[b - 1 + k] - x[b - 2 + k]
        A[r] = msm_cell_rb(up, left, A[r], X[r + 1] - Y[W - r + 1], DL[r + 1], DH[r + 1],
                           Y[W - r] - Y[W - r + 1], arg);
        return;
    }
    A[r] = cell<METRIC, false>(0, 0, up, left, A[r], X[r + 1], X[r], Y[W - r + 1], Y[W - r],
                               METRIC == ERP ? (X[r] - arg) * (X[r] - arg) : 0.0,
                               METRIC == ERP ? (Y[W - r] - arg) * (Y[W - r] - arg) : 0.0, arg);
}

// Slot update of the odd anti-diagonal (B[r], cell (b + 1 + r, b + W - r)).
template <int METRIC, int W>
__device__ __forceinline__ void step_b(int r, double arg, const eld_t<METRIC>* A,
                                       eld_t<METRIC>* B, const double* X, const double* Y,
                                       const double* DL, const double* DH) {
    if constexpr (METRIC == EDR) {  // x[i-1] = X[r + 1], y[j-1] = Y[W - r]
        const int sub = (fabs(X[r + 1] - Y[W - r]) <= arg) ? 0 : 1;
        B[r] = min(min(A[r], A[r + 1]), B[r] - 2 + sub);  // D - (i + j)
        return;
    } else if constexpr (METRIC == LCSS) {  // x[i] = X[r + 2], y[j] = Y[W - r + 1]
        const int hit = (fabs(X[r + 2] - Y[W - r + 1]) <= arg) ? 1 : 0;
        B[r] = max(B[r] + hit, max(A[r], A[r + 1]));
        return;
    } else if constexpr (METRIC == SWALE) {  // x[i-1] = X[r + 1], y[j-1] = Y[W - r]
        const int mv = (fabs(X[r + 1] - Y[W - r]) <= arg) ? 1 : 5;
        B[r] = min(min(A[r], A[r + 1]) + 5, B[r] + mv);
        return;
    }
    if (METRIC == MSM) {
        B[r] = msm_cell_rb(A[r], A[r + 1], B[r], X[r + 2] - Y[W - r + 1], DL[r + 2], DH[r + 2],
                           Y[W - r] - Y[W - r + 1], arg);
        return;
    }
    B[r] = cell<METRIC, false>(0, 0, A[r], A[r + 1], B[r], X[r + 2], X[r + 1], Y[W - r + 1],
                               Y[W - r],
                               METRIC == ERP ? (X[r + 1] - arg) * (X[r + 1] - arg) : 0.0,
                               METRIC == ERP ? (Y[W - r] - arg) * (Y[W - r] - arg) : 0.0, arg);
}

// One iteration (two anti-diagonals).  B[r] needs the new A[r] and A[r + 1]
// and A[r + 1] needs the old B[r], so the halves interleave slot by slot,
// which keeps the values the two cells share (row / column differences)
// live only briefly.
template <int METRIC, int W>
__device__ __forceinline__ void ad_step(double arg, eld_t<METRIC>* A, eld_t<METRIC>* B,
                                        const double* X, const double* Y, const double* DL,
                                        const double* DH) {
    if constexpr (METRIC == MSM) {
        // The cell pair (step_a(r), step_b(r - 1)) reads query row r + 1
        // (X, DL, DH[r + 1]) from LDS; the row of the next pair is loaded
        // during this one.  An empty asm after each pair (its two values in
        // and out, memory clobbered) keeps the compiler from moving the
        // pairs' arithmetic or the row reads across pairs: unconstrained, it
        // hoists every row of the iteration into registers, which spill.
        double xr[W + 2], lr[W + 2], hr[W + 2];
        xr[1] = X[1];
        lr[1] = DL[1];
        hr[1] = DH[1];
#pragma unroll
        for (int r = 0; r <= W; r++) {
            if (r + 2 <= W + 1) {
                xr[r + 2] = X[r + 2];
                lr[r + 2] = DL[r + 2];
                hr[r + 2] = DH[r + 2];
            }
            {
                const eld_t<METRIC> S = band_sentinel<METRIC>();
                const double up = (r >= 1) ? B[r - 1] : S;
                const double left = (r <= W - 1) ? B[r] : S;
                A[r] = msm_cell_rb(up, left, A[r], xr[r + 1] - Y[W - r + 1], lr[r + 1], hr[r + 1],
                                   Y[W - r] - Y[W - r + 1], arg);
            }
            if (r >= 1) {
                const int q = r - 1;
                B[q] = msm_cell_rb(A[q], A[q + 1], B[q], xr[q + 2] - Y[W - q + 1], lr[q + 2],
                                   hr[q + 2], Y[W - q] - Y[W - q + 1], arg);
                asm volatile("" : "+v"(A[r]), "+v"(B[q])::"memory");
            } else {
                asm volatile("" : "+v"(A[r])::"memory");
            }
        }
        return;
    }
    step_a<METRIC, W>(0, arg, A, B, X, Y, DL, DH);
#pragma unroll
    for (int r = 1; r <= W; r++) {
        step_a<METRIC, W>(r, arg, A, B, X, Y, DL, DH);
        step_b<METRIC, W>(r - 1, arg, A, B, X, Y, DL, DH);
    }
}

template <int W>
__device__ __forceinline__ void ring_shift(double* Y, double yn) {
#pragma unroll
    for (int k = 0; k < W + 1; k++) Y[k] = Y[k + 1];
    Y[W + 1] = yn;
}

// Two iterations per pass of the register kernel's loop (the ring moves by
// two, the loop's control runs once) where the band update is short enough
// for the per-iteration overhead to matter.
template <int METRIC>
constexpr bool kEldPair = (METRIC == EDR || METRIC == LCSS || METRIC == SWALE);
// MSM runs kEldFuseK iterations fused (msm_fused): four up to radius 22,
// three beyond, where four would exceed 256 registers and spill (at W = 25,
// 157.6 ms for two, 152.7 for three, 151.1 with 24 registers spilled for four
// on the plain large workload).
template <int METRIC>
constexpr bool kEldFused = (METRIC == MSM);
template <int W>
constexpr int kEldFuseK = (W <= 22) ? 4 : 3;

// K MSM iterations (b .. b + K - 1) in one skewed sweep: step s runs, for
// k = 0 .. K - 1 in order, pair q = s - k (step_a(q), step_b(q - 1)) of
// iteration b + k, whose ring is Y + k and query window X + k.  Pair q of
// iteration b + k reads iteration b + k - 1's B[q - 1], B[q] (written by that
// iteration's pairs q and q + 1, steps s - 1 and s, before it) and A[q] (its
// diagonal), and overwrites A[q] and B[q - 1], which iteration b + k - 1 no
// longer reads: every cell gets exactly the operands of the K ad_step calls
// in sequence.  All pairs of step s take query row b + s (X[s + 1] and its
// clamp bounds DL / DH[s + 1]), and column b + W - r's bounds col(r) serve
// steps r .. r + 2K - 1, so each row is read and each column's bounds
// computed once per K iterations.  Y[W + 2 .. W + K] hold iterations b + 1 ..
// b + K - 1's new samples.  rk[k], k < K - 1: the end cell's slot after
// iteration b + k.  The empty asm statements (values in and out, memory
// clobbered) keep each pair's arithmetic and the next row's reads in place:
// left free, the compiler hoists every row of the sweep into registers, which
// spill.
template <int W, int K>
__device__ __forceinline__ void msm_fused(double arg, double* A, double* B, const double* X,
                                          const double* Y, const double* DL, const double* DH,
                                          double* rk) {
    const double S = kEldSentinel;
    double xr[W + K + 1], lr[W + K + 1], hr[W + K + 1];  // rows 1 .. W + K
    double cl[W + K + 1], ch[W + K + 1];  // col(r) at r + K: min(0, dyn), max(0, dyn)
    auto col = [&](int r) {  // dyn = y_{j-1} - y_j of column j = b + W - r
        const double d = Y[W - r] - Y[W - r + 1];
        cl[r + K] = fmin(d, 0.0);
        ch[r + K] = fmax(d, 0.0);
    };
    // msm_cell with the row's and the column's clamp bounds given
    auto cellf = [&](double up, double left, double diag, double e, int row, int c) {
        const double c1 = diag + fabs(e);
        const double s2 = fabs(fmax(fmin(e, hr[row]), lr[row]));
        const double s3 = fabs(fmax(fmin(e, ch[c + K]), cl[c + K]));
        const double c2 = up + (arg + s2);
        const double c3 = left + (arg + s3);
        return dmin3(c1, c2, c3);
    };
    xr[1] = X[1];
    lr[1] = DL[1];
    hr[1] = DH[1];
#pragma unroll
    for (int r = -(K - 1); r < 0; r++) col(r);
#pragma unroll
    for (int s = 0; s <= W + K - 1; s++) {
        if (s + 2 <= W + K) {
            xr[s + 2] = X[s + 2];
            lr[s + 2] = DL[s + 2];
            hr[s + 2] = DH[s + 2];
        }
        if (s <= W) col(s);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = s - k;
            if (q < 0 || q > W) continue;
            const double up = (q >= 1) ? B[q - 1] : S;
            const double left = (q <= W - 1) ? B[q] : S;
            A[q] = cellf(up, left, A[q], xr[s + 1] - Y[W - q + 1 + k], s + 1, q - k);
            if (q >= 1) {
                B[q - 1] = cellf(A[q - 1], A[q], B[q - 1], xr[s + 1] - Y[W - q + 2 + k], s + 1,
                                 q - 1 - k);
                asm volatile("" : "+v"(A[q]), "+v"(B[q - 1])::"memory");
            } else {
                asm volatile("" : "+v"(A[q])::"memory");
            }
            if (k < K - 1 && q == ((W & 1) ? (W + 1) / 2 : W / 2))
                rk[k] = (W & 1) ? B[(W - 1) / 2] : A[W / 2];
        }
    }
}

// Per-sample lower bound on the cost of consuming candidate sample y whose
// band sees the query envelope [lo, hi] (DTW, ERP, MSM); for EDR 1 if y is
// more than eps from the envelope, for LCSS 1 if it is within eps of it.
// The EDR / LCSS tests are tightened by a relative 1e-12 so that they select
// a superset (EDR) / subset (LCSS) of the positions the GLB pass counted.
template <int METRIC>
__device__ inline double consume_lb(double y, double lo, double hi, double arg) {
    const double e = fmax(fmax(lo - y, y - hi), 0.0);
    if (METRIC == DTW) return e * e;
    if (METRIC == ERP) return fmin(e * e, (y - arg) * (y - arg));
    if (METRIC == EDR) return (e > arg * (1.0 - 1e-12)) ? 1.0 : 0.0;
    if (METRIC == MSM) return fmin(e, arg);
    // SWALE: every sample is consumed by a step that costs at least r
    if (METRIC == SWALE) return kSwaleR;
    return (e <= arg * (1.0 - 1e-12)) ? 1.0 : 0.0;
}

// The bound total of the abandon tests below where no GLB pass handed one
// over: nothing for the distance sums, every sample for LCSS's matches, and
// m r for SWALE (each sample costs at least r, as consume_lb counts).
template <int METRIC>
__device__ inline double no_glb_total(int m) {
    if (METRIC == LCSS) return (double)m;
    if (METRIC == SWALE) return kSwaleR * (double)m;
    return 0.0;
}

// Abandon test after iteration b (the two anti-diagonals in A and B).  total
// is the GLB pass's certified bound over the candidate's samples: a lower
// bound on the sum of their consumption costs (DTW / ERP / MSM, EDR counts
// of certain mismatches) or an upper bound on the samples that can match at
// all (LCSS); p sums consume_lb over samples 0 .. b + KOFF, an upper bound on
// the same quantity over them (rounding: the factor (m + 8) 1.2e-16).  The
// remaining path still consumes every sample beyond the frontier's largest
// column (LCSS: matches beyond the smallest), and its rounded sum loses at
// most (2m + 24) 1.2e-16 relative.
template <int METRIC, int W>
__device__ __forceinline__ bool ad_alive(const eld_t<METRIC>* A, const eld_t<METRIC>* B, int b,
                                         int M, int m, double thr, double total, double p) {
    if constexpr (METRIC == LCSS) {
        int mxi = A[0];
#pragma unroll
        for (int k = 1; k <= W; k++) mxi = max(mxi, A[k]);
#pragma unroll
        for (int k = 0; k < W; k++) mxi = max(mxi, B[k]);
        const double mx = (double)mxi;
        // matches still possible after a cell of anti-diagonal >= 2b+W
        const double rem = fmin((double)((2 * M - (2 * b + W)) / 2 + 1), total - p);
        return !(1.0 - (mx + rem) / (double)m > thr);
    }
    eld_t<METRIC> mna = A[0], mnb = band_sentinel<METRIC>();
#pragma unroll
    for (int k = 1; k <= W; k++) mna = band_min(mna, A[k]);
#pragma unroll
    for (int k = 0; k < W; k++) mnb = band_min(mnb, B[k]);
    // A holds anti-diagonal 2b + W, B the next
    const double mn = fmin(band_value<METRIC>(mna, 2 * b + W),
                           band_value<METRIC>(mnb, 2 * b + W + 1));
    const double cb = fmax(total - p * (1.0 + (double)(m + 8) * 1.2e-16), 0.0);
    return !((mn + cb) * (1.0 - (double)(2 * m + 24) * 1.2e-16) > thr);
}

// src/ElasticMeasures.cpp: sqrt for DTW / ERP (1e18 when not finite), 1 - L/m
// for LCSS, the accumulated cost otherwise.
template <int METRIC>
__device__ inline double eld_result(double r, int m) {
    if (METRIC == DTW || METRIC == ERP) {
        if (!__builtin_isfinite(r)) return kEldSentinel;
        return sqrt(smax(0.0, r));
    }
    if (METRIC == LCSS) return 1 - r / (double)m;
    return r;
}

// Append (g, d) for the lanes with keep set: one atomic per wave.
__device__ inline void wave_append(bool keep, int64_t g, double d, int64_t* out_idx,
                                   double* out_dist, unsigned long long* out_count) {
    const uint64_t mask = __ballot(keep);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(mask);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(out_count, (unsigned long long)__builtin_popcountll(mask));
    base = __shfl(base, leader, 64);
    if (keep) {
        const uint64_t below = lane ? (mask & ((1ull << lane) - 1)) : 0ull;
        const unsigned long long pos = base + __builtin_popcountll(below);
        out_idx[pos] = g;
        out_dist[pos] = d;
    }
}

// Persistent register-band kernel.  Each lane runs one candidate at a time
// through iterations BS .. bend at its own pace; a lane whose candidate ends
// (finished or abandoned) idles until kRefillMin lanes of its wave are idle,
// and then they take the next candidates from the shared queue together.
// in[] lists window indices (target mode, in_total[] the GLB pass's bounds or
// null) or, with rows != null, rows 0 .. n_direct-1 of normalised samples.
// W is the query's radius w: every slot of both anti-diagonals is in the band.
// Two waves per SIMD (three for the integer EDR / LCSS state): the query
// side is read from LDS, so the state that stays in registers (two
// anti-diagonals and the candidate ring) fits half (a third) of the register
// file.
template <int METRIC, int W>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(EldState<METRIC>::waves)))
k_eld(VerifyArgs a, const double* __restrict__ LQ,
                                             const double* __restrict__ UQ,
                                             const int64_t* __restrict__ in,
                                             const double* __restrict__ in_total,
                                             const unsigned long long* __restrict__ in_count,
                                             int64_t n_direct, const double* __restrict__ rows,
                                             int keep_all, int64_t* __restrict__ out_idx,
                                             double* __restrict__ out_dist,
                                             unsigned long long* out_count,
                                             unsigned long long* queue,
                                             unsigned long long* work, int64_t n_min) {
    extern __shared__ double xlds[];
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR || METRIC == SWALE);
    // samples 0 .. b + KOFF lie at or before the frontier's largest column
    // (ERP / EDR cells consume y[j - 1]); LCSS: its smallest column
    constexpr int KOFF = (METRIC == LCSS) ? 0 : (GRID1 ? W - 1 : W);
    // first iteration: anti-diagonals 0 and 1 (W even) or -1 and 0 (W odd)
    constexpr int BS = -((W + 1) / 2);
    const int64_t n = in_count ? (int64_t)*in_count : n_direct;
    if (n < n_min) return;  // few candidates: the wave kernel takes them
    const int m = a.p.m;
    const int M = GRID1 ? m : m - 1;
    const int bend = (2 * M - W) >= 0 ? (2 * M - W) / 2 : -((W - 2 * M + 1) / 2);
    const double arg = a.p.arg, cut = a.p.cutoff;
    double thr = 1e308;
    // abandoning needs non-negative step costs (MSM: c >= 0)
    if (!keep_all && !(METRIC == MSM && arg < 0.0)) {
        if (METRIC == DTW || METRIC == ERP) thr = cut * cut * (1.0 + 1e-12) + 1e-300;
        else thr = cut * (1.0 + 1e-12) + 1e-300;
    }
    const bool check = thr < 1e300;
    const bool raw = rows == nullptr;
    using VT = eld_t<METRIC>;
    const VT S = band_sentinel<METRIC>();
    // the virtual cell (-1, -1): 0, or -1 for EDR with eps < 0; EDR stores
    // D - (i + j) = D + 2.  SWALE: -1, so that cell (0, 0), a match of pad
    // against pad at cost r, is 0
    const VT V = (METRIC == EDR) ? ((arg >= 0.0) ? 2 : 1) : ((METRIC == SWALE) ? -1 : 0);

    const int span = m + 2 * kXPad;
    double* xs = xlds + kXPad;
    double* lqs = xlds + span + kXPad;
    double* uqs = lqs + span;
    // MSM: the rows' clamp bounds min(0, x_i - x_{i-1}), max(0, x_i - x_{i-1}),
    // indexed like xlds (the first padded row: 0, a pad against a pad)
    double* dxl = xlds + 3 * span;
    double* dxh = dxl + span;
    for (int i = threadIdx.x; i < span; i += blockDim.x) {
        const int qi = i - kXPad;
        const bool in_q = qi >= 0 && qi < m;
        xlds[i] = in_q ? a.q.qn[qi] : kPad;
        if (check) {
            xlds[span + i] = in_q ? LQ[qi] : 0.0;
            xlds[2 * span + i] = in_q ? UQ[qi] : 0.0;
        }
        if constexpr (METRIC == MSM) {
            const bool prev_q = qi - 1 >= 0 && qi - 1 < m;
            const double dx = (i == 0) ? 0.0
                                       : (in_q ? a.q.qn[qi] : kPad) -
                                             (prev_q ? a.q.qn[qi - 1] : kPad);
            xlds[3 * span + i] = fmin(dx, 0.0);
            xlds[4 * span + i] = fmax(dx, 0.0);
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    // the query side X[k] = xs[b - 1 + k] is read from LDS (lanes run at
    // their own iteration b); the candidate ring Y stays in registers
    VT A[W + 1], B[W + 1];
    // kEldPair: Y[W + 2], Y[W + 3] take the two new samples; kEldFused: Y[W + 2 ..
    // W + 1 + kEldFuseK] the fused iterations' new samples
    double Y[W + 2 + (kEldFused<METRIC> ? kEldFuseK<W> : 2)];
#pragma unroll
    for (int k = 0; k <= W; k++) A[k] = B[k] = S;
#pragma unroll
    for (int k = 0; k < W + 2; k++) Y[k] = 0.0;
    const double* tp = raw ? a.T : rows;
    double mu = 0.0, sd = 1.0, total = 0.0, p = 0.0;
    double yn[kEldFuseK<W>];  // kEldFused: the next pass's new samples
#pragma unroll
    for (int k = 0; k < kEldFuseK<W>; k++) yn[k] = kPad;
    int64_t ci = 0, g = 0;
    int b = 0, it = 0;
    unsigned live_iters = 0;
    bool busy = false, drained = false;
    // the wave's pool of queue positions [pb, pe) (the same in every lane)
    unsigned long long pb = 0, pe = 0;
    const int refill = a.eld_refill > 0 ? a.eld_refill : kRefillMin<METRIC>;
    while (true) {
        const uint64_t idle = __ballot(!busy);
        const int nidle = __builtin_popcountll(idle);
        if ((pb < pe || !drained) && (nidle >= refill || nidle == 64)) {
            const unsigned long long left = pe - pb;
            const bool take = left < (unsigned long long)nidle && !drained;
            unsigned long long nb = 0;
            if (take) {
                const int leader = __builtin_ctzll(idle);
                if (lane == leader) nb = atomicAdd(queue, kEldChunk);
                nb = __shfl(nb, leader, 64);
                if (nb + kEldChunk >= (unsigned long long)n) drained = true;
            }
            if (!busy) {
                const unsigned long long rank =
                    (unsigned long long)__builtin_popcountll(idle & ((1ull << lane) - 1));
                // the pool's positions first, then the new chunk's
                const bool have = rank < left || take;
                const int64_t c = (int64_t)(rank < left ? pb + rank : nb + (rank - left));
                if (have && c < n) {
                    ci = c;
                    if (raw) {
                        g = in[c];
                        const int64_t si = stat_index(a.G, g);
                        tp = a.T + g;
                        mu = a.MU[si];
                        sd = a.SG[si];
                    } else {
                        g = c;
                        tp = rows + c * (int64_t)m;
                    }
                    total = in_total ? in_total[c] : no_glb_total<METRIC>(m);
                    b = BS;
#pragma unroll
                    for (int k = 0; k <= W; k++) A[k] = B[k] = S;
                    if (W & 1) B[(W - 1) / 2] = V;
                    else A[W / 2] = V;
#pragma unroll
                    for (int k = 0; k < W + 2; k++) {
                        const int j = BS - 1 + k;
                        double v = kPad;
                        if (j >= 0 && j < m) v = raw ? (tp[j] - mu) / sd : tp[j];
                        Y[k] = v;
                    }
                    if constexpr (kEldFused<METRIC>) {  // samples BS + W + 1 + k
#pragma unroll
                        for (int k = 0; k < kEldFuseK<W>; k++) {
                            const int j = BS + W + 1 + k;
                            yn[k] = (j < m) ? (raw ? (tp[j] - mu) / sd : tp[j]) : kPad;
                        }
                    }
                    p = 0.0;
                    if (check) {
#pragma unroll
                        for (int s = 0; s <= BS - 1 + KOFF; s++)
                            if (s < m) p += consume_lb<METRIC>(Y[s - BS + 1], lqs[s], uqs[s], arg);
                    }
                    busy = true;
                }
            }
            if (left >= (unsigned long long)nidle) {
                pb += (unsigned long long)nidle;
            } else if (take) {
                pb = nb + ((unsigned long long)nidle - left);
                pe = nb + kEldChunk;
            } else {
                pb = pe;
            }
        }
        if (!__any(busy)) break;

        // The slots are materialised after each band update: left to itself
        // the compiler sinks the integer updates (EDR / LCSS) past the rest
        // of the iteration and keeps all 2W + 1 comparison masks live in
        // scalar registers, which then spill to vector lanes.
        auto pin = [&]() {
            if constexpr (METRIC == EDR || METRIC == LCSS || METRIC == SWALE) {
#pragma unroll
                for (int k = 0; k <= W; k++) {
                    asm volatile("" : "+v"(A[k]));
                    asm volatile("" : "+v"(B[k]));
                }
            }
        };
        // next candidate sample of iteration bi (the ring's new entry)
        auto sample = [&](int bi) {
            const int jn = bi + W + 1;
            const int jj = min(max(jn, 0), m - 1);
            double yv = tp[jj];
            if (raw) yv = (yv - mu) / sd;
            return (jn == jj) ? yv : kPad;
        };
        if constexpr (kEldFused<METRIC>) {
            // iterations b .. b + K - 1 (msm_fused); a lane whose candidate
            // ends at b + k < b + K - 1 keeps that iteration's value (rk[k]),
            // as in the pair path below.  The next pass's samples are loaded
            // here and normalised after the sweep.
            constexpr int K = kEldFuseK<W>;
#pragma unroll
            for (int k = 0; k + 1 < K; k++) Y[W + 2 + k] = yn[k];
            double tn[K];
#pragma unroll
            for (int k = 0; k < K; k++) tn[k] = tp[min(max(b + K + k + W + 1, 0), m - 1)];
            double rk[K > 1 ? K - 1 : 1];
            msm_fused<W, K>(arg, A, B, xs + (b - 1), Y, dxl + kXPad + (b - 1),
                            dxh + kXPad + (b - 1), rk);
            if (check) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int Ks = b + k + KOFF;
                    if (Ks >= 0 && Ks < m)
                        p += consume_lb<METRIC>(Y[KOFF + 1 + k], lqs[Ks], uqs[Ks], arg);
                }
            }
            Y[W + 1 + K] = yn[K - 1];
#pragma unroll
            for (int k = 0; k < W + 2; k++) Y[k] = Y[k + K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = b + K + k + W + 1;
                yn[k] = (j >= 0 && j < m) ? (raw ? (tn[k] - mu) / sd : tn[k]) : kPad;
            }
            const int kf = bend - b;  // the iteration, if any, this pass ends the candidate in
            const bool fin = busy && kf >= 0 && kf < K;
            live_iters += busy ? (unsigned)(fin ? kf + 1 : K) : 0u;
            if (busy) b += K;
            if (__any(fin)) {
                double rv = (W & 1) ? B[(W - 1) / 2] : A[W / 2];
#pragma unroll
                for (int k = 0; k + 1 < K; k++)
                    if (kf == k) rv = rk[k];
                const double d =
                    eld_result<METRIC>(band_value<METRIC>(rv, 2 * bend + W + (W & 1)), m);
                if (keep_all) {
                    if (fin) out_dist[ci] = d;
                } else {
                    // src/fsmeld.cpp:846-847
                    wave_append(fin && __builtin_isfinite(d) && d <= cut, g, d, out_idx,
                                out_dist, out_count);
                }
                busy = busy && !fin;
            }
            // every second pass (six or eight iterations): the test's 2W + 1
            // minima cost more than the cells a later abandon adds (checking
            // every pass of three: k_eld<MSM,25> 153.7 against 151.7 ms, the
            // evaluated cells 33.6 against 33.9%)
            if (check && ((++it) & 1) == 0)
                busy = busy && ad_alive<METRIC, W>(A, B, b - 1, M, m, thr, total, p);
            continue;
        }
        if constexpr (kEldPair<METRIC>) {
            // iterations b and b + 1; a lane whose candidate ends at b keeps
            // its value (r1) and runs the second half on its stale state
            // (every index stays inside the padded query, the sample index is
            // clamped), then idles like any finished lane
            ad_step<METRIC, W>(arg, A, B, xs + (b - 1), Y, dxl + kXPad + (b - 1),
                               dxh + kXPad + (b - 1));
            pin();
            const double y1 = sample(b);
            if (check) {
                const int K = b + KOFF;
                if (K >= 0 && K < m) p += consume_lb<METRIC>(Y[KOFF + 1], lqs[K], uqs[K], arg);
            }
            Y[W + 2] = y1;
            const bool fin1 = busy && b == bend;
            const VT r1 = (W & 1) ? B[(W - 1) / 2] : A[W / 2];
            ad_step<METRIC, W>(arg, A, B, xs + b, Y + 1, dxl + kXPad + b, dxh + kXPad + b);
            pin();
            const double y2 = sample(b + 1);
            if (check) {
                const int K = b + 1 + KOFF;
                if (K >= 0 && K < m) p += consume_lb<METRIC>(Y[KOFF + 2], lqs[K], uqs[K], arg);
            }
            Y[W + 3] = y2;
#pragma unroll
            for (int k = 0; k < W + 2; k++) Y[k] = Y[k + 2];
            const bool fin = busy && (fin1 || b + 1 == bend);
            live_iters += busy ? (fin1 ? 1u : 2u) : 0u;
            if (busy) b += 2;
            if (__any(fin)) {
                const VT rv = fin1 ? r1 : ((W & 1) ? B[(W - 1) / 2] : A[W / 2]);
                const double d =
                    eld_result<METRIC>(band_value<METRIC>(rv, 2 * bend + W + (W & 1)), m);
                if (keep_all) {
                    if (fin) out_dist[ci] = d;
                } else {
                    // src/fsmeld.cpp:846-847
                    wave_append(fin && __builtin_isfinite(d) && d <= cut, g, d, out_idx,
                                out_dist, out_count);
                }
                busy = busy && !fin;
            }
            // every fourth iteration, as below
            if (check && ((++it) & 1) == 0)
                busy = busy && ad_alive<METRIC, W>(A, B, b - 1, M, m, thr, total, p);
            continue;
        }
        ad_step<METRIC, W>(arg, A, B, xs + (b - 1), Y, dxl + kXPad + (b - 1),
                           dxh + kXPad + (b - 1));
        pin();
        const int jn = b + W + 1;
        const int jj = min(max(jn, 0), m - 1);
        double yv = tp[jj];
        if (raw) yv = (yv - mu) / sd;
        if (check) {  // sample b + KOFF, before the shift at Y[KOFF + 1] (KOFF >= -1)
            const int K = b + KOFF;
            if (K >= 0 && K < m) p += consume_lb<METRIC>(Y[KOFF + 1], lqs[K], uqs[K], arg);
        }
        ring_shift<W>(Y, (jn == jj) ? yv : kPad);
        const bool fin = busy && b == bend;
        live_iters += busy ? 1u : 0u;
        if (busy) b++;
        if (__any(fin)) {
            // cell (M, M) sits on anti-diagonal 2 bend + W (A) or + 1 (B)
            const double d = eld_result<METRIC>(
                band_value<METRIC>((W & 1) ? B[(W - 1) / 2] : A[W / 2], 2 * bend + W + (W & 1)),
                m);
            if (keep_all) {
                if (fin) out_dist[ci] = d;
            } else {
                // src/fsmeld.cpp:846-847
                wave_append(fin && __builtin_isfinite(d) && d <= cut, g, d, out_idx, out_dist,
                            out_count);
            }
            busy = busy && !fin;
        }
        if (check && ((++it) & 3) == 0)
            busy = busy && ad_alive<METRIC, W>(A, B, b - 1, M, m, thr, total, p);
    }
    if (work) {
        unsigned long long v = live_iters;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) atomicAdd(work, v);
    }
}

// ---------------------------------------------------------------- generic ---
// Radii above the register variants: the same per-cell expressions evaluated
// row by row with the band in lane-interleaved global scratch.
struct MemArr {
    double* base;
    int64_t stride;
    __device__ double& operator[](int k) { return base[(int64_t)k * stride]; }
};

// Abandon test every 8 rows when thr < 1e300: every path leaves row i at a
// column j <= i + w (DTW, ERP, EDR, MSM) and still consumes every later
// sample, or enters it at a column >= i - w (LCSS) and has at most one match
// per later matchable sample and per later row (see ad_alive); *dead is set
// and the sentinel returned when the window cannot match.
template <int METRIC>
__device__ double eld_rows(const double* __restrict__ q, const double* __restrict__ qg,
                           const CandSrc& c, int m, int w, double arg, MemArr b, MemArr y,
                           double thr, double total, const double* __restrict__ LQ,
                           const double* __restrict__ UQ, bool* dead) {
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR || METRIC == SWALE);
    const int M = GRID1 ? m : m - 1;
    const int xoff = GRID1 ? 1 : 0;
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
    const int w2 = 2 * w;
    // band slot k = j - i + w; y[k] holds sample j - xoff of the current row
    for (int k = 0; k <= w2 + 1; k++) {
        b[k] = S;
        y[k] = c.at(k - w - xoff);
    }
    const bool check = thr < 1e300;
    double p = 0.0;
    int kp = 0;  // samples 0 .. kp-1 are in p
    *dead = false;
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
        if (check && (i & 7) == 7 && i < M) {
            const int K = (METRIC == LCSS) ? i - w : (GRID1 ? i + w - 1 : i + w);
            for (; kp <= K && kp < m; kp++) p += consume_lb<METRIC>(c.at(kp), LQ[kp], UQ[kp], arg);
            bool alive;
            if (METRIC == LCSS) {
                double mx = b[0];
                for (int k = 1; k <= w2; k++) mx = fmax(mx, b[k]);
                const double rem = fmin((double)(M - i), total - p);
                alive = !(1.0 - (mx + rem) / (double)m > thr);
            } else {
                double mn = b[0];
                for (int k = 1; k <= w2; k++) mn = fmin(mn, b[k]);
                const double cb = fmax(total - p * (1.0 + (double)(m + 8) * 1.2e-16), 0.0);
                alive = !((mn + cb) * (1.0 - (double)(2 * m + 24) * 1.2e-16) > thr);
            }
            if (!alive) {
                *dead = true;
                return S;
            }
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

// Grid-stride kernel of the row-band path (radii above the register
// variants, or queries too long for their LDS staging).
template <int METRIC>
__global__ void __launch_bounds__(256) k_eld_rows(VerifyArgs a, const double* __restrict__ LQ,
                                                  const double* __restrict__ UQ,
                                                  const int64_t* __restrict__ in,
                                                  const double* __restrict__ in_total,
                                                  const unsigned long long* __restrict__ in_count,
                                                  int64_t n_direct,
                                                  const double* __restrict__ rows, int keep_all,
                                                  int64_t* __restrict__ out_idx,
                                                  double* __restrict__ out_dist,
                                                  unsigned long long* out_count,
                                                  double* __restrict__ scratch) {
    const int64_t n = in_count ? (int64_t)*in_count : n_direct;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t iters = (n + stride - 1) / stride;
    const int m = a.p.m, w = a.p.w;
    const double cut = a.p.cutoff;
    double thr = 1e308;
    if (!keep_all && !(METRIC == MSM && a.p.arg < 0.0)) {
        if (METRIC == DTW || METRIC == ERP) thr = cut * cut * (1.0 + 1e-12) + 1e-300;
        else thr = cut * (1.0 + 1e-12) + 1e-300;
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
        MemArr bb{scratch + tid, stride * 2};
        MemArr yy{scratch + tid + stride, stride * 2};
        const double total =
            (in_total && have) ? in_total[ci] : no_glb_total<METRIC>(m);
        bool dead = false;
        const double d = eld_rows<METRIC>(a.q.qn, a.q.XG, c, m, w, a.p.arg, bb, yy, thr, total,
                                          LQ, UQ, &dead);
        if (keep_all) {
            if (have) out_dist[ci] = d;
            continue;
        }
        wave_append(have && !dead && __builtin_isfinite(d) && d <= cut, g, d, out_idx, out_dist,
                    out_count);
    }
}

// ------------------------------------------------------------ wide bands ---
// Data-parallel-primitive moves of a double (two 32-bit DPP movs): quad
// swaps xor 1 / xor 2, row rotations (row_ror:N: lane l reads lane
// (l - N) mod 16 of its row, so row_ror:12 reads l + 4), and the whole-wave
// shifts by one lane of the gfx9 family (wave_shl:1 reads lane l + 1,
// wave_shr:1 lane l - 1; the lane without a source reads 0).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRorNext4 = 0x12C, kDppRor8 = 0x128;
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}

// cell<METRIC, true> with every form evaluated and the boundary one selected:
// the same expressions on the same operands, so the same bits, and no branch.
template <int METRIC>
__device__ __forceinline__ double cell_sel(int i, int j, double up, double left, double diag,
                                           double xi, double xp, double yj, double yp,
                                           double xg, double yg, double arg) {
    const bool i0 = i == 0, j0 = j == 0;
    if (METRIC == DTW) {
        const double S = kEldSentinel;
        const double d = (xi - yj) * (xi - yj);
        const double r = dmin3(i0 ? S : up, j0 ? S : left, (i0 || j0) ? S : diag) + d;
        return (i0 && j0) ? d : r;
    } else if (METRIC == ERP) {
        const double e1 = left + yg, e2 = up + xg;
        const double r = dmin3(diag + (xp - yp) * (xp - yp), e1, e2);
        return i0 ? (j0 ? 0.0 : e1) : (j0 ? e2 : r);
    } else if (METRIC == EDR) {
        const double sub = (fabs(xp - yp) <= arg) ? 0.0 : 1.0;
        const double r = dmin3(up + 1.0, left + 1.0, diag + sub);
        return i0 ? (double)j : (j0 ? (double)i : r);
    } else if (METRIC == SWALE) {
        const double mv = (fabs(xp - yp) <= arg) ? kSwaleR : kSwaleP;
        const double e1 = left + kSwaleP, e2 = up + kSwaleP;
        const double r = dmin3(diag + mv, e2, e1);
        return i0 ? (j0 ? 0.0 : e1) : (j0 ? e2 : r);
    } else if (METRIC == LCSS) {
        return cell<LCSS, true>(i, j, up, left, diag, xi, xp, yj, yp, xg, yg, arg);
    } else {  // MSM: msm_cost in the form of msm_cell (the same bits)
        const double e = xi - yj;
        const double t = fabs(e);
        const double c2 = up + (arg + msm_split(xi - xp, e));
        const double c3 = left + (arg + msm_split(yp - yj, e));
        const double r = dmin3(diag + t, c2, c3);
        return i0 ? (j0 ? t : c3) : (j0 ? c2 : r);
    }
}

// One wavefront per candidate (few candidates, or radii above the register
// variants).  The band slots k = i - j + w in [0, 2w] are spread over the
// lanes, 2 CPL slots per lane (lane l holds k = 2 CPL l + q, q in [0, 2 CPL));
// anti-diagonal d updates the slots of parity (d + w) & 1 from the other
// parity (up k-1, left k+1) and its own old value (diag), so one step needs
// one neighbour value from the adjacent lane (a DPP whole-wave shift).  Cells
// use the reference's boundary forms (cell_sel).  Query and candidate samples
// sit in LDS; a wave takes its next candidate from the work queue when the
// last one ends.
template <int METRIC, int CPL>
__device__ __forceinline__ void wave_step(int par, int d, int w, int M, int m, double arg,
                                          const double* xs, const double* ys, double* V,
                                          int& live_cells) {
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR || METRIC == SWALE);
    constexpr int Q = 2 * CPL;
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
    const int lane = threadIdx.x & 63;
    // neighbour slot from the adjacent lane: the previous lane's last slot
    // (parity 0 updates need it as "up" of q = 0), or the next lane's first
    // (parity 1 updates need it as "left" of q = Q - 1)
    // (the lanes without a source, 0 resp. 63, hold k = 0 resp. k > 2w and
    // never use nb)
    const double nb = (par == 0) ? dpp_f64<kDppWaveShr1>(V[Q - 1]) : dpp_f64<kDppWaveShl1>(V[0]);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        if ((q & 1) != par) continue;
        const int k = Q * lane + q;
        const int sft = k - w;
        const int i = (d + sft) >> 1, j = (d - sft) >> 1;
        const bool inr = k <= 2 * w && i >= 0 && j >= 0 && i <= M && j <= M;
        // every slot reads samples: indices clamped into the staged range
        // [-1, m] (cells outside the band or the matrix are replaced by S)
        const int ic = min(max(i, 0), m), jc = min(max(j, 0), m);
        const double up = (k == 0) ? S : (q >= 1 ? V[q - 1] : nb);
        const double left = (k == 2 * w) ? S : (q + 1 < Q ? V[q + 1] : nb);
        double v;
        if (GRID1) {
            const double xp = xs[ic - 1], yp = ys[jc - 1];
            v = cell_sel<METRIC>(i, j, up, left, V[q], 0.0, xp, 0.0, yp,
                                 (xp - arg) * (xp - arg), (yp - arg) * (yp - arg), arg);
        } else {
            v = cell_sel<METRIC>(i, j, up, left, V[q], xs[ic], xs[ic - 1], ys[jc], ys[jc - 1],
                                 0.0, 0.0, arg);
        }
        live_cells += inr ? 1 : 0;
        V[q] = inr ? v : S;
    }
}

// Anti-diagonals 0 .. 2M of one candidate, in blocks of 8 whose first
// anti-diagonal has parity P0 = w & 1 (slot parity of d0 + w, d0 even).  With
// abandoning, after anti-diagonal d = 8 t + 7 < 2M: the frontier (anti-
// diagonals d - 1 and d) reduced over the nl lanes that hold band slots,
// plus the bound over the samples beyond it, against thr.  Every lane adds
// the new samples' consumption bounds itself (uniform reads), so p is the
// same in all lanes.  Returns false when the candidate is abandoned.
template <int METRIC, int CPL, int P0>
__device__ __forceinline__ bool wave_band(int w, int M, int m, double arg, const double* xs,
                                          const double* ys, double* V, int& live_cells,
                                          bool check, double thr, double total,
                                          const double* __restrict__ LQ,
                                          const double* __restrict__ UQ) {
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR || METRIC == SWALE);
    constexpr int Q = 2 * CPL;
    const int D = 2 * M;
    const int nl = (2 * w + Q) / Q;  // lanes holding slots 0 .. 2w
    double p = 0.0;
    int kp = 0;  // samples 0 .. kp-1 are in p
    int d0 = 0;
    for (; d0 + 8 <= D; d0 += 8) {
#pragma unroll
        for (int s = 0; s < 8; s++)
            wave_step<METRIC, CPL>((P0 + s) & 1, d0 + s, w, M, m, arg, xs, ys, V, live_cells);
        if (!check) continue;
        const int d = d0 + 7;
        // samples at or before the frontier's largest column (LCSS: before
        // its smallest), floor((d + w) / 2) resp. floor((d - 1 - w) / 2)
        const int K = (METRIC == LCSS) ? ((d - 1 - w) >= 0 ? (d - 1 - w) / 2 : -1)
                                       : ((d + w) >> 1) - (GRID1 ? 1 : 0);
        const int kn = min(K, m - 1) + 1;
        for (int s = kp; s < kn; s++) p += consume_lb<METRIC>(ys[s], LQ[s], UQ[s], arg);
        if (kn > kp) kp = kn;
        double e = V[0];
#pragma unroll
        for (int q = 1; q < Q; q++) e = (METRIC == LCSS) ? fmax(e, V[q]) : fmin(e, V[q]);
        // lane 0 ends with the extremum over lanes 0 .. 2^ceil(log2 nl) - 1
        // (lanes past the band hold S): quads, then lane 0 takes quad 1
        // (lane 4), then lanes 8 .. 15 (lane 8)
        auto red = [](double x, double y) { return (METRIC == LCSS) ? fmax(x, y) : fmin(x, y); };
        e = red(e, dpp_f64<kDppXor1>(e));
        if (nl > 2) e = red(e, dpp_f64<kDppXor2>(e));
        if (nl > 4) e = red(e, dpp_f64<kDppRorNext4>(e));
        if (nl > 8) e = red(e, dpp_f64<kDppRor8>(e));
        if (nl > 16) e = red(e, __shfl_xor(e, 16, 64));
        if (nl > 32) e = red(e, __shfl_xor(e, 32, 64));
        bool alive;
        if (METRIC == LCSS) {
            const double rem = fmin((double)((D - (d - 1)) / 2 + 1), total - p);
            alive = !(1.0 - (e + rem) / (double)m > thr);
        } else {
            const double cb = fmax(total - p * (1.0 + (double)(m + 8) * 1.2e-16), 0.0);
            alive = !((e + cb) * (1.0 - (double)(2 * m + 24) * 1.2e-16) > thr);
        }
        if (!__builtin_amdgcn_readfirstlane(alive ? 1 : 0)) return false;
    }
    for (int d = d0; d <= D; d++) {
        if (((d + w) & 1) == 0) wave_step<METRIC, CPL>(0, d, w, M, m, arg, xs, ys, V, live_cells);
        else wave_step<METRIC, CPL>(1, d, w, M, m, arg, xs, ys, V, live_cells);
    }
    return true;
}

template <int METRIC, int CPL>
__global__ void __launch_bounds__(64) k_eld_wave(VerifyArgs a, const double* __restrict__ LQ,
                                                 const double* __restrict__ UQ,
                                                 const int64_t* __restrict__ in,
                                                 const double* __restrict__ in_total,
                                                 const unsigned long long* __restrict__ in_count,
                                                 int64_t n_direct, const double* __restrict__ rows,
                                                 int keep_all, int64_t* __restrict__ out_idx,
                                                 double* __restrict__ out_dist,
                                                 unsigned long long* out_count,
                                                 unsigned long long* queue,
                                                 unsigned long long* work, int64_t n_max) {
    extern __shared__ double wlds[];
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR || METRIC == SWALE);
    constexpr int Q = 2 * CPL;
    const int64_t n = in_count ? (int64_t)*in_count : n_direct;
    if (n >= n_max) return;  // many candidates: the register kernel takes them
    const int m = a.p.m, w = a.p.w;
    const int M = GRID1 ? m : m - 1;
    const double arg = a.p.arg, cut = a.p.cutoff;
    double thr = 1e308;
    if (!keep_all && !(METRIC == MSM && arg < 0.0)) {
        if (METRIC == DTW || METRIC == ERP) thr = cut * cut * (1.0 + 1e-12) + 1e-300;
        else thr = cut * (1.0 + 1e-12) + 1e-300;
    }
    const bool check = thr < 1e300;
    const bool raw = rows == nullptr;
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
    const int lane = threadIdx.x & 63;
    // xs[i], ys[j] for i, j in [-1, m]; out-of-range samples read as 0 and
    // are only ever used by the boundary forms that ignore them
    double* xs = wlds + 1;
    double* ys = wlds + (m + 2) + 1;
    for (int i = lane - 1; i <= m; i += 64) xs[i] = (i >= 0 && i < m) ? a.q.qn[i] : 0.0;
    unsigned long long live_total = 0;
    while (true) {
        unsigned long long c = 0;
        if (lane == 0) c = atomicAdd(queue, 1ull);
        c = __shfl(c, 0, 64);
        if (c >= (unsigned long long)n) break;
        int64_t g;
        const double* tp;
        double mu = 0.0, sd = 1.0;
        if (raw) {
            g = in[c];
            const int64_t si = stat_index(a.G, g);
            tp = a.T + g;
            mu = a.MU[si];
            sd = a.SG[si];
        } else {
            g = (int64_t)c;
            tp = rows + (int64_t)c * m;
        }
        __syncthreads();  // the previous candidate's reads of ys are done
        for (int j = lane - 1; j <= m; j += 64)
            ys[j] = (j >= 0 && j < m) ? (raw ? (tp[j] - mu) / sd : tp[j]) : 0.0;
        __syncthreads();
        const double total = in_total ? in_total[c] : no_glb_total<METRIC>(m);
        double V[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) V[q] = S;
        int live_cells = 0;
        const bool alive =
            (w & 1) ? wave_band<METRIC, CPL, 1>(w, M, m, arg, xs, ys, V, live_cells, check, thr,
                                                total, LQ, UQ)
                    : wave_band<METRIC, CPL, 0>(w, M, m, arg, xs, ys, V, live_cells, check, thr,
                                                total, LQ, UQ);
        live_total += (unsigned long long)live_cells;
        if (!alive) continue;
        // cell (M, M): slot k = w
        const double r = __shfl(V[w % Q], w / Q, 64);
        const double dist = eld_result<METRIC>(r, m);
        if (lane == 0) {
            if (keep_all) {
                out_dist[c] = dist;
            } else if (__builtin_isfinite(dist) && dist <= cut) {  // src/fsmeld.cpp:846-847
                const unsigned long long pos = atomicAdd(out_count, 1ull);
                out_idx[pos] = g;
                out_dist[pos] = dist;
            }
        }
    }
    if (work) {
        for (int o = 32; o > 0; o >>= 1) live_total += __shfl_xor(live_total, o, 64);
        // cells, reported as lane-iterations x (2w + 1) by the caller
        if (lane == 0) atomicAdd(work, live_total / (unsigned long long)(2 * w + 1));
    }
}

// ------------------------------------------------------------ launching ---
// (a template: one definition per translation unit is allowed)
template <int = 0>
__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// query, envelope (abandoning) and, for MSM, the rows' clamp bounds
inline size_t reg_lds_bytes(int metric, int m) {
    return sizeof(double) * (size_t)(m + 2 * kXPad) * (metric == MSM ? 5 : 3);
}

inline size_t wave_lds_bytes(int m) { return sizeof(double) * 2 * (size_t)(m + 2); }

// register band: w <= 32; one wave per candidate: w <= 127; the row-band path
// beyond, or where the LDS staging does not fit
constexpr int kRegMaxW = 25;  // the launcher instantiates k_eld for W = 0 .. kRegMaxW
constexpr int64_t kEldFewCandidates = 1024;
inline bool uses_register_band(int metric, int m, int w) {
    return reg_lds_bytes(metric, m) <= kEldLdsCap && w <= kRegMaxW;
}
inline bool uses_wave_band(int metric, int m, int w) {
    return !uses_register_band(metric, m, w) && w <= 127 && wave_lds_bytes(m) <= kEldLdsCap;
}

template <int METRIC>
hipError_t launch_metric(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                         const double* in_total, const unsigned long long* in_count,
                         int64_t n_direct, const double* rows, bool keep_all, int64_t* out_idx,
                         double* out_dist, unsigned long long* out_count,
                         unsigned long long* queue, unsigned long long* work, double* scratch,
                         int64_t scratch_threads) {
    const int w = a.p.w;
    const dim3 blk(256);
    const int ka = keep_all ? 1 : 0;
    if (!uses_register_band(METRIC, a.p.m, w) && !uses_wave_band(METRIC, a.p.m, w)) {
        if (!scratch) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_eld_rows<METRIC>), dim3((unsigned)(scratch_threads / 256)), blk, 0, st,
                           a, a.q.LQ, a.q.UQ, in, in_total, in_count, n_direct, rows, ka, out_idx,
                           out_dist, out_count, scratch);
        return hipGetLastError();
    }
    if (!queue) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u64<>, dim3(1), dim3(1), 0, st, queue);
    const size_t wl = wave_lds_bytes(a.p.m);
    auto launch_wave = [&](int64_t n_max) {
        // persistent waves; no more than the candidates there can be
        const int64_t bound = in_count ? n_max : n_direct;
        const dim3 wgrid((unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, bound)));
        if (w <= 63)
            hipLaunchKernelGGL((k_eld_wave<METRIC, 1>), wgrid, dim3(64), wl, st, a, a.q.LQ,
                               a.q.UQ, in, in_total, in_count, n_direct, rows, ka, out_idx,
                               out_dist, out_count, queue, work, n_max);
        else
            hipLaunchKernelGGL((k_eld_wave<METRIC, 2>), wgrid, dim3(64), wl, st, a, a.q.LQ,
                               a.q.UQ, in, in_total, in_count, n_direct, rows, ka, out_idx,
                               out_dist, out_count, queue, work, n_max);
    };
    if (uses_wave_band(METRIC, a.p.m, w)) {
        launch_wave(INT64_MAX);
        return hipGetLastError();
    }
    // Register band.  Below kEldFewCandidates candidates a lane per
    // candidate leaves the GPU idle while each lane walks its whole band, so
    // the wave kernel (one wave per candidate) takes them; a count on the
    // device gates both launches, each kernel exiting at once outside its range.
    int64_t n_min = 0;
    if (w <= 127 && wl <= kEldLdsCap) {
        if (in_count) {
            launch_wave(kEldFewCandidates);
            n_min = kEldFewCandidates;
        } else if (n_direct < kEldFewCandidates) {
            launch_wave(INT64_MAX);
            return hipGetLastError();
        }
    }
    // persistent: enough waves to fill every SIMD; later blocks find the
    // queue drained and exit
    const dim3 grid(1024);
    const size_t lds = reg_lds_bytes(METRIC, a.p.m);
    switch (w) {
#define FSM_ELD(WW)                                                                            \
    case WW:                                                                                   \
        hipLaunchKernelGGL((k_eld<METRIC, WW>), grid, blk, lds, st, a, a.q.LQ, a.q.UQ, in,     \
                           in_total, in_count, n_direct, rows, ka, out_idx, out_dist,          \
                           out_count, queue, work, n_min);                                     \
        break;
        FSM_ELD(0) FSM_ELD(1) FSM_ELD(2) FSM_ELD(3) FSM_ELD(4) FSM_ELD(5) FSM_ELD(6) FSM_ELD(7)
        FSM_ELD(8) FSM_ELD(9) FSM_ELD(10) FSM_ELD(11) FSM_ELD(12) FSM_ELD(13) FSM_ELD(14)
        FSM_ELD(15) FSM_ELD(16) FSM_ELD(17) FSM_ELD(18) FSM_ELD(19) FSM_ELD(20) FSM_ELD(21)
        FSM_ELD(22) FSM_ELD(23) FSM_ELD(24) FSM_ELD(25)
#undef FSM_ELD
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// one explicit instantiation per metric, each in its own translation unit
#define FSM_ELD_LAUNCHER(MM)                                                                   \
    template hipError_t launch_metric<MM>(hipStream_t, const VerifyArgs&, const int64_t*,      \
                                          const double*, const unsigned long long*, int64_t,   \
                                          const double*, bool, int64_t*, double*,              \
                                          unsigned long long*, unsigned long long*,            \
                                          unsigned long long*, double*, int64_t);

}  // namespace fsm
