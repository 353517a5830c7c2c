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
// Early abandon (search / exact / cutoff collection only): every optimal path
// crosses one of any two consecutive anti-diagonals and its prefix cost is at
// most the final cost (all step costs are >= 0), so once every band cell of
// two consecutive anti-diagonals exceeds the cutoff (squared for DTW/ERP, with
// a relative margin of 1e-12) the window cannot match and is dropped.  The
// distances of surviving windows are unchanged.
#include "engine_internal.h"

namespace fsm {

__device__ inline double dmin3(double a, double b, double c) {  // std::min({a,b,c})
    double r = a;
    if (b < r) r = b;
    if (c < r) r = c;
    return r;
}
__device__ inline double dmax3(double a, double b, double c) {  // std::max({a,b,c})
    double r = a;
    if (r < b) r = b;
    if (r < c) r = c;
    return r;
}
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

// src/ElasticMeasures.cpp:173-176 -- c if np lies between x and y (inclusive),
// else c + the distance to the nearer one.  Written as c + max(np - hi,
// lo - np, 0): the two differences are the same rounded values the reference
// takes the min of, and c + 0 = c.
__device__ inline double msm_cost(double np, double x, double y, double c) {
    const double hi = x > y ? x : y;
    const double lo = x > y ? y : x;
    double e = np - hi;
    const double f = lo - np;
    if (f > e) e = f;
    if (e < 0.0) e = 0.0;
    return c + e;
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
// xv / yv, with xg = (x[i-1]-g)^2, yg = (y[j-1]-g)^2).
template <int METRIC>
__device__ inline double cell(int i, int j, double up, double left, double diag, double xi,
                              double xp, double yj, double yp, double xg, double yg,
                              double arg) {
    const double S = kEldSentinel;
    if (METRIC == DTW) {
        const double d = (xi - yj) * (xi - yj);
        if (i == 0 && j == 0) return d;
        const double u = (i > 0) ? up : S;
        const double l = (j > 0) ? left : S;
        const double g = (i > 0 && j > 0) ? diag : S;
        return dmin3(u, l, g) + d;
    } else if (METRIC == ERP) {
        if (i == 0 && j == 0) return 0.0;
        if (i == 0) return left + yg;
        if (j == 0) return up + xg;
        const double match = diag + (xp - yp) * (xp - yp);
        return dmin3(match, left + yg, up + xg);
    } else if (METRIC == EDR) {
        if (i == 0 && j == 0) return 0.0;
        if (i == 0) return (double)j;
        if (j == 0) return (double)i;
        const double sub = (fabs(xp - yp) <= arg) ? 0.0 : 1.0;
        return dmin3(up + 1.0, left + 1.0, diag + sub);
    } else if (METRIC == LCSS) {
        const double d1 = (i > 0 && j > 0) ? diag : 0.0;
        const double u = (i > 0) ? up : 0.0;
        const double l = (j > 0) ? left : 0.0;
        return (fabs(xi - yj) <= arg) ? dmax3(d1 + 1, u, l) : dmax3(d1, u, l);
    } else {  // MSM
        if (i == 0 && j == 0) return fabs(xi - yj);
        if (i == 0) return left + msm_cost(yj, xi, yp, arg);
        if (j == 0) return up + msm_cost(xi, xp, yj, arg);
        const double c1 = diag + fabs(xi - yj);
        const double c2 = up + msm_cost(xi, xp, yj, arg);
        const double c3 = left + msm_cost(yj, xi, yp, arg);
        return dmin3(c1, c2, c3);
    }
}

// Anti-diagonal DP for one lane, compile-time band half-width W >= w.
// Iteration b handles anti-diagonals dA = 2b + W (even slots 2r, i = b + r,
// j = b + W - r) and dB = dA + 1 (odd slots 2r+1, i = b + 1 + r, j = b + W - r).
// Rings: X[k] = x[b - 1 + k], Y[k] = y[b - 1 + k], k in [0, W+1]; a cell reads
// x[i] = X[i - b + 1] and x[i-1] = X[i - b] (likewise y), ERP/EDR the latter.
template <int METRIC, int W>
__device__ double eld_ad(const double* __restrict__ q, const double* __restrict__ qg,
                         const CandSrc& c, int m, int w, double arg, double thr, bool active,
                         bool* dead_out) {
    constexpr bool GRID1 = (METRIC == ERP || METRIC == EDR);
    const int M = GRID1 ? m : m - 1;  // last grid index; ERP/EDR cells read x[i-1], y[j-1]
    const double S = (METRIC == LCSS) ? 0.0 : kEldSentinel;
    double A[W + 1], B[W + 1];  // B[W] is a padding slot (always S)
    double X[W + 2], Y[W + 2], XG[W + 2], YG[W + 2];
    const int b0 = -((W + 1) / 2) - 1;
    const int bend = (2 * M - W) >= 0 ? (2 * M - W) / 2 : -((W - 2 * M + 1) / 2);
#pragma unroll
    for (int k = 0; k <= W; k++) {
        A[k] = S;
        B[k] = S;
    }
#pragma unroll
    for (int k = 0; k < W + 2; k++) {
        const int idx = b0 - 1 + k;
        X[k] = qat(q, idx, m);
        Y[k] = c.at(idx);
        if (METRIC == ERP) {
            XG[k] = qat(qg, idx, m);
            YG[k] = (Y[k] - arg) * (Y[k] - arg);
        }
    }
    bool alive = active;
    double result = S;
    const bool check = thr < 1e300 && METRIC != LCSS;
    int it = 0;
    for (int b = b0; b <= bend; b++, it++) {
        // phase A: slots 2r, cells (b + r, b + W - r)
#pragma unroll
        for (int r = 0; r <= W; r++) {
            const int i = b + r, j = b + W - r;
            const int delta = 2 * r - W;
            double v = S;
            if (i >= 0 && j >= 0 && i <= M && j <= M && delta <= w && delta >= -w) {
                const double up = (r >= 1) ? B[r - 1] : S;
                const double left = (r <= W - 1) ? B[r] : S;
                v = cell<METRIC>(i, j, up, left, A[r], X[r + 1], X[r], Y[W - r + 1], Y[W - r],
                                 METRIC == ERP ? XG[r] : 0.0, METRIC == ERP ? YG[W - r] : 0.0,
                                 arg);
            }
            A[r] = v;
        }
        if (!(W & 1) && b == bend) result = A[W / 2];
        // phase B: slots 2r+1, cells (b + 1 + r, b + W - r)
#pragma unroll
        for (int r = 0; r < W; r++) {
            const int i = b + 1 + r, j = b + W - r;
            const int delta = 2 * r + 1 - W;
            double v = S;
            if (i >= 0 && j >= 0 && i <= M && j <= M && delta <= w && delta >= -w) {
                v = cell<METRIC>(i, j, A[r], A[r + 1], B[r], X[r + 2], X[r + 1], Y[W - r + 1],
                                 Y[W - r], METRIC == ERP ? XG[r + 1] : 0.0,
                                 METRIC == ERP ? YG[W - r] : 0.0, arg);
            }
            B[r] = v;
        }
        if ((W & 1) && b == bend) result = B[(W - 1) / 2];
        // shift the rings: X[k] <- x[b + k]
        {
            const int idx = b + W + 1;
            const double xn = qat(q, idx, m);
            const double yn = c.at(idx);
#pragma unroll
            for (int k = 0; k < W + 1; k++) {
                X[k] = X[k + 1];
                Y[k] = Y[k + 1];
                if (METRIC == ERP) {
                    XG[k] = XG[k + 1];
                    YG[k] = YG[k + 1];
                }
            }
            X[W + 1] = xn;
            Y[W + 1] = yn;
            if (METRIC == ERP) {
                XG[W + 1] = qat(qg, idx, m);
                YG[W + 1] = (yn - arg) * (yn - arg);
            }
        }
        if (check && (it & 7) == 7) {
            double mn = A[0];
#pragma unroll
            for (int k = 1; k <= W; k++) mn = fmin(mn, A[k]);
#pragma unroll
            for (int k = 0; k < W; k++) mn = fmin(mn, B[k]);
            if (mn > thr) alive = false;
            if (!__any(alive)) break;
        }
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
                    v = cell<METRIC>(i, j, up, left, b[k], 0.0, xi, 0.0, yj, xg,
                                     (yj - arg) * (yj - arg), arg);
                else
                    v = cell<METRIC>(i, j, up, left, b[k], xi, xp, yj, yp, 0.0, 0.0, arg);
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
// 0..n_direct-1 of normalised candidates.
template <int METRIC, int W>
__global__ void __launch_bounds__(256) k_eld(VerifyArgs a, const int64_t* __restrict__ in,
                                             const unsigned long long* __restrict__ in_count,
                                             int64_t n_direct, const double* __restrict__ rows,
                                             int keep_all, int64_t* __restrict__ out_idx,
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
    if (!keep_all) {
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
        bool dead = false;
        double d;
        if constexpr (W > 0) {
            d = eld_ad<METRIC, W>(a.q.qn, a.q.XG, c, m, w, a.p.arg, thr, have, &dead);
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

template <int METRIC>
static hipError_t launch_metric(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                                const unsigned long long* in_count, int64_t n_direct,
                                const double* rows, bool keep_all, int64_t* out_idx,
                                double* out_dist, unsigned long long* out_count,
                                double* scratch, int64_t scratch_threads) {
    const int w = a.p.w;
    const dim3 blk(256), grid(2048);
    const int ka = keep_all ? 1 : 0;
    if (w <= 6)
        hipLaunchKernelGGL((k_eld<METRIC, 6>), grid, blk, 0, st, a, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch);
    else if (w <= 8)
        hipLaunchKernelGGL((k_eld<METRIC, 8>), grid, blk, 0, st, a, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch);
    else if (w <= 16)
        hipLaunchKernelGGL((k_eld<METRIC, 16>), grid, blk, 0, st, a, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch);
    else if (w <= 25)
        hipLaunchKernelGGL((k_eld<METRIC, 25>), grid, blk, 0, st, a, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch);
    else {
        if (!scratch) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_eld<METRIC, 0>), dim3((unsigned)(scratch_threads / 256)), blk, 0, st, a, in, in_count, n_direct, rows, ka, out_idx, out_dist, out_count, scratch);
    }
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
