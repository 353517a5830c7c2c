// host_verify.cpp -- the exact elastic distance of one window on the host
// (FP64, built with -ffp-contract=off), for the re-verification of the
// windows whose device distance lies within 1e-9 of a decision boundary.
//
// The recurrences are the reference's (src/ElasticMeasures.cpp: dtw :24-61,
// lcss :63-93, erp :95-130, edr :132-165, msm :167-198, swale :200-231),
// each cell with the reference's expression and operand order, evaluated
// over full-width rows whose cells outside the Sakoe-Chiba band keep the
// value the reference's buffers hold there (1e18, LCSS 0).  x is the query,
// y the candidate.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/fsmeld.h"

namespace fsm {

namespace {

constexpr double kOut = 1e18;  // src/ElasticMeasures.cpp:7

inline double sq(double a, double b) { return (a - b) * (a - b); }
inline double min3(double a, double b, double c) { return std::min(std::min(a, b), c); }
inline double max3(double a, double b, double c) { return std::max(std::max(a, b), c); }

// the reference's msm_dist_local: c when np lies between x and y, else c plus
// the distance to the nearer one
inline double split_cost(double np, double x, double y, double c) {
    if ((x <= np && np <= y) || (y <= np && np <= x)) return c;
    return c + std::min(std::abs(np - x), std::abs(np - y));
}

// Two full-width rows.  Row i writes columns [i - w, i + w] (shifted by one
// on the gap grids) and reads its neighbours one column either side, so
// moving to row i clears columns [i - w - 1, i + w + 1] of the buffer it
// reuses (two rows back): every cell a row reads outside its band then holds
// the fill value, at O(w) per row.
struct Rows {
    std::vector<double> a, b;
    double* prev;
    double* cur;
    int n, w;
    Rows(int n_, int w_, double fill)
        : a((size_t)n_, fill), b((size_t)n_, fill), prev(a.data()), cur(b.data()), n(n_), w(w_) {}
    void next(double fill, int i) {
        std::swap(prev, cur);
        const int lo = std::max(0, i - w - 2), hi = std::min(n, i + w + 3);
        std::fill(cur + lo, cur + hi, fill);
    }
};

// Inner loops.  Each row's cells in column order form one dependency chain
// through the left neighbour; everything else a cell needs comes from the
// previous row.  So the loops below take the first column (and the first
// row) apart and run the interior branch-free, with the cell written as
// min(min(up-side terms), left + cost) (LCSS: max): min and max are exact
// and the values are never NaN or -0, so any grouping of the reference's
// min3 / max3 gives its bits, and only the last add and min stay on the
// chain.

double dtw(const double* x, const double* y, int m, int w) {
    Rows r(m, w, kOut);
    r.cur[0] = sq(x[0], y[0]);
    for (int j = 1; j <= std::min(m - 1, w); j++)  // row 0: up and diag are out of the matrix
        r.cur[j] = min3(kOut, r.cur[j - 1], kOut) + sq(x[0], y[j]);
    for (int i = 1; i < m; i++) {
        r.next(kOut, i);
        const double* P = r.prev;
        double* C = r.cur;
        const double xi = x[i];
        int j = std::max(0, i - w);
        const int hi = std::min(m - 1, i + w);
        if (j == 0) {  // column 0: left and diag are out of the matrix
            C[0] = min3(P[0], kOut, kOut) + sq(xi, y[0]);
            j = 1;
        }
        double left = C[j - 1];
        for (; j <= hi; j++) {
            left = std::min(std::min(P[j], P[j - 1]), left) + sq(xi, y[j]);
            C[j] = left;
        }
    }
    const double f = r.cur[m - 1];
    if (!std::isfinite(f)) return kOut;
    return std::sqrt(std::max(0.0, f));
}

double lcss(const double* x, const double* y, int m, int w, double eps) {
    Rows r(m, w, 0.0);
    {  // row 0
        const int hi = std::min(m, w + 1);
        for (int j = 0; j < hi; j++) {
            const bool hit = std::abs(x[0] - y[j]) <= eps;
            r.cur[j] = hit ? 1 : ((j > 0) ? r.cur[j - 1] : 0);
        }
    }
    for (int i = 1; i < m; i++) {
        r.next(0.0, i);
        const double* P = r.prev;
        double* C = r.cur;
        const double xi = x[i];
        int j = std::max(0, i - w);
        const int hi = std::min(m, i + w + 1);
        if (j == 0) {
            C[0] = (std::abs(xi - y[0]) <= eps) ? 1 : P[0];
            j = 1;
        }
        // the band's first cell reads 0 on its left, which the cleared
        // column lo - 1 holds
        double left = C[j - 1];
        for (; j < hi; j++) {
            const double h = (std::abs(xi - y[j]) <= eps) ? 1.0 : 0.0;
            left = std::max(std::max(P[j - 1] + h, P[j]), left);
            C[j] = left;
        }
    }
    return 1 - r.cur[m - 1] / m;
}

// ERP and EDR run on the (m + 1) x (m + 1) grid with a gap row and column
double erp(const double* x, const double* y, int m, int w, double g) {
    Rows r(m + 1, w, kOut);
    std::vector<double> gy((size_t)m);
    for (int j = 0; j < m; j++) gy[j] = sq(y[j], g);
    r.cur[0] = 0.0;
    for (int j = 1; j <= std::min(m, w); j++) r.cur[j] = r.cur[j - 1] + gy[j - 1];
    for (int i = 1; i <= m; i++) {
        r.next(kOut, i);
        const double* P = r.prev;
        double* C = r.cur;
        const double xi = x[i - 1], gx = sq(xi, g);
        if (i <= w) C[0] = P[0] + gx;
        int j = std::max(1, i - w);
        const int hi = std::min(m, i + w);
        double left = C[j - 1];
        for (; j <= hi; j++) {
            const double match = P[j - 1] + sq(xi, y[j - 1]);
            const double del = P[j] + gx;
            left = std::min(std::min(match, del), left + gy[j - 1]);
            C[j] = left;
        }
    }
    const double f = r.cur[m];
    if (!std::isfinite(f)) return kOut;
    return std::sqrt(std::max(0.0, f));
}

double edr(const double* x, const double* y, int m, int w, double eps) {
    Rows r(m + 1, w, kOut);
    r.cur[0] = 0.0;
    for (int j = 1; j <= std::min(m, w); j++) r.cur[j] = (double)j;
    for (int i = 1; i <= m; i++) {
        r.next(kOut, i);
        const double* P = r.prev;
        double* C = r.cur;
        const double xi = x[i - 1];
        if (i <= w) C[0] = (double)i;
        int j = std::max(1, i - w);
        const int hi = std::min(m, i + w);
        double left = C[j - 1];
        for (; j <= hi; j++) {
            const double sub = (std::abs(xi - y[j - 1]) <= eps) ? 0.0 : 1.0;
            left = std::min(std::min(P[j] + 1.0, P[j - 1] + sub), left + 1.0);
            C[j] = left;
        }
    }
    return r.cur[m];
}

// SWALE, src/ElasticMeasures.cpp:200-231 with the costs of include/funcs.h:30-31
double swale(const double* x, const double* y, int m, int w, double eps) {
    constexpr double kP = 5.0, kR = 1.0;
    Rows r(m + 1, w, kOut);
    r.cur[0] = 0.0;
    for (int j = 1; j <= std::min(m, w); j++) r.cur[j] = r.cur[j - 1] + kP;
    for (int i = 1; i <= m; i++) {
        r.next(kOut, i);
        const double* P = r.prev;
        double* C = r.cur;
        const double xi = x[i - 1];
        if (i <= w) C[0] = P[0] + kP;
        int j = std::max(1, i - w);
        const int hi = std::min(m, i + w);
        double left = C[j - 1];
        for (; j <= hi; j++) {
            const double mv = (std::abs(xi - y[j - 1]) <= eps) ? kR : kP;
            left = std::min(std::min(P[j - 1] + mv, P[j] + kP), left + kP);
            C[j] = left;
        }
    }
    return r.cur[m];
}

double msm(const double* x, const double* y, int m, int w, double c) {
    Rows r(m, w, kOut);
    r.cur[0] = std::abs(x[0] - y[0]);
    for (int j = 1; j < m && j <= w; j++) r.cur[j] = r.cur[j - 1] + split_cost(y[j], x[0], y[j - 1], c);
    for (int i = 1; i < m; i++) {
        r.next(kOut, i);
        const double* P = r.prev;
        double* C = r.cur;
        const double xi = x[i], xp = x[i - 1];
        if (i <= w) C[0] = P[0] + split_cost(xi, xp, y[0], c);
        int j = std::max(1, i - w);
        const int hi = std::min(m, i + w + 1);
        double left = C[j - 1];
        for (; j < hi; j++) {
            const double c1 = P[j - 1] + std::abs(xi - y[j]);
            const double c2 = P[j] + split_cost(xi, xp, y[j], c);
            left = std::min(std::min(c1, c2), left + split_cost(y[j], xi, y[j - 1], c));
            C[j] = left;
        }
    }
    return r.cur[m - 1];
}

}  // namespace

// src/fsmeld.cpp:170-183 (exact_elastic_distance)
double host_eld(int metric, const double* x, const double* y, int m, int w, double arg) {
    switch (metric) {
        case FSM_METRIC_DTW: return dtw(x, y, m, w);
        case FSM_METRIC_ERP: return erp(x, y, m, w, arg);
        case FSM_METRIC_EDR: return edr(x, y, m, w, arg);
        case FSM_METRIC_LCSS: return lcss(x, y, m, w, arg);
        case FSM_METRIC_MSM: return msm(x, y, m, w, arg);
        case FSM_METRIC_SWALE: return swale(x, y, m, w, arg);
    }
    return std::nan("");
}

}  // namespace fsm
