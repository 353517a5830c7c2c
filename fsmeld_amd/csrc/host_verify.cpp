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

double dtw(const double* x, const double* y, int m, int w) {
    Rows r(m, w, kOut);
    for (int i = 0; i < m; i++) {
        if (i) r.next(kOut, i);
        for (int j = std::max(0, i - w); j <= std::min(m - 1, i + w); j++) {
            if (i == 0 && j == 0) {
                r.cur[0] = sq(x[0], y[0]);
                continue;
            }
            const double up = i ? r.prev[j] : kOut;
            const double left = j ? r.cur[j - 1] : kOut;
            const double diag = (i && j) ? r.prev[j - 1] : kOut;
            r.cur[j] = min3(up, left, diag) + sq(x[i], y[j]);
        }
    }
    const double f = r.cur[m - 1];
    if (!std::isfinite(f)) return kOut;
    return std::sqrt(std::max(0.0, f));
}

double lcss(const double* x, const double* y, int m, int w, double eps) {
    Rows r(m, w, 0.0);
    for (int i = 0; i < m; i++) {
        if (i) r.next(0.0, i);
        const int lo = std::max(0, i - w), hi = std::min(m, i + w + 1);
        for (int j = lo; j < hi; j++) {
            const bool hit = std::abs(x[i] - y[j]) <= eps;
            double v;
            if (i + j == 0) v = hit ? 1 : 0;
            else if (i == 0) v = hit ? 1 : ((j > lo) ? r.cur[j - 1] : 0);
            else if (j == 0) v = hit ? 1 : r.prev[j];
            else {
                const double v3 = (j > lo) ? r.cur[j - 1] : 0;
                v = hit ? max3(r.prev[j - 1] + 1, r.prev[j], v3)
                        : max3(r.prev[j - 1], r.prev[j], v3);
            }
            r.cur[j] = v;
        }
    }
    return 1 - r.cur[m - 1] / m;
}

// ERP and EDR run on the (m + 1) x (m + 1) grid with a gap row and column
double erp(const double* x, const double* y, int m, int w, double g) {
    Rows r(m + 1, w, kOut);
    r.cur[0] = 0.0;
    for (int j = 1; j <= std::min(m, w); j++) r.cur[j] = r.cur[j - 1] + sq(y[j - 1], g);
    for (int i = 1; i <= m; i++) {
        r.next(kOut, i);
        if (i <= w) r.cur[0] = r.prev[0] + sq(x[i - 1], g);
        for (int j = std::max(1, i - w); j <= std::min(m, i + w); j++) {
            const double match = r.prev[j - 1] + sq(x[i - 1], y[j - 1]);
            const double ins = r.cur[j - 1] + sq(y[j - 1], g);
            const double del = r.prev[j] + sq(x[i - 1], g);
            r.cur[j] = min3(match, ins, del);
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
        if (i <= w) r.cur[0] = (double)i;
        for (int j = std::max(1, i - w); j <= std::min(m, i + w); j++) {
            const double sub = (std::abs(x[i - 1] - y[j - 1]) <= eps) ? 0.0 : 1.0;
            r.cur[j] = min3(r.prev[j] + 1.0, r.cur[j - 1] + 1.0, r.prev[j - 1] + sub);
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
        if (i <= w) r.cur[0] = r.prev[0] + kP;
        for (int j = std::max(1, i - w); j <= std::min(m, i + w); j++) {
            const double mv = (std::abs(x[i - 1] - y[j - 1]) <= eps) ? kR : kP;
            r.cur[j] = min3(r.prev[j - 1] + mv, r.prev[j] + kP, r.cur[j - 1] + kP);
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
        if (i <= w) r.cur[0] = r.prev[0] + split_cost(x[i], x[i - 1], y[0], c);
        for (int j = std::max(1, i - w); j < std::min(m, i + w + 1); j++) {
            const double c1 = r.prev[j - 1] + std::abs(x[i] - y[j]);
            const double c2 = r.prev[j] + split_cost(x[i], x[i - 1], y[j], c);
            const double c3 = r.cur[j - 1] + split_cost(y[j], x[i], y[j - 1], c);
            r.cur[j] = min3(c1, c2, c3);
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
