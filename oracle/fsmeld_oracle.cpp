// fsmeld_oracle.cpp -- CPU oracle for FSMELD's LB_IP -> GLB -> ELD cascade.
//
// TEST INFRASTRUCTURE ONLY.  A CPU restatement of the reference's algorithm,
// used as the checker by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  The product (fsmeld_amd/) never links or calls it.
//
// Every function follows the reference's arithmetic order so that the values
// that decide matches (window mean/std, normalised candidates, exact elastic
// distances) are bit-identical.  Build with -O2 -ffp-contract=off.
//
// Citations are to the reference tree (src/..., include/...).
// Deviations, all documented where they occur:
//   * the sliding inner product (src/sip.cpp, FFTW3) is a direct O(m) sum per
//     window; FFTW's own bits are not reproduced ("parity unpinned" for LB
//     values at the 1e-12 level; the 1e-9 guard of
//     include/fsmeld_internal.h:9-19 absorbs the difference);
//   * the ERP/EDR/LCSS/MSM verifiers evaluate only the Sakoe-Chiba band; the
//     reference fills O(m^2) cells, but out-of-band cells hold the constant
//     sentinel (1e18, or 0 for LCSS) and never win a min/max inside the band,
//     so the band-only DP returns identical bits (checked against the
//     reference's own ElasticMeasures.cpp in tests/test_oracle_ref.py).
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <atomic>
#include <limits>
#include <queue>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { M_DTW = 1, M_EDR = 2, M_ERP = 3, M_LCSS = 4, M_MSM = 5, M_SWALE = 6 };
// include/funcs.h:30-31: SWALE's match reward and mismatch / gap cost
constexpr double kSwaleR = 1.0, kSwaleP = 5.0;

constexpr double kEldSentinel = 1e18;        // src/ElasticMeasures.cpp:7
constexpr double kGlbAbandon = 1008610086.0; // include/def.h:7

inline double sq(double x, double y) { return (x - y) * (x - y); }  // include/def.h:5

// ---------------------------------------------------------------- mvmean ---
// src/mvmean.cpp:9-46: running sum1/sum2 recurrence; non-finite samples are
// kept out of the sums and poison mu/sigma of every window containing them.
void mvmean(const double* a, int len_a, int l, double* miu, double* si) {
    double sum1 = 0, sum2 = 0;
    int nonfinite = 0;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int i = 0; i < l; i++) {
        if (std::isfinite(a[i])) {
            sum1 += a[i];
            sum2 += a[i] * a[i];
        } else {
            ++nonfinite;
        }
    }
    const int ll = len_a - l + 1;
    miu[0] = nonfinite ? nan : sum1 / l;
    si[0] = nonfinite ? nan : std::sqrt(sum2 / l - miu[0] * miu[0]);
    for (int i = 1; i < ll; i++) {
        const bool of = std::isfinite(a[i - 1]);
        const bool nf = std::isfinite(a[i + l - 1]);
        const double out = of ? a[i - 1] : 0.0;
        const double in = nf ? a[i + l - 1] : 0.0;
        nonfinite += !nf - !of;
        sum1 = sum1 - out + in;
        sum2 = sum2 - out * out + in * in;
        miu[i] = nonfinite ? nan : sum1 / l;
        si[i] = nonfinite ? nan : std::sqrt(sum2 / l - miu[i] * miu[i]);
    }
}

// src/fsmeld.cpp:110-129 (without the exclusion part, applied by the drivers
// with global window indices exactly as :139-152 does).
void invalid_mask(const double* t, const double* si, int seg, int m, uint8_t* mask) {
    const int wc = seg - m + 1;
    if (wc <= 0) return;
    int bad = 0;
    for (int k = 0; k < m; k++)
        if (!std::isfinite(t[k])) bad++;
    mask[0] = bad > 0;
    for (int j = 1; j < wc; j++) {
        if (!std::isfinite(t[j - 1])) bad--;
        if (!std::isfinite(t[j + m - 1])) bad++;
        mask[j] = bad > 0 || !std::isfinite(si[j]) || si[j] <= 1e-12;
    }
    mask[0] = mask[0] || !std::isfinite(si[0]) || si[0] <= 1e-12;
}

// src/lower_upper_lemire.cpp:94-126: min/max over the windows that contain
// position i, by monotone deques (one `if`-pop per step, as in the source;
// NaN inputs leave the deques non-monotone exactly as they do there).
void min_max_window(const double* a, int n, int m, double* mn, double* mx) {
    const int kml = n - m + 1;
    std::vector<int> qu(kml), ql(kml);
    int hl = 0, tl = -1, hu = 0, tu = -1;
    for (int i = 0; i < kml; i++) {
        while (hl <= tl && a[ql[tl]] >= a[i]) tl--;
        ql[++tl] = i;
        if (ql[hl] <= i - m) hl++;
        mn[i] = a[ql[hl]];
        while (hu <= tu && a[qu[tu]] <= a[i]) tu--;
        qu[++tu] = i;
        if (qu[hu] <= i - m) hu++;
        mx[i] = a[qu[hu]];
    }
    for (int i = kml; i < n; i++) {
        if (ql[hl] <= i - m) hl++;
        mn[i] = a[ql[hl]];
        if (qu[hu] <= i - m) hu++;
        mx[i] = a[qu[hu]];
    }
}

// src/lower_upper_lemire.cpp:8-44: centred envelope over [i-r, i+r], clipped.
void lower_upper_lemire(const double* a, int n, int r, double* l, double* u) {
    int li = 0, ui = 0;
    const int k = 2 * r + 1;
    std::vector<int> qu(n), ql(n);
    int hl = 0, tl = -1, hu = 0, tu = -1;
    for (int i = 0; i < k - 1 - r; i++) {
        while (hl <= tl && a[ql[tl]] >= a[i]) tl--;
        ql[++tl] = i;
        while (hu <= tu && a[qu[tu]] <= a[i]) tu--;
        qu[++tu] = i;
    }
    for (int i = k - 1 - r; i < n; i++) {
        while (hl <= tl && a[ql[tl]] >= a[i]) tl--;
        ql[++tl] = i;
        while (ql[hl] <= i - k) hl++;
        l[li++] = a[ql[hl]];
        while (hu <= tu && a[qu[tu]] <= a[i]) tu--;
        qu[++tu] = i;
        while (qu[hu] <= i - k) hu++;
        u[ui++] = a[qu[hu]];
    }
    for (int i = n; i < n + r; i++) {
        while (qu[hu] <= i - k) hu++;
        u[ui++] = a[qu[hu]];
        while (ql[hl] <= i - k) hl++;
        l[li++] = a[ql[hl]];
    }
}

// src/lb_normalize.cpp:9-32 (paper Eqs. 33-34).
void normalized_arrays(const double* U, const double* L, double* Un, double* Ln, int len,
                       const double* max_mean, const double* min_mean,
                       const double* max_std, const double* min_std) {
    for (int i = 0; i < len; i++) {
        const double nu = U[i] - min_mean[i];
        Un[i] = (nu >= 0) ? nu / min_std[i] : nu / max_std[i];
        const double nl = L[i] - max_mean[i];
        Ln[i] = (nl <= 0) ? nl / min_std[i] : nl / max_std[i];
    }
}

// ------------------------------------------------------------ thresholds ---
// src/find_t_lp.cpp.  The generic forms (:148-216) are specialised per metric.
struct Unique { double value; int count; };

// src/find_t_lp.cpp:53-67: run-length dedupe of consecutive equal values in
// T[m-1 .. m-1+n).
int unique_runs(const double* T, Unique* out, int m, int n) {
    if (n <= 0) return 0;
    int j = 0;
    for (int i = 0; i < n; i++) {
        if (i == 0 || T[m - 1 + i] != T[m + i - 2]) {
            out[j].value = T[m - 1 + i];
            out[j].count = 1;
            j++;
        } else {
            out[j - 1].count++;
        }
    }
    return j;
}

// src/find_t_lp.cpp:68-81 sorts the unique runs with std::sort.  With NaN
// values that call is undefined behaviour (the comparator is no strict weak
// order); here NaN runs are moved, in order, behind all numbers first, which
// is the reference's behaviour whenever no NaN is present.
void sort_runs(std::vector<Unique>& R, int nr, bool ascending) {
    auto mid = std::stable_partition(R.begin(), R.begin() + nr,
                                     [](const Unique& u) { return !std::isnan(u.value); });
    if (ascending)
        std::sort(R.begin(), mid, [](const Unique& a, const Unique& b) { return a.value < b.value; });
    else
        std::sort(R.begin(), mid, [](const Unique& a, const Unique& b) { return a.value > b.value; });
}

double D_cost(double q, double C, int metric) {  // src/find_t_lp.cpp:148-166
    if (metric == M_SWALE) return kSwaleP;
    if (metric == M_MSM) return C;
    if (metric == M_ERP || metric == M_DTW) return (q - C) * (q - C);
    if (metric == M_EDR || metric == M_LCSS) return 1;
    return 0.0;
}
double tau_u(double q, double C, int metric) {  // src/find_t_lp.cpp:167-176
    if (metric == M_MSM) return q - C;
    if (metric == M_ERP) return q - std::abs(q - C);
    if (metric == M_EDR || metric == M_LCSS || metric == M_SWALE) return q - C;
    return q;
}
double tau_l(double q, double C, int metric) {  // src/find_t_lp.cpp:177-186
    if (metric == M_MSM) return q + C;
    if (metric == M_ERP) return q + std::abs(q - C);
    if (metric == M_EDR || metric == M_LCSS || metric == M_SWALE) return q + C;
    return q;
}
double f_p(const double* mp, int p, int metric, double t) {  // src/find_t_lp.cpp:187-204
    if (metric == M_SWALE) return mp[0] * kSwaleR;
    if (p == 0) {
        if (metric == M_LCSS || metric == M_EDR) return 0;
        return mp[0];
    } else if (p == 1) {
        return mp[1] - mp[0] * t;
    } else if (p == 2) {
        return mp[2] - mp[1] * t * 2 + mp[0] * std::pow(t, 2);
    }
    return 0.0;
}
int first_max(const double* a, int n) {  // src/find_t_lp.cpp:206-216
    if (n <= 0) return -1;
    int mi = 0;
    for (int i = 0; i < n; ++i)
        if (a[i] > a[mi]) mi = i;
    return mi;
}
int metric_p(int metric) {  // src/fsmeld.cpp:586-598
    if (metric == M_ERP || metric == M_DTW) return 2;
    if (metric == M_MSM) return 1;
    return 0;
}

// src/find_t_lp.cpp:218-287
double find_tl_dtw(const double* L, int l, const double* qs, int m, int p) {
    const int n = l - 2 * m + 2;
    std::vector<Unique> R(n);
    std::vector<double> LB(n, 0.0);
    double mp[3] = {0, 0, 0};
    int nr = unique_runs(L, R.data(), m, n);
    sort_runs(R, nr, false);
    int g = 0;
    const double t0 = R[0].value;
    int idx1 = m - 1;
    while (idx1 >= 0 && qs[idx1] >= t0) idx1--;
    for (int i = 0; i <= idx1; i++)
        for (int a = 0; a < p + 1; a++) mp[a] += std::pow(qs[i], a);
    for (int i = 0; i < nr; i++) {
        const double t = R[i].value;
        g = g + R[i].count;
        while (idx1 >= 0) {
            const double v = qs[idx1];
            if (v >= t) {
                for (int a = 0; a < p + 1; a++) mp[a] -= std::pow(v, a);
                idx1--;
            } else {
                break;
            }
        }
        LB[i] = std::abs(f_p(mp, p, M_DTW, t)) * g;
    }
    return R[first_max(LB.data(), nr)].value;
}

// src/find_t_lp.cpp:655-724
double find_tu_dtw(const double* U, int l, const double* qs, int m, int p) {
    const int n = l - 2 * m + 2;
    std::vector<Unique> R(n);
    std::vector<double> LB(n, 0.0);
    double mp[3] = {0, 0, 0};
    int nr = unique_runs(U, R.data(), m, n);
    sort_runs(R, nr, true);
    int g = 0;
    const double t0 = R[0].value;
    int idx1 = 0;
    while (idx1 < m && qs[idx1] <= t0) idx1++;
    for (int i = idx1; i < m; i++)
        for (int a = 0; a < p + 1; a++) mp[a] += std::pow(qs[i], a);
    for (int i = 0; i < nr; i++) {
        const double t = R[i].value;
        g = g + R[i].count;
        while (idx1 < m) {
            const double v = qs[idx1];
            if (v < t) {
                for (int a = 0; a < p + 1; a++) mp[a] -= std::pow(v, a);
                idx1++;
            } else {
                break;
            }
        }
        LB[i] = f_p(mp, p, M_DTW, t) * g;
    }
    return R[first_max(LB.data(), nr)].value;
}

// src/find_t_lp.cpp:288-398 (MSM, EDR, LCSS); :399-509 (SWALE, the same sweep)
double find_tl_generic(const double* L, int l, const double* qs, int m, int p, double C_,
                       int metric) {
    const int n = l - 2 * m + 2;
    std::vector<Unique> R(n);
    std::vector<double> LB(n, 0.0);
    double mp[3] = {0, 0, 0};
    int nr = unique_runs(L, R.data(), m, n);
    sort_runs(R, nr, false);
    int g = 0;
    double C = 0;
    const double t0 = R[0].value;
    int idx1 = m - 1;
    while (idx1 >= 0 && qs[idx1] >= t0) idx1--;
    int idx2 = idx1;
    while (idx2 >= 0) {
        const double v = tau_l(qs[idx2], C_, metric);
        if (v >= t0) {
            for (int a = 0; a < p + 1; a++) mp[a] += std::pow(qs[idx2], a);
            idx2--;
        } else {
            break;
        }
    }
    for (int j = 0; j <= idx2; j++) C += D_cost(qs[j], C_, metric);
    for (int i = 0; i < nr; i++) {
        const double t = R[i].value;
        g = g + R[i].count;
        while (idx2 >= 0) {
            const double v = tau_l(qs[idx2], C_, metric);
            if (v >= t) {
                C = C - D_cost(qs[idx2], C_, metric);
                for (int a = 0; a < p + 1; a++) mp[a] += std::pow(qs[idx2], a);
                idx2--;
            } else {
                break;
            }
        }
        while (idx1 > idx2) {
            const double v = qs[idx1];
            if (v > t) {
                for (int a = 0; a < p + 1; a++) mp[a] -= std::pow(v, a);
                idx1--;
            } else {
                break;
            }
        }
        LB[i] = (C + std::abs(f_p(mp, p, metric, t))) * g;
    }
    return R[first_max(LB.data(), nr)].value;
}

// src/find_t_lp.cpp:726-825 (MSM, EDR, LCSS); :827-927 (SWALE, the same sweep)
double find_tu_generic(const double* U, int l, const double* qs, int m, int p, double C_,
                       int metric) {
    const int n = l - 2 * m + 2;
    std::vector<Unique> R(n);
    std::vector<double> LB(n, 0.0);
    double mp[3] = {0, 0, 0};
    int nr = unique_runs(U, R.data(), m, n);
    sort_runs(R, nr, true);
    int g = 0;
    double C = 0;
    const double t0 = R[0].value;
    int idx1 = 0, idx2 = 0;
    while (idx1 < m && qs[idx1] <= t0) idx1++;
    idx2 = idx1;
    while (idx2 < m) {
        const double v = tau_u(qs[idx2], C_, metric);
        if (v <= t0) {
            for (int a = 0; a < p + 1; a++) mp[a] += std::pow(qs[idx2], a);
            idx2++;
        } else {
            break;
        }
    }
    for (int j = idx2; j < m; j++) C += D_cost(qs[j], C_, metric);
    for (int i = 0; i < nr; i++) {
        const double t = R[i].value;
        g = g + R[i].count;
        while (idx2 < m) {
            const double v = tau_u(qs[idx2], C_, metric);
            if (v <= t) {
                C = C - D_cost(qs[idx2], C_, metric);
                for (int a = 0; a < p + 1; a++) mp[a] += std::pow(qs[idx2], a);
                idx2++;
            } else {
                break;
            }
        }
        while (idx1 < idx2) {
            const double v = qs[idx1];
            if (v < t) {
                for (int a = 0; a < p + 1; a++) mp[a] -= std::pow(v, a);
                idx1++;
            } else {
                break;
            }
        }
        LB[i] = (C + f_p(mp, p, metric, t)) * g;
    }
    return R[first_max(LB.data(), nr)].value;
}

// src/find_t_lp.cpp:25-42 heap items (priority_queue orderings kept as is).
struct MinItem {  // HeapItem: top() is the smallest value
    double value;
    int index;
    bool operator<(const MinItem& o) const { return value > o.value; }
};
struct MaxItem {  // maxHeapItem: top() is the largest value
    double value;
    int index;
    bool operator<(const MaxItem& o) const { return value < o.value; }
};

// src/find_t_lp.cpp:510-582 (ERP lower; iterates the unsorted query)
double find_tl_erp(const double* L, int l, const double* q, int m, int p, double C_) {
    const int n = l - 2 * m + 2;
    std::vector<Unique> R(n);
    std::vector<double> LB(n, 0.0);
    double mp[3] = {0, 0, 0};
    double C = 0.0;
    int g = 0;
    int nr = unique_runs(L, R.data(), m, n);
    sort_runs(R, nr, false);
    const double t0 = R[0].value;
    std::priority_queue<MaxItem> H2, H1;
    for (int i = 0; i < m; i++) {
        const double tau = tau_l(q[i], C_, M_ERP);
        if (q[i] >= t0) {
        } else if (tau >= t0) {
            H1.push({q[i], i});
            for (int a = 0; a < p + 1; a++) mp[a] += std::pow(q[i], a);
        } else {
            H2.push({tau, i});
            C += D_cost(q[i], C_, M_ERP);
        }
    }
    for (int i = 0; i < nr; i++) {
        const double t = R[i].value;
        g = g + R[i].count;
        while (!H2.empty() && H2.top().value >= t) {
            const int oi = H2.top().index;
            const double oq = q[oi];
            H2.pop();
            C -= D_cost(oq, C_, M_ERP);
            H1.push({oq, oi});
            for (int a = 0; a < p + 1; a++) mp[a] += std::pow(oq, a);
        }
        while (!H1.empty() && H1.top().value >= t) {
            const double oq = q[H1.top().index];
            H1.pop();
            for (int a = 0; a < p + 1; a++) mp[a] -= std::pow(oq, a);
        }
        LB[i] = (C + std::abs(f_p(mp, p, M_ERP, t))) * g;
    }
    return R[first_max(LB.data(), nr)].value;
}

// src/find_t_lp.cpp:583-654 (ERP upper)
double find_tu_erp(const double* U, int l, const double* q, int m, int p, double C_) {
    const int n = l - 2 * m + 2;
    std::vector<Unique> R(n);
    std::vector<double> LB(n, 0.0);
    double mp[3] = {0, 0, 0};
    double C = 0.0;
    int g = 0;
    int nr = unique_runs(U, R.data(), m, n);
    sort_runs(R, nr, true);
    const double t0 = R[0].value;
    std::priority_queue<MinItem> H2, H1;
    for (int i = 0; i < m; i++) {
        const double tau = tau_u(q[i], C_, M_ERP);
        if (q[i] <= t0) {
        } else if (tau <= t0) {
            H1.push({q[i], i});
            for (int a = 0; a < p + 1; a++) mp[a] += std::pow(q[i], a);
        } else {
            H2.push({tau, i});
            C += D_cost(q[i], C_, M_ERP);
        }
    }
    for (int i = 0; i < nr; i++) {
        const double t = R[i].value;
        g = g + R[i].count;
        while (!H2.empty() && H2.top().value <= t) {
            const int oi = H2.top().index;
            const double oq = q[oi];
            H2.pop();
            C -= D_cost(oq, C_, M_ERP);
            H1.push({oq, oi});
            for (int a = 0; a < p + 1; a++) mp[a] += std::pow(oq, a);
        }
        while (!H1.empty() && H1.top().value <= t) {
            const double oq = q[H1.top().index];
            H1.pop();
            for (int a = 0; a < p + 1; a++) mp[a] -= std::pow(oq, a);
        }
        LB[i] = (C + f_p(mp, p, M_ERP, t)) * g;
    }
    return R[first_max(LB.data(), nr)].value;
}

// src/fsmeld.cpp:778-790 dispatch.
void thresholds(int metric, const double* Un, const double* Ln, int l, const double* qn,
                const double* qs, int m, double arg, double* tu, double* tl) {
    const int p = metric_p(metric);
    if (metric == M_DTW) {
        *tu = find_tu_dtw(Un, l, qs, m, p);
        *tl = find_tl_dtw(Ln, l, qs, m, p);
    } else if (metric == M_ERP) {
        *tu = find_tu_erp(Un, l, qn, m, p, arg);
        *tl = find_tl_erp(Ln, l, qn, m, p, arg);
    } else {
        *tu = find_tu_generic(Un, l, qs, m, p, arg, metric);
        *tl = find_tl_generic(Ln, l, qs, m, p, arg, metric);
    }
}

// src/q_to_h.cpp:11-134; SWALE :60-82, :128-150.
void q_to_h(int metric, const double* q, int m, double tl, double tu, double C,
            double* hl, double* hu) {
    for (int i = 0; i < m; i++) {
        double vl = 0.0, vu = 0.0;
        switch (metric) {
            case M_MSM:
                if (q[i] < tl) vl = std::min(C, tl - q[i]);
                if (q[i] > tu) vu = std::min(C, q[i] - tu);
                break;
            case M_DTW:
                if (q[i] < tl) vl = (tl - q[i]) * (tl - q[i]);
                if (q[i] > tu) vu = (q[i] - tu) * (q[i] - tu);
                break;
            case M_ERP:
                if (q[i] < tl)
                    vl = ((q[i] - tl) * (q[i] - tl) > (q[i] - C) * (q[i] - C))
                             ? (q[i] - C) * (q[i] - C) : (q[i] - tl) * (q[i] - tl);
                if (q[i] > tu)
                    vu = ((q[i] - tu) * (q[i] - tu) > (q[i] - C) * (q[i] - C))
                             ? (q[i] - C) * (q[i] - C) : (q[i] - tu) * (q[i] - tu);
                break;
            case M_EDR:
            case M_LCSS:
                if (q[i] < tl - C) vl = 1;
                if (q[i] > tu + C) vu = 1;
                break;
            case M_SWALE:
                if (q[i] < tl - C) vl = kSwaleP;
                else if (q[i] < tl) vl = kSwaleR;
                if (q[i] > tu + C) vu = kSwaleP;
                else if (q[i] > tu) vu = kSwaleR;
                break;
        }
        hl[i] = vl;
        hu[i] = vu;
    }
}

// src/lb.cpp:18-34
void s_to_g(const double* Ln, const double* Un, int n, double tl, double tu, double* gl,
            double* gu) {
    for (int i = 0; i < n; i++) {
        gl[i] = (Ln[i] >= tl) ? 1.0 : 0.0;
        gu[i] = (Un[i] <= tu) ? 1.0 : 0.0;
    }
}

// src/sip.cpp:16-52 computes z[j+m-1] = sum_s g[j+s] h[s] by FFT; here a direct sum.
void sip_direct(const double* g, int n, const double* h, int m, double* z) {
    for (int j = 0; j + m <= n; j++) {
        double acc = 0.0;
        for (int s = 0; s < m; s++) acc += g[j + s] * h[s];
        z[j + m - 1] = acc;
    }
}

// src/fsmeld.cpp:84-106 / :912-921 for one window.
double trans1(int metric, double zl, double zu, int m) {
    if (metric == M_ERP || metric == M_DTW) return std::sqrt(std::max(0.0, zl + zu));
    if (metric == M_LCSS) return std::max(0.0, (zl + zu) / m);
    return zl + zu;
}

// include/fsmeld_internal.h:9-19
bool lbip_exceeds_cutoff(double bound, double cutoff, int metric) {
    constexpr double tol = 1e-9;
    if (!std::isfinite(bound) || !(bound > cutoff + tol)) return false;
    if (metric == M_DTW || metric == M_ERP) {
        const double lim = cutoff + tol;
        return bound > std::sqrt(lim * lim + tol);
    }
    return true;
}

// ------------------------------------------------------------------- GLB ---
// src/glb.cpp:8-31 helpers and :200-373 glb_elastic.
inline double edr_subcost(double a, double b, double eps) {
    return (std::abs(a - b) <= eps) ? 0.0 : 1.0;
}
// src/glb.cpp:11-14: the match reward within eps, else the mismatch cost
inline double swale_subcost(double a, double b, double eps) {
    return (std::abs(a - b) <= eps) ? kSwaleR : kSwaleP;
}
inline bool env_match(double v, double lo, double hi, double eps) {
    if (v < lo) return lo - v <= eps;
    if (v > hi) return v - hi <= eps;
    return true;
}

double glb(int metric, const double* Q, int lenQ, const double* T, int lenT, double arg,
           const double* LQ, const double* UQ, int w, double bsf) {
    if (lenQ < 2 || lenT < 2) return 0.0;
    double boundary = 0.0, t2q = 0.0, lcss_t = 0.0;
    switch (metric) {
        case M_DTW:
            boundary = sq(T[0], Q[0]) + sq(T[lenT - 1], Q[lenQ - 1]);
            for (int i = 1; i < lenT - 1; ++i) {
                if (T[i] < LQ[i]) t2q += sq(T[i], LQ[i]);
                else if (T[i] > UQ[i]) t2q += sq(T[i], UQ[i]);
                if (boundary + t2q > bsf * bsf) return kGlbAbandon;
            }
            break;
        case M_ERP:
            boundary = std::min({sq(Q[lenQ - 1], T[lenT - 1]), sq(Q[lenQ - 1], arg),
                                 sq(T[lenT - 1], arg)});
            for (int i = 1; i < lenT - 1; ++i) {
                if (T[i] > UQ[i]) {
                    const double a = sq(T[i], UQ[i]), b = sq(T[i], arg);
                    t2q += (a < b) ? a : b;
                } else if (T[i] < LQ[i]) {
                    const double a = sq(T[i], LQ[i]), b = sq(T[i], arg);
                    t2q += (a < b) ? a : b;
                }
                if (boundary + t2q > bsf * bsf) return kGlbAbandon;
            }
            break;
        case M_EDR: {
            const double e = edr_subcost(Q[lenQ - 1], T[lenT - 1], arg);
            boundary = (e < 1.0) ? e : 1.0;
            for (int i = 1; i < lenT - 1; ++i) {
                if (T[i] > UQ[i]) t2q += edr_subcost(T[i], UQ[i], arg);
                else if (T[i] < LQ[i]) t2q += edr_subcost(T[i], LQ[i], arg);
                if (boundary + t2q > bsf) return kGlbAbandon;
            }
            break;
        }
        case M_LCSS: {
            const double min_len = (lenQ < lenT) ? lenQ : lenT;
            for (int i = 0; i < lenT; ++i)
                if (env_match(T[i], LQ[i], UQ[i], arg)) lcss_t += 1.0;
            if (std::max(0.0, 1.0 - lcss_t / min_len) > bsf) return kGlbAbandon;
            break;
        }
        case M_MSM: {
            const int a = lenT - 1, b = lenQ - 1;
            if ((T[a - 1] >= T[a] && T[a] >= Q[b]) || (T[a - 1] <= T[a] && T[a] <= Q[b]) ||
                (Q[b - 1] <= Q[b] && Q[b] <= T[a]) || (Q[b - 1] >= Q[b] && Q[b] >= T[a])) {
                const double d = std::abs(Q[b] - T[a]);
                boundary = std::abs(Q[0] - T[0]) + ((d < arg) ? d : arg);
            } else {
                boundary = std::abs(Q[0] - T[0]) +
                           std::min({std::abs(Q[b] - T[a]), arg + std::abs(T[a] - T[a - 1]),
                                     arg + std::abs(Q[b] - Q[b - 1])});
            }
            for (int i = 1; i < lenT - 1; ++i) {
                if (T[i] > UQ[i]) {
                    const double d = std::abs(T[i] - UQ[i]);
                    t2q += (d < arg) ? d : arg;
                } else if (T[i] < LQ[i]) {
                    const double d = std::abs(T[i] - LQ[i]);
                    t2q += (d < arg) ? d : arg;
                }
                if (boundary + t2q > bsf) return kGlbAbandon;
            }
            break;
        }
        case M_SWALE: {  // src/glb.cpp:282-297
            const double e = swale_subcost(Q[lenQ - 1], T[lenT - 1], arg);
            boundary = (e < kSwaleP) ? e : kSwaleP;
            for (int i = 1; i < lenT - 1; ++i) {
                if (T[i] > UQ[i]) t2q += swale_subcost(T[i], UQ[i], arg);
                else if (T[i] < LQ[i]) t2q += swale_subcost(T[i], LQ[i], arg);
                if (boundary + t2q > bsf) return kGlbAbandon;
            }
            break;
        }
        default:
            return kGlbAbandon;
    }

    std::vector<double> LT(lenT), UT(lenT);
    lower_upper_lemire(T, lenT, w, LT.data(), UT.data());

    switch (metric) {
        case M_DTW: {
            double q2t = 0.0;
            for (int i = 1; i < lenQ - 1; ++i) {
                if (Q[i] < LT[i]) q2t += sq(Q[i], LT[i]);
                else if (Q[i] > UT[i]) q2t += sq(Q[i], UT[i]);
                if (boundary + q2t > bsf * bsf) return kGlbAbandon;
            }
            return std::sqrt(std::max(0.0, boundary + ((t2q > q2t) ? t2q : q2t)));
        }
        case M_ERP: {
            double q2t = 0.0;
            for (int i = 1; i < lenQ - 1; ++i) {
                if (Q[i] > UT[i]) {
                    const double a = sq(Q[i], UT[i]), b = sq(Q[i], arg);
                    q2t += (a < b) ? a : b;
                } else if (Q[i] < LT[i]) {
                    const double a = sq(Q[i], LT[i]), b = sq(Q[i], arg);
                    q2t += (a < b) ? a : b;
                }
                if (boundary + q2t > bsf * bsf) return kGlbAbandon;
            }
            return std::sqrt(std::max(0.0, boundary + ((t2q > q2t) ? t2q : q2t)));
        }
        case M_EDR: {
            double q2t = 0.0;
            for (int i = 1; i < lenQ - 1; ++i) {
                if (Q[i] > UT[i]) q2t += edr_subcost(Q[i], UT[i], arg);
                else if (Q[i] < LT[i]) q2t += edr_subcost(Q[i], LT[i], arg);
                if (boundary + q2t > bsf) return kGlbAbandon;
            }
            return boundary + ((t2q > q2t) ? t2q : q2t);
        }
        case M_LCSS: {
            double qr = 0.0;
            for (int i = 0; i < lenQ; ++i)
                if (env_match(Q[i], LT[i], UT[i], arg)) qr += 1.0;
            const double min_len = (lenQ < lenT) ? lenQ : lenT;
            return std::max(0.0, 1.0 - ((lcss_t < qr) ? lcss_t : qr) / min_len);
        }
        case M_MSM: {
            double q2t = 0.0;
            for (int i = 1; i < lenQ - 1; ++i) {
                if (Q[i] > UT[i]) {
                    const double d = std::abs(Q[i] - UT[i]);
                    q2t += (d < arg) ? d : arg;
                } else if (Q[i] < LT[i]) {
                    const double d = std::abs(Q[i] - LT[i]);
                    q2t += (d < arg) ? d : arg;
                }
                if (boundary + q2t > bsf) return kGlbAbandon;
            }
            return boundary + ((t2q > q2t) ? t2q : q2t);
        }
        case M_SWALE: {  // src/glb.cpp:356-369
            double q2t = 0.0;
            for (int i = 1; i < lenQ - 1; ++i) {
                if (Q[i] > UT[i]) q2t += swale_subcost(Q[i], UT[i], arg);
                else if (Q[i] < LT[i]) q2t += swale_subcost(Q[i], LT[i], arg);
                if (boundary + q2t > bsf) return kGlbAbandon;
            }
            return boundary + ((t2q > q2t) ? t2q : q2t);
        }
    }
    return kGlbAbandon;
}

// ------------------------------------------------------------------- ELD ---
// src/ElasticMeasures.cpp.  Band-only evaluation, identical per-cell
// expressions and operand order; out-of-band neighbours read the constant the
// reference's arrays hold there.
inline double min3(double a, double b, double c) { return std::min({a, b, c}); }

// Band rows: slot k = j - i + w of a row holds column j; a row carries one
// padding slot on each side (index -1 and W) that keeps the out-of-band value,
// so neighbours outside the band read it without a bounds test.  Buffers are
// per thread.
struct BandRows {
    std::vector<double> a, b;
    double* prev;
    double* cur;
    int W;
    void init(int W_, double pad) {
        W = W_;
        a.assign((size_t)W + 2, pad);
        b.assign((size_t)W + 2, pad);
        prev = a.data() + 1;
        cur = b.data() + 1;
    }
    void clear_cur(double v) { std::fill(cur, cur + W, v); }
    void swap() { std::swap(prev, cur); }
};
thread_local BandRows g_rows;

// src/ElasticMeasures.cpp:24-61 (A = query, B = candidate, n = m)
double eld_dtw(const double* A, const double* B, int m, int r) {
    BandRows& R = g_rows;
    R.init(2 * r + 1, kEldSentinel);
    for (int i = 0; i < m; i++) {
        R.clear_cur(kEldSentinel);
        double* C = R.cur;
        const double* P = R.prev;
        const int j0 = std::max(0, i - r), j1 = std::min(m - 1, i + r);
        for (int j = j0, k = j0 - i + r; j <= j1; j++, k++) {
            if (i == 0 && j == 0) {
                C[k] = sq(A[0], B[0]);
                continue;
            }
            // column j - 1 of this row, column j and j - 1 of the last row;
            // the unwritten slots (j < 0, i < 0, outside the band) hold 1e18
            const double y = C[k - 1];
            const double x = P[k + 1];
            const double z = P[k];
            C[k] = min3(x, y, z) + sq(A[i], B[j]);
        }
        R.swap();
    }
    const double f = R.prev[r];
    if (!std::isfinite(f)) return kEldSentinel;
    return std::sqrt(std::max(0.0, f));
}

// src/ElasticMeasures.cpp:63-93 (Q = query, A = candidate); out-of-band = 0.
double eld_lcss(const double* Q, const double* A, int n, int w, double eps) {
    const int m = n;
    BandRows& R = g_rows;
    R.init(2 * w + 1, 0.0);
    for (int i = 0; i < n; ++i) {
        const int wmin = std::max(0, i - w), wmax = std::min(m, i + w + 1);
        R.clear_cur(0.0);
        double* C = R.cur;
        const double* P = R.prev;
        for (int j = wmin, k = wmin - i + w; j < wmax; ++j, ++k) {
            const bool match = std::abs(Q[i] - A[j]) <= eps;
            double v;
            if (i + j == 0) {
                v = match ? 1 : 0;
            } else if (i == 0) {
                v = match ? 1 : ((j > wmin) ? C[k - 1] : 0);
            } else if (j == 0) {
                v = match ? 1 : P[k + 1];
            } else {
                const double v1 = P[k];
                const double v2 = P[k + 1];
                const double v3 = (j > wmin) ? C[k - 1] : 0;
                v = match ? std::max({v1 + 1, v2, v3}) : std::max({v1, v2, v3});
            }
            C[k] = v;
        }
        R.swap();
    }
    const double res = R.prev[(m - 1) - (n - 1) + w];
    return 1 - res / ((n < m) ? n : m);
}

// src/ElasticMeasures.cpp:95-130 (x = query, y = candidate); rows 0..m.
double eld_erp(const double* x, const double* y, int m, double g, int w) {
    BandRows& R = g_rows;
    R.init(2 * w + 1, kEldSentinel);  // band index k = j - i + w
    R.prev[w] = 0.0;                  // (0,0)
    for (int j = 1; j <= std::min(m, w); ++j) R.prev[j + w] = R.prev[j - 1 + w] + sq(y[j - 1], g);
    for (int i = 1; i <= m; ++i) {
        R.clear_cur(kEldSentinel);
        double* C = R.cur;
        const double* P = R.prev;
        if (i <= w) C[w - i] = P[w - i + 1] + sq(x[i - 1], g);
        const int minw = std::max(1, i - w), maxw = std::min(m, i + w);
        for (int j = minw, k = minw - i + w; j <= maxw; ++j, ++k) {
            const double match = P[k] + sq(x[i - 1], y[j - 1]);
            const double ins = C[k - 1] + sq(y[j - 1], g);
            const double del = P[k + 1] + sq(x[i - 1], g);
            C[k] = min3(match, ins, del);
        }
        R.swap();
    }
    const double res = R.prev[w];
    if (!std::isfinite(res)) return kEldSentinel;
    return std::sqrt(std::max(0.0, res));
}

// src/ElasticMeasures.cpp:132-165 (x = query, y = candidate).
double eld_edr(const double* x, const double* y, int m, double eps, int w) {
    BandRows& R = g_rows;
    R.init(2 * w + 1, kEldSentinel);
    R.prev[w] = 0.0;
    for (int j = 1; j <= std::min(m, w); ++j) R.prev[j + w] = static_cast<double>(j);
    for (int i = 1; i <= m; ++i) {
        R.clear_cur(kEldSentinel);
        double* C = R.cur;
        const double* P = R.prev;
        if (i <= w) C[w - i] = static_cast<double>(i);
        const int minw = std::max(1, i - w), maxw = std::min(m, i + w);
        for (int j = minw, k = minw - i + w; j <= maxw; ++j, ++k) {
            const double sub = (std::abs(x[i - 1] - y[j - 1]) <= eps) ? 0.0 : 1.0;
            C[k] = min3(P[k + 1] + 1.0, C[k - 1] + 1.0, P[k] + sub);
        }
        R.swap();
    }
    return R.prev[w];
}

// src/ElasticMeasures.cpp:167-198 (x = query, y = candidate).
inline double msm_cost(double np, double x, double y, double c) {
    if ((x <= np && np <= y) || (y <= np && np <= x)) return c;
    return c + std::min(std::abs(np - x), std::abs(np - y));
}
double eld_msm(const double* x, const double* y, int m, double c, int w) {
    BandRows& R = g_rows;
    R.init(2 * w + 1, kEldSentinel);
    // row 0
    R.clear_cur(kEldSentinel);
    R.cur[w] = std::abs(x[0] - y[0]);
    for (int j = 1; j < m && j <= w; ++j)
        R.cur[j + w] = R.cur[j - 1 + w] + msm_cost(y[j], x[0], y[j - 1], c);
    R.swap();
    for (int i = 1; i < m; ++i) {
        R.clear_cur(kEldSentinel);
        double* C = R.cur;
        const double* P = R.prev;
        if (i <= w) C[w - i] = P[w - i + 1] + msm_cost(x[i], x[i - 1], y[0], c);
        const int js = std::max(1, i - w), je = std::min(m, i + w + 1);
        for (int j = js, k = js - i + w; j < je; ++j, ++k) {
            const double c1 = P[k] + std::abs(x[i] - y[j]);
            const double c2 = P[k + 1] + msm_cost(x[i], x[i - 1], y[j], c);
            const double c3 = C[k - 1] + msm_cost(y[j], x[i], y[j - 1], c);
            C[k] = min3(c1, c2, c3);
        }
        R.swap();
    }
    return R.prev[w];
}

// src/ElasticMeasures.cpp:200-231 (x = query, y = candidate, p = 5, r = 1):
// the EDR grid with gap cost p and diagonal cost r within eps, else p.
double eld_swale(const double* x, const double* y, int m, double eps, int w) {
    if (eps < 0.0) return std::numeric_limits<double>::quiet_NaN();
    BandRows& R = g_rows;
    R.init(2 * w + 1, kEldSentinel);
    R.prev[w] = 0.0;
    for (int j = 1; j <= std::min(m, w); ++j) R.prev[j + w] = R.prev[j - 1 + w] + kSwaleP;
    for (int i = 1; i <= m; ++i) {
        R.clear_cur(kEldSentinel);
        double* C = R.cur;
        const double* P = R.prev;
        if (i <= w) C[w - i] = P[w - i + 1] + kSwaleP;
        const int minw = std::max(1, i - w), maxw = std::min(m, i + w);
        for (int j = minw, k = minw - i + w; j <= maxw; ++j, ++k) {
            const double mv = (std::abs(x[i - 1] - y[j - 1]) <= eps) ? kSwaleR : kSwaleP;
            C[k] = min3(P[k] + mv, P[k + 1] + kSwaleP, C[k - 1] + kSwaleP);
        }
        R.swap();
    }
    return R.prev[w];
}

// src/fsmeld.cpp:170-183
double eld(int metric, const double* q, const double* t, int m, int w, double arg) {
    switch (metric) {
        case M_SWALE: return eld_swale(q, t, m, arg, w);
        case M_MSM: return eld_msm(q, t, m, arg, w);
        case M_ERP: return eld_erp(q, t, m, arg, w);
        case M_LCSS: return eld_lcss(q, t, m, w, arg);
        case M_EDR: return eld_edr(q, t, m, arg, w);
        case M_DTW: return eld_dtw(q, t, m, w);
        default: return std::numeric_limits<double>::quiet_NaN();
    }
}

// ---------------------------------------------------------------- drivers ---
struct Geometry {
    int m_k = 0;
    int64_t k_m_l = 0, ts_n = 0, forsize = 0, tail = 0;
    int tail_first = 0;
};

// src/fsmeld.cpp:28-37 and :647-661
bool geometry(int64_t N, int m, Geometry* g) {
    long long sel = 1;
    while (sel < 8LL * m) sel *= 2;
    if (N < 2LL * m - 1) return false;
    if (N < sel) sel = N;
    if (sel > INT32_MAX) return false;
    g->m_k = static_cast<int>(sel);
    g->k_m_l = g->m_k - m + 1;
    g->ts_n = N - m + 1;
    g->forsize = g->ts_n / g->k_m_l;
    g->tail = g->ts_n % g->k_m_l;
    g->tail_first = static_cast<int>(g->k_m_l - g->tail);
    return true;
}

struct Query {
    int m = 0;
    std::vector<double> qn, qs, LQ, UQ;
};

// Matching mode, as the reference's set_normalization_mode
// (src/normalization_mode.cpp): 0 z-normalised, 1 raw (paper App. D).
std::atomic<int> g_raw{0};

// src/fsmeld.cpp:622-643 (read order), :721 (query envelope)
void prepare_query(const double* q_raw, int m, int w, Query* Q) {
    if (g_raw.load()) {  // raw matching: the query as it is
        Q->m = m;
        Q->qn.assign(q_raw, q_raw + m);
        Q->qs = Q->qn;
        std::sort(Q->qs.begin(), Q->qs.end());
        Q->LQ.resize(m);
        Q->UQ.resize(m);
        lower_upper_lemire(Q->qn.data(), m, w, Q->LQ.data(), Q->UQ.data());
        return;
    }
    double ex = 0, ex2 = 0;
    for (int i = 0; i < m; i++) {
        ex += q_raw[i];
        ex2 += q_raw[i] * q_raw[i];
    }
    const double mean = ex / m;
    double sd = ex2 / m;
    sd = std::sqrt(sd - mean * mean);
    Q->m = m;
    Q->qn.resize(m);
    for (int i = 0; i < m; i++) Q->qn[i] = (q_raw[i] - mean) / sd;
    Q->qs = Q->qn;
    std::sort(Q->qs.begin(), Q->qs.end());
    Q->LQ.resize(m);
    Q->UQ.resize(m);
    lower_upper_lemire(Q->qn.data(), m, w, Q->LQ.data(), Q->UQ.data());
}

// Per-thread scratch of one macro-segment (src/fsmeld.cpp:690-719).
struct SegmentWork {
    std::vector<double> mu, si, mumin, mumax, simin, simax, U, L, Un, Ln, gl, gu, hl, hu, zl,
        zu, z, tmp;
    std::vector<uint8_t> mask;
    void init(int m_k, int64_t kml, int m) {
        mu.assign(kml, 0); si.assign(kml, 0);
        mumin.assign(m_k, 0); mumax.assign(m_k, 0); simin.assign(m_k, 0); simax.assign(m_k, 0);
        U.assign(m_k, 0); L.assign(m_k, 0); Un.assign(m_k, 0); Ln.assign(m_k, 0);
        gl.assign(m_k, 0); gu.assign(m_k, 0); hl.assign(m, 0); hu.assign(m, 0);
        zl.assign(m_k, 0); zu.assign(m_k, 0); z.assign(kml, 0); tmp.assign(m, 0);
        mask.assign(m_k, 0);
    }
};

bool excluded(int64_t g, int64_t b, int64_t e) { return b >= 0 && e > b && g >= b && g < e; }

// Stats + mask for one segment: src/fsmeld.cpp:753-757 (full), :861-866 (tail).
// Raw matching (paper App. D) estimates no mean or std: every window is
// taken as it is (mean 0, std 1 leave it unchanged under the normalisation
// and its envelope bounds) and only a non-finite sample invalidates it.
void segment_stats(const double* t, const Geometry& G, int m, int64_t block_start,
                   int first, int count, int64_t xb, int64_t xe, SegmentWork& W) {
    std::fill(W.mask.begin(), W.mask.end(), 0);
    if (g_raw.load()) {
        std::fill(W.mu.begin(), W.mu.end(), 0.0);
        std::fill(W.si.begin(), W.si.end(), 1.0);
        int bad = 0;
        for (int k = 0; k < m; k++) bad += !std::isfinite(t[k]);
        for (int64_t j = 0; j < G.k_m_l; j++) {
            if (j) bad += (int)!std::isfinite(t[j + m - 1]) - (int)!std::isfinite(t[j - 1]);
            W.mask[j] = bad > 0;
        }
    } else {
        mvmean(t, G.m_k, m, W.mu.data(), W.si.data());
        invalid_mask(t, W.si.data(), G.m_k, m, W.mask.data());
    }
    for (int j = first; j < first + count; j++)
        if (excluded(block_start + (j - first), xb, xe)) W.mask[j] = 1;
}

// LB_IP for one segment: src/fsmeld.cpp:756-808.
void segment_lbip(const double* t, const Geometry& G, const Query& Q, int metric, int w,
                  double arg, SegmentWork& W, double* tu_out, double* tl_out) {
    const int m = Q.m, mk = G.m_k;
    min_max_window(W.mu.data(), mk, m, W.mumin.data(), W.mumax.data());
    min_max_window(W.si.data(), mk, m, W.simin.data(), W.simax.data());
    lower_upper_lemire(t, mk, w, W.L.data(), W.U.data());
    normalized_arrays(W.U.data(), W.L.data(), W.Un.data(), W.Ln.data(), mk, W.mumax.data(),
                      W.mumin.data(), W.simax.data(), W.simin.data());
    double tu, tl;
    thresholds(metric, W.Un.data(), W.Ln.data(), mk, Q.qn.data(), Q.qs.data(), m, arg, &tu,
               &tl);
    q_to_h(metric, Q.qn.data(), m, tl, tu, arg, W.hl.data(), W.hu.data());
    s_to_g(W.Ln.data(), W.Un.data(), mk, tl, tu, W.gl.data(), W.gu.data());
    sip_direct(W.gl.data(), mk, W.hl.data(), m, W.zl.data());
    sip_direct(W.gu.data(), mk, W.hu.data(), m, W.zu.data());
    for (int64_t j = 0; j < G.k_m_l; j++)
        W.z[j] = trans1(metric, W.zl[j + m - 1], W.zu[j + m - 1], m);
    if (tu_out) *tu_out = tu;
    if (tl_out) *tl_out = tl;
}

void normalize_window(const double* t, int j, int m, double mu, double si, double* out) {
    for (int s = 0; s < m; s++) out[s] = (t[s + j] - mu) / si;  // src/fsmeld.cpp:822-826
}

// A window that passed the filters, to be verified: its global index, its
// start j inside segment seg and the statistics that normalise it.
struct Survivor {
    int64_t index, seg;
    int j;
    double mu, si;
};

struct SegResult {
    uint64_t valid = 0, invalid = 0, k2 = 0, k3 = 0;
    std::vector<Survivor> surv;
};

const double* segment_base(const double* T, int64_t N, const Geometry& G, int64_t seg) {
    return seg == G.forsize ? T + (N - G.m_k) : T + seg * G.k_m_l;
}

// One search segment's filters: src/fsmeld.cpp:812-851 (full) / :926-968
// (tail).  The reference buffers the GLB survivors of a block and verifies
// them afterwards; here they are verified by the caller, in parallel over all
// survivors of the target (the verdict of each is independent of the rest).
void search_segment(const double* T, int64_t N, const Geometry& G, const Query& Q,
                    int metric, int w, double arg, double xi, int64_t xb, int64_t xe,
                    int64_t seg, SegmentWork& W, SegResult& R) {
    const int m = Q.m;
    const bool is_tail = seg == G.forsize;
    const double* t = segment_base(T, N, G, seg);
    const int first = is_tail ? G.tail_first : 0;
    const int count = is_tail ? static_cast<int>(G.tail) : static_cast<int>(G.k_m_l);
    const int64_t base = seg * G.k_m_l;
    segment_stats(t, G, m, base, first, count, xb, xe, W);
    segment_lbip(t, G, Q, metric, w, arg, W, nullptr, nullptr);
    std::vector<double> tt(m);
    for (int j = first; j < first + count; j++) {
        if (W.mask[j]) {
            R.invalid++;
            continue;
        }
        R.valid++;
        if (lbip_exceeds_cutoff(W.z[j], xi, metric)) continue;
        R.k2++;
        normalize_window(t, j, m, W.mu[j], W.si[j], tt.data());
        const double gd = glb(metric, Q.qn.data(), m, tt.data(), m, arg, Q.LQ.data(),
                              Q.UQ.data(), w, xi);
        if (gd <= xi) {
            R.k3++;
            R.surv.push_back({base + (j - first), seg, j, W.mu[j], W.si[j]});
        }
    }
}

// Exact distances of the survivors, in parallel; dist[i] belongs to v[i].
void verify_all(const double* T, int64_t N, const Geometry& G, const Query& Q, int metric,
                int w, double arg, const std::vector<Survivor>& v, int threads,
                std::vector<double>& dist) {
    const int m = Q.m;
    dist.assign(v.size(), 0.0);
#ifdef _OPENMP
    const int nt = threads > 0 ? threads : 1;
#pragma omp parallel num_threads(nt)
#endif
    {
        std::vector<double> tt(m);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < (int64_t)v.size(); i++) {
            normalize_window(segment_base(T, N, G, v[i].seg), v[i].j, m, v[i].mu, v[i].si,
                             tt.data());
            dist[i] = eld(metric, Q.qn.data(), tt.data(), m, w, arg);
        }
    }
}

int check_args(int64_t N, int m, int metric, int w) {
    if (m < 2 || w < 0 || w >= m) return -1;
    if (metric < M_DTW || metric > M_SWALE) return -1;
    if (N < 2LL * m - 1) return -1;
    return 0;
}

void fill_stats(const Geometry& G, orc_stats* st) {
    st->ts_n = G.ts_n;
    st->k_m_l = G.k_m_l;
    st->forsize = G.forsize;
    st->tail = G.tail;
    st->m_k = G.m_k;
}

}  // namespace

// ======================================================================= ABI
extern "C" {

void orc_mvmean(const double* a, int len_a, int l, double* miu, double* si) {
    mvmean(a, len_a, l, miu, si);
}
void orc_invalid_mask(const double* t, const double* si, int seg, int m, uint8_t* mask) {
    invalid_mask(t, si, seg, m, mask);
}
void orc_min_max_window(const double* a, int n, int m, double* mn, double* mx) {
    min_max_window(a, n, m, mn, mx);
}
void orc_lower_upper_lemire(const double* a, int n, int r, double* l, double* u) {
    lower_upper_lemire(a, n, r, l, u);
}
void orc_normalized_arrays(const double* U, const double* L, double* Un, double* Ln, int len,
                           const double* max_mean, const double* min_mean,
                           const double* max_std, const double* min_std) {
    normalized_arrays(U, L, Un, Ln, len, max_mean, min_mean, max_std, min_std);
}
void orc_thresholds(int metric, const double* Un, const double* Ln, int l, const double* qn,
                    const double* qs, int m, double arg, double* tu, double* tl) {
    thresholds(metric, Un, Ln, l, qn, qs, m, arg, tu, tl);
}
void orc_q_to_h(int metric, const double* q, int m, double tl, double tu, double arg,
                double* hl, double* hu) {
    q_to_h(metric, q, m, tl, tu, arg, hl, hu);
}
void orc_s_to_g(const double* Ln, const double* Un, int n, double tl, double tu, double* gl,
                double* gu) {
    s_to_g(Ln, Un, n, tl, tu, gl, gu);
}
void orc_sip_direct(const double* g, int n, const double* h, int m, double* z) {
    sip_direct(g, n, h, m, z);
}
double orc_trans(int metric, double zl, double zu, int m) { return trans1(metric, zl, zu, m); }
void orc_set_normalization_mode(int raw) { g_raw.store(raw ? 1 : 0); }
int orc_lbip_exceeds_cutoff(double bound, double cutoff, int metric) {
    return lbip_exceeds_cutoff(bound, cutoff, metric) ? 1 : 0;
}
double orc_glb(int metric, const double* Q, const double* T, int m, double arg,
               const double* LQ, const double* UQ, int w, double bsf) {
    return glb(metric, Q, m, T, m, arg, LQ, UQ, w, bsf);
}
double orc_eld(int metric, const double* q, const double* t, int m, int w, double arg) {
    return eld(metric, q, t, m, w, arg);
}
void orc_normalize_query(const double* q_raw, int m, double* qn) {
    Query Q;
    prepare_query(q_raw, m, 0, &Q);
    std::memcpy(qn, Q.qn.data(), sizeof(double) * m);
}

// src/fsmeld.cpp:570-1074 (search).  Segments [seg_begin, seg_end) of the
// S = forsize + [tail>0] segments are evaluated (seg_end < 0: all); the tail
// segment has index forsize.  threads > 1 splits segments over OpenMP threads;
// results are concatenated in segment order, which equals the sequential order.
int orc_search(const double* T, int64_t N, const double* q_raw, int m, double xi, int metric,
               int w, double arg, int64_t xb, int64_t xe, int64_t seg_begin, int64_t seg_end,
               int threads, orc_stats* st, int64_t* out, int64_t cap, int64_t* n_out) {
    if (check_args(N, m, metric, w)) return -1;
    Geometry G;
    if (!geometry(N, m, &G)) return -1;
    Query Q;
    prepare_query(q_raw, m, w, &Q);
    const int64_t S = G.forsize + (G.tail > 0 ? 1 : 0);
    if (seg_end < 0 || seg_end > S) seg_end = S;
    if (seg_begin < 0) seg_begin = 0;
    const int64_t nseg = seg_end > seg_begin ? seg_end - seg_begin : 0;
    std::vector<SegResult> res(nseg);
    const int nt = threads > 0 ? threads : 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        SegmentWork W;
        W.init(G.m_k, G.k_m_l, m);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int64_t i = 0; i < nseg; i++)
            search_segment(T, N, G, Q, metric, w, arg, xi, xb, xe, seg_begin + i, W, res[i]);
    }
    std::memset(st, 0, sizeof(*st));
    fill_stats(G, st);
    std::vector<Survivor> all;
    for (auto& r : res) {
        st->valid += r.valid;
        st->invalid += r.invalid;
        st->after_lbip += r.k2;
        st->after_glb += r.k3;
        all.insert(all.end(), r.surv.begin(), r.surv.end());
        std::vector<Survivor>().swap(r.surv);
    }
    std::vector<double> dist;
    verify_all(T, N, G, Q, metric, w, arg, all, nt, dist);
    int64_t n = 0;
    for (size_t i = 0; i < all.size(); i++) {  // survivors are in index order
        if (!(std::isfinite(dist[i]) && dist[i] <= xi)) continue;
        if (n < cap) out[n] = all[i].index;
        n++;
    }
    st->exact_evals = st->after_glb;
    st->matches = static_cast<uint64_t>(n);
    *n_out = n;
    return 0;
}

// src/fsmeld.cpp:1075-1344 (exact, top_k = 0 as the CLI calls it).
int orc_exact(const double* T, int64_t N, const double* q_raw, int m, double xi, int metric,
              int w, double arg, int64_t xb, int64_t xe, int threads, orc_stats* st,
              int64_t* out, double* dists, int64_t cap, int64_t* n_out) {
    if (check_args(N, m, metric, w)) return -1;
    Geometry G;
    if (!geometry(N, m, &G)) return -1;
    Query Q;
    prepare_query(q_raw, m, w, &Q);
    const int64_t S = G.forsize + (G.tail > 0 ? 1 : 0);
    struct R { uint64_t valid = 0, invalid = 0; std::vector<Survivor> win; };
    std::vector<R> res(S);
    const int nt = threads > 0 ? threads : 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        SegmentWork W;
        W.init(G.m_k, G.k_m_l, m);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int64_t seg = 0; seg < S; seg++) {
            const bool is_tail = seg == G.forsize;
            const double* t = segment_base(T, N, G, seg);
            const int first = is_tail ? G.tail_first : 0;
            const int count = is_tail ? static_cast<int>(G.tail) : static_cast<int>(G.k_m_l);
            const int64_t base = seg * G.k_m_l;
            segment_stats(t, G, m, base, first, count, xb, xe, W);
            for (int j = first; j < first + count; j++) {
                if (W.mask[j]) {
                    res[seg].invalid++;
                    continue;
                }
                res[seg].valid++;
                res[seg].win.push_back({base + (j - first), seg, j, W.mu[j], W.si[j]});
            }
        }
    }
    std::memset(st, 0, sizeof(*st));
    fill_stats(G, st);
    std::vector<Survivor> all;
    for (auto& r : res) {
        st->valid += r.valid;
        st->invalid += r.invalid;
        all.insert(all.end(), r.win.begin(), r.win.end());
        std::vector<Survivor>().swap(r.win);
    }
    std::vector<double> dist;
    verify_all(T, N, G, Q, metric, w, arg, all, nt, dist);
    int64_t n = 0;
    for (size_t i = 0; i < all.size(); i++) {
        if (!(std::isfinite(dist[i]) && dist[i] <= xi)) continue;
        if (n < cap) {
            out[n] = all[i].index;
            if (dists) dists[n] = dist[i];
        }
        n++;
    }
    st->exact_evals = st->valid;
    st->matches = static_cast<uint64_t>(n);
    *n_out = n;
    return 0;
}

}  // extern "C"

namespace {

// src/fsmeld.cpp:185-569 (cutoff, filter_by_initial_epsilon = false) over
// segments [s0, s1): the reference's max-heap of (distance, index) pairs with
// strict-< replacement, scanned in index order, LB_IP at DBL_MAX until the
// heap is full and GLB only after that.  The heap it leaves holds the k
// (distance, index)-smallest windows of the range.
struct CutoffRange {
    std::vector<std::pair<double, long long>> best;  // ascending (distance, index)
    uint64_t kk1 = 0, lbp = 0, glbp = 0, ex = 0;
};

void cutoff_range(const double* T, int64_t N, const Geometry& G, const Query& Q, int m, int k,
                  int metric, int w, double arg, int64_t xb, int64_t xe, int64_t s0, int64_t s1,
                  CutoffRange& R) {
    SegmentWork W;
    W.init(G.m_k, G.k_m_l, m);
    std::priority_queue<std::pair<double, long long>> heap;
    double threshold = std::numeric_limits<double>::max();  // src/main.cpp:95
    std::vector<double> tt(m);
    for (int64_t seg = s0; seg < s1; seg++) {
        const bool is_tail = seg == G.forsize;
        const double* t = is_tail ? T + (N - G.m_k) : T + seg * G.k_m_l;
        const int first = is_tail ? G.tail_first : 0;
        const int count = is_tail ? static_cast<int>(G.tail) : static_cast<int>(G.k_m_l);
        const int64_t base = seg * G.k_m_l;
        segment_stats(t, G, m, base, first, count, xb, xe, W);
        segment_lbip(t, G, Q, metric, w, arg, W, nullptr, nullptr);
        for (int j = first; j < first + count; j++) {
            if (W.mask[j]) continue;
            R.kk1++;
            const bool full = heap.size() >= static_cast<size_t>(k);
            const double active = full ? threshold : DBL_MAX;
            if (lbip_exceeds_cutoff(W.z[j], active, metric)) {
                R.lbp++;
                continue;
            }
            normalize_window(t, j, m, W.mu[j], W.si[j], tt.data());
            if (full) {
                const double gd = glb(metric, Q.qn.data(), m, tt.data(), m, arg, Q.LQ.data(),
                                      Q.UQ.data(), w, threshold);
                if (gd > threshold) {
                    R.glbp++;
                    continue;
                }
            }
            const double d = eld(metric, Q.qn.data(), tt.data(), m, w, arg);
            R.ex++;
            if (!std::isfinite(d)) continue;
            const long long idx = base + (j - first);
            if (heap.size() < static_cast<size_t>(k)) {
                heap.push({d, idx});
                if (heap.size() == static_cast<size_t>(k)) threshold = heap.top().first;
            } else if (d < heap.top().first) {
                heap.pop();
                heap.push({d, idx});
                threshold = heap.top().first;
            }
        }
    }
    R.best.clear();
    while (!heap.empty()) {
        R.best.push_back(heap.top());
        heap.pop();
    }
    std::reverse(R.best.begin(), R.best.end());
}

}  // namespace

extern "C" {

// threads <= 1: the reference's sequential scan over all segments.
// threads > 1: contiguous segment ranges scanned as above in parallel; the k
// (distance, index)-smallest of the ranges' heaps are the sequential heap's
// contents (each range's heap holds its k smallest), so ranks and threshold
// are the same; only the pruning counters describe the split schedule.
int orc_cutoff(const double* T, int64_t N, const double* q_raw, int m, int k, int metric,
               int w, double arg, int64_t xb, int64_t xe, int threads, orc_stats* st,
               int64_t* ranks, double* rank_dists) {
    if (check_args(N, m, metric, w) || k < 1) return -1;
    Geometry G;
    if (!geometry(N, m, &G)) return -1;
    Query Q;
    prepare_query(q_raw, m, w, &Q);
    const int64_t S = G.forsize + (G.tail > 0 ? 1 : 0);
    const int nt = threads > 1 ? threads : 1;
    const int64_t parts = nt > 1 ? std::min<int64_t>(S, 4 * (int64_t)nt) : 1;
    std::vector<CutoffRange> res(parts);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
    for (int64_t p = 0; p < parts; p++)
        cutoff_range(T, N, G, Q, m, k, metric, w, arg, xb, xe, S * p / parts,
                     S * (p + 1) / parts, res[p]);
    std::vector<std::pair<double, long long>> v;
    uint64_t kk1 = 0, lbp = 0, glbp = 0, ex = 0;
    for (auto& r : res) {
        v.insert(v.end(), r.best.begin(), r.best.end());
        kk1 += r.kk1;
        lbp += r.lbp;
        glbp += r.glbp;
        ex += r.ex;
    }
    std::sort(v.begin(), v.end());
    if (v.size() > static_cast<size_t>(k)) v.resize(k);
    std::memset(st, 0, sizeof(*st));
    fill_stats(G, st);
    st->valid = kk1;
    st->after_lbip = kk1 - lbp;
    st->after_glb = kk1 - lbp - glbp;
    st->exact_evals = ex;
    if (v.size() < static_cast<size_t>(k)) return -4;  // src/fsmeld.cpp:519-520
    st->threshold = v.back().first;
    for (size_t i = 0; i < v.size(); i++) {
        ranks[i] = v[i].second;
        if (rank_dists) rank_dists[i] = v[i].first;
    }
    st->matches = v.size();
    return 0;
}

// src/fsmeld.cpp:1097-1107, :1273-1290 (FSMELD_exact with top_k = k): the
// k smallest finite distances in a max-heap of values, the threshold its top,
// and every window whose distance is at or below it, ascending by index.
int orc_exact_topk(const double* T, int64_t N, const double* q_raw, int m, int k, int metric,
                   int w, double arg, int64_t xb, int64_t xe, int threads, orc_stats* st,
                   int64_t* out, double* dists, int64_t cap, int64_t* n_out, double* kth) {
    if (k < 1) return -1;
    const int64_t ts_n = N - m + 1;
    if (ts_n < 1) return -1;
    std::vector<int64_t> idx((size_t)ts_n);
    std::vector<double> d((size_t)ts_n);
    int64_t n = 0;
    const int rc = orc_exact(T, N, q_raw, m, std::numeric_limits<double>::infinity(), metric, w,
                             arg, xb, xe, threads, st, idx.data(), d.data(), ts_n, &n);
    if (rc) return rc;
    std::priority_queue<double> heap;
    for (int64_t i = 0; i < n; i++) {  // every finite distance, in index order
        if (heap.size() < static_cast<size_t>(k)) heap.push(d[i]);
        else if (d[i] < heap.top()) {
            heap.pop();
            heap.push(d[i]);
        }
    }
    if (heap.size() < static_cast<size_t>(k)) return -4;
    const double theta = heap.top();
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; i++) {
        if (d[i] > theta) continue;
        if (cnt < cap) {
            out[cnt] = idx[i];
            if (dists) dists[cnt] = d[i];
        }
        cnt++;
    }
    st->threshold = theta;
    st->matches = static_cast<uint64_t>(cnt);
    *n_out = cnt;
    *kth = theta;
    return 0;
}

int orc_window_stats(const double* T, int64_t N, int m, int64_t xb, int64_t xe, double* mu,
                     double* sigma, uint8_t* invalid) {
    if (m < 2 || N < 2LL * m - 1) return -1;
    Geometry G;
    if (!geometry(N, m, &G)) return -1;
    SegmentWork W;
    W.init(G.m_k, G.k_m_l, m);
    const int64_t S = G.forsize + (G.tail > 0 ? 1 : 0);
    for (int64_t seg = 0; seg < S; seg++) {
        const bool is_tail = seg == G.forsize;
        const double* t = is_tail ? T + (N - G.m_k) : T + seg * G.k_m_l;
        const int first = is_tail ? G.tail_first : 0;
        const int count = is_tail ? static_cast<int>(G.tail) : static_cast<int>(G.k_m_l);
        const int64_t base = seg * G.k_m_l;
        segment_stats(t, G, m, base, first, count, xb, xe, W);
        for (int j = first; j < first + count; j++) {
            const int64_t g = base + (j - first);
            mu[g] = W.mu[j];
            sigma[g] = W.si[j];
            invalid[g] = W.mask[j];
        }
    }
    return 0;
}

int orc_lbip_all(const double* T, int64_t N, const double* q_raw, int m, int metric, int w,
                 double arg, double* lb, double* t_u, double* t_l) {
    if (check_args(N, m, metric, w)) return -1;
    Geometry G;
    if (!geometry(N, m, &G)) return -1;
    Query Q;
    prepare_query(q_raw, m, w, &Q);
    SegmentWork W;
    W.init(G.m_k, G.k_m_l, m);
    const int64_t S = G.forsize + (G.tail > 0 ? 1 : 0);
    for (int64_t seg = 0; seg < S; seg++) {
        const bool is_tail = seg == G.forsize;
        const double* t = is_tail ? T + (N - G.m_k) : T + seg * G.k_m_l;
        const int first = is_tail ? G.tail_first : 0;
        const int count = is_tail ? static_cast<int>(G.tail) : static_cast<int>(G.k_m_l);
        const int64_t base = seg * G.k_m_l;
        segment_stats(t, G, m, base, first, count, -1, -1, W);
        double tu, tl;
        segment_lbip(t, G, Q, metric, w, arg, W, &tu, &tl);
        if (t_u) t_u[seg] = tu;
        if (t_l) t_l[seg] = tl;
        for (int j = first; j < first + count; j++)
            lb[base + (j - first)] =
                W.mask[j] ? std::numeric_limits<double>::quiet_NaN() : W.z[j];
    }
    return 0;
}

}  // extern "C"
