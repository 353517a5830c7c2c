// k_verify.hip -- candidate compaction, GLB filter and exact elastic
// distances (ELD) on gfx950.
//
// GLB and ELD run one candidate per lane and evaluate every expression in the
// reference's operand order (src/glb.cpp:200-373, src/ElasticMeasures.cpp),
// so their values are bit-identical to it.  The banded DPs keep the band
// (2w+1 cells) and the normalised candidate samples it touches in registers;
// cells are updated in place in ascending band order, which reads the
// previous row's "up" / "diagonal" values and the current row's "left" value
// exactly as the reference's row arrays do.
#include "engine_internal.h"

namespace fsm {

// include/fsmeld_internal.h:9-19
__device__ inline bool lbip_exceeds_cutoff(double bound, double cutoff, int metric) {
    const double tol = 1e-9;
    if (!__builtin_isfinite(bound) || !(bound > cutoff + tol)) return false;
    if (metric == DTW || metric == ERP) {
        const double lim = cutoff + tol;
        return bound > sqrt(lim * lim + tol);
    }
    return true;
}

// Wavefront-aggregated append of `keep` lanes to out[]; returns nothing.
__device__ inline void wave_append(bool keep, int64_t value, int64_t* out,
                                   unsigned long long* count) {
    const uint64_t mask = __ballot(keep);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(mask);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned long long)__builtin_popcountll(mask));
    base = __shfl(base, leader, 64);
    if (keep) {
        const uint64_t below = lane ? (mask & ((1ull << lane) - 1)) : 0ull;
        out[base + __builtin_popcountll(below)] = value;
    }
}

// ------------------------------------------------------------- compaction ---
__global__ void __launch_bounds__(256) k_compact(Geometry G, QueryParams P,
                                                 const uint8_t* __restrict__ INV,
                                                 const double* __restrict__ LB, int use_lb,
                                                 int64_t* __restrict__ out,
                                                 unsigned long long* counters) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long nvalid = 0;
    // every lane of a wave runs the same number of iterations (ballot safety)
    const int64_t nwave_iters = (G.ts_n + stride - 1) / stride;
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t it = 0; it < nwave_iters; it++, g += stride) {
        bool keep = false;
        if (g < G.ts_n) {
            const bool valid = !INV[stat_index(G, g)] && !is_excluded(P, g);
            nvalid += valid;
            keep = valid && (!use_lb || !lbip_exceeds_cutoff(LB[g], P.cutoff, P.metric));
        }
        wave_append(keep, g, out, &counters[0]);
    }
    for (int o = 32; o > 0; o >>= 1) nvalid += __shfl_xor(nvalid, o, 64);
    if ((threadIdx.x & 63) == 0 && nvalid) atomicAdd(&counters[1], nvalid);
}

hipError_t launch_compact(hipStream_t st, const Geometry& G, const QueryParams& P,
                          const uint8_t* INV, const double* LB, bool use_lb, int64_t* out,
                          unsigned long long* counters) {
    int64_t blocks = (G.ts_n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_compact, dim3((unsigned)blocks), dim3(256), 0, st, G, P, INV, LB,
                       use_lb ? 1 : 0, out, counters);
    return hipGetLastError();
}

// -------------------------------------------------------------------- GLB ---
__device__ inline double sqd(double x, double y) { return (x - y) * (x - y); }  // include/def.h:5
__device__ inline double mn(double x, double y) { return (x < y) ? x : y; }    // include/def.h:9
__device__ inline double mx(double x, double y) { return (x > y) ? x : y; }    // include/def.h:10
__device__ inline double edr_subcost(double a, double b, double eps) {          // src/glb.cpp:8-10
    return (fabs(a - b) <= eps) ? 0.0 : 1.0;
}
__device__ inline bool env_match(double v, double lo, double hi, double eps) {  // src/glb.cpp:17-22
    if (v < lo) return lo - v <= eps;
    if (v > hi) return v - hi <= eps;
    return true;
}
// std::min / std::max of the reference: min(a,b) = (b < a) ? b : a
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }
__device__ inline double smin3(double a, double b, double c) { return smin(smin(a, b), c); }

// Envelope of the normalised candidate, radius w, clipped to the window
// (lower_upper_lemire on t_temp, src/glb.cpp:302-304).  Normalisation
// x -> (x - mu) / sd is monotone for sd > 0, so min / max are taken on raw
// samples and normalised once; interior positions come from the
// target-wide envelope.
struct CandEnvelope {
    const double* t;
    const double* Lr;
    const double* Ur;
    int m, w;
    int pre_r;      // prefix covers t[0 .. pre_r]
    double pre_lo, pre_hi;
    __device__ void init(const double* t_, const double* Lr_, const double* Ur_, int m_, int w_) {
        t = t_; Lr = Lr_; Ur = Ur_; m = m_; w = w_;
        pre_r = 0;
        pre_lo = t[0];
        pre_hi = t[0];
    }
    // raw min / max for position i; calls must use non-decreasing i.
    __device__ void at(int i, double* lo, double* hi) {
        if (i < w) {
            const int r = (i + w < m - 1) ? i + w : m - 1;
            while (pre_r < r) {
                pre_r++;
                pre_lo = nmin(pre_lo, t[pre_r]);
                pre_hi = nmax(pre_hi, t[pre_r]);
            }
            *lo = pre_lo;
            *hi = pre_hi;
        } else if (i > m - 1 - w) {
            double a = t[i - w], b = t[i - w];
            for (int x = i - w + 1; x < m; x++) {
                a = nmin(a, t[x]);
                b = nmax(b, t[x]);
            }
            *lo = a;
            *hi = b;
        } else {
            *lo = Lr[i];
            *hi = Ur[i];
        }
    }
};

template <int METRIC>
__device__ double glb_one(const double* __restrict__ t, double mu, double sd,
                          const double* __restrict__ Q, const double* __restrict__ LQ,
                          const double* __restrict__ UQ, const double* Lr, const double* Ur,
                          int m, int w, double arg, double bsf) {
    if (m < 2) return 0.0;
    const int n = m;
    auto T = [&](int i) { return (t[i] - mu) / sd; };  // src/fsmeld.cpp:822-826
    double boundary = 0.0, t2q = 0.0, lcss_t = 0.0;
    const double Tl = T(n - 1);
    if (METRIC == DTW) {
        boundary = sqd(T(0), Q[0]) + sqd(Tl, Q[n - 1]);
        for (int i = 1; i < n - 1; ++i) {
            const double v = T(i);
            if (v < LQ[i]) t2q += sqd(v, LQ[i]);
            else if (v > UQ[i]) t2q += sqd(v, UQ[i]);
            if (boundary + t2q > bsf * bsf) return kGlbAbandon;
        }
    } else if (METRIC == ERP) {
        boundary = smin3(sqd(Q[n - 1], Tl), sqd(Q[n - 1], arg), sqd(Tl, arg));
        for (int i = 1; i < n - 1; ++i) {
            const double v = T(i);
            if (v > UQ[i]) t2q += mn(sqd(v, UQ[i]), sqd(v, arg));
            else if (v < LQ[i]) t2q += mn(sqd(v, LQ[i]), sqd(v, arg));
            if (boundary + t2q > bsf * bsf) return kGlbAbandon;
        }
    } else if (METRIC == EDR) {
        boundary = mn(edr_subcost(Q[n - 1], Tl, arg), 1.0);
        for (int i = 1; i < n - 1; ++i) {
            const double v = T(i);
            if (v > UQ[i]) t2q += edr_subcost(v, UQ[i], arg);
            else if (v < LQ[i]) t2q += edr_subcost(v, LQ[i], arg);
            if (boundary + t2q > bsf) return kGlbAbandon;
        }
    } else if (METRIC == LCSS) {
        const double min_len = (double)n;
        for (int i = 0; i < n; ++i)
            if (env_match(T(i), LQ[i], UQ[i], arg)) lcss_t += 1.0;
        if (smax(0.0, 1.0 - lcss_t / min_len) > bsf) return kGlbAbandon;
    } else {  // MSM
        const double T2 = T(n - 2), T0 = T(0);
        const int b = n - 1;
        if ((T2 >= Tl && Tl >= Q[b]) || (T2 <= Tl && Tl <= Q[b]) ||
            (Q[b - 1] <= Q[b] && Q[b] <= Tl) || (Q[b - 1] >= Q[b] && Q[b] >= Tl)) {
            boundary = fabs(Q[0] - T0) + mn(fabs(Q[b] - Tl), arg);
        } else {
            boundary = fabs(Q[0] - T0) + smin3(fabs(Q[b] - Tl), arg + fabs(Tl - T2),
                                               arg + fabs(Q[b] - Q[b - 1]));
        }
        for (int i = 1; i < n - 1; ++i) {
            const double v = T(i);
            if (v > UQ[i]) t2q += mn(fabs(v - UQ[i]), arg);
            else if (v < LQ[i]) t2q += mn(fabs(v - LQ[i]), arg);
            if (boundary + t2q > bsf) return kGlbAbandon;
        }
    }

    CandEnvelope E;
    E.init(t, Lr, Ur, m, w);
    auto env = [&](int i, double* lo, double* hi) {
        double a, b;
        E.at(i, &a, &b);
        *lo = (a - mu) / sd;
        *hi = (b - mu) / sd;
    };
    if (METRIC == DTW) {
        double q2t = 0.0;
        for (int i = 1; i < n - 1; ++i) {
            double LT, UT;
            env(i, &LT, &UT);
            if (Q[i] < LT) q2t += sqd(Q[i], LT);
            else if (Q[i] > UT) q2t += sqd(Q[i], UT);
            if (boundary + q2t > bsf * bsf) return kGlbAbandon;
        }
        return sqrt(smax(0.0, boundary + mx(t2q, q2t)));
    } else if (METRIC == ERP) {
        double q2t = 0.0;
        for (int i = 1; i < n - 1; ++i) {
            double LT, UT;
            env(i, &LT, &UT);
            if (Q[i] > UT) q2t += mn(sqd(Q[i], UT), sqd(Q[i], arg));
            else if (Q[i] < LT) q2t += mn(sqd(Q[i], LT), sqd(Q[i], arg));
            if (boundary + q2t > bsf * bsf) return kGlbAbandon;
        }
        return sqrt(smax(0.0, boundary + mx(t2q, q2t)));
    } else if (METRIC == EDR) {
        double q2t = 0.0;
        for (int i = 1; i < n - 1; ++i) {
            double LT, UT;
            env(i, &LT, &UT);
            if (Q[i] > UT) q2t += edr_subcost(Q[i], UT, arg);
            else if (Q[i] < LT) q2t += edr_subcost(Q[i], LT, arg);
            if (boundary + q2t > bsf) return kGlbAbandon;
        }
        return boundary + mx(t2q, q2t);
    } else if (METRIC == LCSS) {
        double qr = 0.0;
        for (int i = 0; i < n; ++i) {
            double LT, UT;
            env(i, &LT, &UT);
            if (env_match(Q[i], LT, UT, arg)) qr += 1.0;
        }
        return smax(0.0, 1.0 - mn(lcss_t, qr) / (double)n);
    } else {
        double q2t = 0.0;
        for (int i = 1; i < n - 1; ++i) {
            double LT, UT;
            env(i, &LT, &UT);
            if (Q[i] > UT) q2t += mn(fabs(Q[i] - UT), arg);
            else if (Q[i] < LT) q2t += mn(fabs(Q[i] - LT), arg);
            if (boundary + q2t > bsf) return kGlbAbandon;
        }
        return boundary + mx(t2q, q2t);
    }
}

template <int METRIC>
__global__ void __launch_bounds__(256) k_glb(VerifyArgs a, const int64_t* __restrict__ in,
                                             const unsigned long long* __restrict__ in_count,
                                             int64_t* __restrict__ out,
                                             unsigned long long* out_count,
                                             double* __restrict__ vals) {
    const int64_t n = (int64_t)*in_count;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t iters = (n + stride - 1) / stride;
    int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t it = 0; it < iters; it++, ci += stride) {
        bool keep = false;
        int64_t g = 0;
        if (ci < n) {
            g = in[ci];
            const int64_t si = stat_index(a.G, g);
            const double v = glb_one<METRIC>(a.T + g, a.MU[si], a.SG[si], a.q.qn, a.q.LQ, a.q.UQ,
                                             a.Lraw + g, a.Uraw + g, a.p.m, a.p.w, a.p.arg,
                                             a.p.cutoff);
            if (vals) vals[ci] = v;
            keep = v <= a.p.cutoff;  // src/fsmeld.cpp:831
        }
        if (out) wave_append(keep, g, out, out_count);
    }
}

static int verify_blocks() { return 2048; }

hipError_t launch_glb(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const unsigned long long* in_count, int64_t* out,
                      unsigned long long* out_count, double* vals) {
    const dim3 g(verify_blocks()), b(256);
    switch (a.p.metric) {
        case DTW: hipLaunchKernelGGL(k_glb<DTW>, g, b, 0, st, a, in, in_count, out, out_count, vals); break;
        case ERP: hipLaunchKernelGGL(k_glb<ERP>, g, b, 0, st, a, in, in_count, out, out_count, vals); break;
        case EDR: hipLaunchKernelGGL(k_glb<EDR>, g, b, 0, st, a, in, in_count, out, out_count, vals); break;
        case LCSS: hipLaunchKernelGGL(k_glb<LCSS>, g, b, 0, st, a, in, in_count, out, out_count, vals); break;
        case MSM: hipLaunchKernelGGL(k_glb<MSM>, g, b, 0, st, a, in, in_count, out, out_count, vals); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// -------------------------------------------------------------------- ELD ---
// Band storage: registers (KB > 0 compile-time slots) or, for radii above
// the register variants, lane-interleaved global scratch.
template <int N>
struct RegArr {
    double v[N];
    __device__ double& operator[](int k) { return v[k]; }
};
struct MemArr {
    double* base;
    int64_t stride;
    __device__ double& operator[](int k) { return base[(int64_t)k * stride]; }
};

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
// src/ElasticMeasures.cpp:173-176
__device__ inline double msm_cost(double np, double x, double y, double c) {
    if ((x <= np && np <= y) || (y <= np && np <= x)) return c;
    return c + smin(fabs(np - x), fabs(np - y));
}

// Candidate sample source: raw target values normalised on the fly, or
// already-normalised rows.
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

// KB > 0: compile-time slot count, fully unrolled, slots above 2w skipped;
// KB == 0: runtime loop over the 2w+1 slots.
#define FSM_BAND_LOOP _Pragma("unroll") for (int k = 0; k < (KB > 0 ? KMAX : w2 + 1); k++) if (k <= w2)

// DTW, src/ElasticMeasures.cpp:24-61.  Cells (i, j), i, j in [0, m);
// band slot k = j - i + w; y[k] holds the candidate sample j of row i.
template <int KB, class Arr>
__device__ double dp_dtw(const double* __restrict__ q, const CandSrc& c, int m, int w, Arr& b,
                         Arr& y) {
    constexpr int KMAX = KB > 0 ? KB : 0x3fffffff;
    const int w2 = 2 * w;
    const double S = kEldSentinel;
    FSM_BAND_LOOP {
        b[k] = S;
        y[k] = c.at(k - w);
    }
    for (int i = 0; i < m; i++) {
        const double x = q[i];
        FSM_BAND_LOOP {
            const int j = i - w + k;
            if (j >= 0 && j < m) {
                const double d = (x - y[k]) * (x - y[k]);
                double v;
                if (i == 0 && j == 0) {
                    v = d;
                } else {
                    const double up = (i > 0 && k < w2) ? b[k + 1] : S;
                    const double left = (j > 0 && k > 0) ? b[k - 1] : S;
                    const double diag = (i > 0 && j > 0) ? b[k] : S;
                    v = dmin3(up, left, diag) + d;
                }
                b[k] = v;
            }
        }
        const double ynew = c.at(i + 1 + w);
        FSM_BAND_LOOP { y[k] = (k < w2) ? y[k + 1] : ynew; }
    }
    const double f = b[w];
    if (!__builtin_isfinite(f)) return S;
    return sqrt(smax(0.0, f));
}

// ERP (EDR when IS_EDR), src/ElasticMeasures.cpp:95-165.  Cells (i, j),
// i, j in [0, m] with the gap row / column; y[k] holds candidate sample j-1.
template <int KB, bool IS_EDR, class Arr>
__device__ double dp_erp_edr(const double* __restrict__ q, const CandSrc& c, int m, int w,
                             double g, Arr& b, Arr& y) {
    constexpr int KMAX = KB > 0 ? KB : 0x3fffffff;
    const int w2 = 2 * w;
    const double S = kEldSentinel;
    // row 0
    FSM_BAND_LOOP { y[k] = c.at(k - w - 1); }
    double run = 0.0;
    FSM_BAND_LOOP {
        const int j = k - w;
        double v = S;
        if (j == 0) {
            v = 0.0;
        } else if (j >= 1 && j <= m && j <= w) {
            v = IS_EDR ? (double)j : run + (y[k] - g) * (y[k] - g);
        }
        if (j >= 0 && j <= m && j <= w) run = v;
        b[k] = v;
    }
    for (int i = 1; i <= m; i++) {
        const double x = q[i - 1];
        const double ynew = c.at(i + w - 1);
        FSM_BAND_LOOP { y[k] = (k < w2) ? y[k + 1] : ynew; }
        const double dxg = (x - g) * (x - g);
        FSM_BAND_LOOP {
            const int j = i - w + k;
            double v = S;
            if (j == 0) {
                v = IS_EDR ? (double)i : b[k + 1] + dxg;
            } else if (j >= 1 && j <= m) {
                const double yv = y[k];
                const double up = (k < w2) ? b[k + 1] : S;
                const double left = (k > 0) ? b[k - 1] : S;
                if (IS_EDR) {
                    const double sub = (fabs(x - yv) <= g) ? 0.0 : 1.0;
                    v = dmin3(up + 1.0, left + 1.0, b[k] + sub);
                } else {
                    const double match = b[k] + (x - yv) * (x - yv);
                    const double ins = left + (yv - g) * (yv - g);
                    const double del = up + dxg;
                    v = dmin3(match, ins, del);
                }
            }
            b[k] = v;
        }
    }
    const double res = b[w];
    if (IS_EDR) return res;
    if (!__builtin_isfinite(res)) return S;
    return sqrt(smax(0.0, res));
}

// LCSS, src/ElasticMeasures.cpp:63-93; out-of-band cells read 0.
template <int KB, class Arr>
__device__ double dp_lcss(const double* __restrict__ q, const CandSrc& c, int m, int w,
                          double eps, Arr& b, Arr& y) {
    constexpr int KMAX = KB > 0 ? KB : 0x3fffffff;
    const int w2 = 2 * w;
    FSM_BAND_LOOP {
        b[k] = 0.0;
        y[k] = c.at(k - w);
    }
    for (int i = 0; i < m; i++) {
        const double x = q[i];
        FSM_BAND_LOOP {
            const int j = i - w + k;
            double v = 0.0;
            if (j >= 0 && j < m) {
                const double diag = (i > 0 && j > 0) ? b[k] : 0.0;
                const double up = (i > 0 && k < w2) ? b[k + 1] : 0.0;
                const double left = (j > 0 && k > 0) ? b[k - 1] : 0.0;
                if (fabs(x - y[k]) <= eps) v = dmax3(diag + 1, up, left);
                else v = dmax3(diag, up, left);
            }
            b[k] = v;
        }
        const double ynew = c.at(i + 1 + w);
        FSM_BAND_LOOP { y[k] = (k < w2) ? y[k + 1] : ynew; }
    }
    const double res = b[w];
    return 1 - res / (double)m;
}

// MSM, src/ElasticMeasures.cpp:167-198.
template <int KB, class Arr>
__device__ double dp_msm(const double* __restrict__ q, const CandSrc& c, int m, int w,
                         double cc, Arr& b, Arr& y) {
    constexpr int KMAX = KB > 0 ? KB : 0x3fffffff;
    const int w2 = 2 * w;
    const double S = kEldSentinel;
    FSM_BAND_LOOP {
        b[k] = S;
        y[k] = c.at(k - w);
    }
    for (int i = 0; i < m; i++) {
        const double xi = q[i];
        const double xp = (i > 0) ? q[i - 1] : 0.0;
        FSM_BAND_LOOP {
            const int j = i - w + k;
            double v = S;
            if (j >= 0 && j < m) {
                const double yj = y[k];
                if (i == 0 && j == 0) {
                    v = fabs(xi - yj);
                } else if (i == 0) {
                    v = b[k - 1] + msm_cost(yj, xi, y[k - 1], cc);
                } else if (j == 0) {
                    v = b[k + 1] + msm_cost(xi, xp, yj, cc);
                } else {
                    const double c1 = b[k] + fabs(xi - yj);
                    const double c2 = ((k < w2) ? b[k + 1] : S) + msm_cost(xi, xp, yj, cc);
                    const double c3 = (k > 0) ? b[k - 1] + msm_cost(yj, xi, y[k - 1], cc) : S;
                    v = dmin3(c1, c2, c3);
                }
            }
            b[k] = v;
        }
        const double ynew = c.at(i + 1 + w);
        FSM_BAND_LOOP { y[k] = (k < w2) ? y[k + 1] : ynew; }
    }
    return b[w];
}

template <int KB, class Arr>
__device__ double eld_one(int metric, const double* q, const CandSrc& c, int m, int w,
                          double arg, Arr& b, Arr& y) {
    switch (metric) {
        case DTW: return dp_dtw<KB>(q, c, m, w, b, y);
        case ERP: return dp_erp_edr<KB, false>(q, c, m, w, arg, b, y);
        case EDR: return dp_erp_edr<KB, true>(q, c, m, w, arg, b, y);
        case LCSS: return dp_lcss<KB>(q, c, m, w, arg, b, y);
        case MSM: return dp_msm<KB>(q, c, m, w, arg, b, y);
    }
    return __builtin_nan("");
}

// in[] lists window indices (target mode) or row numbers (pairs mode:
// in == nullptr, rows 0..n-1 of `rows`).
template <int KB>
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
    int64_t ci = tid;
    for (int64_t it = 0; it < iters; it++, ci += stride) {
        bool keep = false;
        int64_t g = 0;
        double d = 0.0;
        if (ci < n) {
            CandSrc c;
            c.m = m;
            if (rows) {
                g = ci;
                c.t = rows + ci * (int64_t)m;
                c.raw = false;
                c.mu = 0.0;
                c.sd = 1.0;
            } else {
                g = in[ci];
                const int64_t si = stat_index(a.G, g);
                c.t = a.T + g;
                c.mu = a.MU[si];
                c.sd = a.SG[si];
                c.raw = true;
            }
            if constexpr (KB > 0) {
                RegArr<KB + 1> b, y;
                d = eld_one<KB>(a.p.metric, a.q.qn, c, m, w, a.p.arg, b, y);
            } else {
                MemArr b{scratch + tid, stride * 2};
                MemArr y{scratch + tid + stride, stride * 2};
                d = eld_one<0>(a.p.metric, a.q.qn, c, m, w, a.p.arg, b, y);
            }
            if (keep_all) out_dist[ci] = d;
            keep = __builtin_isfinite(d) && d <= a.p.cutoff;  // src/fsmeld.cpp:846-847
        }
        if (!keep_all) {
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
}

int64_t eld_scratch_threads() { return (int64_t)512 * 256; }

template <int KB>
static void launch_eld_kb(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                          const unsigned long long* in_count, int64_t n_direct,
                          const double* rows, bool keep_all, int64_t* out_idx,
                          double* out_dist, unsigned long long* out_count, double* scratch,
                          int blocks) {
    hipLaunchKernelGGL(k_eld<KB>, dim3(blocks), dim3(256), 0, st, a, in, in_count, n_direct,
                       rows, keep_all ? 1 : 0, out_idx, out_dist, out_count, scratch);
}

static hipError_t dispatch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                               const unsigned long long* in_count, int64_t n_direct,
                               const double* rows, bool keep_all, int64_t* out_idx,
                               double* out_dist, unsigned long long* out_count,
                               double* scratch, int64_t scratch_threads) {
    const int w = a.p.w;
    if (w <= 4) launch_eld_kb<9>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, 2048);
    else if (w <= 8) launch_eld_kb<17>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, 2048);
    else if (w <= 16) launch_eld_kb<33>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, 2048);
    else if (w <= 32) launch_eld_kb<65>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist, out_count, scratch, 2048);
    else {
        if (!scratch) return hipErrorInvalidValue;
        launch_eld_kb<0>(st, a, in, in_count, n_direct, rows, keep_all, out_idx, out_dist,
                         out_count, scratch, (int)(scratch_threads / 256));
    }
    return hipGetLastError();
}

hipError_t launch_eld(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const unsigned long long* in_count, bool keep_all, int64_t* out_idx,
                      double* out_dist, unsigned long long* out_count, double* scratch,
                      int64_t scratch_threads) {
    return dispatch_eld(st, a, in, in_count, 0, nullptr, keep_all, out_idx, out_dist,
                        out_count, scratch, scratch_threads);
}

hipError_t launch_eld_pairs(hipStream_t st, int metric, int m, int w, double arg,
                            const double* q, const double* cands, int64_t n, double* out,
                            double* scratch, int64_t scratch_threads) {
    VerifyArgs a{};
    a.q.qn = q;
    a.p.metric = metric;
    a.p.m = m;
    a.p.w = w;
    a.p.arg = arg;
    a.p.cutoff = 0.0;
    return dispatch_eld(st, a, nullptr, nullptr, n, cands, true, nullptr, out, nullptr,
                        scratch, scratch_threads);
}

}  // namespace fsm
