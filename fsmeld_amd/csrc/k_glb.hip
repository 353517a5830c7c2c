// k_glb.hip -- the GLB filter (glb_elastic, src/glb.cpp:200-373) on gfx950.
//
// One candidate per lane.  Only the keep decision of the reference leaves
// this kernel (the stage entry point excepted), and that decision depends on
// three numbers: the boundary term and the final values of the two running
// sums (glb_keeps).  The fast pass (glb_fast) evaluates the sums without
// division and in a branch-free loop, with a rigorous bound on their distance
// from the reference's sums; where the decision is the same at both ends of
// that interval it is the reference's decision.  Otherwise the candidate
// takes the exact path (glb_one), which evaluates every expression in the
// reference's operand order, sums and early abandon included, so its value is
// bit-identical to glb_elastic's.
//
// Inputs are valid windows (finite samples, sd > 1e-12) and a finite query:
// this file is compiled with -fno-honor-nans.
#include "engine_internal.h"

#include <type_traits>

namespace fsm {

__device__ inline double sqd(double x, double y) { return (x - y) * (x - y); }  // include/def.h:5
__device__ inline double mn(double x, double y) { return (x < y) ? x : y; }    // include/def.h:9
__device__ inline double mx(double x, double y) { return (x > y) ? x : y; }    // include/def.h:10
__device__ inline double edr_subcost(double a, double b, double eps) {          // src/glb.cpp:8-10
    return (fabs(a - b) <= eps) ? 0.0 : 1.0;
}
__device__ inline double swale_subcost(double a, double b, double eps) {         // src/glb.cpp:11-14
    return (fabs(a - b) <= eps) ? kSwaleR : kSwaleP;
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
    int pre_r;  // prefix covers t[0 .. pre_r]
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
                pre_lo = fmin(pre_lo, t[pre_r]);
                pre_hi = fmax(pre_hi, t[pre_r]);
            }
            *lo = pre_lo;
            *hi = pre_hi;
        } else if (i > m - 1 - w) {
            double a = t[i - w], b = t[i - w];
            for (int x = i - w + 1; x < m; x++) {
                a = fmin(a, t[x]);
                b = fmax(b, t[x]);
            }
            *lo = a;
            *hi = b;
        } else {
            *lo = Lr[i];
            *hi = Ur[i];
        }
    }
};

// Normalised samples (src/fsmeld.cpp:822-826): the exact path needs the
// correctly rounded quotient fl(fl(t - mu) / sd); a = fl(fl(t - mu) * fl(1 / sd))
// differs from it by at most 3.01 u |a| (u = 2^-53: three roundings of
// relative size u, none of them shared).
struct Norm {
    double mu, sd, rinv;
    __device__ double approx(double x) const { return (x - mu) * rinv; }
    __device__ double exact(double x) const { return (x - mu) / sd; }
};
__device__ inline double nmargin(double a) { return fabs(a) * 8.881784197001252e-16 + 1e-300; }

// T strictly inside [lo, hi] for every rounding of the approximate value a
__device__ inline bool surely_inside(double a, double lo, double hi) {
    const double d = nmargin(a);
    return a - d > lo && a + d < hi;
}

// Boundary term of glb_elastic from the exact normalised samples T0 = t^[0],
// Tl = t^[n-1], T2 = t^[n-2]: src/glb.cpp:217 (DTW), :226-230 (ERP), :242
// (EDR), :264-275 (MSM); LCSS has none.
template <int METRIC>
__device__ inline double glb_boundary(double T0, double Tl, double T2, const double* Q, int n,
                                      double arg) {
    const int b = n - 1;
    if (METRIC == DTW) return sqd(T0, Q[0]) + sqd(Tl, Q[b]);
    if (METRIC == ERP) return smin3(sqd(Q[b], Tl), sqd(Q[b], arg), sqd(Tl, arg));
    if (METRIC == EDR) return mn(edr_subcost(Q[b], Tl, arg), 1.0);
    if (METRIC == SWALE) return mn(swale_subcost(Q[b], Tl, arg), kSwaleP);  // :282-286
    if (METRIC == MSM) {
        if ((T2 >= Tl && Tl >= Q[b]) || (T2 <= Tl && Tl <= Q[b]) ||
            (Q[b - 1] <= Q[b] && Q[b] <= Tl) || (Q[b - 1] >= Q[b] && Q[b] >= Tl))
            return fabs(Q[0] - T0) + mn(fabs(Q[b] - Tl), arg);
        return fabs(Q[0] - T0) +
               smin3(fabs(Q[b] - Tl), arg + fabs(Tl - T2), arg + fabs(Q[b] - Q[b - 1]));
    }
    return 0.0;
}

// ----------------------------------------------------------- exact path ---
template <int METRIC>
__device__ double glb_one(const double* __restrict__ t, double mu, double sd,
                          const double* __restrict__ Q, const double* __restrict__ LQ,
                          const double* __restrict__ UQ, const double* Lr, const double* Ur,
                          int m, int w, double arg, double bsf) {
    if (m < 2) return 0.0;
    const int n = m;
    const Norm N{mu, sd, 1.0 / sd};
    double boundary = 0.0, t2q = 0.0, lcss_t = 0.0;
    const double Tl = N.exact(t[n - 1]);
    if (METRIC != LCSS)
        boundary = glb_boundary<METRIC>(N.exact(t[0]), Tl, N.exact(t[n - 2]), Q, n, arg);
    // The reference tests early abandon after every position; positions that
    // add nothing leave the test unchanged, so it is evaluated once before the
    // loop (its first test) and after every contributing position.
    if (METRIC == DTW) {
        if (n > 2 && boundary + t2q > bsf * bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v < LQ[i]) t2q += sqd(v, LQ[i]);
            else if (v > UQ[i]) t2q += sqd(v, UQ[i]);
            if (boundary + t2q > bsf * bsf) return kGlbAbandon;
        }
    } else if (METRIC == ERP) {
        if (n > 2 && boundary + t2q > bsf * bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v > UQ[i]) t2q += mn(sqd(v, UQ[i]), sqd(v, arg));
            else if (v < LQ[i]) t2q += mn(sqd(v, LQ[i]), sqd(v, arg));
            if (boundary + t2q > bsf * bsf) return kGlbAbandon;
        }
    } else if (METRIC == EDR) {
        if (n > 2 && boundary + t2q > bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v > UQ[i]) t2q += edr_subcost(v, UQ[i], arg);
            else if (v < LQ[i]) t2q += edr_subcost(v, LQ[i], arg);
            if (boundary + t2q > bsf) return kGlbAbandon;
        }
    } else if (METRIC == SWALE) {  // src/glb.cpp:282-297
        if (n > 2 && boundary + t2q > bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v > UQ[i]) t2q += swale_subcost(v, UQ[i], arg);
            else if (v < LQ[i]) t2q += swale_subcost(v, LQ[i], arg);
            if (boundary + t2q > bsf) return kGlbAbandon;
        }
    } else if (METRIC == LCSS) {
        const double min_len = (double)n;
        for (int i = 0; i < n; ++i) {
            const double a = N.approx(t[i]);
            if (surely_inside(a, LQ[i], UQ[i])) {
                lcss_t += 1.0;
                continue;
            }
            const double far = nmargin(a) + 8.881784197001252e-16 * (fabs(LQ[i]) + fabs(UQ[i]) + arg);
            if (a + far < LQ[i] - arg || a - far > UQ[i] + arg) continue;  // surely no match
            if (env_match(N.exact(t[i]), LQ[i], UQ[i], arg)) lcss_t += 1.0;
        }
        if (smax(0.0, 1.0 - lcss_t / min_len) > bsf) return kGlbAbandon;
    } else {  // MSM
        if (n > 2 && boundary + t2q > bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v > UQ[i]) t2q += mn(fabs(v - UQ[i]), arg);
            else if (v < LQ[i]) t2q += mn(fabs(v - LQ[i]), arg);
            if (boundary + t2q > bsf) return kGlbAbandon;
        }
    }

    // candidate envelope (raw), normalised only where a decision needs it;
    // the pass's first abandon test equals pass 1's and has already passed
    CandEnvelope E;
    E.init(t, Lr, Ur, m, w);
    double q2t = 0.0, qr = 0.0;
    const int i0 = (METRIC == LCSS) ? 0 : 1;
    const int i1 = (METRIC == LCSS) ? n : n - 1;
    for (int i = i0; i < i1; ++i) {
        double lo, hi;
        E.at(i, &lo, &hi);
        const double aL = N.approx(lo), aU = N.approx(hi);
        const double q = Q[i];
        const bool inside = q - nmargin(aL) > aL && q + nmargin(aU) < aU;
        if (METRIC == LCSS) {
            if (inside) {
                qr += 1.0;
                continue;
            }
            const double far = nmargin(aL) + nmargin(aU) +
                               8.881784197001252e-16 * (fabs(q) + arg);
            if (q + far < aL - arg || q - far > aU + arg) continue;
            if (env_match(q, N.exact(lo), N.exact(hi), arg)) qr += 1.0;
            continue;
        }
        if (inside) continue;
        const double LT = N.exact(lo), UT = N.exact(hi);
        if (METRIC == DTW) {
            if (q < LT) q2t += sqd(q, LT);
            else if (q > UT) q2t += sqd(q, UT);
            if (boundary + q2t > bsf * bsf) return kGlbAbandon;
        } else if (METRIC == ERP) {
            if (q > UT) q2t += mn(sqd(q, UT), sqd(q, arg));
            else if (q < LT) q2t += mn(sqd(q, LT), sqd(q, arg));
            if (boundary + q2t > bsf * bsf) return kGlbAbandon;
        } else if (METRIC == EDR) {
            if (q > UT) q2t += edr_subcost(q, UT, arg);
            else if (q < LT) q2t += edr_subcost(q, LT, arg);
            if (boundary + q2t > bsf) return kGlbAbandon;
        } else if (METRIC == SWALE) {  // src/glb.cpp:356-369
            if (q > UT) q2t += swale_subcost(q, UT, arg);
            else if (q < LT) q2t += swale_subcost(q, LT, arg);
            if (boundary + q2t > bsf) return kGlbAbandon;
        } else {  // MSM
            if (q > UT) q2t += mn(fabs(q - UT), arg);
            else if (q < LT) q2t += mn(fabs(q - LT), arg);
            if (boundary + q2t > bsf) return kGlbAbandon;
        }
    }
    if (METRIC == DTW || METRIC == ERP) return sqrt(smax(0.0, boundary + mx(t2q, q2t)));
    if (METRIC == LCSS) return smax(0.0, 1.0 - mn(lcss_t, qr) / (double)n);
    return boundary + mx(t2q, q2t);
}

// ------------------------------------------------------------ fast path ---
// The reference's keep decision (glb_elastic then `<= cutoff`,
// src/fsmeld.cpp:827-831) from the boundary B and the final running sums T
// (candidate to query envelope) and Q (query to candidate envelope); for
// LCSS T and Q are the match counts.  glb_elastic abandons (1008610086) once
// B plus a running sum exceeds bsf (bsf^2 for DTW / ERP); running sums never
// decrease, so the final sums decide the same.  Valid for bsf < 1008610086,
// where an abandoned candidate is never kept.  Every operation is monotone,
// so the decision is monotone in T and Q: non-increasing for the distance
// sums, non-decreasing for the LCSS counts.
template <int METRIC>
__device__ inline bool glb_keeps(double B, double T, double Q, int n, double bsf) {
    if (METRIC == DTW || METRIC == ERP) {
        const double b2 = bsf * bsf;
        if (n > 2 && (B + T > b2 || B + Q > b2)) return false;
        return sqrt(smax(0.0, B + mx(T, Q))) <= bsf;
    } else if (METRIC == LCSS) {
        const double nn = (double)n;
        if (smax(0.0, 1.0 - T / nn) > bsf) return false;
        return smax(0.0, 1.0 - mn(T, Q) / nn) <= bsf;
    } else {
        if (n > 2 && (B + T > bsf || B + Q > bsf)) return false;
        return B + mx(T, Q) <= bsf;
    }
}

// Bound on |reference sum - fast sum| for DTW, ERP and MSM.  The fast sum S
// adds at most n terms.  Each term is a function of a distance to an
// envelope whose arguments moved by at most delta (normalisation): MSM's
// min(d, c) is 1-Lipschitz; DTW's d^2 and ERP's min(d^2, (x - g)^2) move by
// at most delta (2K + delta) where K bounds the distances involved.  Both sums
// also carry term rounding (<= 3u relative) and summation in different orders
// (<= (n - 1) u relative each), psi below.
template <int METRIC>
__device__ inline double glb_sum_err(double S, int n, double delta, double K) {
    const double psi = (double)(n + 8) * 1.2e-16;
    const double D = (METRIC == MSM) ? (double)n * delta : (double)n * delta * (2.0 * K + delta);
    return (D * (1.0 + psi) + 2.0 * psi * S * (1.0 + 2.0 * psi)) * 1.01 + 1e-300;
}

// positions between wave-level exit tests: 32 for MSM and DTW, whose
// candidates mostly run both passes to the end (MSM's GLB 21.4 -> 20.7 ms per
// 2^26 query), 16 for the others (8 and 32 measured slower there)
template <int METRIC>
constexpr int kGlbChunk = (METRIC == MSM || METRIC == DTW) ? 32 : 16;
// From this radius on, the bound on |normalised sample| comes from the
// envelope at every (2w + 1)-th position instead of every sample.
constexpr int kGlbSampleW = 4;

// Pass 1's distance of a candidate sample a to the query envelope [L, U] is
// |d| - h with d = fl(fl(t - mu) r - M), M = fl(L + U) / 2, h = fl(U - L) / 2
// (exactly max(L - a, a - U, 0) when positive, in exact arithmetic), h
// replaced by fl(h + eps) for EDR / LCSS and the whole divided by c, i.e.
// fl(|d| fl(1 / c) - fl(h fl(1 / c))) clamped to [0, 1], for MSM.  Bound on
// each term's distance from the reference's term (u = 2^-53, |a| <= amax,
// |L|, |U| <= QM): |d| is within 3.02 u |a| + 2.01 u QM of |a - M| (two
// roundings in a, u QM in M, one after the product), and |d| - h within
// 4.03 u |a| + 4.02 u QM of the exact distance (h: u QM, the difference: u (|a|
// + QM)); MSM's 1 / c and h / c add 2.03 u QM and the clamped fma u c.  The
// reference's own normalisation and subtraction are within 3.02 u |a| + 1.01 u QM
// of the exact distance.  ERP's a - g, fl(fl(t - mu) r - g) against the
// reference's fl(a' - g): 6.04 u |a| + 2.02 u |g|.  So every DTW / ERP / MSM
// term argument is within 7.06 u (amax + QM + |arg|) < 8.2e-16 (amax + QM +
// |arg|); 1e-220 covers results in the subnormal range (c and 1 / c <= 1e100).
__device__ inline double glb_delta1(double amax, double QM, double arg) {
    return 8.2e-16 * (amax + QM + fabs(arg)) + 1e-220;
}

// ERP's term from the distance itself: min(e^2, xg^2) = min(e, |xg|)^2 for
// e, |xg| >= 0, squared and added with one rounding.
__device__ inline void erp_term(double e, double xg, double& s) {
    const double d = fmin(e, fabs(xg));
    s = fma(d, d, s);
}

// Fast decision for one candidate: 1 keep, -1 prune, 0 undecided (take the
// exact path).  Every lane of the wave calls it; loop trip counts are
// uniform and the loops end early only when every lane has been pruned.
// Query arrays reach the kernel as separate read-only, non-aliasing pointers
// so that their wave-uniform reads become scalar loads.
// On keep, *total receives the pass-1 bound the ELD abandon test uses: a
// lower bound on the real-valued sum over samples 1 .. n-2 of the per-sample
// envelope terms (DTW / ERP / MSM), the count of certain mismatches (EDR),
// or an upper bound on the samples that can match at all (LCSS).
//
// EDR / LCSS count positions.  With eps > 0 a position mismatches (EDR) or
// fails to match (LCSS) exactly when R = max(L - a, a - U) - eps > 0 for its
// envelope [L, U] (the clamp at 0 cannot reach eps).  Pass 1 takes R as
// fl(|d| - fl(h + eps)) (glb_delta1's d and h): against the reference's rounded
// quantities (its own normalisation fl(fl(t - mu) / sd) and subtraction) it is
// off by at most 7.05 u |a| + 6.05 u max|q| + 2.01 u eps (|d|: 3.02 u |a| + 2.01
// u QM; h + eps: u QM + u (QM + eps); the difference: u (|a| + 2 QM + eps); the
// reference's side: 3.02 u |a| + 1.01 u QM).  Pass 2 compares max(lo - qr, qr -
// hi) in raw units with fl(eps sd +- eta sd), within 5 u amax + 4 u max|q| + 2 u
// eps + u |mu| / sd.  eta = 1.2e-15 (|a| + max|q| + eps) (pass 2: plus 1.3e-16
// |mu| / sd, scaled by sd) covers both; a position with R > eta is certainly
// out, with R < -eta certainly in, so counting R > eta and R >= -eta gives both
// ends of the reference's count.
template <int METRIC>
__device__ __forceinline__ int glb_fast(const VerifyArgs& A, const double* __restrict__ Q,
                        const double* __restrict__ XG, const double* __restrict__ MQ,
                        const double* __restrict__ HQ, const double* __restrict__ t,
                        const double* __restrict__ Lr, const double* __restrict__ Ur,
                        double mu, double sd, bool active, double* total) {
    constexpr bool COUNT = (METRIC == EDR || METRIC == LCSS);
    const int n = A.p.m, w = A.p.w;
    const double arg = A.p.arg, bsf = A.p.cutoff;
    const double thr = (METRIC == DTW || METRIC == ERP) ? bsf * bsf : bsf;
    const double QM = A.q.qabs;
    const double r = 1.0 / sd;
    const int i0 = (METRIC == LCSS) ? 0 : 1;
    const int i1 = (METRIC == LCSS) ? n : n - 1;
    const int npos = i1 - i0;
    // DTW's pass 2 scales by r^2 and MSM's by r / c once: with sd in this
    // range those factors stay normal (no underflow to 0 against an overflowed
    // sum); other windows take the exact path
    if ((METRIC == DTW || METRIC == MSM) && active && !(sd >= 1e-100 && sd <= 1e100)) return 0;

    double B = 0.0;
    if (METRIC != LCSS)
        B = glb_boundary<METRIC>((t[0] - mu) / sd, (t[n - 1] - mu) / sd, (t[n - 2] - mu) / sd, Q,
                                 n, arg);
    // amax bounds |a| over every sample of the window; the candidate
    // envelope's normalised values are such samples (normalisation is
    // monotone).  With w >= kGlbSampleW it is the normalised window min / max:
    // the envelope at positions w, 3w + 1, ... and n - 1 - w covers the window
    // (each spans 2w + 1 samples), and (v - mu) * r is monotone in v, so the
    // value equals the largest |a| over the samples pass 1 normalises.
    // Otherwise pass 1 tracks it sample by sample.
    const bool sampled = w >= kGlbSampleW && n >= 2 * w + 1;
    double amax;
    if (sampled) {
        double lo = Lr[w], hi = Ur[w];
        for (int p = 3 * w + 1; p < n - 1 - w; p += 2 * w + 1) {
            lo = fmin(lo, Lr[p]);
            hi = fmax(hi, Ur[p]);
        }
        lo = fmin(lo, Lr[n - 1 - w]);
        hi = fmax(hi, Ur[n - 1 - w]);
        amax = fmax(fabs((lo - mu) * r), fabs((hi - mu) * r));
    } else {
        amax = fmax(fabs((t[0] - mu) * r), fabs((t[n - 1] - mu) * r));
    }
    const double kq = fmax(QM, fabs(arg));
    bool dead = !active;

    // pass 1: candidate samples against the query envelope
    double s1 = 0.0;
    // counting metrics: positions with R > eta (cgt) and R >= -eta (cge)
    constexpr double kEtaC = 1.2e-15;
    const double cetaC = kEtaC * (QM + arg) + 1e-300;
    const double eta1 = kEtaC * amax + cetaC;  // with a sampled amax
    int cgt1 = 0, cge1 = 0;
    // MSM sums its terms in units of c (s1, s2), scaled back where they are
    // used: each term is one clamp to [0, 1] of (|d| - h) / c
    const double rho = glb_msm_rho(arg);
    const double sc = (METRIC == MSM) ? arg : 1.0;
    auto term1 = [&](int k, auto track) {
        const double tm = t[k] - mu;
        // the signed distance from the envelope's midpoint, d = a - MQ with
        // a = (t - mu) r, in one rounding after the product
        const double d = fma(tm, r, -MQ[k]);
        double et = eta1;
        if constexpr (decltype(track)::value) {
            const double av = tm * r;
            amax = fmax(amax, fabs(av));
            et = fma(fabs(av), kEtaC, cetaC);
        }
        if constexpr (COUNT) {
            const double R = fabs(d) - HQ[k];  // HQ: the half width widened by eps
            cgt1 += (R > et) ? 1 : 0;
            cge1 += (R >= -et) ? 1 : 0;
        } else if constexpr (METRIC == MSM) {
            s1 += fmin(fmax(fma(fabs(d), rho, -HQ[k]), 0.0), 1.0);
        } else {
            const double e = fmax(fabs(d) - HQ[k], 0.0);
            if constexpr (METRIC == ERP) erp_term(e, fma(tm, r, -arg), s1);
            else s1 = fma(e, e, s1);
        }
    };
    auto check1 = [&](int ie) {
        if (dead) return;
        if (METRIC == LCSS) {  // every position not certainly out can still match
            (void)ie;
            const double tmax = (double)(npos - cgt1);
            dead = smax(0.0, 1.0 - tmax / (double)n) > bsf;
        } else if (METRIC == EDR) {
            dead = B + (double)cgt1 > bsf;
        } else {
            const double S = s1 * sc;
            const double lo = S - glb_sum_err<METRIC>(S, n, glb_delta1(amax, QM, arg), amax + kq);
            dead = B + lo > thr;
        }
    };
    // returns true when every lane of the wave is dead
    auto pass1 = [&](auto track) {
        int i = i0;
        for (; i + kGlbChunk<METRIC> <= i1; i += kGlbChunk<METRIC>) {
#pragma unroll
            for (int u = 0; u < kGlbChunk<METRIC>; u++) term1(i + u, track);
            check1(i + kGlbChunk<METRIC>);
            if (__all(dead)) return true;
        }
        for (; i < i1; i++) term1(i, track);
        return false;
    };
    if (sampled ? pass1(std::false_type{}) : pass1(std::true_type{})) return -1;

    // pass 2: query against the candidate envelope (raw window min / max over
    // positions clipped to the window).  The distance is taken in raw units,
    // x = max(lo - qr, qr - hi) with qr = fl(q sd + mu); e = max(x, 0) r
    // differs from the reference's fl(q - UT) / fl(LT - q) on its normalised
    // envelope by at most 6.02 u amax + 5.02 u QM + 1.01 u |mu| / sd (u = 2^-53):
    // 2 u |a| from the reference's normalisation, 3 u e from the subtraction,
    // r and the product, u |q + mu / sd| from qr, u e for the reference's own
    // subtraction, e <= amax + QM).  DTW sums max(x, 0)^2 and multiplies by r^2
    // once (a few u relative, inside glb_sum_err's psi).  MSM clamps x rc,
    // rc = fl(r fl(1 / c)), to [0, 1]: four roundings on x / sd (4.02 u (amax +
    // QM)) besides the subtraction's, 8.06 u amax + 7.05 u QM + u |mu| / sd in
    // all with the reference's side.
    const double mur = fabs(mu) * r;
    const double delta2 = (METRIC == MSM)
                              ? 9.2e-16 * (amax + QM) + 1.2e-16 * mur + 1e-220
                              : 7e-16 * (amax + QM) + 1.2e-16 * mur + 1e-300;
    // s2: DTW sums the squared raw distances (times r^2 where used), MSM in
    // units of c as pass 1
    double s2 = 0.0;
    const double sc2 = (METRIC == DTW) ? r * r : sc;
    const double rc = r * rho;
    const double epsr = arg * sd;
    const double eta2r = sd * (kEtaC * (amax + QM + arg) + 1.3e-16 * mur) + 1e-300;
    // counting: R = max(lo - qr, qr - hi) - eps sd against +-eta2r, as
    // max(lo - qr, qr - hi) against the two sums
    const double rgt2 = epsr + eta2r, rge2 = epsr - eta2r;
    int cgt2 = 0, cge2 = 0;
    auto term2 = [&](int k, double lo, double hi) {
        const double qr = fma(Q[k], sd, mu);
        const double x = fmax(lo - qr, qr - hi);  // raw units
        if constexpr (COUNT) {
            cgt2 += (x > rgt2) ? 1 : 0;
            cge2 += (x >= rge2) ? 1 : 0;
        } else if constexpr (METRIC == MSM) {
            s2 += fmin(fmax(x * rc, 0.0), 1.0);
        } else if constexpr (METRIC == DTW) {
            const double e = fmax(x, 0.0);
            s2 = fma(e, e, s2);
        } else {  // ERP
            const double e = fmax(x, 0.0) * r;
            s2 += fmin(e * e, XG[k]);
        }
    };
    auto check2 = [&]() {
        if (dead) return;
        if (METRIC == LCSS) {
            const double qmax = (double)(npos - cgt2);
            dead = smax(0.0, 1.0 - mn((double)(npos - cgt1), qmax) / (double)n) > bsf;
        } else if (METRIC == EDR) {
            dead = B + (double)cgt2 > bsf;
        } else {
            const double S = s2 * sc2;
            const double lo = S - glb_sum_err<METRIC>(S, n, delta2, amax + kq);
            dead = B + lo > thr;
        }
    };
    // head: positions i < w see the window prefix t[0 .. min(i + w, n - 1)]
    const int h1 = min(w, i1);
    {
        double plo = t[0], phi = t[0];
        int pr = 0;
        for (int k = i0; k < h1; k++) {
            const int rr = min(k + w, n - 1);
            while (pr < rr) {
                pr++;
                plo = fmin(plo, t[pr]);
                phi = fmax(phi, t[pr]);
            }
            term2(k, plo, phi);
        }
    }
    // interior: w <= i < n - w, the target-wide envelope
    const int m0 = max(i0, w), m1 = min(i1, n - w);
    int k0 = m0;
    for (; k0 + kGlbChunk<METRIC> <= m1; k0 += kGlbChunk<METRIC>) {
#pragma unroll
        for (int u = 0; u < kGlbChunk<METRIC>; u++) term2(k0 + u, Lr[k0 + u], Ur[k0 + u]);
        check2();
        if (__all(dead)) return -1;
    }
    for (; k0 < m1; k0++) term2(k0, Lr[k0], Ur[k0]);
    // tail: positions i >= max(w, n - w) see the suffix t[i - w .. n - 1];
    // walked from the right end so the suffix only grows
    const int tb = max(max(i0, w), n - w);
    {
        double slo = t[n - 1], shi = t[n - 1];
        int sl = n - 1;
        for (int k = i1 - 1; k >= tb; k--) {
            const int ll = k - w;
            while (sl > ll) {
                sl--;
                slo = fmin(slo, t[sl]);
                shi = fmax(shi, t[sl]);
            }
            term2(k, slo, shi);
        }
    }
    check2();
    if (dead) return -1;

    double Tlo, Thi, Qlo, Qhi;
    if (METRIC == EDR) {  // mismatch counts
        Tlo = (double)cgt1;
        Thi = (double)cge1;
        Qlo = (double)cgt2;
        Qhi = (double)cge2;
    } else if (METRIC == LCSS) {  // match counts
        Tlo = (double)(npos - cge1);
        Thi = (double)(npos - cgt1);
        Qlo = (double)(npos - cge2);
        Qhi = (double)(npos - cgt2);
    } else {
        const double K = amax + kq;
        const double S1 = s1 * sc, S2 = s2 * sc2;
        const double e1 = glb_sum_err<METRIC>(S1, n, glb_delta1(amax, QM, arg), K);
        const double e2 = glb_sum_err<METRIC>(S2, n, delta2, K);
        Tlo = fmax(S1 - e1, 0.0);
        Thi = S1 + e1;
        Qlo = fmax(S2 - e2, 0.0);
        Qhi = S2 + e2;
    }
    // pass 1's terms are the reference's rounded terms; the real-valued sum
    // is lower by at most psi relative
    if (METRIC == LCSS) *total = Thi;
    else if (METRIC == EDR) *total = Tlo;
    else *total = Tlo * (1.0 - (double)(n + 8) * 1.2e-16);
    if (METRIC == LCSS) {
        if (glb_keeps<METRIC>(B, Tlo, Qlo, n, bsf)) return 1;
        if (!glb_keeps<METRIC>(B, Thi, Qhi, n, bsf)) return -1;
        return 0;
    }
    if (glb_keeps<METRIC>(B, Thi, Qhi, n, bsf)) return 1;
    if (!glb_keeps<METRIC>(B, Tlo, Qlo, n, bsf)) return -1;
    return 0;
}

template <int METRIC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_glb(VerifyArgs a, const double* __restrict__ Q,
                                             const double* __restrict__ LQ,
                                             const double* __restrict__ UQ,
                                             const double* __restrict__ XG,
                                             const double* __restrict__ MQ,
                                             const double* __restrict__ HQ,
                                             const int64_t* __restrict__ in,
                                             const unsigned long long* __restrict__ in_count,
                                             int64_t* __restrict__ out,
                                             unsigned long long* out_count,
                                             double* __restrict__ out_total,
                                             double* __restrict__ vals,
                                             unsigned long long* __restrict__ queue,
                                             int stage_len) {
    __shared__ unsigned long long sh[256 / 64 + 1];
    __shared__ int64_t rng[2 * (256 / 64)];
    __shared__ unsigned long long chunk;
    // staged samples / envelope of the block's windows (3 x stage_len doubles)
    extern __shared__ double glb_stage[];
    double* Ts = glb_stage;
    double* Ls = Ts + stage_len;
    double* Us = Ls + stage_len;
    const int64_t n = (int64_t)*in_count;
    // the fast decision needs contributions >= 0 (MSM: c = 0 or 1 / c
    // normal, glb_msm_rho), for EDR / LCSS an envelope that counts as a match
    // (eps > 0), and envelope midpoints that cannot overflow; values (vals)
    // need the exact path (SWALE takes the exact path only)
    const bool arg_ok = (METRIC == DTW || METRIC == ERP) ? true
                        : (METRIC == MSM) ? (a.p.arg == 0.0 || glb_msm_rho(a.p.arg) > 0.0)
                                          : a.p.arg > 0.0;
    const bool fast = METRIC != SWALE && vals == nullptr && a.p.cutoff < kGlbAbandon &&
                      a.p.m > 2 && a.q.qabs <= 1e300 && arg_ok;
    // Blocks take runs of blockDim.x consecutive candidates from the queue
    // until it is drained: per-candidate work varies (early exits), so a
    // fixed assignment leaves most of the GPU idle behind the slowest blocks.
    while (true) {
        if (threadIdx.x == 0) chunk = atomicAdd(queue, (unsigned long long)blockDim.x);
        __syncthreads();
        const int64_t base = (int64_t)chunk;
        __syncthreads();  // every thread has read chunk before it is replaced
        if (base >= n) break;
        const int64_t ci = base + threadIdx.x;
        const bool active = ci < n;
        const int64_t g = active ? in[ci] : 0;
        const int64_t si = stat_index(a.G, g);
        const double mu = active ? a.MU[si] : 0.0;
        const double sd = active ? a.SG[si] : 1.0;
        int verdict = 0;
        // without the fast pass's bound: nothing (distances), every sample
        // (LCSS), r per sample (SWALE: every sample costs at least r)
        double total = (METRIC == LCSS)    ? (double)a.p.m
                       : (METRIC == SWALE) ? kSwaleR * (double)a.p.m
                                           : 0.0;
        // When the block's candidate windows span at most stage_len - m
        // windows, their samples and envelope are read once per block into
        // LDS (the passes read each value up to m times).  The list is
        // ascending only within each compaction block's run, so the span
        // comes from the block's min / max window.
        bool staged = false;
        int64_t gfirst = 0;
        if (fast && stage_len > 0) {
            int64_t lo = active ? g : INT64_MAX, hi = active ? g : -1;
            for (int o = 32; o > 0; o >>= 1) {
                const int64_t ol = __shfl_xor(lo, o, 64), oh = __shfl_xor(hi, o, 64);
                lo = ol < lo ? ol : lo;
                hi = oh > hi ? oh : hi;
            }
            const int wid = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) {
                rng[2 * wid] = lo;
                rng[2 * wid + 1] = hi;
            }
            __syncthreads();
            lo = rng[0];
            hi = rng[1];
            for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
                lo = rng[2 * k] < lo ? rng[2 * k] : lo;
                hi = rng[2 * k + 1] > hi ? rng[2 * k + 1] : hi;
            }
            __syncthreads();
            gfirst = lo;
            const int64_t glast = hi;
            staged = glast >= 0 && glast - gfirst + a.p.m <= stage_len;
            if (staged) {
                const int len = (int)(glast - gfirst) + a.p.m;
                for (int x = threadIdx.x; x < len; x += blockDim.x) {
                    Ts[x] = a.T[gfirst + x];
                    Ls[x] = a.Lraw[gfirst + x];
                    Us[x] = a.Uraw[gfirst + x];
                }
                __syncthreads();
            }
        }
        if (fast && staged) {
            const int64_t off = (active ? g : gfirst) - gfirst;
            verdict = glb_fast<METRIC>(a, Q, XG, MQ, HQ, Ts + off, Ls + off, Us + off,
                                       mu, sd, active, &total);
        } else if (fast) {
            verdict = glb_fast<METRIC>(a, Q, XG, MQ, HQ, a.T + g, a.Lraw + g,
                                       a.Uraw + g, mu, sd, active, &total);
        }
        bool keep = verdict > 0;
        if (active && verdict == 0) {
            const double v = glb_one<METRIC>(a.T + g, mu, sd, Q, LQ, UQ, a.Lraw + g, a.Uraw + g,
                                             a.p.m, a.p.w, a.p.arg, a.p.cutoff);
            if (vals) vals[ci] = v;
            keep = v <= a.p.cutoff;  // src/fsmeld.cpp:831
        }
        keep = keep && active;
        if (out) block_append(keep, g, total, out, out_total, out_count, sh);
        if (staged) __syncthreads();  // before the next iteration's staging
    }
}

static int verify_blocks() { return 2048; }
constexpr int kGlbStage = 1664;

__global__ void k_glb_queue_reset(unsigned long long* q) { *q = 0ull; }

hipError_t launch_glb(hipStream_t st, const VerifyArgs& a, const int64_t* in,
                      const unsigned long long* in_count, int64_t* out,
                      unsigned long long* out_count, double* out_total, double* vals,
                      unsigned long long* queue) {
    if (!queue) return hipErrorInvalidValue;
    const dim3 g(a.glb_blocks > 0 ? a.glb_blocks : verify_blocks()), b(256);
    hipLaunchKernelGGL(k_glb_queue_reset, dim3(1), dim3(1), 0, st, queue);
    const QueryTables& q = a.q;
    // staging: 3 arrays of up to kGlbStage doubles (39 KiB, four blocks per CU)
    // when the block's 256 windows leave at least 256 more of the span
    const int stage_len = (a.p.m + 512 <= kGlbStage) ? kGlbStage : 0;
    const size_t lds = sizeof(double) * 3 * (size_t)stage_len;
#define FSM_GLB(MM) \
    hipLaunchKernelGGL(k_glb<MM>, g, b, lds, st, a, q.qn, q.LQ, q.UQ, q.XG, q.MQ, q.HQ, in, \
                       in_count, out, out_count, out_total, vals, queue, stage_len)
    switch (a.p.metric) {
        case DTW: FSM_GLB(DTW); break;
        case ERP: FSM_GLB(ERP); break;
        case EDR: FSM_GLB(EDR); break;
        case LCSS: FSM_GLB(LCSS); break;
        case MSM: FSM_GLB(MSM); break;
        case SWALE: FSM_GLB(SWALE); break;
        default: return hipErrorInvalidValue;
    }
#undef FSM_GLB
    return hipGetLastError();
}

}  // namespace fsm
