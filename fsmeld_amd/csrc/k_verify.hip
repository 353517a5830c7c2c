// k_verify.hip -- candidate compaction and the GLB filter on gfx950.
//
// GLB runs one candidate per lane and evaluates every expression in the
// reference's operand order (src/glb.cpp:200-373), sums included, so its
// values and early-abandon decisions are bit-identical to it.
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

// ------------------------------------------------------------- compaction ---
// Each thread tests kCompactIpt consecutive windows; a block scan orders the
// survivors of its chunk, one atomicAdd per chunk claims their output slots,
// so every chunk's survivors are written ascending and contiguously.
constexpr int kCompactIpt = 16;

__global__ void __launch_bounds__(256) k_compact(Geometry G, QueryParams P,
                                                 const uint8_t* __restrict__ INV,
                                                 const double* __restrict__ LB, int use_lb,
                                                 int64_t* __restrict__ out,
                                                 unsigned long long* counters) {
    __shared__ int sh_tot[4];
    __shared__ unsigned long long sh_base;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t chunk = (int64_t)blockDim.x * kCompactIpt;
    const int64_t nchunks = (G.ts_n + chunk - 1) / chunk;
    unsigned long long nvalid = 0;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int64_t g0 = c * chunk + (int64_t)threadIdx.x * kCompactIpt;
        uint32_t bits = 0;
        for (int e = 0; e < kCompactIpt; e++) {
            const int64_t g = g0 + e;
            if (g < G.ts_n) {
                const bool valid = !INV[stat_index(G, g)] && !is_excluded(P, g);
                nvalid += valid;
                if (valid && (!use_lb || !lbip_exceeds_cutoff(LB[g], P.cutoff, P.metric)))
                    bits |= 1u << e;
            }
        }
        const int cnt = __builtin_popcount(bits);
        int inc = cnt;  // inclusive wave scan
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) sh_tot[wid] = inc;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
            if (k < wid) woff += sh_tot[k];
            tot += sh_tot[k];
        }
        if (threadIdx.x == 0) sh_base = tot ? atomicAdd(&counters[0], (unsigned long long)tot) : 0ull;
        __syncthreads();
        unsigned long long pos = sh_base + (unsigned long long)(woff + inc - cnt);
        while (bits) {
            const int e = __builtin_ctz(bits);
            bits &= bits - 1;
            out[pos++] = g0 + e;
        }
        __syncthreads();
    }
    for (int o = 32; o > 0; o >>= 1) nvalid += __shfl_xor(nvalid, o, 64);
    if (lane == 0 && nvalid) atomicAdd(&counters[1], nvalid);
}

hipError_t launch_compact(hipStream_t st, const Geometry& G, const QueryParams& P,
                          const uint8_t* INV, const double* LB, bool use_lb, int64_t* out,
                          unsigned long long* counters) {
    const int64_t chunk = 256 * kCompactIpt;
    int64_t blocks = (G.ts_n + chunk - 1) / chunk;
    if (blocks > 2048) blocks = 2048;
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

// Normalised samples (src/fsmeld.cpp:822-826).  A position whose normalised
// value lies strictly inside the other series' envelope adds nothing to any
// GLB sum, so the correctly rounded quotient (t - mu) / sd is computed only
// where that cannot be decided from a = fl(fl(t - mu) * fl(1 / sd)).  With
// u = 2^-53, |a - fl((t - mu) / sd)| <= 3.02 u |a|; decisions use 8 u |a|.
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
    // The reference tests early abandon after every position; positions that
    // add nothing leave the test unchanged, so it is evaluated once before the
    // loop (its first test) and after every contributing position.
    if (METRIC == DTW) {
        boundary = sqd(N.exact(t[0]), Q[0]) + sqd(Tl, Q[n - 1]);
        if (n > 2 && boundary + t2q > bsf * bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v < LQ[i]) t2q += sqd(v, LQ[i]);
            else if (v > UQ[i]) t2q += sqd(v, UQ[i]);
            if (boundary + t2q > bsf * bsf) return kGlbAbandon;
        }
    } else if (METRIC == ERP) {
        boundary = smin3(sqd(Q[n - 1], Tl), sqd(Q[n - 1], arg), sqd(Tl, arg));
        if (n > 2 && boundary + t2q > bsf * bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v > UQ[i]) t2q += mn(sqd(v, UQ[i]), sqd(v, arg));
            else if (v < LQ[i]) t2q += mn(sqd(v, LQ[i]), sqd(v, arg));
            if (boundary + t2q > bsf * bsf) return kGlbAbandon;
        }
    } else if (METRIC == EDR) {
        boundary = mn(edr_subcost(Q[n - 1], Tl, arg), 1.0);
        if (n > 2 && boundary + t2q > bsf) return kGlbAbandon;
        for (int i = 1; i < n - 1; ++i) {
            if (surely_inside(N.approx(t[i]), LQ[i], UQ[i])) continue;
            const double v = N.exact(t[i]);
            if (v > UQ[i]) t2q += edr_subcost(v, UQ[i], arg);
            else if (v < LQ[i]) t2q += edr_subcost(v, LQ[i], arg);
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
        const double T2 = N.exact(t[n - 2]), T0 = N.exact(t[0]);
        const int b = n - 1;
        if ((T2 >= Tl && Tl >= Q[b]) || (T2 <= Tl && Tl <= Q[b]) ||
            (Q[b - 1] <= Q[b] && Q[b] <= Tl) || (Q[b - 1] >= Q[b] && Q[b] >= Tl)) {
            boundary = fabs(Q[0] - T0) + mn(fabs(Q[b] - Tl), arg);
        } else {
            boundary = fabs(Q[0] - T0) + smin3(fabs(Q[b] - Tl), arg + fabs(Tl - T2),
                                               arg + fabs(Q[b] - Q[b - 1]));
        }
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

template <int METRIC>
__global__ void __launch_bounds__(256) k_glb(VerifyArgs a, const int64_t* __restrict__ in,
                                             const unsigned long long* __restrict__ in_count,
                                             int64_t* __restrict__ out,
                                             unsigned long long* out_count,
                                             double* __restrict__ vals) {
    __shared__ unsigned long long sh[256 / 64 + 1];
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
        if (out) block_append(keep, g, 0.0, out, nullptr, out_count, sh);
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

}  // namespace fsm
