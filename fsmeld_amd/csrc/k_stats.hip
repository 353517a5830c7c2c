// k_stats.hip -- per-window mean / std / invalid flag, and the target's
// sliding envelope, on gfx950.
//
// Window statistics must equal the reference's bits: every window's mean and
// std come from the running-sum recurrence of the macro-segment that
// evaluates it (src/mvmean.cpp:9-46), so one lane owns one segment and runs
// that recurrence in the reference's order; the per-window division and
// square root run afterwards in a parallel pass.  Build with -ffp-contract=off.
#include <algorithm>

#include "engine_internal.h"

namespace fsm {

// One lane per macro-segment runs the running-sum recurrence of
// src/mvmean.cpp:9-46 in the reference's order (its order fixes the bits);
// only the sums and the non-finite count are loop-carried, so this pass
// stores them (sum1 in MU, sum2 in SG; sum1 is replaced by the marker kBadSum
// while the window holds a non-finite sample) and k_window_finish turns every
// window's sums into mean / std / invalid flag in parallel
// (src/mvmean.cpp:30-44, the invalid test of src/fsmeld.cpp:110-129), with
// the same expressions.  Outputs are segment-major [S][kml].
//
// Long segments leave few lanes (2^24 samples at m = 2048 are 1170
// segments), so each lane hides its own load latency: the samples of the
// next kStatsU steps are loaded while the current kStatsU steps run (two
// register buffers), and the invalid flag rides in sum1's slot, two stores a
// step instead of three.
constexpr int kStatsU = 32;
// Segments per 64-lane wave (the other lanes exit at once).  Every lane
// streams its own segment, so each memory instruction touches one cache line
// per active lane; half-filled waves spread the segments over twice as many
// SIMDs (2^26 samples at m = 512: 0.90 -> 0.79 ms; 2^24 at m_k = 4096 /
// 8192 / 16384: 13-22% less; 16 or 8 lanes lose at m = 512, where the waves
// then outnumber the SIMDs, and gain at most 5% on the longest segments).
constexpr int kStatsLanes = 32;
// a quiet NaN whose payload no arithmetic produces (operations on finite
// values give the default NaN)
constexpr unsigned long long kBadSum = 0x7ff8dead0000beefull;

__global__ void __launch_bounds__(64) k_window_sums(const double* __restrict__ T, Geometry G,
                                                    double* __restrict__ MU,
                                                    double* __restrict__ SG, int lpw) {
    if ((int)threadIdx.x >= lpw) return;
    const int64_t s = (int64_t)blockIdx.x * lpw + threadIdx.x;
    if (s >= G.S) return;
    const double* a = (s < G.forsize) ? T + s * G.kml : T + (G.N - G.m_k);
    double* s1o = MU + s * G.kml;
    double* s2o = SG + s * G.kml;
    const int l = G.m;
    const double badv = __longlong_as_double((long long)kBadSum);
    double sum1 = 0.0, sum2 = 0.0;
    int bad = 0;
    auto first = [&](double v) {
        if (__builtin_isfinite(v)) {
            sum1 += v;
            sum2 += v * v;
        } else {
            ++bad;
        }
    };
    int i0 = 0;
    for (; i0 + kStatsU <= l; i0 += kStatsU) {
        double v[kStatsU];
#pragma unroll
        for (int u = 0; u < kStatsU; u++) v[u] = a[i0 + u];
#pragma unroll
        for (int u = 0; u < kStatsU; u++) first(v[u]);
    }
    for (; i0 < l; i0++) first(a[i0]);
    s1o[0] = bad > 0 ? badv : sum1;
    s2o[0] = sum2;
    const int64_t kml = G.kml;
    auto step = [&](int64_t i, double vo, double vi) {
        const bool of = __builtin_isfinite(vo);
        const bool nf = __builtin_isfinite(vi);
        const double out = of ? vo : 0.0;
        const double in = nf ? vi : 0.0;
        bad += (int)!nf - (int)!of;
        sum1 = sum1 - out + in;
        sum2 = sum2 - out * out + in * in;
        s1o[i] = bad > 0 ? badv : sum1;
        s2o[i] = sum2;
    };
    // window i drops sample i - 1 and takes sample i + l - 1
    double ao[kStatsU], ai[kStatsU], bo[kStatsU], bi[kStatsU];
    auto load = [&](double* vo, double* vi, int64_t base) {
#pragma unroll
        for (int u = 0; u < kStatsU; u++) {
            vo[u] = a[base + u - 1];
            vi[u] = a[base + u + l - 1];
        }
    };
    auto run = [&](const double* vo, const double* vi, int64_t base) {
#pragma unroll
        for (int u = 0; u < kStatsU; u++) step(base + u, vo[u], vi[u]);
    };
    int64_t i = 1;
    if (i + kStatsU <= kml) load(ao, ai, i);
    // a holds steps [i, i + U) here; b's loads fly while a's steps run, and
    // a's next loads while b's run
    while (i + 2 * kStatsU <= kml) {
        load(bo, bi, i + kStatsU);
        run(ao, ai, i);
        if (i + 3 * kStatsU <= kml) load(ao, ai, i + 2 * kStatsU);
        run(bo, bi, i + kStatsU);
        i += 2 * kStatsU;
    }
    if (i + kStatsU <= kml) {
        run(ao, ai, i);
        i += kStatsU;
    }
    for (; i < kml; i++) step(i, a[i - 1], a[i + l - 1]);
}

// mean = sum1 / m, std = sqrt(sum2 / m - mean^2), NaN when the window holds a
// non-finite sample; invalid when so, or std is not finite or <= 1e-12.  A
// thread takes four consecutive windows: 16-byte loads and stores of the
// sums, one 4-byte store of their flags (the buffers start 16-byte aligned).
__global__ void __launch_bounds__(256) k_window_finish(int64_t n, int m, double* __restrict__ MU,
                                                       double* __restrict__ SG,
                                                       uint8_t* __restrict__ INV) {
    const double dl = (double)m;
    const double nan = __builtin_nan("");
    auto one = [&](double s1, double s2, double* mi, double* sg) {
        const bool bad = (unsigned long long)__double_as_longlong(s1) == kBadSum;
        *mi = bad ? nan : s1 / dl;
        *sg = bad ? nan : sqrt(s2 / dl - *mi * *mi);
        return (uint32_t)(bad || !__builtin_isfinite(*sg) || *sg <= 1e-12);
    };
    const int64_t n4 = n / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4;
         q += (int64_t)gridDim.x * blockDim.x) {
        double2* m2 = reinterpret_cast<double2*>(MU + 4 * q);
        double2* s2 = reinterpret_cast<double2*>(SG + 4 * q);
        const double2 a0 = m2[0], a1 = m2[1], b0 = s2[0], b1 = s2[1];
        double mi[4], sg[4];
        const uint32_t f = one(a0.x, b0.x, &mi[0], &sg[0]) | (one(a0.y, b0.y, &mi[1], &sg[1]) << 8) |
                           (one(a1.x, b1.x, &mi[2], &sg[2]) << 16) |
                           (one(a1.y, b1.y, &mi[3], &sg[3]) << 24);
        m2[0] = double2{mi[0], mi[1]};
        m2[1] = double2{mi[2], mi[3]};
        s2[0] = double2{sg[0], sg[1]};
        s2[1] = double2{sg[2], sg[3]};
        *reinterpret_cast<uint32_t*>(INV + 4 * q) = f;
    }
    // the last n mod 4 windows
    const int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double mi, sg;
        INV[i] = (uint8_t)one(MU[i], SG[i], &mi, &sg);
        MU[i] = mi;
        SG[i] = sg;
    }
}

// Raw matching (paper App. D): no mean / std; a window is invalid only when
// it holds a non-finite sample.  Mean 0 and std 1 make every later stage's
// normalisation (x - 0) / 1 = x exact, so the cascade runs on raw values.
__global__ void __launch_bounds__(256) k_window_finish_raw(int64_t n, double* __restrict__ MU,
                                                           double* __restrict__ SG,
                                                           uint8_t* __restrict__ INV) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const bool bad = (unsigned long long)__double_as_longlong(MU[i]) == kBadSum;
        MU[i] = 0.0;
        SG[i] = 1.0;
        INV[i] = bad;
    }
}

hipError_t launch_window_stats(hipStream_t st, const double* T, const Geometry& G, double* MU,
                               double* SG, uint8_t* INV, hipEvent_t between, bool raw) {
    const int bs = 64;
    const int lpw = kStatsLanes;
    const int64_t blocks = (G.S + lpw - 1) / lpw;
    hipLaunchKernelGGL(k_window_sums, dim3((unsigned)blocks), dim3(bs), 0, st, T, G, MU, SG, lpw);
    if (between) {
        const hipError_t e = hipEventRecord(between, st);
        if (e != hipSuccess) return e;
    }
    const int64_t n = G.S * G.kml;
    if (((uintptr_t)MU | (uintptr_t)SG | (uintptr_t)INV) & 15) return hipErrorInvalidValue;
    const int64_t fb = std::min<int64_t>((n + 255) / 256, 8192);
    if (raw)
        hipLaunchKernelGGL(k_window_finish_raw, dim3((unsigned)fb), dim3(256), 0, st, n, MU, SG,
                           INV);
    else
        hipLaunchKernelGGL(k_window_finish, dim3((unsigned)fb), dim3(256), 0, st, n, G.m, MU, SG,
                           INV);
    return hipGetLastError();
}

// Centred envelope of the whole target, radius w, clipped at its ends and
// NaN-propagating: L[p] = min T[max(0,p-w) .. min(N-1,p+w)], U = max.
// For NaN-free ranges this equals lower_upper_lemire (src/lower_upper_lemire.cpp:8-44).
// A block stages a tile plus halo in LDS and builds the min/max over blocks
// of 2^K samples by K doubling passes; each output combines two overlapping
// 2^K blocks.
constexpr int kEnvTile = 1024;
constexpr int kEnvThreads = 256;

__global__ void __launch_bounds__(kEnvThreads) k_envelope(const double* __restrict__ T,
                                                          int64_t N, int w, int K,
                                                          double* __restrict__ L,
                                                          double* __restrict__ U) {
    extern __shared__ double lds[];
    const int span = kEnvTile + 2 * w;
    double* A = lds;         // running min over [x, x + 2^p)
    double* B = lds + span;  // running max
    const int64_t p0 = (int64_t)blockIdx.x * kEnvTile;
    const double inf = __builtin_inf();
    for (int x = threadIdx.x; x < span; x += blockDim.x) {
        const int64_t src = p0 - w + x;
        const bool in = src >= 0 && src < N;
        const double v = in ? T[src] : 0.0;
        A[x] = in ? v : inf;
        B[x] = in ? v : -inf;
    }
    __syncthreads();
    for (int p = 0; p < K; p++) {
        const int d = 1 << p;
        double ra[8], rb[8];  // span <= 8 * kEnvThreads is checked by the launcher
        int c = 0;
        for (int x = threadIdx.x; x < span; x += blockDim.x, c++) {
            const int y = x + d;
            ra[c] = (y < span) ? nmin(A[x], A[y]) : A[x];
            rb[c] = (y < span) ? nmax(B[x], B[y]) : B[x];
        }
        __syncthreads();
        c = 0;
        for (int x = threadIdx.x; x < span; x += blockDim.x, c++) {
            A[x] = ra[c];
            B[x] = rb[c];
        }
        __syncthreads();
    }
    const int len = 2 * w + 1;
    const int off = len - (1 << K);
    for (int x = threadIdx.x; x < kEnvTile; x += blockDim.x) {
        const int64_t pos = p0 + x;
        if (pos >= N) break;
        // window [pos - w, pos + w] is LDS [x, x + 2w]
        L[pos] = nmin(A[x], A[x + off]);
        U[pos] = nmax(B[x], B[x + off]);
    }
}

// Wide radii (2w + 1 > kEnvTile): the windows of the kEnvTile outputs
// [p0, p0 + T) of a block all contain the middle run C = [p0 + T - 1 - w,
// p0 + w].  Window p0 + d is C plus a suffix [p0 - w + d, p0 + T - 2 - w] of
// the left run and a prefix [p0 + w + 1, p0 + d + w] of the right run (T - 1
// samples each), so a block reduces C once (2w + 2 - T samples, coalesced)
// and scans the two runs in LDS.  Samples outside [0, N) are the identity.
__global__ void __launch_bounds__(kEnvThreads) k_envelope_wide(const double* __restrict__ T,
                                                               int64_t N, int w,
                                                               double* __restrict__ L,
                                                               double* __restrict__ U) {
    constexpr int TT = kEnvTile;
    constexpr int PER = TT / kEnvThreads;
    // [0]: left run, slot d = sample p0 - w + d (slot T - 1: identity),
    //      scanned into suffixes; [1]: right run, slot d = sample p0 + w + d
    //      (slot 0: identity), scanned into prefixes
    __shared__ double smin[2][TT], smax[2][TT];
    __shared__ double rmin[kEnvThreads / 64], rmax[kEnvThreads / 64];
    const int64_t p0 = (int64_t)blockIdx.x * TT;
    const double inf = __builtin_inf();
    const int tid = threadIdx.x;
    for (int d = tid; d < TT; d += kEnvThreads) {
        const int64_t sl = p0 - w + d, sr = p0 + w + d;
        const bool inl = d < TT - 1 && sl >= 0 && sl < N;
        const bool inr = d > 0 && sr >= 0 && sr < N;
        const double vl = inl ? T[sl] : 0.0, vr = inr ? T[sr] : 0.0;
        smin[0][d] = inl ? vl : inf;
        smax[0][d] = inl ? vl : -inf;
        smin[1][d] = inr ? vr : inf;
        smax[1][d] = inr ? vr : -inf;
    }
    // middle run, clipped to the target
    double cmin = inf, cmax = -inf;
    {
        int64_t c0 = p0 + TT - 1 - w, c1 = p0 + w;  // inclusive
        if (c0 < 0) c0 = 0;
        if (c1 > N - 1) c1 = N - 1;
        for (int64_t c = c0 + tid; c <= c1; c += kEnvThreads) {
            const double v = T[c];
            cmin = nmin(cmin, v);
            cmax = nmax(cmax, v);
        }
        for (int off = 32; off >= 1; off >>= 1) {
            cmin = nmin(cmin, __shfl_xor(cmin, off));
            cmax = nmax(cmax, __shfl_xor(cmax, off));
        }
        if ((tid & 63) == 0) {
            rmin[tid >> 6] = cmin;
            rmax[tid >> 6] = cmax;
        }
    }
    __syncthreads();
    cmin = rmin[0];
    cmax = rmax[0];
    for (int k = 1; k < kEnvThreads / 64; k++) {
        cmin = nmin(cmin, rmin[k]);
        cmax = nmax(cmax, rmax[k]);
    }
    // Hillis-Steele scans: suffixes of the left run, prefixes of the right
    for (int off = 1; off < TT; off <<= 1) {
        double a[PER], b[PER], c[PER], e[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int d = tid + u * kEnvThreads;
            const bool hs = d + off < TT, hp = d - off >= 0;
            a[u] = hs ? nmin(smin[0][d], smin[0][d + off]) : smin[0][d];
            b[u] = hs ? nmax(smax[0][d], smax[0][d + off]) : smax[0][d];
            c[u] = hp ? nmin(smin[1][d - off], smin[1][d]) : smin[1][d];
            e[u] = hp ? nmax(smax[1][d - off], smax[1][d]) : smax[1][d];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int d = tid + u * kEnvThreads;
            smin[0][d] = a[u];
            smax[0][d] = b[u];
            smin[1][d] = c[u];
            smax[1][d] = e[u];
        }
        __syncthreads();
    }
    for (int d = tid; d < TT; d += kEnvThreads) {
        const int64_t pos = p0 + d;
        if (pos >= N) break;
        L[pos] = nmin(nmin(smin[0][d], cmin), smin[1][d]);
        U[pos] = nmax(nmax(smax[0][d], cmax), smax[1][d]);
    }
}

hipError_t launch_envelope(hipStream_t st, const double* T, int64_t N, int w, double* L,
                           double* U) {
    if (w < 0) return hipErrorInvalidValue;
    const int span = kEnvTile + 2 * w;
    if (span > 8 * kEnvThreads) {
        // here 2w + 1 > kEnvTile, so every block's windows share a middle run
        const int64_t blocks = (N + kEnvTile - 1) / kEnvTile;
        hipLaunchKernelGGL(k_envelope_wide, dim3((unsigned)blocks), dim3(kEnvThreads), 0, st, T,
                           N, w, L, U);
        return hipGetLastError();
    }
    int K = 0;
    while ((2 << K) <= 2 * w + 1) K++;  // 2^K <= 2w+1 < 2^(K+1)
    const int64_t blocks = (N + kEnvTile - 1) / kEnvTile;
    const size_t lds = sizeof(double) * 2 * (size_t)span;
    hipLaunchKernelGGL(k_envelope, dim3((unsigned)blocks), dim3(kEnvThreads), lds, st, T, N, w,
                       K, L, U);
    return hipGetLastError();
}

}  // namespace fsm
