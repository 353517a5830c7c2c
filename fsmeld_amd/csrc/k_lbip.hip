// k_lbip.hip -- LB_IP stage on gfx950: one workgroup per macro-segment.
//
// Per segment (src/fsmeld.cpp:756-808):
//   1. min/max of window mean and std over the windows containing each
//      position (min_max_window, src/lower_upper_lemire.cpp:94-126): van Herk
//      prefix/suffix scans over blocks of m windows;
//   2. centred envelope L/U of the raw segment (lower_upper_lemire, :8-44),
//      taken from the target-wide envelope except within w of the edges;
//   3. normalised envelopes (src/lb_normalize.cpp:9-32);
//   4. thresholds t_u, t_l: the argmax over the segment's envelope values of
//      h(t) * G(t), the objective of src/find_t_lp.cpp (sorted values via a
//      bitonic sort, h(t) in closed form over the sorted query);
//   5. penalties h (src/q_to_h.cpp) in exact 64-bit fixed point, rounded
//      down, and their prefix sums CH;
//   6. masks g (src/lb.cpp) as bitmaps built with wavefront ballots;
//   7. per window j: z = sum over runs [a, b) of set mask bits inside
//      [j, j+m) of CH[b-j] - CH[a-j] -- the sliding inner product of
//      src/sip.cpp evaluated exactly, so the bound is a certified lower bound
//      of the FP64 penalties; then trans (src/fsmeld.cpp:84-106).
// Min / max / compare steps are exact, so for NaN-free segments steps 1-3
// reproduce the reference's bits.  Near NaN samples the min / max propagate
// NaN, which only clears mask bits (admissible).
#include <algorithm>

#include "engine_internal.h"

namespace fsm {

constexpr int kLdsKeys = 8192;

__host__ __device__ inline int npad_of(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Workgroup size per segment: 256 threads up to 4096 keys, 1024 beyond, so
// that every sort up to 16384 keys runs in registers with at most 16 keys per
// thread, and long segments spread their per-position work over more waves.
inline int lbip_threads(int npad) { return npad <= 4096 ? 256 : 1024; }

// The segment's sort keys live in LDS up to 8192 of them, 16384 with a
// 1024-thread workgroup (one per CU: 128 KiB of its 160), else in the
// block's global scratch.  The penalty prefix sums and mask bitmaps of the
// later steps reuse the keys' LDS space: the keys are dead once both
// thresholds are known.
__host__ __device__ inline bool keys_in_lds(int npad, int nt) {
    return npad <= (nt >= 1024 ? 16384 : kLdsKeys);
}

size_t lbip_slab_bytes(const Geometry& G) {
    const int nthr = G.m_k - 2 * G.m + 2;
    const int npad = npad_of(nthr);
    size_t b = sizeof(double) * (8 * (size_t)G.kml + 2 * (size_t)G.m_k);
    if (!keys_in_lds(npad, lbip_threads(npad))) b += sizeof(double) * (size_t)npad;
    return (b + 255) & ~(size_t)255;
}

size_t lbip_lds_bytes(const Geometry& G) {
    const int nthr = G.m_k - 2 * G.m + 2;
    const int npad = npad_of(nthr);
    const int nw = (G.m_k + 63) / 64;
    const size_t late = sizeof(int64_t) * 2 * (size_t)(G.m + 1) + sizeof(uint64_t) * 2 * (size_t)nw;
    const size_t k = keys_in_lds(npad, lbip_threads(npad)) ? sizeof(double) * (size_t)npad : 0;
    return std::max(late, k);
}

// ---------------------------------------------------------------- helpers ---
__device__ inline int upper_bound(const double* a, int n, double x) {  // first a[i] > x
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo;
}
__device__ inline int lower_bound(const double* a, int n, double x) {  // first a[i] >= x
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] >= x) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// sum over qs[a..b) of (q - t)^2 from prefix sums
__device__ inline double quad_range(const QueryTables& Q, int a, int b, double t) {
    const double s1 = Q.S1[b] - Q.S1[a];
    const double s2 = Q.S2[b] - Q.S2[a];
    const double c = (double)(b - a);
    return s2 - s1 * t * 2 + c * (t * t);
}

// Objective h(t) for the upper threshold (src/find_t_lp.cpp:583-825):
// DTW sum_{q>t}(q-t)^2; ERP sum_{q>t} min((q-t)^2,(q-g)^2);
// MSM sum_{q>t} min(q-t, c); EDR/LCSS #{q - eps > t}.
__device__ double h_upper(const QueryTables& Q, int metric, int m, double C, double t) {
    switch (metric) {
        case DTW: {
            const int i = upper_bound(Q.qs, m, t);
            return quad_range(Q, i, m, t);
        }
        case ERP: {
            const int i = upper_bound(Q.qs, m, t);
            if (t < C) {
                const double x = (t + C) * 0.5;
                const int f = upper_bound(Q.qs, m, x);
                return (Q.SD[m] - Q.SD[f]) + quad_range(Q, i, f > i ? f : i, t);
            }
            return quad_range(Q, i, m, t);
        }
        case MSM: {
            const int i1 = upper_bound(Q.qs, m, t);
            const int i2 = upper_bound(Q.TU, m, t);
            const int lo = i1 < i2 ? i1 : i2;
            const double mid = (Q.S1[i2] - Q.S1[lo]) - (double)(i2 - lo) * t;
            return C * (double)(m - i2) + mid;
        }
        default:  // EDR, LCSS
            return (double)(m - upper_bound(Q.TU, m, t));
    }
}

// Objective h(t) for the lower threshold (src/find_t_lp.cpp:218-582).
__device__ double h_lower(const QueryTables& Q, int metric, int m, double C, double t) {
    switch (metric) {
        case DTW: {
            const int i = lower_bound(Q.qs, m, t);
            return fabs(quad_range(Q, 0, i, t));
        }
        case ERP: {
            const int i = lower_bound(Q.qs, m, t);
            if (t > C) {
                const double x = (t + C) * 0.5;
                const int f = lower_bound(Q.qs, m, x);
                const int ff = f < i ? f : i;
                return Q.SD[ff] + fabs(quad_range(Q, ff, i, t));
            }
            return fabs(quad_range(Q, 0, i, t));
        }
        case MSM: {
            const int is = lower_bound(Q.TL, m, t);
            const int iq = lower_bound(Q.qs, m, t);
            const int hi = iq > is ? iq : is;
            const double mid = (Q.S1[hi] - Q.S1[is]) - (double)(hi - is) * t;
            return C * (double)is + fabs(mid);
        }
        default:
            return (double)lower_bound(Q.TL, m, t);
    }
}

// src/q_to_h.cpp:11-134 for one query point.
__device__ inline void penalties(int metric, double q, double tl, double tu, double C,
                                 double* hl, double* hu) {
    double vl = 0.0, vu = 0.0;
    switch (metric) {
        case MSM:
            if (q < tl) vl = fmin(C, tl - q);
            if (q > tu) vu = fmin(C, q - tu);
            break;
        case DTW:
            if (q < tl) vl = (tl - q) * (tl - q);
            if (q > tu) vu = (q - tu) * (q - tu);
            break;
        case ERP:
            if (q < tl)
                vl = ((q - tl) * (q - tl) > (q - C) * (q - C)) ? (q - C) * (q - C)
                                                                : (q - tl) * (q - tl);
            if (q > tu)
                vu = ((q - tu) * (q - tu) > (q - C) * (q - C)) ? (q - C) * (q - C)
                                                                : (q - tu) * (q - tu);
            break;
        default:  // EDR, LCSS
            if (q < tl - C) vl = 1.0;
            if (q > tu + C) vu = 1.0;
            break;
    }
    *hl = vl;
    *hu = vu;
}

// std::min(C, x) of the reference returns C when equal; fmin returns the
// same value, and penalties are never NaN for finite t.

// min / max of a window statistic over windows [lo, hi] (hi - lo < m) from
// the van Herk prefix value at hi (ph) and suffix value at lo (sl)
__device__ inline double stat_range(double ph, double sl, int64_t lo, int64_t hi, int m,
                                    bool is_min) {
    const int64_t blo = lo / m, bhi = hi / m;
    if (blo != bhi) return is_min ? nmin(sl, ph) : nmax(sl, ph);
    if (lo == blo * m) return ph;
    return sl;
}

// Bitonic sorts of the segment's envelope values, ascending.  NaN values are
// replaced by +inf or -inf before sorting (the caller counts them out), so the
// values compare totally and each comparator is one min and one max.  Every
// comparator is ascending: the first step of each merge pairs index i with
// its mirror i ^ (k - 1) in the k-block, the later steps with i ^ j.
// Equal values (+0 / -0 included) may leave in either order; the thresholds
// read values and ranks only, as std::sort leaves equal keys unordered.

// min / max of two values that are never NaN (a compare and two selects,
// without fmin's NaN quieting)
__device__ __forceinline__ double vmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double vmax(double a, double b) { return (b < a) ? a : b; }

// In LDS or global scratch, any power-of-two n.
__device__ void bitonic_sort(double* keys, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const bool mirror = j == (k >> 1);
            for (int q = threadIdx.x; q < (n >> 1); q += blockDim.x) {
                const int i = ((q / j) * (j << 1)) + (q % j);  // bit j of i clear
                const int p = mirror ? (i ^ (k - 1)) : (i + j);
                const double a = keys[i], b = keys[p];
                keys[i] = vmin(a, b);
                keys[p] = vmax(a, b);
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}

// n = blockDim.x * R values in registers: thread t holds [t R, t R + R).
// Distances below R stay inside a thread, below 64 R pair lanes of one wave
// (a mirror step reads the partner lane's register R - 1 - r), and only
// larger ones exchange through `keys` (LDS or global scratch) with a barrier.
// Leaves the sorted values in `keys`.
// In-thread step at distance J (MIRROR: the first step of a 2J merge);
// register indices are compile-time constants.
template <int R, int J, bool MIRROR>
__device__ __forceinline__ void cx_local(double* v) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        constexpr int M = MIRROR ? 2 * J - 1 : J;
        const int p = r ^ M;
        if ((r & J) == 0 && p < R) {
            const double a = v[r], b = v[p];
            v[r] = vmin(a, b);
            v[p] = vmax(a, b);
        }
    }
}

template <int R, int J>
__device__ __forceinline__ void cx_local(double* v, bool mirror) {
    if constexpr (J < R) {
        if (mirror) cx_local<R, J, true>(v);
        else cx_local<R, J, false>(v);
    }
}

template <int R>
__device__ void bitonic_sort_reg(double* keys, int n) {
    const int t = threadIdx.x;
    const int base = t * R;
    double v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = keys[base + r];
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const bool mirror = j == (k >> 1);
            if (j >= 64 * R) {
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; r++) keys[base + r] = v[r];
                __threadfence_block();
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int i = base + r, p = mirror ? (i ^ (k - 1)) : (i ^ j);
                    const double u = keys[p];
                    v[r] = (i < p) ? vmin(v[r], u) : vmax(v[r], u);
                }
            } else if (j >= R) {
                const bool lower = (t & (j / R)) == 0;
                if (mirror) {  // registers r and R - 1 - r trade with the partner lane
                    const int lm = k / R - 1;
#pragma unroll
                    for (int r = 0; r < R / 2; r++) {
                        const double a = __shfl_xor(v[R - 1 - r], lm, 64);
                        const double b = __shfl_xor(v[r], lm, 64);
                        v[r] = lower ? vmin(v[r], a) : vmax(v[r], a);
                        v[R - 1 - r] = lower ? vmin(v[R - 1 - r], b) : vmax(v[R - 1 - r], b);
                    }
                } else {
                    const int lm = j / R;
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const double u = __shfl_xor(v[r], lm, 64);
                        v[r] = lower ? vmin(v[r], u) : vmax(v[r], u);
                    }
                }
            } else {
                switch (j) {
                    case 1: cx_local<R, 1>(v, mirror); break;
                    case 2: cx_local<R, 2>(v, mirror); break;
                    case 4: cx_local<R, 4>(v, mirror); break;
                    case 8: cx_local<R, 8>(v, mirror); break;
                    default: break;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) keys[base + r] = v[r];
    __threadfence_block();
    __syncthreads();
}

// R = n / blockDim.x values per thread in registers, or 0: the LDS / scratch sort
template <int R>
__device__ void sort_keys(double* keys, int n) {
    if constexpr (R > 0) bitonic_sort_reg<R>(keys, n);
    else bitonic_sort(keys, n);
}

// register-sort class of a padded key count for nt threads (one kernel
// instantiation each): keys per thread, 0 for the LDS / scratch sort
inline int sort_class(int npad, int nt) {
    const int r = npad / nt;
    return (npad % nt == 0 && (r == 2 || r == 4 || r == 8 || r == 16)) ? r : 0;
}

// Block argmax with "first index wins" among equal scores; NaN never wins.
struct ArgBest {
    double v;
    int64_t i;
};
__device__ inline ArgBest better_max(ArgBest a, ArgBest b) {
    if (b.v != b.v) return a;
    if (a.v != a.v) return b;
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}
__device__ inline ArgBest better_min(ArgBest a, ArgBest b) {
    if (b.v != b.v) return a;
    if (a.v != a.v) return b;
    if (b.v < a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}

template <bool MAX>
__device__ ArgBest block_best(ArgBest x, ArgBest* sh) {
    for (int o = 32; o > 0; o >>= 1) {
        ArgBest y;
        y.v = __shfl_xor(x.v, o, 64);
        y.i = __shfl_xor(x.i, o, 64);
        x = MAX ? better_max(x, y) : better_min(x, y);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wid] = x;
    __syncthreads();
    ArgBest r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++)
        r = MAX ? better_max(r, sh[k]) : better_min(r, sh[k]);
    __syncthreads();
    return r;
}

// Block-wide sum of one int per thread (ends with a barrier).
__device__ int block_count(int c, int* sh) {
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
    __syncthreads();
    return r;
}

__device__ double block_sum(double x, double* sh) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wid] = x;
    __syncthreads();
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
    __syncthreads();
    return r;
}

// Exclusive scan of per-thread int64 totals; returns this thread's offset.
__device__ int64_t block_exscan(int64_t x, int64_t* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t inc = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    int64_t off = 0;
    for (int k = 0; k < wid; k++) off += sh[k];
    __syncthreads();
    return off + inc - x;
}

// First position p in [x, end) whose mask bit equals `want`, else end.
__device__ inline int find_bit(const uint64_t* bm, int x, int end, bool want) {
    while (x < end) {
        const int wi = x >> 6, b = x & 63;
        uint64_t word = bm[wi];
        if (!want) word = ~word;
        word >>= b;
        if (word) {
            const int r = x + (int)__builtin_ctzll(word);
            return r < end ? r : end;
        }
        x = (wi + 1) << 6;
    }
    return end;
}

// sum over runs [a, b) of set bits inside [j, j+m) of CH[b-j] - CH[a-j]
__device__ inline int64_t run_sum(const uint64_t* bm, const int64_t* CH, int j, int m) {
    const int end = j + m;
    int64_t acc = 0;
    int x = j;
    while (x < end) {
        const int a = find_bit(bm, x, end, true);
        if (a >= end) break;
        const int b = find_bit(bm, a, end, false);
        acc += CH[b - j] - CH[a - j];
        x = b;
    }
    return acc;
}

// ------------------------------------------------------------------ kernel ---
// Up to 8 values per thread the kernel fits four waves per SIMD, 16 three.
template <int R, int NT>
__global__ void __launch_bounds__(NT)
    __attribute__((amdgpu_waves_per_eu(NT == 256 ? (R <= 8 ? 4 : 3) : 1))) k_lbip(LbipArgs a) {
    extern __shared__ unsigned char lds_raw[];
    __shared__ ArgBest sh_best[NT / 64];
    __shared__ double sh_sum[NT / 64];
    __shared__ int64_t sh_scan[NT / 64];
    __shared__ int sh_cnt[NT / 64];

    const Geometry G = a.G;
    const QueryTables Q = a.q;
    const int m = G.m, mk = G.m_k, w = a.p.w, metric = a.p.metric;
    const double C = a.p.arg;
    const int kml = (int)G.kml;
    const int nthr = mk - 2 * m + 2;
    const int npad = npad_of(nthr);
    const int nw = (mk + 63) / 64;

    int64_t* CHl = (int64_t*)lds_raw;
    int64_t* CHu = CHl + (m + 1);
    uint64_t* GLw = (uint64_t*)(CHu + (m + 1));
    uint64_t* GUw = GLw + nw;
    double* slab = (double*)((char*)a.slab + (size_t)blockIdx.x * a.slab_bytes_per_block);
    double* PQ = slab;  // 8 arrays of kml: (2*stat + dir)
    double* Un = slab + 8 * (size_t)kml;
    double* Ln = Un + mk;
    double* keys = keys_in_lds(npad, NT) ? (double*)lds_raw : (Ln + mk);

    for (int64_t s = blockIdx.x; s < G.S; s += gridDim.x) {
        const bool tail = (s == G.forsize);
        const int64_t seg0 = tail ? G.N - mk : s * G.kml;
        const double* t = a.T + seg0;
        const double* MU = a.MU + s * G.kml;
        const double* SG = a.SG + s * G.kml;
        const uint8_t* INV = a.INV + s * G.kml;
        const int jfirst = tail ? G.tail_first : 0;
        const int64_t gbase = s * G.kml;

        // 1. van Herk prefix (dir 0) / suffix (dir 1) scans over blocks of m
        //    windows: one wave per (statistic, direction, block), 64 windows
        //    per step, min and max of the same statistic in one pass.  The
        //    NaN-propagating min / max are associative, so the values equal a
        //    sequential scan's.
        {
            const int nb = (kml + m - 1) / m;
            const int lane = threadIdx.x & 63;
            const double inf = __builtin_inf();
            for (int item = threadIdx.x >> 6; item < 4 * nb; item += NT / 64) {
                const int sg = item / (2 * nb);  // 0: mean, 1: std
                const int rem = item - sg * 2 * nb;
                const int dir = rem / nb;
                const int b = rem - dir * nb;
                const double* src = sg ? SG : MU;
                double* omin = PQ + (size_t)(2 * (2 * sg) + dir) * kml;
                double* omax = PQ + (size_t)(2 * (2 * sg + 1) + dir) * kml;
                const int j0 = b * m;
                const int j1 = (j0 + m < kml) ? j0 + m : kml;
                double cmin = inf, cmax = -inf;
                // the loads of up to kScanG chunks are issued before the
                // chunk-to-chunk carry chain that consumes them
                constexpr int kScanG = 4;
                for (int c0 = 0; c0 < j1 - j0; c0 += 64 * kScanG) {
                    double v[kScanG];
#pragma unroll
                    for (int u = 0; u < kScanG; u++) {
                        const int c = c0 + 64 * u;
                        // prefix: ascending windows; suffix: descending windows
                        const int j = (dir == 0) ? j0 + c + lane : j1 - 1 - c - lane;
                        const bool in = (dir == 0) ? (j < j1) : (j >= j0);
                        v[u] = in ? src[j] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < kScanG; u++) {
                        const int c = c0 + 64 * u;
                        if (c >= j1 - j0) break;
                        const int j = (dir == 0) ? j0 + c + lane : j1 - 1 - c - lane;
                        const bool in = (dir == 0) ? (j < j1) : (j >= j0);
                        double vmin = in ? v[u] : inf, vmax = in ? v[u] : -inf;
                        for (int o = 1; o < 64; o <<= 1) {
                            const double pmin = __shfl_up(vmin, o, 64);
                            const double pmax = __shfl_up(vmax, o, 64);
                            if (lane >= o) {
                                vmin = nmin(pmin, vmin);
                                vmax = nmax(pmax, vmax);
                            }
                        }
                        vmin = nmin(cmin, vmin);
                        vmax = nmax(cmax, vmax);
                        if (in) {
                            omin[j] = vmin;
                            omax[j] = vmax;
                        }
                        cmin = __shfl(vmin, 63, 64);
                        cmax = __shfl(vmax, 63, 64);
                    }
                }
            }
        }
        __threadfence_block();
        __syncthreads();

        // 2-3. envelopes, stat ranges, normalised envelopes; upper keys.
        //     Each thread takes kEnvG positions at a time and issues their
        //     loads (envelope, the eight prefix / suffix values) first.
        constexpr int kEnvG = 2;
        const double pinf = __builtin_inf();
        int ucnt = 0;
        for (int i0 = threadIdx.x; i0 < mk; i0 += kEnvG * NT) {
            double Lg[kEnvG], Ug[kEnvG], PS[kEnvG][8];
#pragma unroll
            for (int e = 0; e < kEnvG; e++) {
                const int i = i0 + e * NT;
                if (i >= mk) break;
                const bool edge = i < w || i >= mk - w;
                Lg[e] = edge ? 0.0 : a.Lraw[seg0 + i];
                Ug[e] = edge ? 0.0 : a.Uraw[seg0 + i];
                const int64_t lo = i - m + 1 > 0 ? i - m + 1 : 0;
                const int64_t hi = i < kml - 1 ? i : kml - 1;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    PS[e][2 * q] = PQ[(size_t)(2 * q) * kml + hi];          // prefix at hi
                    PS[e][2 * q + 1] = PQ[(size_t)(2 * q + 1) * kml + lo];  // suffix at lo
                }
            }
#pragma unroll
            for (int e = 0; e < kEnvG; e++) {
            const int i = i0 + e * NT;
            if (i >= mk) break;
            double Lv = Lg[e], Uv = Ug[e];
            if (i < w || i >= mk - w) {
                const int lo = i - w > 0 ? i - w : 0;
                const int hi = i + w < mk - 1 ? i + w : mk - 1;
                Lv = t[lo];
                Uv = t[lo];
                for (int x = lo + 1; x <= hi; x++) {
                    Lv = nmin(Lv, t[x]);
                    Uv = nmax(Uv, t[x]);
                }
            }
            const int64_t lo = i - m + 1 > 0 ? i - m + 1 : 0;
            const int64_t hi = i < kml - 1 ? i : kml - 1;
            const double mumin = stat_range(PS[e][0], PS[e][1], lo, hi, m, true);
            const double mumax = stat_range(PS[e][2], PS[e][3], lo, hi, m, false);
            const double sgmin = stat_range(PS[e][4], PS[e][5], lo, hi, m, true);
            const double sgmax = stat_range(PS[e][6], PS[e][7], lo, hi, m, false);
            const double nu = Uv - mumin;
            const double un = (nu >= 0) ? nu / sgmin : nu / sgmax;
            const double nl = Lv - mumax;
            const double ln = (nl <= 0) ? nl / sgmin : nl / sgmax;
            Un[i] = un;
            Ln[i] = ln;
            if (i >= m - 1 && i <= mk - m) {
                keys[i - (m - 1)] = (un != un) ? pinf : un;
                ucnt += (un == un);
            }
            }
        }
        for (int p = nthr + threadIdx.x; p < npad; p += blockDim.x) keys[p] = pinf;
        __threadfence_block();
        __syncthreads();

        // 4a. upper threshold: ascending order, G = rank + 1; the first nval
        //     sorted values are the non-NaN ones (NaN and padding sort last)
        const int nval_u = block_count(ucnt, sh_cnt);
        sort_keys<R>(keys, npad);
        double t_u;
        {
            const int nval = nval_u;
            ArgBest best{__builtin_nan(""), 0};
            for (int p = threadIdx.x; p < nval; p += blockDim.x) {
                const double tv = keys[p];
                const double sc = h_upper(Q, metric, m, C, tv) * (double)(p + 1);
                best = better_max(best, ArgBest{sc, p});
            }
            best = block_best<true>(best, sh_best);
            int64_t idx = best.i;
            if (nval > 0) {
                const double t0 = keys[0];
                const double s0 = h_upper(Q, metric, m, C, t0);
                if (s0 != s0) idx = 0;  // src/find_t_lp.cpp:206-216 keeps index 0
            }
            t_u = (nval > 0) ? keys[idx] : __builtin_nan("");
        }
        __syncthreads();

        // 4b. lower threshold: descending order, G = rank-from-top + 1; NaN
        //     and padding sort first
        int lcnt = 0;
        for (int i = threadIdx.x; i < nthr; i += blockDim.x) {
            const double ln = Ln[i + m - 1];
            keys[i] = (ln != ln) ? -pinf : ln;
            lcnt += (ln == ln);
        }
        for (int p = nthr + threadIdx.x; p < npad; p += blockDim.x) keys[p] = -pinf;
        __threadfence_block();
        __syncthreads();
        const int nval_l = block_count(lcnt, sh_cnt);
        sort_keys<R>(keys, npad);
        double t_l;
        {
            const int nval = nval_l;
            ArgBest best{__builtin_nan(""), 0};
            for (int r = threadIdx.x; r < nval; r += blockDim.x) {
                const double tv = keys[npad - 1 - r];
                const double sc = h_lower(Q, metric, m, C, tv) * (double)(r + 1);
                best = better_max(best, ArgBest{sc, r});
            }
            best = block_best<true>(best, sh_best);
            int64_t idx = best.i;
            if (nval > 0) {
                const double s0 = h_lower(Q, metric, m, C, keys[npad - 1]);
                if (s0 != s0) idx = 0;
            }
            t_l = (nval > 0) ? keys[npad - 1 - idx] : __builtin_nan("");
        }
        if (threadIdx.x == 0) {
            if (a.TUo) a.TUo[s] = t_u;
            if (a.TLo) a.TLo[s] = t_l;
        }

        // 5. penalties in fixed point: hI = floor(h * 2^F), exact prefix sums
        double hs = 0.0;
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            double hl, hu;
            penalties(metric, Q.qn[i], t_l, t_u, C, &hl, &hu);
            hs += hl + hu;
        }
        hs = block_sum(hs, sh_sum);
        const bool finite = __builtin_isfinite(hs);
        int F = 0;
        if (finite && hs > 0.0) {
            F = 60 - ilogb(hs);
            if (F > 1000) F = 1000;
        }
        {
            const int per = (m + NT - 1) / NT;
            const int i0 = threadIdx.x * per;
            const int i1 = (i0 + per < m) ? i0 + per : m;
            int64_t tl_sum = 0, tu_sum = 0;
            for (int i = i0; i < i1; i++) {
                double hl, hu;
                penalties(metric, Q.qn[i], t_l, t_u, C, &hl, &hu);
                if (finite) {
                    tl_sum += (int64_t)floor(ldexp(hl, F));
                    tu_sum += (int64_t)floor(ldexp(hu, F));
                }
            }
            const int64_t ol = block_exscan(tl_sum, sh_scan);
            const int64_t ou = block_exscan(tu_sum, sh_scan);
            int64_t cl = ol, cu = ou;
            if (threadIdx.x == 0) {
                CHl[0] = 0;
                CHu[0] = 0;
            }
            for (int i = i0; i < i1; i++) {
                double hl, hu;
                penalties(metric, Q.qn[i], t_l, t_u, C, &hl, &hu);
                if (finite) {
                    cl += (int64_t)floor(ldexp(hl, F));
                    cu += (int64_t)floor(ldexp(hu, F));
                }
                CHl[i + 1] = cl;
                CHu[i + 1] = cu;
            }
        }

        // 6. mask bitmaps: g_l = L_n >= t_l, g_u = U_n <= t_u
        {
            const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
            for (int wi = wid; wi < nw; wi += NT / 64) {
                const int pos = wi * 64 + lane;
                const bool in = pos < mk;
                const bool gl = in && (Ln[pos] >= t_l);
                const bool gu = in && (Un[pos] <= t_u);
                const uint64_t bl = __ballot(gl);
                const uint64_t bu = __ballot(gu);
                if (lane == 0) {
                    GLw[wi] = bl;
                    GUw[wi] = bu;
                }
            }
        }
        __syncthreads();

        // 7. per-window bound
        const double scale = ldexp(1.0, -F);
        ArgBest rep{__builtin_nan(""), -1};
        for (int j = jfirst + threadIdx.x; j < kml; j += blockDim.x) {
            const int64_t g = gbase + (j - jfirst);
            double lb;
            if (INV[j]) {
                lb = __builtin_nan("");
            } else if (!finite) {
                lb = __builtin_nan("");
            } else {
                const int64_t z = run_sum(GLw, CHl, j, m) + run_sum(GUw, CHu, j, m);
                const double zd = __ll2double_rd(z) * scale;
                if (metric == DTW || metric == ERP) lb = sqrt(fmax(0.0, zd));
                else if (metric == LCSS) lb = fmax(0.0, zd / (double)m);
                else lb = zd;
                if (!is_excluded(a.p, g)) rep = better_min(rep, ArgBest{lb, g});
            }
            a.LB[g] = lb;
        }
        if (a.rep_lb) {
            rep = block_best<false>(rep, sh_best);
            if (threadIdx.x == 0) {
                a.rep_lb[s] = rep.v;
                a.rep_idx[s] = (rep.v == rep.v) ? rep.i : -1;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_lbip(hipStream_t st, const LbipArgs& a) {
    const size_t lds = lbip_lds_bytes(a.G);
    const int npad = npad_of(a.G.m_k - 2 * a.G.m + 2);
    const int nt = lbip_threads(npad);
    const dim3 g((unsigned)a.blocks), b(nt);
#define FSM_LB(RR, TT) hipLaunchKernelGGL((k_lbip<RR, TT>), g, b, lds, st, a)
    if (nt == 256) {
        switch (sort_class(npad, 256)) {
            case 2: FSM_LB(2, 256); break;
            case 4: FSM_LB(4, 256); break;
            case 8: FSM_LB(8, 256); break;
            case 16: FSM_LB(16, 256); break;
            default: FSM_LB(0, 256); break;
        }
    } else {
        switch (sort_class(npad, 1024)) {
            case 8: FSM_LB(8, 1024); break;
            case 16: FSM_LB(16, 1024); break;
            default: FSM_LB(0, 1024); break;
        }
    }
#undef FSM_LB
    return hipGetLastError();
}

}  // namespace fsm
