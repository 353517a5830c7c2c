// k_verify.hip -- candidate compaction after LB_IP on gfx950.
#include <cstring>

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
// A block takes chunks of 256 * kCompactIpt windows; at step e its threads
// test the 256 consecutive windows base + 256 e + t (coalesced loads), and
// each wave's ballot marks the kept ones.  The (e, wave) masks, in that order,
// list the chunk's windows ascending, so a scan of their popcounts and one
// atomicAdd per chunk give every survivor its slot: each chunk's survivors
// are written ascending and contiguously.
constexpr int kCompactIpt = 16;

__global__ void __launch_bounds__(256) k_compact(Geometry G, QueryParams P,
                                                 const uint8_t* __restrict__ INV,
                                                 const double* __restrict__ LB, int use_lb,
                                                 int64_t* __restrict__ out,
                                                 unsigned long long* counters) {
    constexpr int kMasks = kCompactIpt * (256 / 64);
    static_assert(kMasks == 64, "one wave scans the masks");
    __shared__ uint64_t sh_mask[kMasks];
    __shared__ int sh_pre[kMasks];
    __shared__ unsigned long long sh_base;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t chunk = (int64_t)blockDim.x * kCompactIpt;
    const int64_t nchunks = (G.ts_n + chunk - 1) / chunk;
    const uint64_t below = (1ull << lane) - 1;  // lanes before this one
    unsigned long long nvalid = 0;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int64_t base = c * chunk;
        uint32_t bits = 0;  // bit e: window base + 256 e + threadIdx.x is kept
        for (int e = 0; e < kCompactIpt; e++) {
            const int64_t g = base + (int64_t)blockDim.x * e + threadIdx.x;
            bool keep = false;
            if (g < G.ts_n) {
                const bool valid = !INV[stat_index(G, g)] && !is_excluded(P, g);
                nvalid += valid;
                keep = valid && (!use_lb || !lbip_exceeds_cutoff(LB[g], P.cutoff, P.metric));
            }
            const uint64_t mask = __ballot(keep);
            if (lane == 0) sh_mask[e * (256 / 64) + wid] = mask;
            bits |= (uint32_t)keep << e;
        }
        __syncthreads();
        if (threadIdx.x < kMasks) {  // one wave: kMasks == 64
            const int cnt = __builtin_popcountll(sh_mask[threadIdx.x]);
            int inc = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(inc, o, 64);
                if (lane >= o) inc += y;
            }
            sh_pre[threadIdx.x] = inc - cnt;
            if (threadIdx.x == kMasks - 1)
                sh_base = inc ? atomicAdd(&counters[0], (unsigned long long)inc) : 0ull;
        }
        __syncthreads();
        while (bits) {
            const int e = __builtin_ctz(bits);
            bits &= bits - 1;
            const int k = e * (256 / 64) + wid;
            const unsigned long long pos =
                sh_base + (unsigned long long)sh_pre[k] + __builtin_popcountll(sh_mask[k] & below);
            out[pos] = base + (int64_t)blockDim.x * e + threadIdx.x;
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

// ------------------------------------------------- cutoff: tighter bound ---
// Entries of list[0 .. *count) whose sign * keys[i] <= tau, appended to out.
__global__ void __launch_bounds__(256) k_select_le(const int64_t* __restrict__ list,
                                                   const double* __restrict__ keys,
                                                   const unsigned long long* __restrict__ count,
                                                   double sign, double tau,
                                                   int64_t* __restrict__ out,
                                                   unsigned long long* out_count) {
    __shared__ unsigned long long sh[256 / 64 + 1];
    const int64_t n = (int64_t)*count;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t iters = (n + stride - 1) / stride;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t it = 0; it < iters; it++, i += stride) {
        const bool keep = i < n && sign * keys[i] <= tau;
        block_append(keep, keep ? list[i] : 0, 0.0, out, nullptr, out_count, sh);
    }
}

// --------------------------------------------------- radix select (top-k) ---
// Order-preserving 64-bit key of a double (finite values; larger key, larger
// value, -0 below +0).
__device__ inline uint64_t order_key(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}

// hist[d] += number of finite keys[0 .. *count) whose order key has the bits
// `prefix` under `pmask` and digit d = (key >> shift) & (2^width - 1),
// width <= 11.
constexpr int kDigitBins = 2048;
__global__ void __launch_bounds__(256) k_digit_hist(const double* __restrict__ keys,
                                                    const unsigned long long* __restrict__ count,
                                                    int shift, int width, uint64_t pmask,
                                                    uint64_t prefix,
                                                    unsigned long long* __restrict__ hist) {
    __shared__ unsigned int h[kDigitBins];
    for (int b = threadIdx.x; b < kDigitBins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const int64_t n = (int64_t)*count;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double x = keys[i];
        if (!__builtin_isfinite(x)) continue;
        const uint64_t k = order_key(x);
        if ((k & pmask) == prefix) atomicAdd(&h[(k >> shift) & ((1u << width) - 1)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kDigitBins; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

__global__ void k_zero_hist(unsigned long long* hist) {
    for (int b = threadIdx.x; b < kDigitBins; b += blockDim.x) hist[b] = 0ull;
}

// (list[i], keys[i]) for i < *count with keys[i] <= tau, appended to out / out_keys.
__global__ void __launch_bounds__(256) k_select_le_kv(const int64_t* __restrict__ list,
                                                      const double* __restrict__ keys,
                                                      const unsigned long long* __restrict__ count,
                                                      double tau, int64_t* __restrict__ out,
                                                      double* __restrict__ out_keys,
                                                      unsigned long long* out_count) {
    __shared__ unsigned long long sh[256 / 64 + 1];
    const int64_t n = (int64_t)*count;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t iters = (n + stride - 1) / stride;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t it = 0; it < iters; it++, i += stride) {
        const bool keep = i < n && keys[i] <= tau;
        block_append(keep, keep ? list[i] : 0, keep ? keys[i] : 0.0, out, out_keys, out_count,
                     sh);
    }
}

hipError_t kth_smallest(hipStream_t st, const double* keys, const unsigned long long* count,
                        int64_t k, unsigned long long* hist, double* kth, int64_t* n_finite) {
    // digits of the order key from the top: 11 bits five times, then 9
    const int shifts[6] = {53, 42, 31, 20, 9, 0};
    const int widths[6] = {11, 11, 11, 11, 11, 9};
    uint64_t prefix = 0, pmask = 0;
    int64_t rank = k - 1;  // 0-based rank among the finite keys
    *n_finite = 0;
    for (int p = 0; p < 6; p++) {
        hipLaunchKernelGGL(k_zero_hist, dim3(1), dim3(256), 0, st, hist);
        hipLaunchKernelGGL(k_digit_hist, dim3(1024), dim3(256), 0, st, keys, count, shifts[p],
                           widths[p], pmask, prefix, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        unsigned long long h[kDigitBins];
        e = hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (p == 0) {
            unsigned long long tot = 0;
            for (int b = 0; b < kDigitBins; b++) tot += h[b];
            *n_finite = (int64_t)tot;
            if ((int64_t)tot <= rank) return hipSuccess;  // fewer than k finite keys
        }
        int b = 0;
        for (; b < (1 << widths[p]); b++) {
            if ((int64_t)h[b] > rank) break;
            rank -= (int64_t)h[b];
        }
        const uint64_t dmask = ((1ull << widths[p]) - 1) << shifts[p];
        prefix |= ((uint64_t)b << shifts[p]) & dmask;
        pmask |= dmask;
    }
    const uint64_t u = (prefix >> 63) ? (prefix & ~(1ull << 63)) : ~prefix;
    std::memcpy(kth, &u, sizeof(u));
    return hipSuccess;
}

hipError_t launch_select_le_kv(hipStream_t st, const int64_t* list, const double* keys,
                               const unsigned long long* count, double tau, int64_t* out,
                               double* out_keys, unsigned long long* out_count) {
    hipLaunchKernelGGL(k_select_le_kv, dim3(1024), dim3(256), 0, st, list, keys, count, tau, out,
                       out_keys, out_count);
    return hipGetLastError();
}

hipError_t launch_select_le(hipStream_t st, const int64_t* list, const double* keys,
                            const unsigned long long* count, double sign, double tau,
                            int64_t* out, unsigned long long* out_count) {
    hipLaunchKernelGGL(k_select_le, dim3(1024), dim3(256), 0, st, list, keys, count, sign, tau,
                       out, out_count);
    return hipGetLastError();
}

}  // namespace fsm
