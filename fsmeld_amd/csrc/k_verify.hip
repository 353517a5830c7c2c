// k_verify.hip -- candidate compaction after LB_IP on gfx950.
#include <cstdint>
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
// A block takes chunks of 256 * kCompactIpt windows; thread t tests the
// kCompactIpt consecutive windows base + kCompactIpt t .. + kCompactIpt - 1
// with vector loads (16 invalid flags in one 16-byte load, the bounds in
// 16-byte pairs), so the chunk's windows are in thread order.  A block scan
// of the threads' counts and one atomicAdd per chunk give every survivor its
// slot; the survivors pass through LDS so that the stores are contiguous.
// Each chunk's survivors are written ascending and contiguously.
constexpr int kCompactIpt = 16;

__global__ void __launch_bounds__(256) k_compact(Geometry G, QueryParams P,
                                                 const uint8_t* __restrict__ INV,
                                                 const double* __restrict__ LB, int use_lb,
                                                 int64_t* __restrict__ out,
                                                 unsigned long long* counters) {
    __shared__ int64_t sh_idx[256 * kCompactIpt];
    __shared__ int sh_wsum[256 / 64];
    __shared__ unsigned long long sh_base;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t chunk = 256 * kCompactIpt;
    const int64_t nchunks = (G.ts_n + chunk - 1) / chunk;
    const int64_t full = G.forsize * G.kml;  // windows below: stat_index(g) = g
    unsigned long long nvalid = 0;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int64_t g0 = c * chunk + (int64_t)tid * kCompactIpt;
        // flags and bounds of the thread's windows (1 / +inf past the end)
        uint8_t inv[kCompactIpt];
        double lb[kCompactIpt];
        if (g0 + kCompactIpt <= G.ts_n && g0 + kCompactIpt <= full) {
            const uint4 v = *reinterpret_cast<const uint4*>(INV + g0);  // g0 % 16 == 0
            const uint32_t wds[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int i = 0; i < kCompactIpt; i++) inv[i] = (uint8_t)(wds[i >> 2] >> (8 * (i & 3)));
        } else {
#pragma unroll
            for (int i = 0; i < kCompactIpt; i++) {
                const int64_t g = g0 + i;
                inv[i] = g < G.ts_n ? INV[stat_index(G, g)] : (uint8_t)1;
            }
        }
        if (use_lb) {
            if (g0 + kCompactIpt <= G.ts_n) {
                const double2* l2 = reinterpret_cast<const double2*>(LB + g0);
#pragma unroll
                for (int i = 0; i < kCompactIpt / 2; i++) {
                    const double2 v = l2[i];
                    lb[2 * i] = v.x;
                    lb[2 * i + 1] = v.y;
                }
            } else {
#pragma unroll
                for (int i = 0; i < kCompactIpt; i++)
                    lb[i] = g0 + i < G.ts_n ? LB[g0 + i] : 0.0;
            }
        }
        uint32_t keep = 0;
#pragma unroll
        for (int i = 0; i < kCompactIpt; i++) {
            const int64_t g = g0 + i;
            const bool valid = g < G.ts_n && !inv[i] && !is_excluded(P, g);
            nvalid += valid;
            const bool kp =
                valid && (!use_lb || !lbip_exceeds_cutoff(lb[i], P.cutoff, P.metric));
            keep |= (uint32_t)kp << i;
        }
        const int cnt = __builtin_popcount(keep);
        int inc = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) sh_wsum[wid] = inc;
        __syncthreads();
        int off = inc - cnt, tot = 0;
#pragma unroll
        for (int k = 0; k < 256 / 64; k++) {
            if (k < wid) off += sh_wsum[k];
            tot += sh_wsum[k];
        }
        if (tid == 0) sh_base = tot ? atomicAdd(&counters[0], (unsigned long long)tot) : 0ull;
        for (uint32_t b = keep; b; b &= b - 1) sh_idx[off++] = g0 + __builtin_ctz(b);
        __syncthreads();
        const unsigned long long base = sh_base;
        for (int x = tid; x < tot; x += 256) out[base + (unsigned long long)x] = sh_idx[x];
        __syncthreads();  // before the next chunk's scan and staging
    }
    for (int o = 32; o > 0; o >>= 1) nvalid += __shfl_xor(nvalid, o, 64);
    if (lane == 0 && nvalid) atomicAdd(&counters[1], nvalid);
}

hipError_t launch_compact(hipStream_t st, const Geometry& G, const QueryParams& P,
                          const uint8_t* INV, const double* LB, bool use_lb, int64_t* out,
                          unsigned long long* counters) {
    // the vector loads read the flags in 16-byte and the bounds in 16-byte
    // pieces from the buffers' starts
    if (((uintptr_t)INV & 15) || (use_lb && ((uintptr_t)LB & 15))) return hipErrorInvalidValue;
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
