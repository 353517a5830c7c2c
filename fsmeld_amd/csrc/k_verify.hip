// k_verify.hip -- candidate compaction after LB_IP on gfx950.
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
// Every step-th key of keys[0 .. n) into out (count = ceil(n / step)).
__global__ void k_gather_stride(const double* __restrict__ keys, int64_t n, int64_t step,
                                double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i * step < n) out[i] = keys[i * step];
}

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

hipError_t launch_gather_stride(hipStream_t st, const double* keys, int64_t n, int64_t step,
                                double* out) {
    const int64_t cnt = (n + step - 1) / step;
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_stride, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                       keys, n, step, out);
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
