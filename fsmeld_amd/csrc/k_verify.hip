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

}  // namespace fsm
