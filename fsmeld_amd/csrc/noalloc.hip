// noalloc.hip -- fsm_eld_pairs_noalloc: the exact distances of
// fsm_eld_pairs for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops that forbid hipMalloc).
//
// The caller passes the query, the candidate rows and the output in device
// memory and a stream.  The only other device memory the kernels need -- the
// query's gap-cost table (m doubles) and the work-queue counter of the
// register / wave band kernels -- lives in `scratch`, a file-scope device
// array of kScratchBytes that every GPU holds for the life of the process.
// Its device address is looked up once per device (hipGetSymbolAddress) and
// only that address is ever passed to kernels.  Every call on a GPU shares
// that GPU's array, so calls that overlap on one GPU race; the counter is
// cleared by the ELD launcher's own fill kernel before each use, and the gap
// costs are rewritten in full by every call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/fsmeld.h"
#include "engine_internal.h"

namespace fsm {

constexpr size_t kCounterBytes = 256;  // the queue counter, padded to the alignment
constexpr size_t kScratchBytes = kCounterBytes + sizeof(double) * FSM_NOALLOC_MAX_M;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

// the current device's copy of `scratch` (its device address)
hipError_t scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_addr[dev] = p;
    }
    *out = g_addr[dev];
    return hipSuccess;
}

// ERP's gap cost of each query sample, (q - g)^2, as fsm_eld_pairs computes it
__global__ void k_gap_costs(const double* __restrict__ q, int m, double g,
                            double* __restrict__ xg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) xg[i] = (q[i] - g) * (q[i] - g);
}

int fail(int code, const char* msg) { return report_error(code, msg); }

}  // namespace

}  // namespace fsm

extern "C" int fsm_eld_pairs_noalloc(int metric, int m, int w, double arg, const double* q,
                                     const double* cands, int64_t n, double* out,
                                     void* stream) {
    using namespace fsm;
    if (n < 0 || (n > 0 && (!q || !cands || !out)))
        return fail(FSM_ERR_ARG, "null device pointer or negative count");
    if (m < 2 || w < 0 || w >= m || metric < FSM_METRIC_DTW || metric > FSM_METRIC_SWALE ||
        (metric == FSM_METRIC_SWALE && arg < 0))
        return fail(FSM_ERR_ARG, "invalid m, w, metric or tolerance");
    if (!std::isfinite(arg)) return fail(FSM_ERR_ARG, "Invalid finite metric parameter");
    if (m > FSM_NOALLOC_MAX_M)
        return fail(FSM_ERR_ARG, "m exceeds FSM_NOALLOC_MAX_M, the scratch array's query length");
    if (eld_uses_scratch(metric, m, w))
        return fail(FSM_ERR_ARG,
                    "(m, w) needs the row-band kernel, whose per-thread band rows do not fit the "
                    "scratch array (w <= 127 and m <= 4094 are within it)");
    if (n == 0) return FSM_OK;
    void* base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) return fail(FSM_ERR_HIP, hipGetErrorString(e));
    unsigned long long* queue = static_cast<unsigned long long*>(base);
    double* xg = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kCounterBytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_gap_costs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, q, m,
                       arg, xg);
    if ((e = hipGetLastError()) != hipSuccess) return fail(FSM_ERR_HIP, hipGetErrorString(e));
    VerifyArgs va{};
    va.q.qn = q;
    va.q.XG = xg;
    va.p.metric = metric;
    va.p.m = m;
    va.p.w = w;
    va.p.arg = arg;
    e = launch_eld_pairs(st, va, cands, n, out, queue, nullptr, 0);
    if (e != hipSuccess) return fail(FSM_ERR_HIP, hipGetErrorString(e));
    return FSM_OK;
}
