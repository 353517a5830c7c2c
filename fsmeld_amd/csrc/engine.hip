// engine.hip -- host side of the C ABI in include/fsmeld.h: device-resident
// targets, per-query orchestration of the kernels, result objects.
//
// Query flow (one HIP stream per target, one host synchronisation per call):
//   search: window stats (cached per m) -> envelope (cached per w) -> LB_IP
//           -> compact by the LB_IP test -> GLB -> ELD -> sorted matches
//   exact : window stats -> compact valid windows -> ELD
//   cutoff: search's stages plus a representative pass: the smallest-LB
//           window of every segment is verified first; the k-th smallest of
//           those exact distances is an upper bound xi0 of the k-th smallest
//           distance overall, a search at xi0 returns every window that can
//           rank, and the k smallest by (distance, index) are the answer --
//           the order the reference's max-heap leaves (src/fsmeld.cpp:376-418).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <mutex>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fsmeld.h"
#include "engine_internal.h"

using namespace fsm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? FSM_ERR_OOM : FSM_ERR_HIP,        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Device buffer that grows on demand.  Sizes are rounded up (64 KiB, or
// 64 MiB from 64 MiB on) so that requests that differ a little -- the window
// statistics of two lengths, S k_m_l entries each -- reuse the allocation:
// hipFree waits for the device, and a free / malloc per query stalled
// concurrent handles.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t g = bytes >= (64u << 20) ? (64u << 20) : (64u << 10);
        const size_t want = (bytes + g - 1) / g * g;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace

bool fsm::make_geometry(int64_t N, int m, Geometry* G) {
    long long sel = 1;
    while (sel < 8LL * m) sel *= 2;  // src/fsmeld.cpp:28-37
    if (N < 2LL * m - 1) return false;
    if (N < sel) sel = N;
    if (sel > (1LL << 30)) return false;
    G->N = N;
    G->m = m;
    G->m_k = (int)sel;
    G->kml = G->m_k - m + 1;
    G->ts_n = N - m + 1;
    G->forsize = G->ts_n / G->kml;
    G->tail = G->ts_n % G->kml;
    G->tail_first = (int)(G->kml - G->tail);
    G->S = G->forsize + (G->tail > 0 ? 1 : 0);
    return true;
}

namespace {

// Centred clipped envelope on the host for the query (lower_upper_lemire,
// src/lower_upper_lemire.cpp:8-44; the query holds only finite values).
void host_envelope(const std::vector<double>& a, int r, std::vector<double>& L,
                   std::vector<double>& U) {
    const int n = (int)a.size();
    L.assign(n, 0.0);
    U.assign(n, 0.0);
    std::vector<int> ql(n), qu(n);
    int hl = 0, tl = -1, hu = 0, tu = -1, out = 0;
    auto push = [&](int i) {
        while (hl <= tl && a[ql[tl]] >= a[i]) tl--;
        ql[++tl] = i;
        while (hu <= tu && a[qu[tu]] <= a[i]) tu--;
        qu[++tu] = i;
    };
    const int k = 2 * r + 1;
    for (int i = 0; i < n + r; i++) {
        if (i < n) push(i);
        if (i >= r) {
            while (ql[hl] <= i - k) hl++;
            while (qu[hu] <= i - k) hu++;
            L[out] = a[ql[hl]];
            U[out] = a[qu[hu]];
            out++;
        }
    }
}

struct HostQuery {
    std::vector<double> qn, LQ, UQ, qs, S1, S2, SD, TU, TL, XG, MQ, HQ;
};


// src/fsmeld.cpp:622-643 (normalisation and sorted copy), :721 (envelope).
// Raw matching (paper App. D) takes the query as it is.
void prepare_query(const double* q_raw, int m, int w, double arg, int metric, HostQuery* H,
                   bool raw = false) {
    H->qn.resize(m);
    if (raw) {
        for (int i = 0; i < m; i++) H->qn[i] = q_raw[i];
    } else {
        double ex = 0.0, ex2 = 0.0;
        for (int i = 0; i < m; i++) {
            ex += q_raw[i];
            ex2 += q_raw[i] * q_raw[i];
        }
        const double mean = ex / m;
        double sd = ex2 / m;
        sd = std::sqrt(sd - mean * mean);
        for (int i = 0; i < m; i++) H->qn[i] = (q_raw[i] - mean) / sd;
    }
    host_envelope(H->qn, w, H->LQ, H->UQ);
    H->XG.resize(m);
    for (int i = 0; i < m; i++) H->XG[i] = (H->qn[i] - arg) * (H->qn[i] - arg);  // dist(x, g)
    // envelope midpoints and half widths, the half width widened by eps
    // (EDR / LCSS) or divided by c (MSM)
    H->MQ.resize(m);
    H->HQ.resize(m);
    const double rho = glb_msm_rho(arg);
    for (int i = 0; i < m; i++) {
        H->MQ[i] = (H->LQ[i] + H->UQ[i]) * 0.5;
        const double h = (H->UQ[i] - H->LQ[i]) * 0.5;
        H->HQ[i] = (metric == EDR || metric == LCSS) ? h + arg : (metric == MSM ? h * rho : h);
    }
    H->qs = H->qn;
    std::sort(H->qs.begin(), H->qs.end());
    H->S1.assign(m + 1, 0.0);
    H->S2.assign(m + 1, 0.0);
    H->SD.assign(m + 1, 0.0);
    H->TU.assign(m, 0.0);
    H->TL.assign(m, 0.0);
    for (int i = 0; i < m; i++) {
        const double q = H->qs[i];
        H->S1[i + 1] = H->S1[i] + q;
        H->S2[i + 1] = H->S2[i] + q * q;
        H->SD[i + 1] = H->SD[i] + (q - arg) * (q - arg);
        H->TU[i] = q - arg;
        H->TL[i] = q + arg;
    }
}

// Argument checks of src/main.cpp:46-99 that concern the engine.
int validate(const fsm_query_desc* d, int64_t N, const double* q, int64_t qlen, bool need_mode) {
    if (!d) return fail(FSM_ERR_ARG, "null query descriptor");
    if (need_mode && (d->mode < FSM_MODE_SEARCH || d->mode > FSM_MODE_EXACT_TOPK))
        return fail(FSM_ERR_ARG, "Unknown command");
    if (d->m < 2 || d->m > (1 << 27))
        return fail(FSM_ERR_ARG, "m must be at least 2 and fit the FFT indexing range.");
    if (d->metric < FSM_METRIC_DTW || d->metric > FSM_METRIC_SWALE)
        return fail(FSM_ERR_ARG, "Unsupported metric");
    if (d->w < 0 || d->w >= d->m) return fail(FSM_ERR_ARG, "w must satisfy 0 <= w < m.");
    if (!std::isfinite(d->arg)) return fail(FSM_ERR_ARG, "Invalid finite metric parameter");
    if ((d->metric == FSM_METRIC_EDR || d->metric == FSM_METRIC_LCSS ||
         d->metric == FSM_METRIC_MSM || d->metric == FSM_METRIC_SWALE) && d->arg < 0)
        return fail(FSM_ERR_ARG, "Matching tolerance and MSM cost must be non-negative.");
    if (N < 2LL * d->m - 1)
        return fail(FSM_ERR_ARG, "Target must contain at least 2*m-1 complete double values.");
    if (!q || qlen < d->m) return fail(FSM_ERR_ARG, "Query must begin with m finite numeric values.");
    double sum = 0, sq = 0;
    for (int i = 0; i < d->m; i++) {
        if (!std::isfinite(q[i]))
            return fail(FSM_ERR_ARG, "Query must begin with m finite numeric values.");
        sum += q[i];
        sq += q[i] * q[i];
    }
    const double mean = sum / d->m;
    const double sigma = std::sqrt(sq / d->m - mean * mean);
    if (d->raw != 0 && d->raw != 1) return fail(FSM_ERR_ARG, "Unknown normalization mode");
    if (d->lbip_f32 != 0 && d->lbip_f32 != 1) return fail(FSM_ERR_ARG, "lbip_f32 must be 0 or 1");
    if (!d->raw && (!std::isfinite(sigma) || sigma <= 1e-12))
        return fail(FSM_ERR_ARG,
                    "Query cannot be z-normalized: non-finite or near-zero standard deviation.");
    const bool excl = !(d->exclude_begin == -1 && d->exclude_end == -1);
    if (excl && (d->exclude_begin < 0 || d->exclude_end <= d->exclude_begin))
        return fail(FSM_ERR_ARG, "Exclusion requires 0 <= start < end.");
    if (need_mode && (d->mode == FSM_MODE_CUTOFF || d->mode == FSM_MODE_EXACT_TOPK)) {
        if (d->k < 1 || (int64_t)d->k > N - d->m + 1)
            return fail(FSM_ERR_ARG, "k exceeds the candidate count or supported integer range.");
    } else if (need_mode) {
        if (!std::isfinite(d->cutoff)) return fail(FSM_ERR_ARG, "Invalid finite cutoff");
        if (d->cutoff < 0) return fail(FSM_ERR_ARG, "cutoff must be non-negative.");
    }
    return FSM_OK;
}

// Elapsed time between two recorded events.  A failure (an event this query
// did not record) would stay behind as the thread's last HIP error and be
// returned by a later launch's hipGetLastError, so it is cleared here.
double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return (double)ms;
}

}  // namespace

// Window statistics for one m and the target envelope for one w depend on
// the target alone, so all handles of a target share them.  A key computes
// into the handle's own buffers (reused from query to query, as when nothing
// repeats) unless it was requested within the last kCacheReuse requests of
// its kind: then it gets a shared entry, made on the stream of the handle that
// needs it and published with an event the other handles' streams wait on.
// (A new entry's allocation costs more than the statistics it saves once, so
// keys that come back only after many others stay private.)  Beyond the byte
// cap (half of device memory) a new entry takes over the buffers of the least
// recently used entry no handle is using, so a cache that is full allocates
// and frees nothing.
struct CacheEntry {
    DevBuf a, b, c;  // statistics: mean, std, invalid flag; envelope: L, U
    hipEvent_t ready = nullptr;
    uint64_t bytes = 0, tick = 0;
    ~CacheEntry() {
        a.release();
        b.release();
        c.release();
        if (ready) (void)hipEventDestroy(ready);
    }
};
struct TargetCache {
    // launch shapes the autotuner chose (fsm_tuning_set / fsm_autotune); every
    // value 0 means the built-in rule.  Written while no query runs.
    fsm_tuning tune{};
    std::mutex mu;
    std::map<int, std::shared_ptr<CacheEntry>> stats, env;  // by m, by w
    std::map<int, uint64_t> seen_stats, seen_env;            // last request, by key
    uint64_t req_stats = 0, req_env = 0;                     // requests so far
    uint64_t tick = 0, bytes = 0, cap = 0;
    // the least recently used entry of mp no handle is using, taken out of
    // the cache (null if every entry is in use)
    std::shared_ptr<CacheEntry> take_idle(std::map<int, std::shared_ptr<CacheEntry>>& mp) {
        auto best = mp.end();
        for (auto it = mp.begin(); it != mp.end(); ++it)
            if (it->second.use_count() == 1 &&
                (best == mp.end() || it->second->tick < best->second->tick))
                best = it;
        if (best == mp.end()) return nullptr;
        std::shared_ptr<CacheEntry> e = std::move(best->second);
        mp.erase(best);
        bytes -= e->bytes;
        return e;
    }
    // free every entry no handle is using
    void drop_idle() {
        for (auto* mp : {&stats, &env})
            for (auto it = mp->begin(); it != mp->end();) {
                if (it->second.use_count() == 1) {
                    bytes -= it->second->bytes;
                    it = mp->erase(it);
                } else {
                    ++it;
                }
            }
    }
};

// A handle owns a stream, events and every working buffer; the device copy of
// the target's samples and the cache of statistics and envelopes are shared by
// all handles made from one another with fsm_target_share (and freed with the
// last of them).
struct fsm_target {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t N = 0;
    std::shared_ptr<DevBuf> T;
    std::shared_ptr<TargetCache> cache;
    int stats_m = -1;                // m, or -m for raw matching's statistics
    std::shared_ptr<CacheEntry> st;  // MU, SG, INV of stats_m
    int env_w = -1;
    std::shared_ptr<CacheEntry> env;  // Lr, Ur of env_w
    CacheEntry own_st, own_env;       // this handle's buffers for first requests
    DevBuf lb, list_a, list_b, out_idx, out_dist, slab, rep_lb, rep_idx, tu, tl, qtab,
        scratch, counters, aux, hist;
    hipEvent_t ev[8] = {};
    // false: every query recomputes its window statistics and envelope into
    // this handle's own buffers (fsm_target_reuse_stats)
    bool reuse = true;
    // fsm_query_batch's handles, kept between calls
    std::mutex pool_mu;
    std::vector<fsm_target*> pool;
};

struct fsm_result {
    std::vector<int64_t> idx;
    std::vector<double> dist;
    fsm_stats st{};
};

namespace {

int upload_query(fsm_target* t, const HostQuery& H, int m, QueryTables* Q) {
    const size_t stride = (size_t)m + 1;
    HIP_TRY(t->qtab.ensure(sizeof(double) * stride * 12));
    std::vector<double> buf(stride * 12, 0.0);
    const std::vector<double>* parts[12] = {&H.qn, &H.LQ, &H.UQ, &H.qs, &H.S1, &H.S2,
                                            &H.SD, &H.TU, &H.TL, &H.XG, &H.MQ, &H.HQ};
    for (int k = 0; k < 12; k++)
        std::copy(parts[k]->begin(), parts[k]->end(), buf.begin() + k * stride);
    HIP_TRY(hipMemcpyAsync(t->qtab.p, buf.data(), sizeof(double) * buf.size(),
                           hipMemcpyHostToDevice, t->stream));
    double* d = t->qtab.as<double>();
    Q->qn = d;
    Q->LQ = d + stride;
    Q->UQ = d + 2 * stride;
    Q->qs = d + 3 * stride;
    Q->S1 = d + 4 * stride;
    Q->S2 = d + 5 * stride;
    Q->SD = d + 6 * stride;
    Q->TU = d + 7 * stride;
    Q->TL = d + 8 * stride;
    Q->XG = d + 9 * stride;
    Q->MQ = d + 10 * stride;
    Q->HQ = d + 11 * stride;
    Q->qabs = 0.0;
    for (double v : H.qn) Q->qabs = std::max(Q->qabs, std::fabs(v));
    // the staging vector must outlive the async copy
    HIP_TRY(hipStreamSynchronize(t->stream));
    return FSM_OK;
}

// Key k of map mp for this handle: a shared entry (this stream waits for it),
// or one made by fill() on this stream, shared when k was requested before
// and otherwise in the handle's own buffers (own, wrapped without ownership).
// *made tells whether fill ran.
constexpr uint64_t kCacheReuse = 64;
template <class F>
int cached(fsm_target* t, std::map<int, std::shared_ptr<CacheEntry>> TargetCache::*mp,
           std::map<int, uint64_t> TargetCache::*seen, uint64_t TargetCache::*req, int k,
           uint64_t bytes, CacheEntry* own, F fill, std::shared_ptr<CacheEntry>* out,
           bool* made) {
    TargetCache& C = *t->cache;
    std::shared_ptr<CacheEntry> e;
    *made = false;
    {
        std::lock_guard<std::mutex> lk(C.mu);
        const uint64_t now = ++(C.*req);
        auto sn = (C.*seen).find(k);
        const bool recent = sn != (C.*seen).end() && now - sn->second <= kCacheReuse;
        (C.*seen)[k] = now;
        auto it = (C.*mp).find(k);
        if (it != (C.*mp).end()) {
            e = it->second;
            e->tick = ++C.tick;
        } else if (recent && bytes <= C.cap) {
            // Idle entries make room, least recently used first (one of the
            // same kind is recycled, the others freed).  With every entry in
            // use the key stays private, so the cache never exceeds its cap.
            std::shared_ptr<CacheEntry> reuse;
            while (C.bytes + bytes > C.cap) {
                std::shared_ptr<CacheEntry> x = C.take_idle(C.*mp);
                if (!x) x = C.take_idle(mp == &TargetCache::stats ? C.env : C.stats);
                if (!x) break;
                if (!reuse && x->ready) reuse = std::move(x);
            }
            if (C.bytes + bytes <= C.cap) {
                // made, launched and its event recorded under the lock:
                // whoever finds it afterwards waits on a recorded event
                e = reuse ? std::move(reuse) : std::make_shared<CacheEntry>();
                if (!e->ready) HIP_TRY(hipEventCreateWithFlags(&e->ready, hipEventDisableTiming));
                const int rc = fill(*e);
                if (rc) return rc;
                HIP_TRY(hipEventRecord(e->ready, t->stream));
                e->bytes = bytes;
                e->tick = ++C.tick;
                (C.*mp)[k] = e;
                C.bytes += bytes;
                *made = true;
            }
        }
    }
    if (!e) {  // first request: the handle's own buffers
        const int rc = fill(*own);
        if (rc) return rc;
        *out = std::shared_ptr<CacheEntry>(std::shared_ptr<CacheEntry>(), own);
        *made = true;
        return FSM_OK;
    }
    if (!*made) HIP_TRY(hipStreamWaitEvent(t->stream, e->ready, 0));
    *out = std::move(e);
    // (the handle keeps its own buffers, which fsm_query_bytes counts, for
    // its next private length: freeing them here and allocating them again
    // made every switch between shared and private entries two device-wide
    // waits)
    return FSM_OK;
}

// ev[6] marks the end of the statistics' first kernel when they are computed.
// Raw matching's statistics (mean 0, std 1, invalid = a non-finite sample)
// are cached under the key -m.
int ensure_stats(fsm_target* t, const Geometry& G, bool* computed, bool raw = false) {
    *computed = false;
    const int key = raw ? -G.m : G.m;
    if (t->reuse && t->st && t->stats_m == key) return FSM_OK;
    const size_t n = (size_t)G.S * (size_t)G.kml;
    auto fill = [&](CacheEntry& e) {
        HIP_TRY(e.a.ensure(sizeof(double) * n));
        HIP_TRY(e.b.ensure(sizeof(double) * n));
        HIP_TRY(e.c.ensure(n));
        HIP_TRY(launch_window_stats(t->stream, t->T->as<double>(), G, e.a.as<double>(),
                                    e.b.as<double>(), e.c.as<uint8_t>(), t->ev[6], raw));
        return FSM_OK;
    };
    t->st.reset();  // before own buffers may be refilled
    if (!t->reuse) {
        const int rc = fill(t->own_st);
        if (rc) return rc;
        t->st = std::shared_ptr<CacheEntry>(std::shared_ptr<CacheEntry>(), &t->own_st);
        *computed = true;
    } else {
        const int rc = cached(t, &TargetCache::stats, &TargetCache::seen_stats,
                              &TargetCache::req_stats, key, 17 * (uint64_t)n, &t->own_st, fill,
                              &t->st, computed);
        if (rc) return rc;
    }
    t->stats_m = key;
    return FSM_OK;
}

int ensure_env(fsm_target* t, int w, bool* computed = nullptr) {
    if (computed) *computed = false;
    if (t->reuse && t->env && t->env_w == w) return FSM_OK;
    auto fill = [&](CacheEntry& e) {
        HIP_TRY(e.a.ensure(sizeof(double) * (size_t)t->N));
        HIP_TRY(e.b.ensure(sizeof(double) * (size_t)t->N));
        HIP_TRY(launch_envelope(t->stream, t->T->as<double>(), t->N, w, e.a.as<double>(),
                                e.b.as<double>()));
        return FSM_OK;
    };
    bool made = false;
    t->env.reset();
    if (!t->reuse) {
        const int rc = fill(t->own_env);
        if (rc) return rc;
        t->env = std::shared_ptr<CacheEntry>(std::shared_ptr<CacheEntry>(), &t->own_env);
        made = true;
    } else {
        const int rc = cached(t, &TargetCache::env, &TargetCache::seen_env,
                              &TargetCache::req_env, w, 16 * (uint64_t)t->N, &t->own_env, fill,
                              &t->env, &made);
        if (rc) return rc;
    }
    t->env_w = w;
    if (computed) *computed = made;
    return FSM_OK;
}

// lb_out / tu_out / tl_out: device buffers that take the bounds and
// thresholds instead of the target's own (the device-memory stage entry)
int run_lbip(fsm_target* t, const Geometry& G, const QueryTables& Q, const QueryParams& P,
             bool with_rep, double* lb_out = nullptr, double* tu_out = nullptr,
             double* tl_out = nullptr) {
    HIP_TRY(t->lb.ensure(sizeof(double) * (size_t)G.ts_n));
    HIP_TRY(t->tu.ensure(sizeof(double) * (size_t)G.S));
    HIP_TRY(t->tl.ensure(sizeof(double) * (size_t)G.S));
    HIP_TRY(t->rep_lb.ensure(sizeof(double) * (size_t)G.S));
    HIP_TRY(t->rep_idx.ensure(sizeof(int64_t) * (size_t)G.S));
    const size_t per = lbip_slab_bytes(G);
    int blocks = (int)std::min<int64_t>(G.S, per > (1u << 20) ? 1024 : 2048);
    HIP_TRY(t->slab.ensure(per * (size_t)blocks));
    LbipArgs a{};
    a.T = t->T->as<double>();
    a.MU = t->st->a.as<double>();
    a.SG = t->st->b.as<double>();
    a.INV = t->st->c.as<uint8_t>();
    a.Lraw = t->env ? t->env->a.as<double>() : nullptr;
    a.Uraw = t->env ? t->env->b.as<double>() : nullptr;
    a.q = Q;
    a.p = P;
    a.G = G;
    a.LB = t->lb.as<double>();
    a.TUo = t->tu.as<double>();
    a.TLo = t->tl.as<double>();
    a.rep_lb = with_rep ? t->rep_lb.as<double>() : nullptr;
    a.rep_idx = with_rep ? t->rep_idx.as<int64_t>() : nullptr;
    if (lb_out) {
        a.LB = lb_out;
        a.TUo = tu_out;
        a.TLo = tl_out;
    }
    a.slab = t->slab.p;
    a.slab_bytes_per_block = per;
    a.blocks = blocks;
    const int cls = lbip_tune_class(G);
    a.nt = cls >= 0 ? t->cache->tune.lbip_threads[cls] : 0;
    HIP_TRY(launch_lbip(t->stream, a));
    return FSM_OK;
}

VerifyArgs verify_args(fsm_target* t, const Geometry& G, const QueryTables& Q,
                       const QueryParams& P) {
    VerifyArgs v{};
    v.T = t->T->as<double>();
    v.MU = t->st->a.as<double>();
    v.SG = t->st->b.as<double>();
    v.Lraw = t->env ? t->env->a.as<double>() : nullptr;
    v.Uraw = t->env ? t->env->b.as<double>() : nullptr;
    v.q = Q;
    v.p = P;
    v.G = G;
    const fsm_tuning& tu = t->cache->tune;
    v.glb_blocks = tu.glb_blocks;
    v.eld_refill = (P.metric >= 1 && P.metric <= 6) ? tu.eld_refill[P.metric] : 0;
    return v;
}

int ensure_eld_scratch(fsm_target* t, int metric, int m, int w) {
    if (!eld_uses_scratch(metric, m, w)) return FSM_OK;
    const int64_t thr = eld_scratch_threads();
    HIP_TRY(t->scratch.ensure(sizeof(double) * 2 * (size_t)(2 * w + 2) * (size_t)thr));
    return FSM_OK;
}

QueryParams params_of(const fsm_query_desc* d) {
    QueryParams P{};
    P.metric = d->metric;
    P.m = d->m;
    P.w = d->w;
    P.arg = d->arg;
    P.cutoff = d->cutoff;
    const bool excl = !(d->exclude_begin == -1 && d->exclude_end == -1);
    P.xb = excl ? d->exclude_begin : -1;
    P.xe = excl ? d->exclude_end : -1;
    P.lbip_f32 = d->lbip_f32 != 0;
    return P;
}

int ensure_eld_scratch(fsm_target* t, int metric, int m, int w);

// Tighter cutoff bound (see run_filter_verify): selects the max(8k, 4096)
// GLB survivors of smallest certified bound (lb[0 .. C[2]); largest for LCSS,
// whose bound counts possible matches), ties included, by an exact radix
// select of that order statistic -- so the set, and the bound that follows,
// do not depend on the order in which GLB appended the survivors -- computes
// their exact distances and lowers P->cutoff to their k-th smallest when that
// is smaller.  Uses list_a, out_dist, hist and counters 6, 7.
int tighten_cutoff(fsm_target* t, const Geometry& G, const QueryTables& Q, QueryParams* P,
                   int k, uint64_t* evals) {
    unsigned long long* C = t->counters.as<unsigned long long>();
    unsigned long long n_l = 0;
    HIP_TRY(hipMemcpyAsync(&n_l, C + 2, sizeof(n_l), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    const int64_t want = std::max<int64_t>(8 * (int64_t)k, 4096);
    if ((int64_t)n_l <= 4 * want) return FSM_OK;
    // tau: the want-th smallest bound (LCSS: the want-th largest, the
    // (n - want + 1)-th smallest of the n finite ones)
    const double sign = (P->metric == LCSS) ? -1.0 : 1.0;
    HIP_TRY(t->hist.ensure(sizeof(unsigned long long) * 2048));
    double tau = 0.0;
    int64_t nfin = 0;
    HIP_TRY(kth_smallest(t->stream, t->lb.as<double>(), C + 2, want,
                         t->hist.as<unsigned long long>(), &tau, &nfin));
    if (nfin < 4 * want) return FSM_OK;
    if (sign < 0)
        HIP_TRY(kth_smallest(t->stream, t->lb.as<double>(), C + 2, nfin - want + 1,
                             t->hist.as<unsigned long long>(), &tau, &nfin));
    unsigned long long zero[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(C + 6, zero, sizeof(zero), hipMemcpyHostToDevice, t->stream));
    HIP_TRY(launch_select_le(t->stream, t->list_b.as<int64_t>(), t->lb.as<double>(), C + 2, sign,
                             tau, t->list_a.as<int64_t>(), C + 6));
    unsigned long long nsel = 0;
    HIP_TRY(hipMemcpyAsync(&nsel, C + 6, sizeof(nsel), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if ((int64_t)nsel < k || (int64_t)nsel > 64 * want) return FSM_OK;
    HIP_TRY(t->out_dist.ensure(sizeof(double) * (size_t)std::max<unsigned long long>(nsel, 1)));
    int rc = ensure_eld_scratch(t, P->metric, P->m, P->w);
    if (rc) return rc;
    VerifyArgs v = verify_args(t, G, Q, *P);
    HIP_TRY(launch_eld(t->stream, v, t->list_a.as<int64_t>(), nullptr, C + 6, true, nullptr,
                       t->out_dist.as<double>(), nullptr, C + 7, nullptr, t->scratch.as<double>(),
                       eld_scratch_threads()));
    std::vector<double> dist((size_t)nsel);
    HIP_TRY(hipMemcpyAsync(dist.data(), t->out_dist.p, sizeof(double) * nsel,
                           hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (evals) *evals = nsel;
    std::vector<double> fin;
    fin.reserve(dist.size());
    for (double x : dist)
        if (std::isfinite(x)) fin.push_back(x);
    if ((int64_t)fin.size() < k) return FSM_OK;
    std::nth_element(fin.begin(), fin.begin() + (k - 1), fin.end());
    // widened by a relative 1e-9: a window whose distance equals the bound
    // exactly must survive GLB, whose rounded sums can exceed the exact
    // distance by an ulp where both are mathematically equal (w = 0)
    const double xi1 = fin[(size_t)(k - 1)] * (1.0 + 1e-9) + 1e-300;
    if (xi1 < P->cutoff) P->cutoff = xi1;
    return FSM_OK;
}

// Near-boundary re-verification: a window whose device distance lies within
// kNearRel (relative) of a decision boundary is recomputed on the host in
// FP64 (host_eld) and the host value decides.
constexpr double kNearRel = 1e-9;
inline double near_upper(double c) { return c * (1.0 + kNearRel) + 1e-300; }
inline bool near(double d, double c) { return std::fabs(d - c) <= kNearRel * std::fabs(c) + 1e-300; }

// Host distances of windows idx[i] for the i with near(dist[i], boundary):
// samples and window statistics are read back from the device.  Returns the
// number rechecked (*changed: how many the host value moved).
int recheck_near(fsm_target* t, const Geometry& G, const HostQuery& H, const fsm_query_desc* d,
                 double boundary, std::vector<int64_t>& idx, std::vector<double>& dist,
                 uint64_t* rechecked, uint64_t* changed) {
    const int m = d->m;
    std::vector<double> raw((size_t)m), y((size_t)m);
    for (size_t i = 0; i < idx.size(); i++) {
        if (!near(dist[i], boundary)) continue;
        const int64_t g = idx[i], si = stat_index(G, g);
        double mu = 0.0, sd = 1.0;
        HIP_TRY(hipMemcpy(raw.data(), t->T->as<double>() + g, sizeof(double) * m,
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&mu, t->st->a.as<double>() + si, sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&sd, t->st->b.as<double>() + si, sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < m; k++) y[k] = (raw[k] - mu) / sd;  // src/fsmeld.cpp:822-826
        const double h = host_eld(d->metric, H.qn.data(), y.data(), m, d->w, d->arg);
        ++*rechecked;
        if (std::memcmp(&h, &dist[i], sizeof(h)) != 0) ++*changed;
        dist[i] = h;
    }
    return FSM_OK;
}

// counters: [0] compacted, [1] valid, [2] after GLB, [3] ELD kept, [8] ELD work
// queue, [9] ELD lane-iterations on live candidates, [10] GLB work queue.  The LB buffer is free once compaction has read it; GLB writes each
// survivor's consumption bound there for the ELD abandon test.
// cutoff mode (tighten_k > 0): when GLB leaves many candidates at the first
// bound, the tighten_k-th smallest distance among the few thousand survivors
// of smallest certified GLB bound is a tighter upper bound xi1 of the k-th
// smallest distance; the survivors are filtered again by GLB at xi1 before
// ELD.  *cutoff_used receives the bound ELD ran with, *extra_evals the
// distances computed for it.
int run_filter_verify(fsm_target* t, const Geometry& G, const QueryTables& Q,
                      const QueryParams& P0, bool use_lb, bool use_glb, int tighten_k = 0,
                      double* cutoff_used = nullptr, uint64_t* extra_evals = nullptr) {
    QueryParams P = P0;
    if (cutoff_used) *cutoff_used = P.cutoff;
    if (extra_evals) *extra_evals = 0;
    unsigned long long* C = t->counters.as<unsigned long long>();
    HIP_TRY(hipMemsetAsync(C, 0, sizeof(unsigned long long) * 16, t->stream));
    HIP_TRY(t->list_a.ensure(sizeof(int64_t) * (size_t)G.ts_n));
    HIP_TRY(t->list_b.ensure(sizeof(int64_t) * (size_t)G.ts_n));
    HIP_TRY(t->out_idx.ensure(sizeof(int64_t) * (size_t)G.ts_n));
    HIP_TRY(t->out_dist.ensure(sizeof(double) * (size_t)G.ts_n));
    HIP_TRY(launch_compact(t->stream, G, P, t->st->c.as<uint8_t>(),
                           use_lb ? t->lb.as<double>() : nullptr, use_lb,
                           t->list_a.as<int64_t>(), C));
    HIP_TRY(hipEventRecord(t->ev[3], t->stream));
    VerifyArgs v = verify_args(t, G, Q, P);
    const int64_t* eld_in = t->list_a.as<int64_t>();
    const unsigned long long* eld_cnt = C;
    const double* eld_total = nullptr;
    if (use_glb) {
        HIP_TRY(t->lb.ensure(sizeof(double) * (size_t)G.ts_n));
        HIP_TRY(launch_glb(t->stream, v, t->list_a.as<int64_t>(), C, t->list_b.as<int64_t>(),
                           C + 2, t->lb.as<double>(), nullptr, C + 10));
        eld_in = t->list_b.as<int64_t>();
        eld_cnt = C + 2;
        eld_total = t->lb.as<double>();
        if (tighten_k > 0) {
            int rc = tighten_cutoff(t, G, Q, &P, tighten_k, extra_evals);
            if (rc) return rc;
            if (P.cutoff < P0.cutoff) {
                // GLB again over the first survivors at the tighter bound
                v = verify_args(t, G, Q, P);
                HIP_TRY(t->aux.ensure(sizeof(double) * (size_t)G.ts_n));
                HIP_TRY(hipMemsetAsync(C + 5, 0, sizeof(unsigned long long), t->stream));
                HIP_TRY(launch_glb(t->stream, v, t->list_b.as<int64_t>(), C + 2,
                                   t->list_a.as<int64_t>(), C + 5, t->aux.as<double>(), nullptr,
                                   C + 10));
                eld_in = t->list_a.as<int64_t>();
                eld_cnt = C + 5;
                eld_total = t->aux.as<double>();
            }
            if (cutoff_used) *cutoff_used = P.cutoff;
        }
    }
    HIP_TRY(hipEventRecord(t->ev[4], t->stream));
    int rc = ensure_eld_scratch(t, P.metric, P.m, P.w);
    if (rc) return rc;
    // ELD keeps windows up to the cutoff widened by kNearRel, so the ones just
    // above it reach the host re-verification (GLB above decided at the
    // cutoff itself, as the reference's cascade does)
    v.p.cutoff = near_upper(P.cutoff);
    HIP_TRY(launch_eld(t->stream, v, eld_in, eld_total, eld_cnt, false, t->out_idx.as<int64_t>(),
                       t->out_dist.as<double>(), C + 3, C + 8, C + 9, t->scratch.as<double>(),
                       eld_scratch_threads()));
    HIP_TRY(hipEventRecord(t->ev[5], t->stream));
    return FSM_OK;
}

// Exact top-k (src/fsmeld.cpp:1097-1107, :1273-1290): every valid window's
// exact distance (aux, in the order of the valid list in list_a), the k-th
// smallest finite one as *kth, and the windows at or below it appended to
// out_idx / out_dist with their count in counters[3].  *n_finite < k means
// fewer than k finite distances (the reference then reports none).
int run_exact_topk(fsm_target* t, const Geometry& G, const QueryTables& Q, const QueryParams& P,
                   int k, double* kth, int64_t* n_finite) {
    unsigned long long* C = t->counters.as<unsigned long long>();
    HIP_TRY(hipMemsetAsync(C, 0, sizeof(unsigned long long) * 16, t->stream));
    HIP_TRY(t->list_a.ensure(sizeof(int64_t) * (size_t)G.ts_n));
    HIP_TRY(t->aux.ensure(sizeof(double) * (size_t)G.ts_n));
    HIP_TRY(t->out_idx.ensure(sizeof(int64_t) * (size_t)G.ts_n));
    HIP_TRY(t->out_dist.ensure(sizeof(double) * (size_t)G.ts_n));
    HIP_TRY(t->hist.ensure(sizeof(unsigned long long) * 2048));
    HIP_TRY(launch_compact(t->stream, G, P, t->st->c.as<uint8_t>(), nullptr, false,
                           t->list_a.as<int64_t>(), C));
    HIP_TRY(hipEventRecord(t->ev[3], t->stream));
    int rc = ensure_eld_scratch(t, P.metric, P.m, P.w);
    if (rc) return rc;
    VerifyArgs v = verify_args(t, G, Q, P);
    HIP_TRY(launch_eld(t->stream, v, t->list_a.as<int64_t>(), nullptr, C, true, nullptr,
                       t->aux.as<double>(), nullptr, C + 8, nullptr, t->scratch.as<double>(),
                       eld_scratch_threads()));
    HIP_TRY(hipEventRecord(t->ev[4], t->stream));
    HIP_TRY(kth_smallest(t->stream, t->aux.as<double>(), C, k, t->hist.as<unsigned long long>(),
                         kth, n_finite));
    if (*n_finite < k) return FSM_OK;
    HIP_TRY(launch_select_le_kv(t->stream, t->list_a.as<int64_t>(), t->aux.as<double>(), C, *kth,
                                t->out_idx.as<int64_t>(), t->out_dist.as<double>(), C + 3));
    HIP_TRY(hipEventRecord(t->ev[5], t->stream));
    return FSM_OK;
}

}  // namespace

int fsm::report_error(int code, const char* msg) { return fail(code, msg); }

extern "C" {

int fsm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* fsm_last_error(void) { return g_last_error.c_str(); }

namespace {

// A handle with its own stream, events and counters over device samples T.
int new_handle(int64_t n, std::shared_ptr<DevBuf> T, std::shared_ptr<TargetCache> cache,
               fsm_target** out) {
    auto* t = new fsm_target();
    t->N = n;
    t->T = std::move(T);
    t->cache = std::move(cache);
    auto cleanup = [&](int rc) {
        fsm_target_close(t);
        return rc;
    };
    if (hipSetDevice(0) != hipSuccess) return cleanup(fail(FSM_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FSM_ERR_HIP, "hipStreamCreate failed"));
    for (auto& e : t->ev)
        if (hipEventCreate(&e) != hipSuccess)
            return cleanup(fail(FSM_ERR_HIP, "hipEventCreate failed"));
    if (t->counters.ensure(sizeof(unsigned long long) * 16) != hipSuccess)
        return cleanup(fail(FSM_ERR_OOM, "device allocation failed"));
    *out = t;
    return FSM_OK;
}

}  // namespace

int fsm_target_open(const double* host, int64_t n, fsm_target** out) {
    if (!out || !host || n <= 0) return fail(FSM_ERR_ARG, "invalid target");
    *out = nullptr;
    if (fsm_device_count() < 1) return fail(FSM_ERR_NODEV, "no HIP device");
    fsm_target* t = nullptr;
    auto cache = std::make_shared<TargetCache>();
    size_t mem_free = 0, mem_total = 0;
    if (hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) mem_total = 0;
    cache->cap = (uint64_t)mem_total / 2;
    int rc = new_handle(n, std::shared_ptr<DevBuf>(new DevBuf(), [](DevBuf* b) {
                            b->release();
                            delete b;
                        }), std::move(cache), &t);
    if (rc) return rc;
    if (t->T->ensure(sizeof(double) * (size_t)n) != hipSuccess) {
        fsm_target_close(t);
        return fail(FSM_ERR_OOM, "device allocation of the target failed");
    }
    if (hipMemcpy(t->T->p, host, sizeof(double) * (size_t)n, hipMemcpyHostToDevice) !=
        hipSuccess) {
        fsm_target_close(t);
        return fail(FSM_ERR_HIP, "target upload failed");
    }
    *out = t;
    return FSM_OK;
}

int fsm_target_share(fsm_target* base, fsm_target** out) {
    if (!base || !out) return fail(FSM_ERR_ARG, "null argument");
    *out = nullptr;
    return new_handle(base->N, base->T, base->cache, out);
}

int fsm_target_open_file(const char* path, fsm_target** out) {
    if (!path || !out) return fail(FSM_ERR_ARG, "invalid path");
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return fail(FSM_ERR_IO, "Cannot open target file.");
    const std::streamoff bytes = f.tellg();
    if (bytes <= 0 || bytes % (std::streamoff)sizeof(double))
        return fail(FSM_ERR_ARG, "Target must contain at least 2*m-1 complete double values.");
    std::vector<double> v((size_t)(bytes / (std::streamoff)sizeof(double)));
    f.seekg(0);
    f.read(reinterpret_cast<char*>(v.data()), bytes);
    if (f.gcount() != bytes) return fail(FSM_ERR_IO, "short read of target file");
    return fsm_target_open(v.data(), (int64_t)v.size(), out);
}

int64_t fsm_target_length(const fsm_target* t) { return t ? t->N : 0; }

void fsm_target_close(fsm_target* t) {
    if (!t) return;
    for (fsm_target* h : t->pool) fsm_target_close(h);
    t->pool.clear();
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    t->st.reset();
    t->env.reset();
    t->cache.reset();
    DevBuf* bufs[] = {&t->lb,
                      &t->list_a, &t->list_b, &t->out_idx, &t->out_dist, &t->slab,
                      &t->rep_lb, &t->rep_idx, &t->tu, &t->tl, &t->qtab, &t->scratch,
                      &t->counters, &t->aux, &t->hist};
    for (DevBuf* b : bufs) b->release();
    for (auto& e : t->ev)
        if (e) (void)hipEventDestroy(e);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    t->T.reset();
    delete t;
}

int fsm_query(fsm_target* t, const fsm_query_desc* d, const double* q, int64_t qlen,
              fsm_result** out) {
    const auto wall0 = std::chrono::steady_clock::now();
    if (!t || !out) return fail(FSM_ERR_ARG, "null argument");
    *out = nullptr;
    int rc = validate(d, t->N, q, qlen, true);
    if (rc) return rc;
    Geometry G;
    if (!make_geometry(t->N, d->m, &G))
        return fail(FSM_ERR_ARG, "Macro segment exceeds the supported FFT length.");
    HostQuery H;
    prepare_query(q, d->m, d->w, d->arg, d->metric, &H, d->raw != 0);
    QueryTables Q{};
    rc = upload_query(t, H, d->m, &Q);
    if (rc) return rc;
    QueryParams P = params_of(d);
    std::unique_ptr<fsm_result> R(new fsm_result());
    fsm_stats& st = R->st;
    st.ts_n = G.ts_n;
    st.k_m_l = G.kml;
    st.forsize = G.forsize;
    st.tail = G.tail;
    st.m_k = G.m_k;
    if (!lbip_f32_supported(G)) P.lbip_f32 = 0;
    st.lbip_f32 = P.lbip_f32 && (d->mode == FSM_MODE_SEARCH || d->mode == FSM_MODE_CUTOFF);
    auto bail = [&](int code) { return code; };
    unsigned long long C[16] = {};

    HIP_TRY(hipEventRecord(t->ev[0], t->stream));
    bool stats_new = false, env_new = false;
    rc = ensure_stats(t, G, &stats_new, d->raw != 0);
    if (rc) return bail(rc);
    HIP_TRY(hipEventRecord(t->ev[1], t->stream));

    double topk_kth = 0.0;
    if (d->mode == FSM_MODE_EXACT_TOPK) {
        HIP_TRY(hipEventRecord(t->ev[2], t->stream));
        int64_t nfin = 0;
        rc = run_exact_topk(t, G, Q, P, d->k, &topk_kth, &nfin);
        if (rc) return bail(rc);
        if (nfin < d->k)
            return bail(fail(FSM_ERR_FEWER_THAN_K,
                             "Fewer finite exact distances than requested k."));
        HIP_TRY(hipStreamSynchronize(t->stream));
        HIP_TRY(hipMemcpy(C, t->counters.p, sizeof(C), hipMemcpyDeviceToHost));
        st.valid = C[1];
        st.invalid = (uint64_t)G.ts_n - C[1];
        st.after_lbip = C[0];
        st.after_glb = C[0];
        st.exact_evals = C[0];
    } else if (d->mode == FSM_MODE_EXACT) {
        HIP_TRY(hipEventRecord(t->ev[2], t->stream));
        rc = run_filter_verify(t, G, Q, P, false, false);
        if (rc) return bail(rc);
        HIP_TRY(hipStreamSynchronize(t->stream));
        HIP_TRY(hipMemcpy(C, t->counters.p, sizeof(C), hipMemcpyDeviceToHost));
        st.valid = C[1];
        st.invalid = (uint64_t)G.ts_n - C[1];
        st.after_lbip = C[0];
        st.after_glb = C[0];
        st.exact_evals = C[0];
    } else {
        rc = ensure_env(t, d->w, &env_new);
        if (rc) return bail(rc);
        HIP_TRY(hipEventRecord(t->ev[7], t->stream));
        rc = run_lbip(t, G, Q, P, d->mode == FSM_MODE_CUTOFF);
        if (rc) return bail(rc);
        HIP_TRY(hipEventRecord(t->ev[2], t->stream));
        uint64_t rep_evals = 0;
        if (d->mode == FSM_MODE_CUTOFF) {
            // valid count and representatives
            unsigned long long* Cd = t->counters.as<unsigned long long>();
            HIP_TRY(hipMemsetAsync(Cd, 0, sizeof(unsigned long long) * 8, t->stream));
            HIP_TRY(t->list_a.ensure(sizeof(int64_t) * (size_t)G.ts_n));
            HIP_TRY(launch_compact(t->stream, G, P, t->st->c.as<uint8_t>(), nullptr, false,
                                   t->list_a.as<int64_t>(), Cd));
            std::vector<int64_t> reps((size_t)G.S);
            HIP_TRY(hipMemcpyAsync(reps.data(), t->rep_idx.p, sizeof(int64_t) * G.S,
                                   hipMemcpyDeviceToHost, t->stream));
            HIP_TRY(hipMemcpyAsync(C, t->counters.p, sizeof(C), hipMemcpyDeviceToHost,
                                   t->stream));
            HIP_TRY(hipStreamSynchronize(t->stream));
            const uint64_t nvalid = C[1];
            if (nvalid < (uint64_t)d->k)
                return bail(fail(FSM_ERR_FEWER_THAN_K,
                                 "Fewer valid, non-excluded candidates than requested k."));
            std::vector<int64_t> rl;
            rl.reserve(reps.size());
            for (int64_t r : reps)
                if (r >= 0) rl.push_back(r);
            double xi0 = std::numeric_limits<double>::infinity();
            if ((int64_t)rl.size() >= d->k) {
                HIP_TRY(t->list_b.ensure(sizeof(int64_t) * (size_t)G.ts_n));
                HIP_TRY(t->out_dist.ensure(sizeof(double) * (size_t)G.ts_n));
                unsigned long long nrep = rl.size();
                HIP_TRY(hipMemcpyAsync(t->list_b.p, rl.data(), sizeof(int64_t) * rl.size(),
                                       hipMemcpyHostToDevice, t->stream));
                HIP_TRY(hipMemcpyAsync(Cd + 4, &nrep, sizeof(nrep), hipMemcpyHostToDevice,
                                       t->stream));
                rc = ensure_eld_scratch(t, d->metric, d->m, d->w);
                if (rc) return bail(rc);
                VerifyArgs v = verify_args(t, G, Q, P);
                HIP_TRY(launch_eld(t->stream, v, t->list_b.as<int64_t>(), nullptr, Cd + 4, true,
                                   nullptr, t->out_dist.as<double>(), nullptr, Cd + 8,
                                   nullptr, t->scratch.as<double>(), eld_scratch_threads()));
                std::vector<double> rd(rl.size());
                HIP_TRY(hipMemcpyAsync(rd.data(), t->out_dist.p, sizeof(double) * rd.size(),
                                       hipMemcpyDeviceToHost, t->stream));
                HIP_TRY(hipStreamSynchronize(t->stream));
                rep_evals = rl.size();
                std::vector<double> fin;
                for (double x : rd)
                    if (std::isfinite(x)) fin.push_back(x);
                if ((int64_t)fin.size() >= d->k) {
                    std::nth_element(fin.begin(), fin.begin() + (d->k - 1), fin.end());
                    // widened as xi1 is (tighten_cutoff): the window whose
                    // distance is the k-th must survive LB_IP and GLB, whose
                    // rounded sums can exceed an equal exact distance by an
                    // ulp (w = 0); the answer is cut to k afterwards
                    xi0 = fin[d->k - 1] * (1.0 + 1e-9) + 1e-300;
                }
            }
            P.cutoff = xi0;
        }
        uint64_t tight_evals = 0;
        double used = P.cutoff;
        rc = run_filter_verify(t, G, Q, P, true, true, d->mode == FSM_MODE_CUTOFF ? d->k : 0,
                               &used, &tight_evals);
        if (rc) return bail(rc);
        HIP_TRY(hipStreamSynchronize(t->stream));
        unsigned long long C2[16] = {};
        HIP_TRY(hipMemcpy(C2, t->counters.p, sizeof(C2), hipMemcpyDeviceToHost));
        std::memcpy(C, C2, sizeof(C));
        const bool tightened = d->mode == FSM_MODE_CUTOFF && used < P.cutoff;
        st.valid = C[1];
        st.invalid = (uint64_t)G.ts_n - C[1];
        st.after_lbip = C[0];
        st.after_glb = tightened ? C[5] : C[2];
        st.exact_evals = st.after_glb + rep_evals + tight_evals;
    }
    const uint64_t nres = C[3];
    R->idx.resize(nres);
    R->dist.resize(nres);
    if (nres) {
        HIP_TRY(hipMemcpy(R->idx.data(), t->out_idx.p, sizeof(int64_t) * nres,
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(R->dist.data(), t->out_dist.p, sizeof(double) * nres,
                          hipMemcpyDeviceToHost));
    }
    // near the boundary the host decides: search / exact at the cutoff (the
    // device kept windows up to near_upper(cutoff)); cutoff and top-k at the
    // k-th distance
    {
        double boundary = d->cutoff;
        if (d->mode == FSM_MODE_CUTOFF && !R->dist.empty()) {
            std::vector<double> s2(R->dist);
            const size_t kk = std::min<size_t>((size_t)d->k, s2.size()) - 1;
            std::nth_element(s2.begin(), s2.begin() + kk, s2.end());
            boundary = s2[kk];
        } else if (d->mode == FSM_MODE_EXACT_TOPK) {
            boundary = topk_kth;
        }
        rc = recheck_near(t, G, H, d, boundary, R->idx, R->dist, &st.near_rechecked,
                          &st.near_changed);
        if (rc) return bail(rc);
        if (d->mode == FSM_MODE_SEARCH || d->mode == FSM_MODE_EXACT ||
            d->mode == FSM_MODE_EXACT_TOPK) {
            const double lim = (d->mode == FSM_MODE_EXACT_TOPK) ? topk_kth : d->cutoff;
            size_t o = 0;
            for (size_t i = 0; i < R->idx.size(); i++)
                if (R->dist[i] <= lim) {  // src/fsmeld.cpp:846-847
                    R->idx[o] = R->idx[i];
                    R->dist[o] = R->dist[i];
                    o++;
                }
            R->idx.resize(o);
            R->dist.resize(o);
        }
    }
    const size_t nkept = R->idx.size();
    // order: ascending index (search / exact), (distance, index) for cutoff
    std::vector<size_t> ord(nkept);
    for (size_t i = 0; i < nkept; i++) ord[i] = i;
    if (d->mode == FSM_MODE_CUTOFF) {
        std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
            if (R->dist[a] != R->dist[b]) return R->dist[a] < R->dist[b];
            return R->idx[a] < R->idx[b];
        });
        if (ord.size() > (size_t)d->k) ord.resize(d->k);
    } else {
        std::sort(ord.begin(), ord.end(),
                  [&](size_t a, size_t b) { return R->idx[a] < R->idx[b]; });
    }
    std::vector<int64_t> idx(ord.size());
    std::vector<double> dist(ord.size());
    for (size_t i = 0; i < ord.size(); i++) {
        idx[i] = R->idx[ord[i]];
        dist[i] = R->dist[ord[i]];
    }
    R->idx.swap(idx);
    R->dist.swap(dist);
    if (d->mode == FSM_MODE_CUTOFF) {
        if ((int64_t)R->idx.size() < d->k)
            return bail(fail(FSM_ERR_FEWER_THAN_K,
                             "Fewer valid, non-excluded candidates than requested k."));
        st.threshold = R->dist.back();
    } else if (d->mode == FSM_MODE_EXACT_TOPK) {
        st.threshold = topk_kth;
    }
    st.matches = R->idx.size();
    st.eld_cells = (uint64_t)C[9] * (uint64_t)(2 * d->w + 1);
    st.ms_stats = stats_new ? ms_between(t->ev[0], t->ev[1]) : 0.0;
    st.ms_stats_finish = stats_new ? ms_between(t->ev[6], t->ev[1]) : 0.0;
    st.ms_env = env_new ? ms_between(t->ev[1], t->ev[7]) : 0.0;
    const bool cascade = d->mode == FSM_MODE_SEARCH || d->mode == FSM_MODE_CUTOFF;
    st.ms_lbip = cascade ? ms_between(t->ev[7], t->ev[2]) : 0.0;
    if (d->mode == FSM_MODE_EXACT_TOPK) {
        st.ms_glb = 0.0;
        st.ms_eld = ms_between(t->ev[3], t->ev[4]);
    } else if (d->mode == FSM_MODE_EXACT) {
        st.ms_glb = 0.0;
        st.ms_eld = ms_between(t->ev[2], t->ev[5]);
    } else {
        st.ms_glb = ms_between(t->ev[3], t->ev[4]);
        st.ms_eld = ms_between(t->ev[4], t->ev[5]);
    }
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                           wall0).count();
    *out = R.release();
    return FSM_OK;
}

int fsm_result_stats(const fsm_result* r, fsm_stats* out) {
    if (!r || !out) return fail(FSM_ERR_ARG, "null argument");
    *out = r->st;
    return FSM_OK;
}
int64_t fsm_result_count(const fsm_result* r) { return r ? (int64_t)r->idx.size() : 0; }
const int64_t* fsm_result_indices(const fsm_result* r) { return r ? r->idx.data() : nullptr; }
const double* fsm_result_distances(const fsm_result* r) { return r ? r->dist.data() : nullptr; }
void fsm_result_free(fsm_result* r) { delete r; }

// ------------------------------------------------------ stage entry points ---
int fsm_stage_window_stats(fsm_target* t, int m, int64_t xb, int64_t xe, double* mu,
                           double* sigma, uint8_t* invalid) {
    if (!t || !mu || !sigma || !invalid) return fail(FSM_ERR_ARG, "null argument");
    Geometry G;
    if (m < 2 || !make_geometry(t->N, m, &G)) return fail(FSM_ERR_ARG, "invalid m");
    bool fresh;
    int rc = ensure_stats(t, G, &fresh);
    if (rc) return rc;
    const size_t n = (size_t)G.S * (size_t)G.kml;
    std::vector<double> hm(n), hs(n);
    std::vector<uint8_t> hi(n);
    HIP_TRY(hipStreamSynchronize(t->stream));
    HIP_TRY(hipMemcpy(hm.data(), t->st->a.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hs.data(), t->st->b.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hi.data(), t->st->c.p, n, hipMemcpyDeviceToHost));
    const bool excl = xb >= 0 && xe > xb;
    for (int64_t g = 0; g < G.ts_n; g++) {
        const int64_t s = stat_index(G, g);
        mu[g] = hm[s];
        sigma[g] = hs[s];
        invalid[g] = hi[s] || (excl && g >= xb && g < xe);
    }
    return FSM_OK;
}

int fsm_stage_lbip(fsm_target* t, const fsm_query_desc* d, const double* q, double* lb,
                   double* t_u, double* t_l) {
    if (!t || !lb) return fail(FSM_ERR_ARG, "null argument");
    int rc = validate(d, t->N, q, d ? d->m : 0, false);
    if (rc) return rc;
    Geometry G;
    if (!make_geometry(t->N, d->m, &G)) return fail(FSM_ERR_ARG, "invalid m");
    HostQuery H;
    prepare_query(q, d->m, d->w, d->arg, d->metric, &H, d->raw != 0);
    QueryTables Q{};
    if ((rc = upload_query(t, H, d->m, &Q))) return rc;
    QueryParams P = params_of(d);
    bool fresh;
    if ((rc = ensure_stats(t, G, &fresh, d->raw != 0))) return rc;
    if ((rc = ensure_env(t, d->w))) return rc;
    if ((rc = run_lbip(t, G, Q, P, false))) return rc;
    HIP_TRY(hipStreamSynchronize(t->stream));
    HIP_TRY(hipMemcpy(lb, t->lb.p, sizeof(double) * (size_t)G.ts_n, hipMemcpyDeviceToHost));
    if (t_u) HIP_TRY(hipMemcpy(t_u, t->tu.p, sizeof(double) * (size_t)G.S, hipMemcpyDeviceToHost));
    if (t_l) HIP_TRY(hipMemcpy(t_l, t->tl.p, sizeof(double) * (size_t)G.S, hipMemcpyDeviceToHost));
    return FSM_OK;
}

enum StageKind { STAGE_GLB_VALUES, STAGE_ELD_VALUES, STAGE_GLB_KEEP };

static int stage_verify(fsm_target* t, const fsm_query_desc* d, const double* q,
                        const int64_t* win, int64_t n, void* out, StageKind kind) {
    const bool glb = kind != STAGE_ELD_VALUES;
    if (!t || !out || (n > 0 && !win) || n < 0) return fail(FSM_ERR_ARG, "null argument");
    int rc = validate(d, t->N, q, d ? d->m : 0, false);
    if (rc) return rc;
    Geometry G;
    if (!make_geometry(t->N, d->m, &G)) return fail(FSM_ERR_ARG, "invalid m");
    for (int64_t i = 0; i < n; i++)
        if (win[i] < 0 || win[i] >= G.ts_n) return fail(FSM_ERR_ARG, "window index out of range");
    if (n == 0) return FSM_OK;
    HostQuery H;
    prepare_query(q, d->m, d->w, d->arg, d->metric, &H, d->raw != 0);
    QueryTables Q{};
    if ((rc = upload_query(t, H, d->m, &Q))) return rc;
    QueryParams P = params_of(d);
    bool fresh;
    if ((rc = ensure_stats(t, G, &fresh, d->raw != 0))) return rc;
    if (glb && (rc = ensure_env(t, d->w))) return rc;
    if ((rc = ensure_eld_scratch(t, d->metric, d->m, d->w))) return rc;
    HIP_TRY(t->list_a.ensure(sizeof(int64_t) * (size_t)n));
    HIP_TRY(t->list_b.ensure(sizeof(int64_t) * (size_t)n));
    HIP_TRY(t->out_dist.ensure(sizeof(double) * (size_t)n));
    unsigned long long cnt[2] = {(unsigned long long)n, 0ull};
    unsigned long long* Cd = t->counters.as<unsigned long long>();
    HIP_TRY(hipMemcpyAsync(t->list_a.p, win, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                           t->stream));
    HIP_TRY(hipMemcpyAsync(Cd + 6, cnt, sizeof(cnt), hipMemcpyHostToDevice, t->stream));
    VerifyArgs v = verify_args(t, G, Q, P);
    if (kind == STAGE_GLB_VALUES)
        HIP_TRY(launch_glb(t->stream, v, t->list_a.as<int64_t>(), Cd + 6, nullptr, nullptr,
                           nullptr, t->out_dist.as<double>(), Cd + 10));
    else if (kind == STAGE_GLB_KEEP)
        HIP_TRY(launch_glb(t->stream, v, t->list_a.as<int64_t>(), Cd + 6, t->list_b.as<int64_t>(),
                           Cd + 7, nullptr, nullptr, Cd + 10));
    else
        HIP_TRY(launch_eld(t->stream, v, t->list_a.as<int64_t>(), nullptr, Cd + 6, true, nullptr,
                           t->out_dist.as<double>(), nullptr, Cd + 8, nullptr,
                           t->scratch.as<double>(), eld_scratch_threads()));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (kind != STAGE_GLB_KEEP) {
        HIP_TRY(hipMemcpy(out, t->out_dist.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        return FSM_OK;
    }
    HIP_TRY(hipMemcpy(cnt, Cd + 6, sizeof(cnt), hipMemcpyDeviceToHost));
    std::vector<int64_t> kept(cnt[1]);
    if (cnt[1])
        HIP_TRY(hipMemcpy(kept.data(), t->list_b.p, sizeof(int64_t) * cnt[1], hipMemcpyDeviceToHost));
    std::sort(kept.begin(), kept.end());
    uint8_t* flags = static_cast<uint8_t*>(out);
    for (int64_t i = 0; i < n; i++)
        flags[i] = std::binary_search(kept.begin(), kept.end(), win[i]) ? 1 : 0;
    return FSM_OK;
}

int fsm_stage_glb(fsm_target* t, const fsm_query_desc* d, const double* q, const int64_t* win,
                  int64_t n, double* out) {
    return stage_verify(t, d, q, win, n, out, STAGE_GLB_VALUES);
}

int fsm_stage_glb_filter(fsm_target* t, const fsm_query_desc* d, const double* q,
                         const int64_t* win, int64_t n, uint8_t* keep) {
    return stage_verify(t, d, q, win, n, keep, STAGE_GLB_KEEP);
}

int fsm_stage_eld(fsm_target* t, const fsm_query_desc* d, const double* q, const int64_t* win,
                  int64_t n, double* out) {
    return stage_verify(t, d, q, win, n, out, STAGE_ELD_VALUES);
}

// ------------------------------------- stage entry points, device memory ---
}  // extern "C"

namespace {

// The target's stream waits for the work the caller enqueued on its stream,
// and the caller's stream then waits for the stage's work: events only, no
// host synchronisation.
struct StreamJoin {
    hipStream_t cs, ts;
    hipEvent_t e = nullptr;
    StreamJoin(void* c, hipStream_t t) : cs((hipStream_t)c), ts(t) {}
    hipError_t enter() {
        hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (r == hipSuccess) r = hipEventRecord(e, cs);
        if (r == hipSuccess) r = hipStreamWaitEvent(ts, e, 0);
        return r;
    }
    hipError_t leave() {
        hipError_t r = hipEventRecord(e, ts);
        if (r == hipSuccess) r = hipStreamWaitEvent(cs, e, 0);
        return r;
    }
    ~StreamJoin() {
        if (e) (void)hipEventDestroy(e);
    }
};

// per-window statistics in global window order (mvmean's order), exclusion
// [xb, xe) marked invalid as the host entry does
__global__ void k_stats_gather(Geometry G, const double* __restrict__ MU,
                               const double* __restrict__ SG, const uint8_t* __restrict__ INV,
                               int64_t xb, int64_t xe, double* mu, double* sg, uint8_t* inv) {
    const bool excl = xb >= 0 && xe > xb;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G.ts_n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = stat_index(G, g);
        mu[g] = MU[s];
        sg[g] = SG[s];
        inv[g] = INV[s] || (excl && g >= xb && g < xe);
    }
}

// *bad = 1 when a window index lies outside [0, ts_n)
__global__ void k_check_windows(const int64_t* __restrict__ win, int64_t n, int64_t ts_n,
                                unsigned long long* bad) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (win[i] < 0 || win[i] >= ts_n) *bad = 1ull;
}

__global__ void k_set_counters(unsigned long long* c, unsigned long long c0,
                               unsigned long long c1) {
    c[0] = c0;
    c[1] = c1;
}

__global__ void k_zero_words(unsigned* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// bit g of bits for each kept window g, then keep[i] = bit win[i]
__global__ void k_mark_kept(const int64_t* __restrict__ kept,
                            const unsigned long long* __restrict__ count, unsigned* bits) {
    const int64_t n = (int64_t)*count;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicOr(&bits[kept[i] >> 5], 1u << (kept[i] & 31));
}
__global__ void k_keep_flags(const int64_t* __restrict__ win, int64_t n,
                             const unsigned* __restrict__ bits, uint8_t* keep) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        keep[i] = (bits[win[i] >> 5] >> (win[i] & 31)) & 1u;
}

inline dim3 grid_for(int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)));
}

}  // namespace

extern "C" {

int fsm_stage_window_stats_dev(fsm_target* t, int m, int64_t xb, int64_t xe, double* mu,
                               double* sigma, uint8_t* invalid, void* stream) {
    if (!t || !mu || !sigma || !invalid) return fail(FSM_ERR_ARG, "null argument");
    Geometry G;
    if (m < 2 || !make_geometry(t->N, m, &G)) return fail(FSM_ERR_ARG, "invalid m");
    StreamJoin j(stream, t->stream);
    HIP_TRY(j.enter());
    bool fresh;
    int rc = ensure_stats(t, G, &fresh);
    if (rc) return rc;
    hipLaunchKernelGGL(k_stats_gather, grid_for(G.ts_n), dim3(256), 0, t->stream, G,
                       t->st->a.as<double>(), t->st->b.as<double>(), t->st->c.as<uint8_t>(), xb,
                       xe, mu, sigma, invalid);
    HIP_TRY(hipGetLastError());
    HIP_TRY(j.leave());
    return FSM_OK;
}

int fsm_stage_lbip_dev(fsm_target* t, const fsm_query_desc* d, const double* q, double* lb,
                       double* t_u, double* t_l, void* stream) {
    if (!t || !lb) return fail(FSM_ERR_ARG, "null argument");
    int rc = validate(d, t->N, q, d ? d->m : 0, false);
    if (rc) return rc;
    Geometry G;
    if (!make_geometry(t->N, d->m, &G)) return fail(FSM_ERR_ARG, "invalid m");
    StreamJoin j(stream, t->stream);
    HIP_TRY(j.enter());
    HostQuery H;
    prepare_query(q, d->m, d->w, d->arg, d->metric, &H, d->raw != 0);
    QueryTables Q{};
    if ((rc = upload_query(t, H, d->m, &Q))) return rc;
    QueryParams P = params_of(d);
    bool fresh;
    if ((rc = ensure_stats(t, G, &fresh, d->raw != 0))) return rc;
    if ((rc = ensure_env(t, d->w))) return rc;
    if ((rc = run_lbip(t, G, Q, P, false, lb, t_u, t_l))) return rc;
    HIP_TRY(j.leave());
    return FSM_OK;
}

static int stage_verify_dev(fsm_target* t, const fsm_query_desc* d, const double* q,
                            const int64_t* win, int64_t n, void* out, StageKind kind,
                            void* stream) {
    const bool glb = kind != STAGE_ELD_VALUES;
    if (!t || !out || (n > 0 && !win) || n < 0) return fail(FSM_ERR_ARG, "null argument");
    int rc = validate(d, t->N, q, d ? d->m : 0, false);
    if (rc) return rc;
    Geometry G;
    if (!make_geometry(t->N, d->m, &G)) return fail(FSM_ERR_ARG, "invalid m");
    if (n == 0) return FSM_OK;
    StreamJoin j(stream, t->stream);
    HIP_TRY(j.enter());
    unsigned long long* Cd = t->counters.as<unsigned long long>();
    // the window indices are checked on the device before any kernel reads
    // through them (one word read back)
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(1), 0, t->stream, Cd + 11, 0ull, 0ull);
    hipLaunchKernelGGL(k_check_windows, grid_for(n), dim3(256), 0, t->stream, win, n, G.ts_n,
                       Cd + 11);
    HIP_TRY(hipGetLastError());
    unsigned long long bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, Cd + 11, sizeof(bad), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (bad) return fail(FSM_ERR_ARG, "window index out of range");
    HostQuery H;
    prepare_query(q, d->m, d->w, d->arg, d->metric, &H, d->raw != 0);
    QueryTables Q{};
    if ((rc = upload_query(t, H, d->m, &Q))) return rc;
    QueryParams P = params_of(d);
    bool fresh;
    if ((rc = ensure_stats(t, G, &fresh, d->raw != 0))) return rc;
    if (glb && (rc = ensure_env(t, d->w))) return rc;
    if ((rc = ensure_eld_scratch(t, d->metric, d->m, d->w))) return rc;
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(1), 0, t->stream, Cd + 6,
                       (unsigned long long)n, 0ull);
    HIP_TRY(hipGetLastError());
    VerifyArgs v = verify_args(t, G, Q, P);
    if (kind == STAGE_GLB_VALUES) {
        HIP_TRY(launch_glb(t->stream, v, win, Cd + 6, nullptr, nullptr, nullptr,
                           static_cast<double*>(out), Cd + 10));
    } else if (kind == STAGE_ELD_VALUES) {
        HIP_TRY(launch_eld(t->stream, v, win, nullptr, Cd + 6, true, nullptr,
                           static_cast<double*>(out), nullptr, Cd + 8, nullptr,
                           t->scratch.as<double>(), eld_scratch_threads()));
    } else {
        const int64_t nbits = (G.ts_n + 31) / 32;
        HIP_TRY(t->list_b.ensure(sizeof(int64_t) * (size_t)n));
        HIP_TRY(t->lb.ensure(sizeof(unsigned) * (size_t)nbits));  // the bitmap
        unsigned* bits = t->lb.as<unsigned>();
        HIP_TRY(launch_glb(t->stream, v, win, Cd + 6, t->list_b.as<int64_t>(), Cd + 7, nullptr,
                           nullptr, Cd + 10));
        hipLaunchKernelGGL(k_zero_words, grid_for(nbits), dim3(256), 0, t->stream, bits, nbits);
        hipLaunchKernelGGL(k_mark_kept, grid_for(n), dim3(256), 0, t->stream,
                           t->list_b.as<int64_t>(), Cd + 7, bits);
        hipLaunchKernelGGL(k_keep_flags, grid_for(n), dim3(256), 0, t->stream, win, n, bits,
                           static_cast<uint8_t*>(out));
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(j.leave());
    return FSM_OK;
}

int fsm_stage_glb_dev(fsm_target* t, const fsm_query_desc* d, const double* q,
                      const int64_t* win, int64_t n, double* out, void* stream) {
    return stage_verify_dev(t, d, q, win, n, out, STAGE_GLB_VALUES, stream);
}

int fsm_stage_glb_filter_dev(fsm_target* t, const fsm_query_desc* d, const double* q,
                             const int64_t* win, int64_t n, uint8_t* keep, void* stream) {
    return stage_verify_dev(t, d, q, win, n, keep, STAGE_GLB_KEEP, stream);
}

int fsm_stage_eld_dev(fsm_target* t, const fsm_query_desc* d, const double* q,
                      const int64_t* win, int64_t n, double* out, void* stream) {
    return stage_verify_dev(t, d, q, win, n, out, STAGE_ELD_VALUES, stream);
}

int fsm_eld_pairs(int metric, int m, int w, double arg, const double* q, const double* cands,
                  int64_t n, double* out) {
    if (!q || !cands || !out || n < 0) return fail(FSM_ERR_ARG, "null argument");
    if (m < 2 || w < 0 || w >= m || metric < FSM_METRIC_DTW || metric > FSM_METRIC_SWALE)
        return fail(FSM_ERR_ARG, "invalid m, w or metric");
    if (metric == FSM_METRIC_SWALE && arg < 0)
        return fail(FSM_ERR_ARG, "Matching tolerance must be non-negative.");
    if (n == 0) return FSM_OK;
    if (!std::isfinite(arg)) return fail(FSM_ERR_ARG, "Invalid finite metric parameter");
    for (int i = 0; i < m; i++)
        if (!std::isfinite(q[i])) return fail(FSM_ERR_ARG, "query values must be finite");
    for (int64_t i = 0; i < (int64_t)m * n; i++)
        if (!std::isfinite(cands[i])) return fail(FSM_ERR_ARG, "candidate values must be finite");
    if (fsm_device_count() < 1) return fail(FSM_ERR_NODEV, "no HIP device");
    HIP_TRY(hipSetDevice(0));
    DevBuf dq, dxg, dc, dout, dscr, dqueue;
    const int64_t thr = eld_scratch_threads();
    auto done = [&](int rc) {
        dq.release();
        dxg.release();
        dc.release();
        dout.release();
        dscr.release();
        dqueue.release();
        return rc;
    };
    if (dq.ensure(sizeof(double) * m) || dxg.ensure(sizeof(double) * m) ||
        dc.ensure(sizeof(double) * (size_t)m * n) ||
        dout.ensure(sizeof(double) * n) || dqueue.ensure(sizeof(unsigned long long)))
        return done(fail(FSM_ERR_OOM, "device allocation failed"));
    if (eld_uses_scratch(metric, m, w) &&
        dscr.ensure(sizeof(double) * 2 * (size_t)(2 * w + 2) * (size_t)thr))
        return done(fail(FSM_ERR_OOM, "device allocation failed"));
    std::vector<double> xg(m);
    for (int i = 0; i < m; i++) xg[i] = (q[i] - arg) * (q[i] - arg);
    if (hipMemcpy(dq.p, q, sizeof(double) * m, hipMemcpyHostToDevice) ||
        hipMemcpy(dxg.p, xg.data(), sizeof(double) * m, hipMemcpyHostToDevice) ||
        hipMemcpy(dc.p, cands, sizeof(double) * (size_t)m * n, hipMemcpyHostToDevice))
        return done(fail(FSM_ERR_HIP, "upload failed"));
    VerifyArgs va{};
    va.q.qn = dq.as<double>();
    va.q.XG = dxg.as<double>();
    va.p.metric = metric;
    va.p.m = m;
    va.p.w = w;
    va.p.arg = arg;
    if (launch_eld_pairs(nullptr, va, dc.as<double>(), n, dout.as<double>(),
                         dqueue.as<unsigned long long>(), dscr.as<double>(), thr) != hipSuccess)
        return done(fail(FSM_ERR_HIP, "ELD launch failed"));
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, dout.p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(FSM_ERR_HIP, "ELD failed"));
    return done(FSM_OK);
}

}  // extern "C"

int fsm_query_batch(fsm_target* t, int64_t n, const fsm_query_desc* descs,
                    const double* const* queries, const int64_t* query_lens, int concurrency,
                    fsm_result** results, int* status) {
    if (!t || n < 0 || (n > 0 && (!descs || !queries || !query_lens || !results || !status)))
        return fail(FSM_ERR_ARG, "null argument");
    if (concurrency < 1) return fail(FSM_ERR_ARG, "concurrency must be at least 1");
    for (int64_t i = 0; i < n; i++) {
        results[i] = nullptr;
        status[i] = FSM_ERR_ARG;
    }
    if (n == 0) return FSM_OK;
    const int k = (int)std::min<int64_t>(concurrency, n);
    std::vector<fsm_target*> hs;
    {
        std::lock_guard<std::mutex> lk(t->pool_mu);
        while ((int)t->pool.size() < k) {
            fsm_target* h = nullptr;
            int rc = fsm_target_share(t, &h);
            if (rc) return rc;
            t->pool.push_back(h);
        }
        hs.assign(t->pool.begin(), t->pool.begin() + k);
    }
    // each worker thread owns one handle (stream and buffers) and takes the
    // next query as soon as its previous one has finished
    std::atomic<int64_t> next{0};
    auto work = [&](fsm_target* h) {
        (void)hipSetDevice(0);
        for (int64_t i = next++; i < n; i = next++)
            status[i] = fsm_query(h, &descs[i], queries[i], query_lens[i], &results[i]);
    };
    std::vector<std::thread> th;
    for (int j = 1; j < k; j++) th.emplace_back(work, hs[j]);
    work(hs[0]);
    for (auto& x : th) x.join();
    return FSM_OK;
}

int fsm_tuning_get(const fsm_target* t, fsm_tuning* out) {
    if (!t || !out) return fail(FSM_ERR_ARG, "null argument");
    *out = t->cache->tune;
    return FSM_OK;
}

int fsm_tuning_set(fsm_target* t, const fsm_tuning* in) {
    if (!t || !in) return fail(FSM_ERR_ARG, "null argument");
    for (int v : in->lbip_threads)
        if (v != 0 && v != 128 && v != 256 && v != 512 && v != 1024)
            return fail(FSM_ERR_ARG, "LB_IP workgroup sizes are 128, 256, 512 or 1024");
    for (int v : in->eld_refill)
        if (v < 0 || v > 64) return fail(FSM_ERR_ARG, "ELD refill thresholds lie in [0, 64]");
    if (in->glb_blocks < 0 || in->glb_blocks > (1 << 20))
        return fail(FSM_ERR_ARG, "GLB workgroups lie in [0, 2^20]");
    t->cache->tune = *in;
    return FSM_OK;
}

uint64_t fsm_query_bytes(int64_t n, int metric, int m, int w) {
    Geometry G;
    if (m < 2 || w < 0 || w >= m || !make_geometry(n, m, &G)) return 0;
    const uint64_t rows = (uint64_t)G.S * (uint64_t)G.kml;
    const size_t per = lbip_slab_bytes(G);
    const uint64_t blocks = (uint64_t)std::min<int64_t>(G.S, per > (1u << 20) ? 1024 : 2048);
    // each buffer as DevBuf allocates it (rounded up)
    auto r = [](uint64_t x) {
        const uint64_t g = x >= (64u << 20) ? (64u << 20) : (64u << 10);
        return (x + g - 1) / g * g;
    };
    uint64_t b = 2 * r(8 * rows) + r(rows)          // window mean, std, invalid flag
                 + 2 * r(8 * (uint64_t)n)           // target envelope
                 + 6 * r(8 * (uint64_t)G.ts_n)      // LB, two candidate lists, matches,
                                                    // cutoff's second GLB bounds
                 + 4 * r(8 * (uint64_t)G.S)         // segment thresholds, representatives
                 + r(per * blocks)                  // LB_IP per-block scratch
                 + r(96 * (uint64_t)(m + 1)) + r(128) + r(16384);  // query tables, counters,
                                                                   // radix-select counts
    if (eld_uses_scratch(metric, m, w))
        b += r(sizeof(double) * 2 * (uint64_t)(2 * w + 2) * (uint64_t)eld_scratch_threads());
    return b;
}

int fsm_target_reuse_stats(fsm_target* t, int enable) {
    if (!t) return fail(FSM_ERR_ARG, "null target");
    t->reuse = enable != 0;
    t->stats_m = -1;
    t->env_w = -1;
    return FSM_OK;
}

int fsm_target_cache_limit(fsm_target* t, uint64_t bytes) {
    if (!t || !t->cache) return fail(FSM_ERR_ARG, "null target");
    std::lock_guard<std::mutex> lk(t->cache->mu);
    t->cache->cap = bytes;
    // entries beyond the new cap go as soon as no handle is using them
    while (t->cache->bytes > bytes) {
        if (t->cache->take_idle(t->cache->stats)) continue;
        if (!t->cache->take_idle(t->cache->env)) break;
    }
    return FSM_OK;
}

void fsm_target_trim(fsm_target* t) {
    if (!t) return;
    (void)hipStreamSynchronize(t->stream);
    DevBuf* bufs[] = {&t->lb, &t->list_a, &t->list_b, &t->out_idx, &t->out_dist, &t->slab,
                      &t->rep_lb, &t->rep_idx, &t->tu, &t->tl, &t->qtab, &t->scratch, &t->aux,
                      &t->hist};
    for (DevBuf* b : bufs) b->release();
    t->st.reset();
    t->env.reset();
    t->stats_m = -1;
    t->env_w = -1;
    t->own_st.a.release();
    t->own_st.b.release();
    t->own_st.c.release();
    t->own_env.a.release();
    t->own_env.b.release();
    // and the target's cached statistics / envelopes no handle is using
    if (t->cache) {
        std::lock_guard<std::mutex> lk(t->cache->mu);
        t->cache->drop_idle();
    }
}

// ------------------------------------------------------ out-of-core query ---
// fsm_query_file: the reference's file streaming (src/fsmeld.cpp:740-746,
// read_tail_segment :159-168) for targets larger than the device budget.
// The target is cut into chunks of whole macro-segments: a chunk of segments
// [s0, s1) is the sample range [s0 k_m_l, (s1 - 1) k_m_l + m_k), and the last
// chunk runs to the end of the file, so its tail segment is the reference's
// EOF-aligned one.  Run as a target of its own, a chunk has exactly the
// global segmentation (no tail except in the last chunk), so every window
// gets the same statistics, bounds and distance as in a whole-target query;
// its indices are shifted by s0 k_m_l.  Two handles alternate: while one
// chunk's cascade runs, a reader thread fills a pinned buffer with the next
// chunk and it is uploaded on a copy stream into the other handle.
namespace {

struct PinnedBuf {
    double* p = nullptr;
    size_t cap = 0;  // doubles
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc((void**)&p, sizeof(double) * n, hipHostMallocDefault);
        if (e == hipSuccess) cap = n;
        return e;
    }
};

struct ChunkPlan {
    int64_t off, n;  // sample range [off, off + n)
};

// Chunks of `per` full segments each; the last also takes the tail segment
// and at least one full segment (a chunk shorter than m_k would change the
// segment length).
std::vector<ChunkPlan> plan_chunks(const Geometry& G, int64_t per) {
    std::vector<ChunkPlan> v;
    if (G.forsize <= 0) {  // a single (tail) segment covers the target
        v.push_back({0, G.N});
        return v;
    }
    for (int64_t s0 = 0; s0 < G.forsize; s0 += per) {
        int64_t s1 = std::min(G.forsize, s0 + per);
        if (G.forsize - s1 < 1 && s1 < G.forsize) s1 = G.forsize;
        const bool last = s1 >= G.forsize;
        const int64_t off = s0 * G.kml;
        const int64_t end = last ? G.N : (s1 - 1) * G.kml + G.m_k;
        v.push_back({off, end - off});
        if (last) break;
    }
    return v;
}

// device bytes of one chunk handle: its samples and its query's buffers
uint64_t chunk_bytes(int64_t n, const fsm_query_desc* d) {
    return 8 * (uint64_t)n + fsm_query_bytes(n, d->metric, d->m, d->w);
}

struct Hit {
    double d;
    int64_t g;
};

}  // namespace

extern "C" int fsm_query_file(const char* path, const fsm_query_desc* d, const double* q,
                              int64_t qlen, uint64_t budget, fsm_result** out) {
    const auto wall0 = std::chrono::steady_clock::now();
    if (!path || !out || !d) return fail(FSM_ERR_ARG, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return fail(FSM_ERR_IO, "Cannot open target file.");
    const std::streamoff bytes = f.tellg();
    if (bytes <= 0 || bytes % (std::streamoff)sizeof(double))
        return fail(FSM_ERR_ARG, "Target must contain at least 2*m-1 complete double values.");
    const int64_t N = (int64_t)(bytes / (std::streamoff)sizeof(double));
    int rc = validate(d, N, q, qlen, true);
    if (rc) return rc;
    Geometry G;
    if (!make_geometry(N, d->m, &G))
        return fail(FSM_ERR_ARG, "Macro segment exceeds the supported FFT length.");
    if (fsm_device_count() < 1) return fail(FSM_ERR_NODEV, "no HIP device");

    // the most full segments per chunk whose two handles fit the budget
    int64_t per = std::max<int64_t>(G.forsize, 1);
    auto fits = [&](int64_t p) {
        uint64_t mx = 0;
        for (const ChunkPlan& c : plan_chunks(G, p)) mx = std::max(mx, chunk_bytes(c.n, d));
        return 2 * mx <= budget;
    };
    if (!fits(per)) {
        int64_t lo = 1, hi = per;
        if (!fits(lo))
            return fail(FSM_ERR_OOM, "the device budget holds no chunk of one macro-segment");
        while (hi - lo > 1) {
            const int64_t mid = lo + (hi - lo) / 2;
            (fits(mid) ? lo : hi) = mid;
        }
        per = lo;
    }
    const std::vector<ChunkPlan> chunks = plan_chunks(G, per);
    size_t cap = 0;
    for (const ChunkPlan& c : chunks) cap = std::max(cap, (size_t)c.n);

    // two handles (own samples, no shared cache) and two pinned buffers
    fsm_target* h[2] = {nullptr, nullptr};
    PinnedBuf pin[2];
    hipStream_t copy = nullptr;
    hipEvent_t up[2] = {nullptr, nullptr};
    auto cleanup = [&](int code) {
        if (copy) (void)hipStreamSynchronize(copy);
        for (auto& e : up)
            if (e) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
        for (auto* x : h)
            if (x) fsm_target_close(x);
        return code;
    };
    for (int s = 0; s < 2 && s < (int)chunks.size(); s++) {
        auto cache = std::make_shared<TargetCache>();  // cap 0: nothing shared
        rc = new_handle(0, std::shared_ptr<DevBuf>(new DevBuf(), [](DevBuf* b) {
                            b->release();
                            delete b;
                        }), std::move(cache), &h[s]);
        if (rc) return cleanup(rc);
        h[s]->reuse = false;
        if (h[s]->T->ensure(sizeof(double) * cap) != hipSuccess)
            return cleanup(fail(FSM_ERR_OOM, "device allocation of a target chunk failed"));
        if (pin[s].ensure(cap) != hipSuccess)
            return cleanup(fail(FSM_ERR_OOM, "pinned host allocation failed"));
        if (hipEventCreateWithFlags(&up[s], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(FSM_ERR_HIP, "hipEventCreate failed"));
    }
    if (hipStreamCreateWithFlags(&copy, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FSM_ERR_HIP, "hipStreamCreate failed"));

    // reading chunk c into pin[c % 2] (after that buffer's last upload)
    bool read_ok[2] = {true, true}, recorded[2] = {false, false};
    auto read_chunk = [&](size_t c) {
        const int s = (int)(c % 2);
        if (recorded[s]) (void)hipEventSynchronize(up[s]);
        std::ifstream in(path, std::ios::binary);
        in.seekg((std::streamoff)chunks[c].off * (std::streamoff)sizeof(double));
        in.read(reinterpret_cast<char*>(pin[s].p),
                (std::streamsize)(sizeof(double) * (size_t)chunks[c].n));
        read_ok[s] = (bool)in;
    };
    auto upload = [&](size_t c) -> int {
        const int s = (int)(c % 2);
        if (!read_ok[s]) return fail(FSM_ERR_IO, "short read of target file");
        HIP_TRY(hipMemcpyAsync(h[s]->T->p, pin[s].p, sizeof(double) * (size_t)chunks[c].n,
                               hipMemcpyHostToDevice, copy));
        HIP_TRY(hipEventRecord(up[s], copy));
        recorded[s] = true;
        h[s]->N = chunks[c].n;
        h[s]->stats_m = -1;
        h[s]->env_w = -1;
        return FSM_OK;
    };

    std::unique_ptr<fsm_result> R(new fsm_result());
    fsm_stats& st = R->st;
    st.ts_n = G.ts_n;
    st.k_m_l = G.kml;
    st.forsize = G.forsize;
    st.tail = G.tail;
    st.m_k = G.m_k;
    std::vector<Hit> hits;  // search / exact: in order; cutoff / top-k: candidates
    double theta = std::numeric_limits<double>::infinity();
    const bool is_cut = d->mode == FSM_MODE_CUTOFF, is_topk = d->mode == FSM_MODE_EXACT_TOPK;

    read_chunk(0);
    if ((rc = upload(0))) return cleanup(rc);
    std::thread reader;
    if (chunks.size() > 1) reader = std::thread(read_chunk, (size_t)1);
    for (size_t c = 0; c < chunks.size(); c++) {
        const int s = (int)(c % 2);
        if (c + 1 < chunks.size()) {  // next chunk: read, then upload while this one runs
            reader.join();
            if ((rc = upload(c + 1))) return cleanup(rc);
            if (c + 2 < chunks.size()) reader = std::thread(read_chunk, c + 2);
        }
        if (hipStreamWaitEvent(h[s]->stream, up[s], 0) != hipSuccess) {
            if (reader.joinable()) reader.join();
            return cleanup(fail(FSM_ERR_HIP, "hipStreamWaitEvent failed"));
        }
        // this chunk's query: windows [off, off + ts_n_c) of the target
        const int64_t off = chunks[c].off, nwin = chunks[c].n - d->m + 1;
        fsm_query_desc dc = *d;
        const bool excl = !(d->exclude_begin == -1 && d->exclude_end == -1);
        dc.exclude_begin = dc.exclude_end = -1;
        if (excl) {
            const int64_t b = std::max<int64_t>(d->exclude_begin - off, 0);
            const int64_t e = std::min<int64_t>(d->exclude_end - off, nwin);
            if (b < e) {
                dc.exclude_begin = b;
                dc.exclude_end = e;
            }
        }
        const bool known = (is_cut || is_topk) && (int64_t)hits.size() >= d->k;
        if (known) {  // later chunks: every window at or below the k-th so far
            dc.mode = is_cut ? FSM_MODE_SEARCH : FSM_MODE_EXACT;
            dc.cutoff = theta;
        }
        fsm_result* r = nullptr;
        const bool few = (is_cut || is_topk) && !known && nwin < d->k;
        rc = few ? FSM_ERR_FEWER_THAN_K : fsm_query(h[s], &dc, q, qlen, &r);
        if (rc == FSM_ERR_FEWER_THAN_K && (is_cut || is_topk)) {
            // fewer than k windows here: all of them take part
            dc.mode = FSM_MODE_EXACT;
            dc.cutoff = std::numeric_limits<double>::max();
            rc = fsm_query(h[s], &dc, q, qlen, &r);
        }
        if (rc) {
            if (reader.joinable()) reader.join();
            return cleanup(rc);
        }
        std::unique_ptr<fsm_result> rr(r);
        const fsm_stats& cs = rr->st;
        st.valid += cs.valid;
        st.invalid += cs.invalid;
        st.after_lbip += cs.after_lbip;
        st.after_glb += cs.after_glb;
        st.exact_evals += cs.exact_evals;
        st.eld_cells += cs.eld_cells;
        st.ms_stats += cs.ms_stats;
        st.ms_stats_finish += cs.ms_stats_finish;
        st.ms_env += cs.ms_env;
        st.ms_lbip += cs.ms_lbip;
        st.ms_glb += cs.ms_glb;
        st.ms_eld += cs.ms_eld;
        for (size_t i = 0; i < rr->idx.size(); i++) hits.push_back({rr->dist[i], rr->idx[i] + off});
        if (is_cut || is_topk) {
            // keep the k (distance, index)-smallest (top-k: and their ties)
            std::sort(hits.begin(), hits.end(), [](const Hit& a, const Hit& b) {
                return a.d != b.d ? a.d < b.d : a.g < b.g;
            });
            if ((int64_t)hits.size() >= d->k) {
                theta = hits[(size_t)d->k - 1].d;
                size_t keep = (size_t)d->k;
                if (is_topk)
                    while (keep < hits.size() && hits[keep].d <= theta) keep++;
                hits.resize(keep);
            }
        }
    }
    if (reader.joinable()) reader.join();
    rc = cleanup(FSM_OK);
    if ((is_cut || is_topk) && (int64_t)hits.size() < d->k)
        return fail(FSM_ERR_FEWER_THAN_K,
                    is_cut ? "Fewer valid, non-excluded candidates than requested k."
                           : "Fewer finite exact distances than requested k.");
    if (is_topk)  // ascending index, as the reference lists them
        std::sort(hits.begin(), hits.end(), [](const Hit& a, const Hit& b) { return a.g < b.g; });
    for (const Hit& x : hits) {
        R->idx.push_back(x.g);
        R->dist.push_back(x.d);
    }
    if (is_cut || is_topk) st.threshold = theta;
    st.matches = R->idx.size();
    st.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0)
            .count();
    *out = R.release();
    return rc;
}
