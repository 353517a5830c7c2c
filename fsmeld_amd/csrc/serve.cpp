// serve.cpp -- the serving loop of include/fsmeld.h (fsm_server_*), built on
// the public C ABI only: each worker thread drives its own handle
// (fsm_target_share), so queries run concurrently on separate HIP streams.
//
// Memory: a quarter of the budget is the target's shared statistics /
// envelope cache (its cap is set to that); the workers' buffers share the
// rest.  Scheduling: first come, first served.  A free worker takes the
// oldest queued query when the memory held by every worker (the largest
// estimate it has served since it last released its buffers) plus this
// query's estimate fits the workers' share; otherwise idle workers release
// their cached buffers (fsm_target_trim) and the query waits for running
// ones to finish.  A worker that holds nothing and finds the oldest query
// too large wakes the others, so an idle worker whose buffers block it
// releases them.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fsmeld.h"

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { QUEUED, RUNNING, DONE, CANCELLED };

struct Request {
    fsm_query_desc desc{};
    std::vector<double> query;
    uint64_t bytes = 0;
    ReqState state = QUEUED;
    bool cancel = false;
    fsm_result* result = nullptr;
    int status = 0;
    Clock::time_point submitted, finished;
};

}  // namespace

struct fsm_server {
    fsm_target* base = nullptr;
    uint64_t budget = 0;  // the workers' share of mem_budget
    std::vector<fsm_target*> handles;
    std::vector<uint64_t> held;  // per worker: bytes its handle may hold
    std::vector<bool> busy;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Request> reqs;
    uint64_t next_ticket = 1;
    int64_t running = 0;
    bool stop = false;
    std::vector<std::thread> threads;

    uint64_t held_by_others(size_t j) const {
        uint64_t s = 0;
        for (size_t k = 0; k < held.size(); k++)
            if (k != j) s += held[k];
        return s;
    }

    // handles run on the device they were opened on (device 0)
    void worker(size_t j) {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (stop && queue.empty()) return;
            if (!queue.empty()) {
                const uint64_t id = queue.front();
                Request& r = reqs.at(id);
                const uint64_t need = std::max(held[j], r.bytes);
                if (held_by_others(j) + need <= budget) {
                    queue.pop_front();
                    r.state = RUNNING;
                    held[j] = need;
                    busy[j] = true;
                    running++;
                    const fsm_query_desc d = r.desc;
                    const std::vector<double>* q = &r.query;
                    lk.unlock();
                    fsm_result* res = nullptr;
                    const int st = fsm_query(handles[j], &d, q->data(), (int64_t)q->size(), &res);
                    lk.lock();
                    Request& rr = reqs.at(id);
                    rr.finished = Clock::now();
                    rr.status = st;
                    if (rr.cancel) {
                        fsm_result_free(res);
                        rr.state = CANCELLED;
                    } else {
                        rr.result = res;
                        rr.state = DONE;
                    }
                    busy[j] = false;
                    running--;
                    done_cv.notify_all();
                    work_cv.notify_all();
                    continue;
                }
                // does not fit: release this idle worker's cached buffers
                if (held[j] > 0) {
                    lk.unlock();
                    fsm_target_trim(handles[j]);
                    lk.lock();
                    held[j] = 0;
                    work_cv.notify_all();
                    continue;
                }
                // nothing of its own to release: wake the idle workers that
                // hold buffers (they release them, or fit the query)
                for (size_t k = 0; k < held.size(); k++)
                    if (k != j && !busy[k] && held[k] > 0) {
                        work_cv.notify_all();
                        break;
                    }
            }
            work_cv.wait(lk);
        }
    }
};

int fsm_server_create(fsm_target* t, uint64_t mem_budget, int max_workers, fsm_server** out) {
    if (!t || !out || max_workers < 1 || mem_budget == 0) return FSM_ERR_ARG;
    *out = nullptr;
    auto s = std::make_unique<fsm_server>();
    s->base = t;
    const uint64_t cache_share = mem_budget / 4;
    s->budget = mem_budget - cache_share;
    const int lrc = fsm_target_cache_limit(t, cache_share);
    if (lrc) return lrc;
    for (int j = 0; j < max_workers; j++) {
        fsm_target* h = nullptr;
        const int rc = fsm_target_share(t, &h);
        if (rc) {
            for (fsm_target* x : s->handles) fsm_target_close(x);
            return rc;
        }
        s->handles.push_back(h);
    }
    s->held.assign(s->handles.size(), 0);
    s->busy.assign(s->handles.size(), false);
    fsm_server* p = s.get();
    for (size_t j = 0; j < p->handles.size(); j++) p->threads.emplace_back([p, j] { p->worker(j); });
    *out = s.release();
    return FSM_OK;
}

int fsm_server_submit(fsm_server* s, const fsm_query_desc* d, const double* q, int64_t len,
                      uint64_t* ticket) {
    if (!s || !d || !q || len < 0 || !ticket) return FSM_ERR_ARG;
    const uint64_t bytes =
        fsm_query_bytes(fsm_target_length(s->base), d->metric, d->m, d->w);
    if (bytes == 0) return FSM_ERR_ARG;
    if (bytes > s->budget) return FSM_ERR_OOM;
    Request r;
    r.desc = *d;
    r.query.assign(q, q + len);
    r.bytes = bytes;
    r.submitted = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return FSM_ERR_ARG;
    const uint64_t id = s->next_ticket++;
    s->reqs.emplace(id, std::move(r));
    s->queue.push_back(id);
    *ticket = id;
    s->work_cv.notify_all();
    return FSM_OK;
}

int fsm_server_poll(fsm_server* s, uint64_t ticket, int wait_ms, fsm_result** out, int* status,
                    double* latency_ms) {
    if (!s || !out || !status) return FSM_ERR_ARG;
    *out = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return FSM_ERR_ARG;
    auto finished = [&] {
        const ReqState st = s->reqs.at(ticket).state;
        return st == DONE || st == CANCELLED;
    };
    if (wait_ms < 0) s->done_cv.wait(lk, finished);
    else if (wait_ms > 0) s->done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms), finished);
    Request& r = s->reqs.at(ticket);
    if (r.state == CANCELLED) {
        s->reqs.erase(ticket);
        return FSM_ERR_CANCELLED;
    }
    if (r.state != DONE) return FSM_PENDING;
    *out = r.result;
    *status = r.status;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r.finished - r.submitted).count();
    s->reqs.erase(ticket);
    return FSM_OK;
}

int fsm_server_cancel(fsm_server* s, uint64_t ticket) {
    if (!s) return FSM_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return FSM_ERR_ARG;
    Request& r = it->second;
    if (r.state == QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), ticket));
        r.state = CANCELLED;
        s->done_cv.notify_all();
        s->work_cv.notify_all();
    } else if (r.state == RUNNING) {
        r.cancel = true;
    } else if (r.state == DONE) {
        fsm_result_free(r.result);
        r.result = nullptr;
        r.state = CANCELLED;
    }
    return FSM_OK;
}

int fsm_server_info(fsm_server* s, int* workers, int64_t* queued, int64_t* running) {
    if (!s) return FSM_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (workers) *workers = (int)s->handles.size();
    if (queued) *queued = (int64_t)s->queue.size();
    if (running) *running = s->running;
    return FSM_OK;
}

void fsm_server_destroy(fsm_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (uint64_t id : s->queue) s->reqs.at(id).state = CANCELLED;
        s->queue.clear();
        s->work_cv.notify_all();
        s->done_cv.notify_all();
    }
    for (auto& th : s->threads) th.join();
    for (auto& kv : s->reqs)
        if (kv.second.result) fsm_result_free(kv.second.result);
    for (fsm_target* h : s->handles) fsm_target_close(h);
    delete s;
}
