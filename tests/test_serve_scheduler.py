"""The serving loop's scheduler (fsmeld_amd/csrc/serve.cpp) on the CPU,
linked against a stub engine whose queries sleep instead of running on a GPU.

Checks that admission never deadlocks when the idle worker holding memory is
not the one woken first, that the workers' share of the budget is three
quarters of it, and that cancel / poll behave as fsmeld.h documents."""
import ctypes
import os
import subprocess
import threading
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STUB = r'''
#include <chrono>
#include <cstdint>
#include <mutex>
#include <thread>
#include "fsmeld.h"
struct fsm_target { int64_t n; uint64_t held; };
struct fsm_result { int m; };
static std::mutex mu;
static uint64_t live = 0, peak = 0, cache_cap = 0;
extern "C" {
int fsm_target_share(fsm_target* b, fsm_target** out) { *out = new fsm_target{b->n, 0}; return 0; }
void fsm_target_close(fsm_target* t) { delete t; }
int64_t fsm_target_length(const fsm_target* t) { return t->n; }
int fsm_target_cache_limit(fsm_target*, uint64_t b) { cache_cap = b; return 0; }
void fsm_target_trim(fsm_target* t) {
    std::lock_guard<std::mutex> lk(mu);
    live -= t->held;
    t->held = 0;
}
uint64_t fsm_query_bytes(int64_t, int, int m, int) { return (uint64_t)m * 1000; }
int fsm_query(fsm_target* t, const fsm_query_desc* d, const double*, int64_t, fsm_result** out) {
    {
        std::lock_guard<std::mutex> lk(mu);
        const uint64_t b = (uint64_t)d->m * 1000;
        if (b > t->held) { live += b - t->held; t->held = b; }
        if (live > peak) peak = live;
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(d->w));
    *out = new fsm_result{d->m};
    return 0;
}
void fsm_result_free(fsm_result* r) { delete r; }
uint64_t stub_peak() { return peak; }
uint64_t stub_cache_cap() { return cache_cap; }
fsm_target* stub_target() { return new fsm_target{1000, 0}; }
}
'''


@pytest.fixture(scope="module")
def sched(tmp_path_factory):
    d = tmp_path_factory.mktemp("sched")
    (d / "stub.cpp").write_text(STUB)
    so = d / "libsched.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-pthread",
                    "-I", os.path.join(ROOT, "include"), "-o", str(so), str(d / "stub.cpp"),
                    os.path.join(ROOT, "fsmeld_amd", "csrc", "serve.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    VP = ctypes.c_void_p
    lib.stub_target.restype = VP
    lib.stub_peak.restype = ctypes.c_uint64
    lib.stub_cache_cap.restype = ctypes.c_uint64
    lib.fsm_server_create.argtypes = [VP, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(VP)]
    lib.fsm_server_submit.argtypes = [VP, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
    lib.fsm_server_poll.argtypes = [VP, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(VP),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    lib.fsm_server_cancel.argtypes = [VP, ctypes.c_uint64]
    lib.fsm_server_destroy.argtypes = [VP]
    return lib


class Desc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("metric", ctypes.c_int), ("m", ctypes.c_int),
                ("w", ctypes.c_int), ("arg", ctypes.c_double), ("cutoff", ctypes.c_double),
                ("k", ctypes.c_int), ("exclude_begin", ctypes.c_int64),
                ("exclude_end", ctypes.c_int64)]


def submit(lib, s, m, sleep_ms):
    d = Desc(0, 1, m, sleep_ms, 0.0, 1.0, 0, -1, -1)
    q = (ctypes.c_double * 1)(0.0)
    tk = ctypes.c_uint64()
    rc = lib.fsm_server_submit(s, ctypes.byref(d), q, 1, ctypes.byref(tk))
    return rc, tk.value


def poll(lib, s, tk, wait_ms):
    r, st, lat = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_double()
    return lib.fsm_server_poll(s, tk, wait_ms, ctypes.byref(r), ctypes.byref(st),
                               ctypes.byref(lat))


def test_admission_wakes_the_idle_holder(sched):
    """Workers' share 3000 (budget 4000): one query of 2000 fits at a time.
    After each query its worker stays idle holding 2000, so the next one only
    fits once that worker releases or runs it; every query must finish."""
    lib = sched
    t = lib.stub_target()
    s = ctypes.c_void_p()
    assert lib.fsm_server_create(t, 4000, 3, ctypes.byref(s)) == 0
    assert lib.stub_cache_cap() == 1000
    for rnd in range(60):
        rc, tk = submit(lib, s, 2, 1 + rnd % 3)
        assert rc == 0
        assert poll(lib, s, tk, 5000) == 0, f"round {rnd}: query never admitted"
    # several at once, and one that can never fit the workers' share
    tks = [submit(lib, s, 2, 2)[1] for _ in range(8)]
    assert all(poll(lib, s, tk, 10000) == 0 for tk in tks)
    assert submit(lib, s, 4, 1)[0] == -2   # 4000 > 3000: FSM_ERR_OOM
    assert lib.stub_peak() <= 3000
    lib.fsm_server_destroy(s)


def test_cancel_and_pending_poll(sched):
    lib = sched
    t = lib.stub_target()
    s = ctypes.c_void_p()
    assert lib.fsm_server_create(t, 4000, 1, ctypes.byref(s)) == 0
    _, first = submit(lib, s, 1, 200)
    _, queued = submit(lib, s, 1, 1)
    t0 = time.perf_counter()
    assert poll(lib, s, first, 0) == 1          # FSM_PENDING, at once
    assert time.perf_counter() - t0 < 0.1
    assert lib.fsm_server_cancel(s, queued) == 0
    assert poll(lib, s, queued, -1) == -7       # FSM_ERR_CANCELLED
    assert poll(lib, s, first, -1) == 0
    assert poll(lib, s, first, 0) == -1         # retired ticket
    lib.fsm_server_destroy(s)
