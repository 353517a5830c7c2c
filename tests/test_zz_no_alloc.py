"""fsm_eld_pairs_noalloc (device pointers, a caller's stream, no device
allocation inside) gives the same bits as fsm_eld_pairs, on two calls in a
row with the second one smaller, for every metric and both kernel paths;
sizes beyond its static scratch are rejected with FSM_ERR_ARG."""
import ctypes

import numpy as np
import pytest

import fsmeld_amd as F

pytestmark = pytest.mark.gpu

# the HIP runtime the engine library itself is linked against
HIP = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
VP = ctypes.c_void_p
HIP.hipMalloc.argtypes = [ctypes.POINTER(VP), ctypes.c_size_t]
HIP.hipFree.argtypes = [VP]
HIP.hipMemcpy.argtypes = [VP, VP, ctypes.c_size_t, ctypes.c_int]
HIP.hipStreamCreate.argtypes = [ctypes.POINTER(VP)]
HIP.hipStreamDestroy.argtypes = [VP]
HIP.hipStreamSynchronize.argtypes = [VP]
LIB = ctypes.CDLL(F.lib_path)
LIB.fsm_eld_pairs_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      VP, VP, ctypes.c_int64, VP, VP]
H2D, D2H = 1, 2


def dev(a):
    p = VP()
    assert HIP.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
    assert HIP.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def noalloc(metric, q, c, w, arg, stream):
    q = np.ascontiguousarray(q, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    n = c.shape[0]
    dq, dc, do = dev(q), dev(c), dev(np.zeros(n))
    try:
        rc = LIB.fsm_eld_pairs_noalloc(F.METRICS[metric], q.size, w, arg, dq, dc, n, do, stream)
        assert rc == 0, F._lib.fsm_last_error()
        assert HIP.hipStreamSynchronize(stream) == 0
        out = np.empty(n)
        assert HIP.hipMemcpy(out.ctypes.data, do, out.nbytes, D2H) == 0
        return out
    finally:
        for p in (dq, dc, do):
            HIP.hipFree(p)


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("erp", 0.3), ("edr", 0.25),
                                        ("lcss", 0.25), ("msm", 0.5)])
def test_noalloc_equals_eld_pairs(gpu, metric, arg):
    rng = np.random.default_rng(F.METRICS[metric])
    stream = VP()
    assert HIP.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # register band (n >= 1024) then, smaller, wave kernel (n < 1024);
        # wide band (wave band) then smaller again
        for m, w, n in [(200, 12, 3000), (64, 5, 300), (500, 40, 1500), (33, 33 - 1, 50)]:
            q = rng.normal(size=m)
            c = rng.normal(size=(n, m))
            c[:5] = q + rng.normal(size=(5, m)) * 1e-3
            got = noalloc(metric, q, c, w, arg, stream)
            want = F.eld_pairs(metric, q, c, w, arg)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (m, w, n)
        # the null stream works too
        q = rng.normal(size=40)
        c = rng.normal(size=(20, 40))
        got = noalloc(metric, q, c, 4, arg, None)
        assert np.array_equal(got.view(np.uint64), F.eld_pairs(metric, q, c, 4, arg).view(np.uint64))
    finally:
        HIP.hipStreamDestroy(stream)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    one = VP(8)   # never dereferenced: the arguments are refused first
    assert LIB.fsm_eld_pairs_noalloc(1, 4097, 3, 0.0, one, one, 1, one, None) == -1
    assert LIB.fsm_eld_pairs_noalloc(1, 300, 200, 0.0, one, one, 1, one, None) == -1
    assert LIB.fsm_eld_pairs_noalloc(1, 64, 64, 0.0, one, one, 1, one, None) == -1
    assert LIB.fsm_eld_pairs_noalloc(1, 64, 3, 0.0, None, None, 0, None, None) == 0
