"""Stage entry points on device memory (fsm_stage_*_dev): the window list and
the outputs live in device memory and the work is ordered on a caller's
stream.  Every output equals the host entry's, bit for bit, on a target with
NaN samples and constant runs, for every metric; the window list is written
by an asynchronous copy on the caller's stream just before the call, so the
stage must wait for that stream; an out-of-range window index is refused
before any kernel reads through it."""
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
HIP.hipMemcpyAsync.argtypes = [VP, VP, ctypes.c_size_t, ctypes.c_int, VP]
HIP.hipStreamCreate.argtypes = [ctypes.POINTER(VP)]
HIP.hipStreamDestroy.argtypes = [VP]
HIP.hipStreamSynchronize.argtypes = [VP]
H2D, D2H, D2D = 1, 2, 3

METRICS = [("dtw", 0.0), ("erp", 0.3), ("edr", 0.25), ("lcss", 0.25), ("msm", 0.5),
           ("swale", 0.25)]
M, W = 128, 12


class Dev:
    """A device buffer of nbytes (freed by free())."""

    def __init__(self, nbytes, host=None):
        self.p = VP()
        self.n = max(int(nbytes), 8)
        assert HIP.hipMalloc(ctypes.byref(self.p), self.n) == 0
        if host is not None:
            assert HIP.hipMemcpy(self.p, host.ctypes.data, host.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert HIP.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        HIP.hipFree(self.p)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.fixture(scope="module")
def tgt():
    t = F.synth_target(21, 1 << 16, 0.001, 5.0, 4, 200, 3)  # spikes, runs and NaNs
    q, _ = F.synth_query(22, t, M, 0.1)
    return t, F.Target(t), q


@pytest.fixture
def stream():
    s = VP()
    assert HIP.hipStreamCreate(ctypes.byref(s)) == 0
    yield s
    HIP.hipStreamDestroy(s)


def test_window_stats_dev(gpu, tgt, stream):
    t, T, _ = tgt
    n = t.size - M + 1
    mu, sg, inv = Dev(8 * n), Dev(8 * n), Dev(n)
    try:
        T.window_stats_dev(M, mu.p.value, sg.p.value, inv.p.value, stream=stream.value,
                           exclude=(1000, 3000))
        assert HIP.hipStreamSynchronize(stream) == 0
        hmu, hsg, hinv = T.window_stats(M, exclude=(1000, 3000))
        assert np.array_equal(bits(mu.get(np.float64, n)), bits(hmu))
        assert np.array_equal(bits(sg.get(np.float64, n)), bits(hsg))
        assert np.array_equal(inv.get(np.uint8, n).astype(bool), hinv)
        assert hinv[1000:3000].all() and hinv.sum() > 2000
    finally:
        for b in (mu, sg, inv):
            b.free()


@pytest.mark.parametrize("metric,arg", METRICS)
def test_lbip_dev(gpu, tgt, stream, metric, arg):
    t, T, q = tgt
    n = t.size - M + 1
    S = F.geometry(t.size, M)["S"]
    lb, tu, tl = Dev(8 * n), Dev(8 * S), Dev(8 * S)
    try:
        T.lbip_dev(q, M, metric, W, arg, lb.p.value, tu.p.value, tl.p.value,
                   stream=stream.value)
        assert HIP.hipStreamSynchronize(stream) == 0
        hlb, htu, htl = T.lbip(q, M, metric, W, arg)
        assert np.array_equal(bits(lb.get(np.float64, n)), bits(hlb))
        assert np.array_equal(bits(tu.get(np.float64, S)), bits(htu))
        assert np.array_equal(bits(tl.get(np.float64, S)), bits(htl))
    finally:
        for b in (lb, tu, tl):
            b.free()


@pytest.mark.parametrize("metric,arg", METRICS)
def test_glb_eld_dev(gpu, tgt, stream, metric, arg):
    t, T, q = tgt
    _, _, inv = T.window_stats(M)
    win = np.flatnonzero(~inv)[::23].astype(np.int64)
    n = win.size
    d_hold = T.eld(q, M, metric, W, arg, win)
    # the median GLB value (no abandoning at 1e300): about half kept
    bsf = float(np.median(T.glb(q, M, metric, W, arg, win, 1e300)))
    stage, dwin = Dev(8 * n, win), Dev(8 * n)
    glb, keep, eld = Dev(8 * n), Dev(n), Dev(8 * n)
    try:
        # the list reaches dwin by a copy queued on the caller's stream
        assert HIP.hipMemcpyAsync(dwin.p, stage.p, 8 * n, D2D, stream) == 0
        T.glb_dev(q, M, metric, W, arg, dwin.p.value, n, glb.p.value, bsf, stream=stream.value)
        T.glb_keep_dev(q, M, metric, W, arg, dwin.p.value, n, keep.p.value, bsf,
                       stream=stream.value)
        T.eld_dev(q, M, metric, W, arg, dwin.p.value, n, eld.p.value, stream=stream.value)
        assert HIP.hipStreamSynchronize(stream) == 0
        assert np.array_equal(bits(glb.get(np.float64, n)),
                              bits(T.glb(q, M, metric, W, arg, win, bsf)))
        hkeep = T.glb_keep(q, M, metric, W, arg, win, bsf)
        assert np.array_equal(keep.get(np.uint8, n).astype(bool), hkeep)
        assert 0 < hkeep.sum() < n
        assert np.array_equal(bits(eld.get(np.float64, n)), bits(d_hold))
    finally:
        for b in (stage, dwin, glb, keep, eld):
            b.free()


def test_dev_window_out_of_range(gpu, tgt, stream):
    t, T, q = tgt
    win = np.array([0, 5, t.size - M + 1], dtype=np.int64)  # the last is one past ts_n
    dwin, out = Dev(8 * win.size, win), Dev(8 * win.size)
    try:
        with pytest.raises(F.FsmeldError, match="out of range"):
            T.eld_dev(q, M, "dtw", W, 0.0, dwin.p.value, win.size, out.p.value,
                      stream=stream.value)
        # the null stream works too, and the target stays usable
        ok = np.array([0, 5, 9], dtype=np.int64)
        assert HIP.hipMemcpy(dwin.p, ok.ctypes.data, ok.nbytes, H2D) == 0
        T.eld_dev(q, M, "dtw", W, 0.0, dwin.p.value, ok.size, out.p.value, stream=0)
        assert HIP.hipStreamSynchronize(None) == 0
        assert np.array_equal(bits(out.get(np.float64, ok.size)),
                              bits(T.eld(q, M, "dtw", W, 0.0, ok)))
    finally:
        dwin.free()
        out.free()
