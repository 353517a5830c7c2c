"""Precision of LB_IP at the long-query shapes (m = 128 and 1024, w = 5% of m)
on a prefix of the plain large target.

LB_IP's sliding sums are exact 64-bit fixed-point sums of the FP64 penalties
rounded down (a certified lower bound, error below 2^-F m of the penalty
scale); the oracle evaluates the same bound with FP64 direct sums.  Bar:
where both chose the same thresholds the values agree within 1e-9 relative
(the reference's FFT guard, include/fsmeld_internal.h:9-19), no bound exceeds
an exact distance, and search returns exactly the exact scan's matches."""
import numpy as np
import pytest

import fsmeld_amd as F
import oracle as O
from fsmeld_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prefix():
    t = workloads.plain_large(n=1 << 26)["target"][: 1 << 22].copy()
    return t, F.Target(t)


@pytest.mark.parametrize("m,w", [(128, 6), (1024, 51)])
@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("msm", 0.5), ("lcss", 0.25)])
def test_lbip_precision_long_queries(gpu, prefix, m, w, metric, arg):
    t, T = prefix
    q, _ = F.synth_query(3 * m + 1, t, m, 0.1)
    lb, tu, tl = T.lbip(q, m, metric, w, arg)
    # the oracle on the first 2^20 samples: its full segments are the target's
    sub = t[: 1 << 20]
    olb, otu, otl = O.lbip_all(sub, q, m, metric, w, arg)
    G = F.geometry(sub.size, m)
    fs, kml = G["forsize"], G["k_m_l"]
    full = fs * kml
    seg = np.arange(full) // kml
    # Segments holding a window whose std is NaN (a constant run longer than
    # m makes sum(x^2)/m - mean^2 a tiny negative number) are excluded: there
    # the engine's window min / max propagate NaN and the reference's deques
    # are order-dependent (bounds differ, both admissible).
    _, sg, _ = O.window_stats(sub, m)
    nan_seg = np.zeros(fs, bool)
    np.logical_or.at(nan_seg, seg, ~np.isfinite(sg[:full]))
    same = (tu[:fs] == otu[:fs]) & (tl[:fs] == otl[:fs]) & ~nan_seg
    sel = same[seg] & np.isfinite(olb[:full])
    assert same.mean() >= 0.9
    a, b = lb[:full][sel], olb[:full][sel]
    rel = np.abs(a - b) / np.maximum(1.0, np.abs(b))
    print(f"m={m} w={w} {metric}: {nan_seg.mean():.3f} of segments hold a NaN-std window; "
          f"thresholds equal in {same.mean():.3f}; LB rel. error max {rel.max():.3g} "
          f"mean {rel.mean():.3g} over {sel.sum()} windows")
    assert rel.max() <= 1e-9
    # admissible: never above the exact distance
    _, _, inv = O.window_stats(t, m)
    win = np.flatnonzero(~inv)[::4099]
    d = T.eld(q, m, metric, w, arg, win)
    assert np.all(lb[win] <= d + 1e-9 * np.maximum(1.0, d))
    # identity of the match sets at the k = 10 cutoff
    r = T.cutoff(q, m, 10, metric, w, arg)
    s = T.search(q, m, r.threshold, metric, w, arg)
    e = T.exact(q, m, r.threshold, metric, w, arg)
    missing = np.setdiff1d(e.indices, s.indices)
    assert np.setdiff1d(s.indices, e.indices).size == 0
    assert np.all(e.distances[np.isin(e.indices, missing)] == r.threshold)


def _f32_gamma(m):
    nr = 4 * ((m + 18) // 4)  # products (and additions) per side and window
    return nr * 2.0 ** -23 / (1 - nr * 2.0 ** -23)


def _zdom(lb, metric, m):
    """Bound -> the sum z_l + z_u it came from (src/fsmeld.cpp:84-106)."""
    if metric in ("dtw", "erp"):
        return lb * lb
    return lb * m if metric == "lcss" else lb


def _check_f32_bounds(T, q, m, metric, w, arg, oracle_prefix=None):
    """The FP32 matrix-core bounds against the exact fixed-point bounds of the
    same thresholds: above them by no more than the fixed-point sums' own
    rounding (their 2^-F grain, far below the reference's 1e-9 guard), below
    them by no more than the certified slack; with oracle_prefix (the target's
    first samples), also against the oracle's FP64 direct sums over it, in the
    segments whose thresholds agree."""
    lb, tu, tl = T.lbip(q, m, metric, w, arg)
    lb32, tu32, tl32 = T.lbip(q, m, metric, w, arg, lbip_f32=True)
    assert np.array_equal(tu, tu32, equal_nan=True) and np.array_equal(tl, tl32, equal_nan=True)
    fin = np.isfinite(lb)
    assert np.array_equal(fin, np.isfinite(lb32))
    za, zb = _zdom(lb32[fin], metric, m), _zdom(lb[fin], metric, m)
    g = _f32_gamma(m)
    assert np.all(za <= zb * (1 + 1e-12) + 1e-9), "FP32 bound above the exact sums"
    assert np.all(za >= zb * (1 - 2.5 * g) - 1e-30), "FP32 bound looser than its slack"
    if oracle_prefix is not None:
        t = oracle_prefix
        olb, otu, otl = O.lbip_all(t, q, m, metric, w, arg)
        G = F.geometry(t.size, m)
        fs, kml = G["forsize"], G["k_m_l"]
        seg = np.arange(fs * kml) // kml
        # (segments holding a NaN-std window excluded, as in the test above)
        _, sg, _ = O.window_stats(t, m)
        nan_seg = np.zeros(fs, bool)
        np.logical_or.at(nan_seg, seg, ~np.isfinite(sg[:fs * kml]))
        same = (tu[:fs] == otu[:fs]) & (tl[:fs] == otl[:fs]) & ~nan_seg
        sel = same[seg] & np.isfinite(olb[:fs * kml])
        zo = _zdom(olb[:fs * kml][sel], metric, m)
        z32 = _zdom(lb32[:fs * kml][sel], metric, m)
        assert np.all(z32 <= zo * (1 + 1e-9) + 1e-9), "FP32 bound above the FP64 oracle"
        assert np.all(z32 >= zo * (1 - 2.5 * g) - 1e-9), "FP32 bound looser than its slack"
    rel = (zb - za) / np.maximum(zb, 1e-300)
    return float(rel.max()), float(rel.mean()), g



@pytest.mark.parametrize("m,w", [(128, 6), (1024, 51)])
@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("erp", 0.0), ("edr", 0.25),
                                        ("lcss", 0.25), ("msm", 0.5)])
def test_lbip_f32_mfma_certified(gpu, prefix, m, w, metric, arg):
    """FP32 matrix-core LB_IP (config 3): certified against the exact bounds,
    and cutoff / search answers identical to the exact path's and to the
    exact scan's."""
    t, T = prefix
    q, _ = F.synth_query(3 * m + 1, t, m, 0.1)
    mx, mean, g = _check_f32_bounds(T, q, m, metric, w, arg, oracle_prefix=t[:1 << 20])
    print(f"m={m} {metric}: FP32 bound below the exact one by max {mx:.3g} mean {mean:.3g} "
          f"(slack {g:.3g})")
    c = T.cutoff(q, m, 10, metric, w, arg)
    c32 = T.cutoff(q, m, 10, metric, w, arg, lbip_f32=True)
    assert c32.stats["lbip_f32"] == 1 and c.stats["lbip_f32"] == 0
    assert c32.threshold == c.threshold and np.array_equal(c32.indices, c.indices)
    s32 = T.search(q, m, c.threshold, metric, w, arg, lbip_f32=True)
    s = T.search(q, m, c.threshold, metric, w, arg)
    assert np.array_equal(s32.indices, s.indices)
    assert np.array_equal(s32.distances, s.distances)
    e = T.exact(q, m, c.threshold, metric, w, arg)
    missing = np.setdiff1d(e.indices, s32.indices)
    assert np.setdiff1d(s32.indices, e.indices).size == 0
    assert np.all(e.distances[np.isin(e.indices, missing)] == c.threshold)


@pytest.mark.parametrize("n,m,w,metric,arg,exclude", [
    (20000, 2, 0, "dtw", 0.0, None),
    (20000, 16, 1, "msm", 0.5, (300, 900)),
    (50001, 100, 5, "erp", 0.3, None),
    (70007, 333, 20, "lcss", 0.2, (1000, 5000)),
    (123457, 1000, 999, "edr", 0.1, None),
])
def test_lbip_f32_mfma_shapes(gpu, n, m, w, metric, arg, exclude):
    """Odd lengths, tail segments, radius extremes and exclusions: the FP32
    bounds stay certified, and search equals the exact path's search."""
    t = F.synth_target(n % 97 + 1, n, 0.001, 5.0, 3, 2 * m + 10, 2)
    q, _ = F.synth_query(m + 7, t, m, 0.1)
    with F.Target(t) as T:
        _check_f32_bounds(T, q, m, metric, w, arg)
        c = T.cutoff(q, m, 3, metric, w, arg, exclude=exclude)
        s = T.search(q, m, c.threshold, metric, w, arg, exclude=exclude)
        s32 = T.search(q, m, c.threshold, metric, w, arg, exclude=exclude, lbip_f32=True)
        assert np.array_equal(s.indices, s32.indices)
        if m <= 333:  # (the oracle's O(m^2) verifiers at m = 1000 take minutes)
            want, _ = O.search(t, q, m, c.threshold, metric, w, arg, exclude=exclude,
                               threads=8)
            assert np.array_equal(s32.indices, want)


def test_lbip_f32_falls_back_beyond_8192(gpu):
    """m_k above 8192 keeps the fixed-point sums and says so."""
    t = F.synth_target(5, 1 << 16, 0.0, 0.0, 0, 0, 0)
    m = 1100
    q, _ = F.synth_query(9, t, m, 0.1)
    with F.Target(t) as T:
        r = T.search(q, m, 1.0, "dtw", 10, 0.0, lbip_f32=True)
        assert r.stats["lbip_f32"] == 0
