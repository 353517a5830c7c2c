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
