"""Pins every other oracle function with checks a plausible mistake would
fail: brute force on small inputs, closed forms, hand-computed known
answers, textbook reductions and end-to-end invariants of the cascade."""
import math

import numpy as np
import pytest

import fsmeld_amd as F
import oracle as O

METRICS = [("dtw", 0.0), ("erp", 0.0), ("edr", 0.25), ("lcss", 0.25), ("msm", 0.5),
           ("swale", 0.25)]


# ---------------------------------------------------------------- mvmean ---
def test_mvmean_two_pass_and_nan_semantics():
    rng = np.random.default_rng(1)
    a = np.cumsum(rng.normal(size=600)) + 50.0
    l = 37
    mu, si = O.mvmean(a, l)
    win = np.lib.stride_tricks.sliding_window_view(a, l)
    assert np.allclose(mu, win.mean(axis=1), rtol=1e-12, atol=1e-12)
    assert np.allclose(si, win.std(axis=1), rtol=1e-9, atol=1e-9)   # population std (ddof=0)
    # first window: a plain left-to-right sum
    s = 0.0
    for x in a[:l]:
        s += x
    assert mu[0] == s / l
    # non-finite samples poison exactly the windows that contain them, then recover
    b = a.copy()
    b[100] = np.nan
    b[300] = np.inf
    mu2, si2 = O.mvmean(b, l)
    bad = np.zeros(mu2.size, bool)
    for p in (100, 300):
        bad[max(0, p - l + 1): p + 1] = True
    assert np.array_equal(np.isnan(mu2), bad) and np.array_equal(np.isnan(si2), bad)
    assert np.allclose(mu2[~bad], win.mean(axis=1)[~bad], rtol=1e-9)


def test_invalid_mask_definition():
    rng = np.random.default_rng(2)
    a = np.cumsum(rng.normal(size=400))
    a[50:150] = a[50]           # constant run: windows inside it have ~zero std
    a[300] = np.nan
    m = 20
    mu, si = O.mvmean(a, m)
    inv = O.invalid_mask(a, si, m)
    win = np.lib.stride_tricks.sliding_window_view(a, m)
    nonfinite = ~np.isfinite(win).all(axis=1)
    want = nonfinite | ~np.isfinite(si) | (si <= 1e-12)
    assert np.array_equal(inv, want)
    assert inv[60] and inv[300] and not inv[0]


# --------------------------------------------------------------- windows ---
def test_min_max_window_brute_force():
    rng = np.random.default_rng(3)
    for n, m in [(50, 7), (64, 1), (40, 40), (100, 33)]:
        kml = n - m + 1
        a = rng.normal(size=kml)
        mn, mx = O.min_max_window(a, n, m)
        for i in range(n):
            lo, hi = max(0, i - m + 1), min(i, kml - 1)
            assert mn[i] == a[lo:hi + 1].min() and mx[i] == a[lo:hi + 1].max()


def test_lemire_envelope_brute_force():
    rng = np.random.default_rng(4)
    for n, r in [(30, 0), (30, 3), (100, 25), (12, 11)]:
        a = rng.normal(size=n)
        lo, hi = O.lemire(a, r)
        for i in range(n):
            s = a[max(0, i - r): min(n, i + r + 1)]
            assert lo[i] == s.min() and hi[i] == s.max()


def test_normalized_arrays_closed_form():
    U = np.array([2.0, -1.0, 0.5])
    L = np.array([1.0, -3.0, 0.5])
    mxm, mnm = np.array([1.5, 0.0, 0.25]), np.array([0.5, -0.5, 0.0])
    mxs, mns = np.array([2.0, 4.0, 1.0]), np.array([1.0, 2.0, 0.5])
    un, ln = O.normalized_arrays(U, L, mxm, mnm, mxs, mns)
    # U_n = (U - min_mean) / (min_std if >= 0 else max_std); L_n symmetric (paper Eqs. 33-34)
    assert un.tolist() == [1.5, -0.125, 1.0]
    assert ln.tolist() == [-0.5, -1.5, 0.25]


# ------------------------------------------------------------ thresholds ---
def h_upper(metric, q, t, C):
    if metric == "dtw":
        return sum((x - t) ** 2 for x in q if x > t)
    if metric == "erp":
        return sum(min((x - t) ** 2, (x - C) ** 2) for x in q if x > t)
    if metric == "msm":
        return sum(min(x - t, C) for x in q if x > t)
    if metric == "swale":   # p = 5 beyond eps, r = 1 within it (find_t_lp.cpp:827-927)
        return sum(5 if x - C > t else 1 for x in q if x >= t)
    return sum(1 for x in q if x - C > t)


def h_lower(metric, q, t, C):
    if metric == "dtw":
        return sum((t - x) ** 2 for x in q if x < t)
    if metric == "erp":
        return sum(min((x - t) ** 2, (x - C) ** 2) for x in q if x < t)
    if metric == "msm":
        return sum(min(t - x, C) for x in q if x < t)
    if metric == "swale":   # find_t_lp.cpp:399-509
        return sum(5 if x + C < t else 1 for x in q if x <= t)
    return sum(1 for x in q if x + C < t)


@pytest.mark.parametrize("metric,arg", METRICS)
def test_thresholds_brute_force_objective(metric, arg):
    """t_u = argmax_t h_u(t) * #{U_n <= t}, first in ascending order;
    t_l = argmax_t h_l(t) * #{L_n >= t}, first in descending order, over the
    envelope values at positions [m-1, l-m] (src/find_t_lp.cpp)."""
    rng = np.random.default_rng(hash(metric) % 1000)
    for trial in range(15):
        m = int(rng.integers(3, 12))
        l = int(rng.integers(2 * m, 6 * m))
        Un = rng.normal(size=l) + 0.5
        Ln = rng.normal(size=l) - 0.5
        Un[5:9] = Un[5]      # consecutive duplicates (run-length dedupe)
        q = rng.normal(size=m)
        tu, tl = O.thresholds(metric, Un, Ln, q, arg)
        vals_u = Un[m - 1: l - m + 1]
        vals_l = Ln[m - 1: l - m + 1]
        cand_u = sorted(set(vals_u.tolist()))
        score_u = [h_upper(metric, q, t, arg) * np.sum(vals_u <= t) for t in cand_u]
        best_u = cand_u[int(np.argmax(score_u))]
        cand_l = sorted(set(vals_l.tolist()), reverse=True)
        score_l = [h_lower(metric, q, t, arg) * np.sum(vals_l >= t) for t in cand_l]
        best_l = cand_l[int(np.argmax(score_l))]
        # near-ties may resolve differently under rounding; require exact argmax otherwise
        su = sorted(score_u)
        if len(su) < 2 or su[-1] - su[-2] > 1e-9 * max(1.0, su[-1]):
            assert tu == best_u, (trial, tu, best_u)
        sl = sorted(score_l)
        if len(sl) < 2 or sl[-1] - sl[-2] > 1e-9 * max(1.0, sl[-1]):
            assert tl == best_l, (trial, tl, best_l)


def test_q_to_h_known_answers():
    q = np.array([-2.0, 0.0, 1.5])
    # DTW: (t_l - q)^2 below t_l, (q - t_u)^2 above t_u
    hl, hu = O.q_to_h("dtw", q, 0.5, 1.0, 0.0)
    assert hl.tolist() == [6.25, 0.25, 0.0] and hu.tolist() == [0.0, 0.0, 0.25]
    # ERP g=0.25: min((q-t)^2, (q-g)^2)
    hl, hu = O.q_to_h("erp", q, 0.5, 1.0, 0.25)
    assert hl.tolist() == [5.0625, 0.0625, 0.0] and hu.tolist() == [0.0, 0.0, 0.25]
    # MSM c=1: min(c, |q - t|)
    hl, hu = O.q_to_h("msm", q, 0.5, 1.0, 1.0)
    assert hl.tolist() == [1.0, 0.5, 0.0] and hu.tolist() == [0.0, 0.0, 0.5]
    # EDR / LCSS eps=0.25: indicator of q beyond t -/+ eps
    for metric in ("edr", "lcss"):
        hl, hu = O.q_to_h(metric, q, 0.5, 1.0, 0.25)
        assert hl.tolist() == [1.0, 1.0, 0.0] and hu.tolist() == [0.0, 0.0, 1.0]
    # SWALE eps=0.75: 5 beyond t -/+ eps, 1 between t and t -/+ eps, else 0
    # (q = -2: below 0.5 - 0.75; q = 0: in [-0.25, 0.5); q = 1.5: in (1, 1.75])
    hl, hu = O.q_to_h("swale", q, 0.5, 1.0, 0.75)
    assert hl.tolist() == [5.0, 1.0, 0.0] and hu.tolist() == [0.0, 0.0, 1.0]


def test_sip_direct_is_a_correlation():
    rng = np.random.default_rng(5)
    g = (rng.random(200) < 0.4).astype(float)
    h = rng.random(17)
    z = O.sip_direct(g, h)
    want = np.correlate(g, h, mode="valid")
    assert np.allclose(z[16:], want, rtol=1e-13, atol=1e-13)


def test_lbip_cutoff_guard_known_answers():
    assert not O.lbip_exceeds_cutoff(float("nan"), 1.0, "msm")
    assert not O.lbip_exceeds_cutoff(float("inf"), 1.0, "msm")
    assert not O.lbip_exceeds_cutoff(1.0 + 5e-10, 1.0, "msm")    # within the 1e-9 guard
    assert O.lbip_exceeds_cutoff(1.0 + 2e-9, 1.0, "msm")
    # DTW/ERP compare in squared units near zero: sqrt((0+1e-9)^2 + 1e-9) ~ 3.16e-5
    assert not O.lbip_exceeds_cutoff(2e-5, 0.0, "dtw")
    assert O.lbip_exceeds_cutoff(4e-5, 0.0, "dtw")


# ---------------------------------------------------------------- GLB/ELD ---
def z(x):
    return (x - x.mean()) / x.std()


@pytest.mark.parametrize("metric,arg", METRICS)
def test_glb_is_a_lower_bound_of_eld(metric, arg):
    rng = np.random.default_rng(11)
    for trial in range(150):
        m = int(rng.integers(2, 60))
        w = int(rng.integers(0, m))
        q = z(np.cumsum(rng.normal(size=m)) + 1e-3 * np.arange(m))
        t = z(np.cumsum(rng.normal(size=m)) + 1e-3 * np.arange(m))
        LQ, UQ = O.lemire(q, w)
        g = O.glb(metric, q, t, arg, LQ, UQ, w, 1e300)
        d = O.eld(metric, q, t, w, arg)
        assert g <= d + 1e-12, (trial, g, d)
        # identical series: nothing outside the envelopes; SWALE's boundary
        # still charges its match cost r = 1 for the last pair
        assert O.glb(metric, q, q, arg, LQ, UQ, w, 1e300) == (1.0 if metric == "swale" else 0.0)
    # early abandon returns the sentinel of include/def.h:7
    m, w = 40, 4
    q = z(np.cumsum(rng.normal(size=m)))
    t = -q
    LQ, UQ = O.lemire(q, w)
    if metric != "lcss" or O.glb(metric, q, t, arg, LQ, UQ, w, 1e300) > 0:
        assert O.glb(metric, q, t, arg, LQ, UQ, w, 1e-6) == 1008610086.0


def full_dtw(a, b):
    n, m = len(a), len(b)
    D = np.full((n + 1, m + 1), np.inf)
    D[0, 0] = 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            D[i, j] = (a[i - 1] - b[j - 1]) ** 2 + min(D[i - 1, j], D[i, j - 1], D[i - 1, j - 1])
    return math.sqrt(D[n, m])


def test_eld_special_cases():
    rng = np.random.default_rng(12)
    for _ in range(30):
        m = int(rng.integers(2, 40))
        q, t = rng.normal(size=m), rng.normal(size=m)
        for metric, arg in METRICS:
            # identical series: 0, except SWALE, whose matches cost r = 1 each
            want = float(m) if metric == "swale" else 0.0
            assert O.eld(metric, q, q, int(rng.integers(0, m)), arg) == want
        # w = 0 reduces every measure to its lock-step form
        assert np.isclose(O.eld("dtw", q, t, 0, 0.0), np.sqrt(np.sum((q - t) ** 2)), rtol=1e-12)
        assert np.isclose(O.eld("erp", q, t, 0, 0.3), np.sqrt(np.sum((q - t) ** 2)), rtol=1e-12)
        assert np.isclose(O.eld("msm", q, t, 0, 0.5), np.sum(np.abs(q - t)), rtol=1e-12)
        assert O.eld("edr", q, t, 0, 0.4) == np.sum(np.abs(q - t) > 0.4)
        assert O.eld("lcss", q, t, 0, 0.4) == 1 - np.sum(np.abs(q - t) <= 0.4) / m
        # a tolerance covering every pair makes EDR / LCSS zero
        big = float(np.max(np.abs(q[:, None] - t[None, :])))
        assert O.eld("edr", q, t, int(rng.integers(0, m)), big) == 0.0
        assert O.eld("lcss", q, t, int(rng.integers(0, m)), big) == 0.0
        # w >= m-1 is unconstrained DTW
        assert np.isclose(O.eld("dtw", q, t, m - 1, 0.0), full_dtw(q, t), rtol=1e-12)


# ---------------------------------------------------------------- drivers ---
def test_geometry_table():
    # macro-segmentation of the benchmark configurations
    g = F.geometry(1 << 20, 128)
    assert (g["m_k"], g["k_m_l"], g["ts_n"], g["S"], g["tail"]) == (1024, 897, 1048449, 1169, 753)
    g = F.geometry(1 << 26, 512)
    assert (g["m_k"], g["k_m_l"], g["ts_n"], g["S"], g["tail"]) == (4096, 3585, 67108353, 18720, 738)
    g = F.geometry(700, 128)
    assert (g["m_k"], g["S"], g["tail"]) == (700, 1, 0)


def test_segmentation_covers_every_window_once_and_tail_uses_its_own_segment():
    t = F.synth_target(9, 5000, 0.0, 0.0, 0, 0, 0)
    m = 40
    G = F.geometry(t.size, m)
    assert G["tail"] > 0
    mu, sg, inv = O.window_stats(t, m)
    assert mu.size == G["ts_n"] and np.isfinite(mu).all()
    # full segments: windows of segment s come from the recurrence started at s*k_m_l
    kml = G["k_m_l"]
    seg_mu, _ = O.mvmean(t[kml: kml + G["m_k"]], m)
    assert np.array_equal(mu[kml: 2 * kml], seg_mu)
    # tail windows come from the EOF-aligned segment's own recurrence
    tail_mu, _ = O.mvmean(t[-G["m_k"]:], m)
    assert np.array_equal(mu[G["forsize"] * kml:], tail_mu[G["tail_first"]:])


@pytest.mark.parametrize("metric,arg", METRICS)
def test_search_equals_exact_and_cutoff_is_kth(metric, arg):
    t = F.synth_target(13, 6000, 0.001, 5.0, 3, 150, 2)
    m, w = 32, 3
    q, _ = F.synth_query(13, t, m, 0.1)
    thr, ranks, rd, st = O.cutoff(t, q, m, 7, metric, w, arg)
    s_ids, s_st = O.search(t, q, m, thr, metric, w, arg)
    e_ids, e_d, e_st = O.exact(t, q, m, thr, metric, w, arg)
    assert np.array_equal(s_ids, e_ids)
    assert s_st["valid"] == e_st["valid"] and s_st["invalid"] + s_st["valid"] == s_st["ts_n"]
    a_ids, a_d, _ = O.exact(t, q, m, 1e300, metric, w, arg)
    order = np.lexsort((a_ids, a_d))[:7]
    assert np.array_equal(a_ids[order], ranks) and a_d[order][-1] == thr
    assert np.array_equal(e_ids, np.sort(a_ids[a_d <= thr]))


def test_cutoff_ties_break_by_index_and_exclusion():
    # A series with period k_m_l makes every full macro-segment identical, so
    # windows at the same in-segment offset get bit-identical mean, std and
    # exact distance: the ranking must order those ties by index.
    m, w = 24, 2
    G = F.geometry(6 * 233, m)
    kml = G["k_m_l"]
    assert kml == 233
    base = F.synth_target(14, kml, 0.0, 0.0, 0, 0, 0)
    t = np.tile(base, 6)[: 5 * kml + m - 1]   # ts_n = 5 k_m_l: no EOF-aligned tail segment
    assert F.geometry(t.size, m)["tail"] == 0
    q = base[100:124] + np.linspace(0.0, 0.05, 24)
    thr, ranks, rd, _ = O.cutoff(t, q, m, 4, "dtw", w, 0.0)
    assert np.all(rd == rd[0])
    assert np.array_equal(ranks, [100, 100 + kml, 100 + 2 * kml, 100 + 3 * kml])
    _, ranks2, _, _ = O.cutoff(t, q, m, 3, "dtw", w, 0.0, exclude=(90, 340))
    assert np.array_equal(ranks2, [100 + 2 * kml, 100 + 3 * kml, 100 + 4 * kml])
    ids, _ = O.search(t, q, m, thr, "dtw", w, 0.0, exclude=(90, 340))
    assert ids.size and not np.any((ids >= 90) & (ids < 340))


def test_fewer_than_k_raises():
    t = F.synth_target(15, 300, 0.0, 0.0, 0, 0, 0)
    t[150:] = np.nan
    with pytest.raises(LookupError):
        O.cutoff(t, t[:40] + 0.1, 40, 200, "dtw", 2, 0.0)
