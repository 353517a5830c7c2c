"""GPU parity: every stage and every command of the HIP engine against the
CPU oracle on the same seeded inputs.

Bar: bit-exact for window statistics, invalid flags, GLB values, exact
distances, match sets, cutoff thresholds and rankings.  LB_IP values are
compared where both sides chose the same thresholds (within 1e-9 relative,
the FFT guard of include/fsmeld_internal.h:9-19) and are checked to be
admissible (never above an exact distance)."""
import numpy as np
import pytest

import fsmeld_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

METRICS = [("dtw", 0.0), ("erp", 0.0), ("edr", 0.25), ("lcss", 0.25), ("msm", 0.5),
           ("swale", 0.25)]


def bits(x):
    return np.asarray(x, dtype=np.float64).view(np.uint64)


def same_bits(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.all((bits(a) == bits(b)) | both_nan)


def make_target(seed, n, nans=3, runs=4, spikes=0.001):
    return F.synth_target(seed, n, spikes, 5.0, runs, 200, nans)


@pytest.fixture(scope="module")
def small():
    t = make_target(3, 1 << 15)
    return t, F.Target(t)


@pytest.mark.parametrize("m", [16, 64, 128])
def test_window_stats_bitexact(gpu, small, m):
    t, T = small
    mu, sg, inv = T.window_stats(m)
    omu, osg, oinv = O.window_stats(t, m)
    assert same_bits(mu, omu)
    assert same_bits(sg, osg)
    assert np.array_equal(inv, oinv)
    assert inv.sum() > 0  # the NaNs and constant runs produce invalid windows


def test_window_stats_exclusion(gpu, small):
    t, T = small
    mu, sg, inv = T.window_stats(64, exclude=(1000, 5000))
    _, _, oinv = O.window_stats(t, 64, exclude=(1000, 5000))
    assert np.array_equal(inv, oinv)
    assert inv[1000:5000].all()


@pytest.mark.parametrize("metric,arg", METRICS)
@pytest.mark.parametrize("m,w", [(2, 0), (7, 1), (33, 0), (40, 6), (64, 25), (50, 33), (24, 23),
                                 (64, 30), (2100, 8), (3000, 8), (160, 100), (130, 70),
                                 (4500, 10)])
def test_eld_pairs_bitexact(gpu, metric, arg, m, w):
    rng = np.random.default_rng(m * 100 + w)
    q = rng.normal(size=m)
    c = rng.normal(size=(97, m))
    c[:10] = q + rng.normal(size=(10, m)) * 1e-3   # near matches
    c[10:13] = q                                    # identical
    c[13:20] = np.round(c[13:20], 1)                # ties
    got = F.eld_pairs(metric, q, c, w, arg)
    want = np.array([O.eld(metric, q, row, w, arg) for row in c])
    assert same_bits(got, want)


@pytest.mark.parametrize("metric,arg", METRICS)
@pytest.mark.parametrize("m,w", [(40, 0), (33, 3), (64, 12), (64, 25)])
def test_eld_register_and_wave_kernels_agree(gpu, small, metric, arg, m, w):
    """Below 1024 candidates the one-wave-per-candidate kernel computes the
    distances, from 1024 on the lane-per-candidate register kernel: the same
    rows and windows through both give the same bits, and the register
    kernel's agree with the oracle."""
    t, T = small
    rng = np.random.default_rng(m * 7 + w)
    q = rng.normal(size=m)
    c = rng.normal(size=(1100, m))
    c[:8] = q + rng.normal(size=(8, m)) * 1e-3
    many = F.eld_pairs(metric, q, c, w, arg)            # register kernel
    few = F.eld_pairs(metric, q, c[:600], w, arg)       # wave kernel
    assert same_bits(many[:600], few)
    want = np.array([O.eld(metric, q, row, w, arg) for row in c[:40]])
    assert same_bits(many[:40], want)
    qw, _ = F.synth_query(6, t, m, 0.1)
    _, _, inv = O.window_stats(t, m)
    win = np.flatnonzero(~inv)[::7][:1100]
    assert win.size >= 1024
    dm = T.eld(qw, m, metric, w, arg, win)
    df = T.eld(qw, m, metric, w, arg, win[:600])
    assert same_bits(dm[:600], df)


@pytest.mark.parametrize("metric,arg", METRICS)
def test_eld_and_glb_windows_bitexact(gpu, small, metric, arg):
    t, T = small
    m, w = 64, 6
    q, _ = F.synth_query(5, t, m, 0.1)
    mu, sg, inv = O.window_stats(t, m)
    win = np.flatnonzero(~inv)[::97][:300]
    d = T.eld(q, m, metric, w, arg, win)
    qn = O.normalize_query(q, m)
    LQ, UQ = O.lemire(qn, w)
    want_d = np.array([O.eld(metric, qn, (t[g:g + m] - mu[g]) / sg[g], w, arg) for g in win])
    assert same_bits(d, want_d)
    for bsf in (np.median(want_d), np.quantile(want_d, 0.1), 1e30):
        g = T.glb(q, m, metric, w, arg, win, bsf)
        want_g = np.array([O.glb(metric, qn, (t[x:x + m] - mu[x]) / sg[x], arg, LQ, UQ, w, bsf)
                           for x in win])
        assert same_bits(g, want_g)
        assert np.all(want_g[want_g < 1e9] <= want_d[want_g < 1e9] + 1e-12)


@pytest.mark.parametrize("metric,arg", METRICS)
@pytest.mark.parametrize("m,w", [(64, 6), (50, 30), (40, 0), (3, 1)])
def test_glb_filter_decisions(gpu, small, metric, arg, m, w):
    """The search filter's keep decisions equal `glb <= bsf` of the oracle,
    with bsf set exactly on, and one ulp below, some windows' GLB values so
    that the decision interval of the fast pass straddles the threshold."""
    glb_decisions(small, metric, arg, m, w)


@pytest.mark.parametrize("metric,arg", [("msm", 0.0), ("msm", 1e-3), ("msm", 3.0), ("msm", 1e-120),
                                        ("erp", 0.7), ("erp", -1.3), ("edr", 1e-3),
                                        ("lcss", 2.0)])
@pytest.mark.parametrize("m,w", [(64, 6), (40, 0)])
def test_glb_filter_decisions_args(gpu, small, metric, arg, m, w):
    """The same at the ends of the metric arguments: MSM with c = 0 (every
    envelope term 0), tiny and large c, and c too small for the fast pass's
    1 / c (the exact path decides); ERP with g != 0; EDR / LCSS with a tiny
    and a wide eps."""
    glb_decisions(small, metric, arg, m, w)


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("msm", 0.5), ("erp", 0.0)])
def test_glb_filter_decisions_huge_values(gpu, small, metric, arg):
    """A target scaled by 1e120 (window std far above 1e100): DTW and MSM
    windows leave the fast pass for the exact path, whose decisions still
    equal the oracle's."""
    t, _ = small
    big = np.ascontiguousarray(t * 1e120)
    glb_decisions((big, F.Target(big)), metric, arg, 64, 6)


def glb_decisions(small, metric, arg, m, w):
    t, T = small
    q, _ = F.synth_query(11, t, m, 0.1)
    mu, sg, inv = O.window_stats(t, m)
    win = np.flatnonzero(~inv)[::13][:2500]
    qn = O.normalize_query(q, m)
    LQ, UQ = O.lemire(qn, w)
    cand = [(t[g:g + m] - mu[g]) / sg[g] for g in win]
    full = np.array([O.glb(metric, qn, c, arg, LQ, UQ, w, 1e30) for c in cand])
    picks = np.sort(full)[[len(full) // 50, len(full) // 5, len(full) // 2]]
    for bsf in list(picks) + [np.nextafter(p, -np.inf) for p in picks]:
        keep = T.glb_keep(q, m, metric, w, arg, win, bsf)
        want = np.array([O.glb(metric, qn, c, arg, LQ, UQ, w, bsf) for c in cand]) <= bsf
        bad = np.flatnonzero(keep != want)
        assert bad.size == 0, f"bsf={bsf!r}: windows {win[bad][:5]} differ"


@pytest.mark.parametrize("metric,arg", METRICS)
@pytest.mark.parametrize("m,w", [(64, 0), (48, 1), (64, 4), (64, 12), (96, 27), (100, 30), (80, 36),
                                 (160, 70), (300, 130)])
def test_exact_command_abandon_is_safe(gpu, small, metric, arg, m, w):
    """The exact command abandons windows by DP frontier + suffix bound; with
    the cutoff set exactly on windows' distances it must return precisely the
    windows whose distance (all valid windows, no abandoning) is <= cutoff."""
    t, T = small
    q, _ = F.synth_query(17, t, m, 0.1)
    _, _, inv = O.window_stats(t, m)
    win = np.flatnonzero(~inv)
    d = T.eld(q, m, metric, w, arg, win)
    ds = np.sort(d)
    for cut in (ds[0], ds[5], ds[len(ds) // 1000], ds[len(ds) // 100], ds[len(ds) // 20]):
        e = T.exact(q, m, cut, metric, w, arg)
        want = win[d <= cut]
        assert np.array_equal(e.indices, want), f"cut={cut!r}"
        assert same_bits(e.distances, d[d <= cut])


@pytest.mark.parametrize("metric,arg", METRICS)
@pytest.mark.parametrize("m,w,logn", [(32, 3, 14), (512, 25, 17), (1024, 51, 18), (2048, 102, 19)])
def test_lbip_values(gpu, metric, arg, m, w, logn):
    """One case per LB_IP workgroup shape and key count: 256 threads (m = 32,
    512), 1024 threads with 8192 (m = 1024) and 16384 (m = 2048) keys in LDS."""
    lbip_values_case(metric, arg, m, w, logn)


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("msm", 0.5), ("edr", 0.25)])
@pytest.mark.parametrize("m,w,logn", [(100, 5, 16), (200, 10, 16), (500, 25, 17), (700, 35, 18),
                                      (1500, 75, 19)])
def test_lbip_values_unaligned(gpu, metric, arg, m, w, logn):
    """Lengths that are not multiples of 64: a tile's window ranges start and
    end inside 64-window chunks (the partial-chunk suffixes and the sparse
    table's uneven middles), for every segment class from m_k = 1024 to 16384."""
    lbip_values_case(metric, arg, m, w, logn)


def lbip_values_case(metric, arg, m, w, logn):
    t = make_target(8, 1 << logn, nans=0, runs=0, spikes=0.0)  # NaN-free: deques exact
    T = F.Target(t)
    q, _ = F.synth_query(8, t, m, 0.1)
    lb, tu, tl = T.lbip(q, m, metric, w, arg)
    olb, otu, otl = O.lbip_all(t, q, m, metric, w, arg)
    same_t = (tu == otu) & (tl == otl)
    assert same_t.mean() >= 0.9, f"thresholds agree in {same_t.mean():.3f} of segments"
    G = F.geometry(t.size, m)
    seg = np.minimum(np.arange(G["ts_n"]) // G["k_m_l"], G["S"] - 1)
    sel = same_t[seg]
    a, b = lb[sel], olb[sel]
    assert np.all(np.abs(a - b) <= 1e-9 * np.maximum(1.0, np.abs(b)))
    # admissibility: LB_IP never exceeds the exact distance (beyond the guard)
    _, _, exact_inv = O.window_stats(t, m)
    win = np.flatnonzero(~exact_inv)[::37]
    d = T.eld(q, m, metric, w, arg, win)
    assert np.all(lb[win] <= d + 1e-9)


@pytest.mark.parametrize("metric", ["edr", "lcss"])
@pytest.mark.parametrize("m,w,logn", [(32, 3, 14), (512, 25, 17), (1024, 51, 18), (2100, 105, 19)])
def test_lbip_flat_objective(gpu, metric, m, w, logn):
    """eps = 10 lies beyond the query's range, so h(t) = 0 at every envelope
    value: every score ties at 0, rank 0 wins, and the thresholds are each
    segment's smallest U_n and largest L_n.  No bucket of values can be set
    aside, so this runs the full sort of the segment's values (in LDS; in
    global scratch for m = 2100)."""
    t = make_target(9, 1 << logn, nans=0, runs=0, spikes=0.0)
    T = F.Target(t)
    q, _ = F.synth_query(9, t, m, 0.1)
    lb, tu, tl = T.lbip(q, m, metric, w, 10.0)
    olb, otu, otl = O.lbip_all(t, q, m, metric, w, 10.0)
    assert np.array_equal(tu, otu)
    assert np.array_equal(tl, otl)
    assert np.array_equal(lb, olb, equal_nan=True)


def run_all_modes(t, q, m, metric, w, arg, k, exclude=None):
    T = F.Target(t)
    thr, ranks, rd, _ = O.cutoff(t, q, m, k, metric, w, arg, exclude=exclude)
    g = T.cutoff(q, m, k, metric, w, arg, exclude=exclude)
    assert bits(g.threshold) == bits(thr)
    assert np.array_equal(g.indices, ranks)
    assert same_bits(g.distances, rd)
    s = T.search(q, m, thr, metric, w, arg, exclude=exclude)
    e = T.exact(q, m, thr, metric, w, arg, exclude=exclude)
    os_ids, ost = O.search(t, q, m, thr, metric, w, arg, exclude=exclude)
    oe_ids, oe_d, oest = O.exact(t, q, m, thr, metric, w, arg, exclude=exclude)
    assert np.array_equal(s.indices, os_ids)
    assert np.array_equal(e.indices, oe_ids)
    assert same_bits(e.distances, oe_d)
    # search returns exactly the exact scan's matches, except where the
    # reference's own cascade drops a window whose distance equals the cutoff
    # (an ulp-level GLB > ELD when the bound is tight, e.g. DTW with w = 0);
    # the engine reproduces that decision, so the oracle shows it too.
    missing = np.setdiff1d(e.indices, s.indices)
    assert np.array_equal(np.setdiff1d(oe_ids, os_ids), missing)
    assert np.setdiff1d(s.indices, e.indices).size == 0
    assert np.all(e.distances[np.isin(e.indices, missing)] == thr)
    assert same_bits(s.distances, e.distances[np.isin(e.indices, s.indices)])
    assert s.stats["valid"] == ost["valid"] == oest["valid"]
    assert s.stats["invalid"] == ost["invalid"]
    return s, ost


@pytest.mark.parametrize("metric,arg", METRICS)
def test_commands_match_oracle(gpu, metric, arg):
    t = make_target(21, 1 << 15)
    q, _ = F.synth_query(21, t, 64, 0.1)
    s, ost = run_all_modes(t, q, 64, metric, 4, arg, 10)
    # LB_IP / GLB survivor counts track the reference restatement closely
    assert abs(int(s.stats["after_lbip"]) - int(ost["after_lbip"])) <= 0.02 * ost["valid"] + 5


@pytest.mark.parametrize("case", ["tail0", "single_segment", "w0", "m2", "exclusion", "big_w"])
def test_command_edge_cases(gpu, case):
    if case == "tail0":      # ts_n a multiple of k_m_l: no tail segment
        m = 32
        G = F.geometry(1 << 14, m)
        n = G["k_m_l"] * 20 + m - 1
        t = make_target(31, n)
        q, _ = F.synth_query(31, t, m, 0.1)
        assert F.geometry(n, m)["tail"] == 0
        run_all_modes(t, q, m, "dtw", 3, 0.0, 5)
    elif case == "single_segment":   # N < 8m: one macro-segment of length N
        t = make_target(32, 700, nans=1, runs=1)
        q, _ = F.synth_query(32, t, 128, 0.1)
        run_all_modes(t, q, 128, "erp", 6, 0.0, 5)
    elif case == "w0":
        t = make_target(33, 1 << 13)
        q, _ = F.synth_query(33, t, 40, 0.1)
        for metric, arg in METRICS:
            run_all_modes(t, q, 40, metric, 0, arg, 3)
    elif case == "m2":
        t = make_target(34, 4000, nans=0, runs=0)
        q = np.array([0.0, 1.0])
        run_all_modes(t, q, 2, "msm", 1, 0.5, 4)
    elif case == "exclusion":
        t = make_target(35, 1 << 14)
        q, off = F.synth_query(35, t, 64, 0.0)
        run_all_modes(t, q, 64, "dtw", 3, 0.0, 8, exclude=(off - 100, off + 100))
    elif case == "big_w":   # radius above the register band variants
        t = make_target(36, 1 << 13, nans=0, runs=1)
        q, _ = F.synth_query(36, t, 96, 0.1)
        for metric, arg in METRICS:
            run_all_modes(t, q, 96, metric, 40, arg, 3)


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("edr", 0.25), ("msm", 0.5)])
@pytest.mark.parametrize("m,w", [(3000, 40), (4100, 8), (2600, 130)])
def test_long_queries_search_equals_exact(gpu, metric, arg, m, w):
    """Queries past the register sort (2^15 keys per segment: the LDS / scratch
    bitonic sort with 1024 threads) and past the register ELD band (wave and
    row-band kernels): search returns exactly the exact scan's windows with the
    same distance bits, cutoff's k-th distance bounds them, and no LB_IP bound
    exceeds its window's exact distance."""
    t = make_target(50 + m % 7, 1 << 16, nans=2, runs=3)
    q, _ = F.synth_query(51, t, m, 0.1)
    T = F.Target(t)
    c = T.cutoff(q, m, 5, metric, w, arg)
    s = T.search(q, m, c.threshold, metric, w, arg)
    e = T.exact(q, m, c.threshold, metric, w, arg)
    assert s.indices.size >= 5
    assert np.all(np.isin(c.indices, e.indices))
    missing = np.setdiff1d(e.indices, s.indices)
    assert np.setdiff1d(s.indices, e.indices).size == 0
    assert np.all(e.distances[np.isin(e.indices, missing)] == c.threshold)
    assert same_bits(s.distances, e.distances[np.isin(e.indices, s.indices)])
    lb, _, _ = T.lbip(q, m, metric, w, arg)
    _, _, inv = O.window_stats(t, m)
    win = np.flatnonzero(~inv)[::211]
    d = T.eld(q, m, metric, w, arg, win)
    assert np.all(~(lb[win] > d + 1e-9 * np.maximum(1.0, d)))


@pytest.mark.parametrize("metric,arg", METRICS)
def test_cutoff_large_target(gpu, metric, arg):
    """cutoff on 2^18 windows: enough GLB survivors at the first bound that
    the engine tightens it before verifying; the ranking must still equal the
    oracle's heap-based restatement bit for bit."""
    t = make_target(29, 1 << 18, nans=5, runs=8)
    q, _ = F.synth_query(29, t, 64, 0.3)
    thr, ranks, rd, _ = O.cutoff(t, q, 64, 5, metric, 4, arg)
    g = F.Target(t).cutoff(q, 64, 5, metric, 4, arg)
    assert bits(g.threshold) == bits(thr)
    assert np.array_equal(g.indices, ranks)
    assert same_bits(g.distances, rd)


def test_fewer_than_k(gpu):
    t = make_target(40, 600, nans=0, runs=0)
    t[:] = np.nan
    t[100:400] = np.cumsum(np.random.default_rng(0).normal(size=300))
    q, _ = F.synth_query(40, t[100:400].copy(), 64, 0.1)
    T = F.Target(t)
    with pytest.raises(F.FsmeldError) as e:
        T.cutoff(q, 64, 300, "dtw", 3, 0.0)   # 237 valid windows, k <= N - m + 1
    assert e.value.code == -4
    with pytest.raises(LookupError):
        O.cutoff(t, q, 64, 300, "dtw", 3, 0.0)


def test_plain_small_config(gpu):
    """The plain small configuration at full size: 2^20 random walk, 10
    constant runs, 10 NaNs, m=128, DTW w=6, cutoff from k=10."""
    from fsmeld_amd import workloads
    c = workloads.plain_small()
    t, q = c["target"], c["query"]
    T = F.Target(t)
    thr, ranks, rd, _ = O.cutoff(t, q, 128, 10, "dtw", 6, 0.0)
    g = T.cutoff(q, 128, 10, "dtw", 6, 0.0)
    assert bits(g.threshold) == bits(thr) and np.array_equal(g.indices, ranks)
    s = T.search(q, 128, thr, "dtw", 6, 0.0)
    e = T.exact(q, 128, thr, "dtw", 6, 0.0)
    os_ids, _ = O.search(t, q, 128, thr, "dtw", 6, 0.0, threads=8)
    assert np.array_equal(s.indices, os_ids)
    assert np.array_equal(e.indices, os_ids)
    assert c["offset"] in set(s.indices.tolist())


def test_eld_matches_reference_golden_vectors(gpu):
    """GPU verifiers against values produced by the reference's own
    src/ElasticMeasures.cpp (tests/golden/eld_ref_vectors.json)."""
    import json
    import os
    cases = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                        "eld_ref_vectors.json")))["cases"]
    for c in cases:
        q = np.array(c["q"])
        got = F.eld_pairs(c["metric"], q, np.array(c["t"])[None, :], c["w"], c["arg"])[0]
        assert float(got).hex() == c["d"], c


@pytest.mark.parametrize("metric", ["dtw", "erp"])
def test_cutoff_k1_w0_first_bound_is_widened(gpu, metric):
    """k = 1, w = 0: the best window's distance is the first bound itself and
    GLB's reordered sum can exceed it by an ulp; the bound is widened so that
    window survives (on the oracle, seeds 9 and 24 of DTW lose it otherwise)."""
    for seed in (0, 3, 9, 17, 24, 31):
        t = make_target(seed, 1 << 13)
        q, _ = F.synth_query(seed, t, 40, 0.1)
        thr, ranks, rd, _ = O.cutoff(t, q, 40, 1, metric, 0, 0.0)
        g = F.Target(t).cutoff(q, 40, 1, metric, 0, 0.0)
        assert bits(g.threshold) == bits(thr), seed
        assert np.array_equal(g.indices, ranks), seed


@pytest.mark.parametrize("metric,arg", METRICS)
def test_exact_topk_matches_oracle(gpu, metric, arg):
    """FSMELD_exact with top_k = k by brute force on the GPU (every valid
    window's distance, a device radix select of the k-th): threshold bits,
    the windows at or below it (ties included) and their distance bits equal
    the oracle's heap restatement; also with an exclusion and a wide band."""
    t = make_target(81, 1 << 14)
    T = F.Target(t)
    for m, w, k, excl in [(48, 3, 9, None), (64, 40, 4, (2000, 6000)), (32, 0, 25, None)]:
        q, _ = F.synth_query(81 + m, t, m, 0.2)
        thr, ids, d, _ = O.exact_topk(t, q, m, k, metric, w, arg, exclude=excl, threads=8)
        g = T.exact_topk(q, m, k, metric, w, arg, exclude=excl)
        assert bits(g.threshold) == bits(thr), (m, w)
        assert np.array_equal(g.indices, ids), (m, w)
        assert same_bits(g.distances, d), (m, w)
        assert g.indices.size >= k
    with pytest.raises(F.FsmeldError) as e:
        F.Target(t[:200]).exact_topk(q[:32], 32, 170, metric, 3, arg)   # 169 windows
    assert e.value.code in (-1, -4)


@pytest.mark.parametrize("metric,arg", METRICS)
def test_raw_mode_matches_oracle(gpu, metric, arg):
    """Raw (non-normalised) matching, paper App. D: cutoff threshold bits and
    ranking, search and exact windows, distance bits, exact top-k and the
    LB_IP bounds equal the oracle's raw mode; a constant query, which
    z-normalised matching rejects, is accepted."""
    t = make_target(91, 1 << 14, nans=3, runs=4)
    T = F.Target(t)
    m, w, k = 48, 4, 6
    q, _ = F.synth_query(91, t, m, 0.05)   # a finite window of the target, plus noise
    with O.raw_mode():
        thr, ranks, rd, _ = O.cutoff(t, q, m, k, metric, w, arg)
        os_ids, ost = O.search(t, q, m, thr, metric, w, arg)
        oe_ids, oe_d, _ = O.exact(t, q, m, thr, metric, w, arg)
        kth, tk_ids, tk_d, _ = O.exact_topk(t, q, m, k, metric, w, arg)
        olb, _, _ = O.lbip_all(t, q, m, metric, w, arg)
    g = T.cutoff(q, m, k, metric, w, arg, raw=True)
    assert bits(g.threshold) == bits(thr)
    assert np.array_equal(g.indices, ranks) and same_bits(g.distances, rd)
    s = T.search(q, m, thr, metric, w, arg, raw=True)
    e = T.exact(q, m, thr, metric, w, arg, raw=True)
    assert np.array_equal(s.indices, os_ids)
    assert np.array_equal(e.indices, oe_ids) and same_bits(e.distances, oe_d)
    assert s.stats["valid"] == ost["valid"] and s.stats["invalid"] == ost["invalid"]
    tk = T.exact_topk(q, m, k, metric, w, arg, raw=True)
    assert bits(tk.threshold) == bits(kth) and np.array_equal(tk.indices, tk_ids)
    lb, _, _ = T.lbip(q, m, metric, w, arg, raw=True)
    ok = np.isfinite(olb)
    assert np.array_equal(np.isfinite(lb), ok)
    # equal wherever both chose the same thresholds (near-ties in the
    # threshold search can resolve differently, as in z-normalised matching)
    close = np.abs(lb[ok] - olb[ok]) <= 1e-9 * np.maximum(1.0, np.abs(olb[ok]))
    assert close.mean() >= 0.9
    flat = np.full(m, 0.5)
    with O.raw_mode():
        fthr, franks, _, _ = O.cutoff(t, flat, m, 3, metric, w, arg)
    fg = T.cutoff(flat, m, 3, metric, w, arg, raw=True)
    assert bits(fg.threshold) == bits(fthr) and np.array_equal(fg.indices, franks)


@pytest.mark.parametrize("metric,arg", METRICS)
def test_near_cutoff_windows_are_reverified_on_the_host(gpu, metric, arg):
    """With the cutoff exactly on a window's distance and one ulp either side,
    the windows within 1e-9 of it are recomputed on the host (their count is
    reported, and the host values equal the device's); the results equal the
    oracle's at each cutoff."""
    t = make_target(93, 1 << 14)
    T = F.Target(t)
    m, w = 48, 4
    q, _ = F.synth_query(93, t, m, 0.2)
    c = T.cutoff(q, m, 6, metric, w, arg)
    assert c.stats["near_rechecked"] >= 1 and c.stats["near_changed"] == 0
    d5 = c.distances[4]
    for cut in (d5, np.nextafter(d5, -np.inf), np.nextafter(d5, np.inf)):
        s = T.search(q, m, cut, metric, w, arg)
        e = T.exact(q, m, cut, metric, w, arg)
        assert s.stats["near_rechecked"] >= 1 and s.stats["near_changed"] == 0
        assert e.stats["near_rechecked"] >= 1 and e.stats["near_changed"] == 0
        os_ids, _ = O.search(t, q, m, cut, metric, w, arg)
        oe_ids, oe_d, _ = O.exact(t, q, m, cut, metric, w, arg)
        assert np.array_equal(s.indices, os_ids), cut
        assert np.array_equal(e.indices, oe_ids) and same_bits(e.distances, oe_d), cut
