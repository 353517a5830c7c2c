"""Pins the oracle's GLB, trans, s_to_g, query normalisation and parallel
cutoff with checks a plausible mistake fails:

* GLB (src/glb.cpp:200-373): hand-computed known answers on 4-sample
  inputs for every metric, each written out below, including both MSM
  endpoint branches, ERP with a nonzero gap, LCSS's two-direction min and
  cases where each direction dominates; the w = 0 DTW reduction to the
  Euclidean distance; and a brute-force restatement from the definition
  (envelopes as plain window min / max, boundary + max of the two
  directions) on random inputs, together with the proof that the tests
  separate it from its plausible mutants (min for max, no boundary, the
  other MSM branch, one direction only).
* trans (src/fsmeld.cpp:84-106), s_to_g (src/lb.cpp:18-34) and the query
  z-normalisation (src/fsmeld.cpp:622-639): closed forms and known answers.
* orc_cutoff with threads > 1 equals the sequential restatement."""
import math

import numpy as np
import pytest

import oracle as O

BIG = 1e300


def env(a, w):
    """Clipped sliding min / max of radius w by brute force."""
    n = len(a)
    lo = [min(a[max(0, i - w):min(n, i + w + 1)]) for i in range(n)]
    hi = [max(a[max(0, i - w):min(n, i + w + 1)]) for i in range(n)]
    return lo, hi


def glb(metric, Q, T, arg, w, variant=None):
    """GLB by definition (no early abandon): boundary term + max of the two
    directional sums, each position outside the other side's envelope paying
    the metric's term.  `variant` builds a plausible mutant."""
    Q, T = [float(x) for x in Q], [float(x) for x in T]
    n = len(Q)
    LQ, UQ = env(Q, w)
    LT, UT = env(T, w)

    def sq(a, b):
        return (a - b) * (a - b)

    def direction(X, L, U):
        s = 0.0
        if metric == "lcss":   # count matches within eps of the envelope, all positions
            for i in range(n):
                d = L[i] - X[i] if X[i] < L[i] else (X[i] - U[i] if X[i] > U[i] else 0.0)
                s += 1.0 if d <= arg else 0.0
            return s
        for i in range(1, n - 1):
            if X[i] > U[i]:
                e = U[i]
            elif X[i] < L[i]:
                e = L[i]
            else:
                continue
            if metric == "dtw":
                s += sq(X[i], e)
            elif metric == "erp":
                s += min(sq(X[i], e), sq(X[i], arg))
            elif metric == "edr":
                s += 0.0 if abs(X[i] - e) <= arg else 1.0
            elif metric == "msm":
                s += min(abs(X[i] - e), arg)
        return s

    t2q = direction(T, LQ, UQ)
    q2t = direction(Q, LT, UT)
    if metric == "lcss":
        pick = max if variant == "min_max" else min
        reward = t2q if variant == "one_direction" else pick(t2q, q2t)
        return max(0.0, 1.0 - reward / n)
    if metric == "dtw":
        b = sq(T[0], Q[0]) + sq(T[-1], Q[-1])
    elif metric == "erp":
        b = min(sq(Q[-1], T[-1]), sq(Q[-1], arg), sq(T[-1], arg))
    elif metric == "edr":
        b = min(0.0 if abs(Q[-1] - T[-1]) <= arg else 1.0, 1.0)
    else:  # msm
        simple = ((T[-2] >= T[-1] >= Q[-1]) or (T[-2] <= T[-1] <= Q[-1]) or
                  (Q[-2] <= Q[-1] <= T[-1]) or (Q[-2] >= Q[-1] >= T[-1]))
        if variant == "msm_branch":
            simple = not simple
        end = abs(Q[-1] - T[-1])
        if simple:
            b = abs(Q[0] - T[0]) + min(end, arg)
        else:
            b = abs(Q[0] - T[0]) + min(end, arg + abs(T[-1] - T[-2]), arg + abs(Q[-1] - Q[-2]))
    if variant == "no_boundary":
        b = 0.0
    pick = min if variant == "min_max" else max
    tot = b + (t2q if variant == "one_direction" else pick(t2q, q2t))
    return math.sqrt(max(0.0, tot)) if metric in ("dtw", "erp") else tot


def oracle_glb(metric, Q, T, arg, w, bsf=BIG):
    Q, T = np.asarray(Q, float), np.asarray(T, float)
    LQ, UQ = O.lemire(Q, w)
    return O.glb(metric, Q, T, arg, LQ, UQ, w, bsf)


# Case A: Q = [0, 2, -1, 1], T = [1, -1, 3, 0], w = 1.
#   L_Q = [0, -1, -1, -1], U_Q = [2, 2, 2, 1]; L_T = [-1, -1, -1, 0], U_T = [1, 3, 3, 3].
#   Inner positions 1, 2 only (the endpoints are the boundary term), except LCSS.
#   T -> Q: T1 = -1 inside [-1, 2]; T2 = 3 above U_Q2 = 2 by 1.   Q -> T: Q1, Q2 inside.
# Case B: Q = [0, 3, 0, 1], T = [1, 0, 1, 0], w = 1 (the Q -> T direction dominates).
#   L_Q = [0, 0, 0, 0], U_Q = [3, 3, 3, 1]; L_T = [0, 0, 0, 0], U_T = [1, 1, 1, 1].
#   T -> Q: T1 = 0, T2 = 1 inside.   Q -> T: Q1 = 3 above U_T1 = 1 by 2; Q2 = 0 inside.
A = ([0, 2, -1, 1], [1, -1, 3, 0])
B = ([0, 3, 0, 1], [1, 0, 1, 0])
KNOWN = [
    # DTW A: boundary (1-0)^2 + (0-1)^2 = 2; T->Q (3-2)^2 = 1; Q->T 0 -> sqrt(2 + 1)
    ("dtw", A, 0.0, math.sqrt(3.0)),
    # DTW B: boundary 1 + 1 = 2; T->Q 0; Q->T (3-1)^2 = 4 -> sqrt(6)
    ("dtw", B, 0.0, math.sqrt(6.0)),
    # ERP A, g = 0.5: boundary min((1-0)^2, (1-.5)^2, (0-.5)^2) = .25;
    #   T->Q min((3-2)^2, (3-.5)^2) = 1 -> sqrt(1.25)
    ("erp", A, 0.5, math.sqrt(1.25)),
    # ERP A, g = 2.5: boundary min(1, 2.25, 6.25) = 1; T->Q min(1, (3-2.5)^2) = .25
    ("erp", A, 2.5, math.sqrt(1.25)),
    # ERP B, g = 0: boundary min(1, 1, 0) = 0; Q->T min((3-1)^2, (3-0)^2) = 4 -> 2
    ("erp", B, 0.0, 2.0),
    # EDR A, eps = .5: boundary |1-0| > .5 -> 1; T->Q |3-2| > .5 -> 1; Q->T 0 -> 1 + 1
    ("edr", A, 0.5, 2.0),
    # EDR A, eps = 1: every difference is within eps -> 0
    ("edr", A, 1.0, 0.0),
    # EDR B, eps = .5: boundary 1; Q->T |3-1| > .5 -> 1 -> 2
    ("edr", B, 0.5, 2.0),
    # LCSS A, eps = .5 (all 4 positions): T matches Q's envelope at 0, 1, 3
    #   (T2 = 3 is 1 above U_Q2): 3; Q matches T's envelope at all 4: 4;
    #   1 - min(3, 4) / 4
    ("lcss", A, 0.5, 0.25),
    # LCSS with the roles swapped: T->Q 4, Q->T 3 -> the min picks the other side
    ("lcss", (A[1], A[0]), 0.5, 0.25),
    # MSM A, c = .5: no endpoint chain holds (T2=3 >= T3=0 >= Q3=1 fails, ...), so
    #   boundary |0-1| + min(|1-0|, .5 + |0-3|, .5 + |1-(-1)|) = 2; T->Q min(1, .5) = .5
    ("msm", A, 0.5, 2.5),
    # MSM, T = [1, -1, 3, 2]: T2 = 3 >= T3 = 2 >= Q3 = 1 holds, so
    #   boundary |0-1| + min(|1-2|, .5) = 1.5; T->Q min(|3-2|, .5) = .5 -> 2
    ("msm", ([0, 2, -1, 1], [1, -1, 3, 2]), 0.5, 2.0),
    # MSM B, c = .5: boundary |0-1| + min(1, .5 + 1, .5 + 1) = 2; Q->T min(2, .5) -> 2.5
    ("msm", B, 0.5, 2.5),
    # MSM B, c = 5: boundary 1 + min(1, 6, 6) = 2; Q->T min(2, 5) = 2 -> 4
    ("msm", B, 5.0, 4.0),
]


@pytest.mark.parametrize("metric,case,arg,want", KNOWN)
def test_glb_known_answers(metric, case, arg, want):
    Q, T = case
    assert oracle_glb(metric, Q, T, arg, 1) == want
    assert glb(metric, Q, T, arg, 1) == want


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("erp", 0.3), ("edr", 0.2),
                                        ("lcss", 0.2), ("msm", 0.5)])
def test_glb_abandons_with_the_sentinel(metric, arg):
    # LCSS abandons only on its first (T -> Q) direction, which case A makes lossy
    Q, T = A if metric == "lcss" else B
    full = oracle_glb(metric, Q, T, arg, 1)
    assert full > 0
    # below the value (squared scale for DTW / ERP): abandoned, include/def.h:7
    assert oracle_glb(metric, Q, T, arg, 1, bsf=full * 0.5) == 1008610086.0
    # at the value itself: kept, except that DTW / ERP compare against bsf^2,
    # which can round below the squared sum (here sqrt(6)^2 < 6)
    assert oracle_glb(metric, Q, T, arg, 1, bsf=full * (1 + 1e-12)) == full
    if metric not in ("dtw", "erp"):
        assert oracle_glb(metric, Q, T, arg, 1, bsf=full) == full


def test_glb_dtw_w0_is_the_euclidean_distance():
    """With w = 0 both envelopes are the series themselves, so both
    directions sum (q_i - t_i)^2 over the inner positions and the boundary
    adds the two ends: GLB = sqrt(sum (q - t)^2)."""
    rng = np.random.default_rng(5)
    for m in (2, 3, 7, 64, 301):
        q, t = rng.normal(size=m), rng.normal(size=m)
        g = oracle_glb("dtw", q, t, 0.0, 0)
        assert math.isclose(g, math.sqrt(np.sum((q - t) ** 2)), rel_tol=1e-13)


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("erp", 0.0), ("erp", 0.7), ("edr", 0.3),
                                        ("lcss", 0.3), ("msm", 0.5), ("msm", 3.0)])
def test_glb_brute_force_definition(metric, arg):
    rng = np.random.default_rng(hash((metric, arg)) % 2**32)
    # (LCSS has no boundary term, only MSM has endpoint branches)
    separated = {"min_max": False, "no_boundary": metric == "lcss", "one_direction": False,
                 "msm_branch": metric != "msm"}
    for trial in range(300):
        m = int(rng.integers(3, 12))
        w = int(rng.integers(0, m))
        q = np.round(rng.normal(size=m), 2)
        t = np.round(rng.normal(size=m), 2) if trial % 2 else np.round(q + rng.normal(size=m) * 0.3, 2)
        got = oracle_glb(metric, q, t, arg, w)
        want = glb(metric, q, t, arg, w)
        assert math.isclose(got, want, rel_tol=1e-14, abs_tol=1e-14), (metric, q, t, w)
        for v in separated:
            if not separated[v] and not math.isclose(got, glb(metric, q, t, arg, w, v),
                                                     rel_tol=1e-12, abs_tol=1e-12):
                separated[v] = True
    # every mutant gave a different value somewhere: the comparison above
    # would catch that mistake in the oracle
    assert all(separated.values()), separated


# ------------------------------------------------------------------ trans ---
def test_trans_known_answers():
    # DTW / ERP: sqrt of the clamped sum; negative sums (FFT roundoff) clamp to 0
    assert O.trans("dtw", 3.0, 1.0, 8) == 2.0
    assert O.trans("erp", 5.0, 4.0, 8) == 3.0
    assert O.trans("dtw", -5.0, 1.0, 8) == 0.0
    # MSM / EDR: the plain sum, no root and no clamp
    assert O.trans("msm", 3.0, 1.0, 8) == 4.0
    assert O.trans("edr", 3.0, 1.0, 8) == 4.0
    assert O.trans("msm", -3.0, 1.0, 8) == -2.0
    # LCSS: the sum over m, clamped at 0
    assert O.trans("lcss", 3.0, 1.0, 8) == 0.5
    assert O.trans("lcss", -9.0, 1.0, 8) == 0.0


def test_s_to_g_is_inclusive_at_the_thresholds():
    Ln = np.array([-1.0, 0.5, 0.5000000000000001, np.nan, 2.0])
    Un = np.array([-1.0, 0.5, 0.4999999999999999, np.nan, 2.0])
    gl, gu = O.s_to_g(Ln, Un, 0.5, 0.5)
    assert gl.tolist() == [0.0, 1.0, 1.0, 0.0, 1.0]   # L_n >= t_l
    assert gu.tolist() == [1.0, 1.0, 1.0, 0.0, 0.0]   # U_n <= t_u


def test_normalize_query_closed_form():
    # q = 1..4: mean 2.5, E[q^2] = 7.5, population sd sqrt(7.5 - 6.25) = sqrt(1.25)
    qn = O.normalize_query(np.array([1.0, 2.0, 3.0, 4.0, 99.0]), 4)   # reads the first m only
    want = (np.array([1.0, 2.0, 3.0, 4.0]) - 2.5) / math.sqrt(1.25)
    assert np.array_equal(qn, want)
    rng = np.random.default_rng(9)
    q = rng.normal(size=257) * 3 + 11
    qn = O.normalize_query(q, 257)
    assert abs(qn.mean()) < 1e-13 and abs(qn.std() - 1.0) < 1e-12   # ddof = 0, not 1
    assert abs(qn.std(ddof=1) - 1.0) > 1e-3


# ---------------------------------------------------------------- cutoff ---
@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("lcss", 0.25), ("msm", 0.5)])
def test_parallel_cutoff_equals_sequential(metric, arg):
    import fsmeld_amd as F
    t = F.synth_target(13, 1 << 14, 0.001, 5.0, 3, 200, 2)
    q, _ = F.synth_query(13, t, 48, 0.2)
    thr, ranks, rd, _ = O.cutoff(t, q, 48, 9, metric, 3, arg)
    thr2, ranks2, rd2, _ = O.cutoff(t, q, 48, 9, metric, 3, arg, threads=4)
    assert thr == thr2 and np.array_equal(ranks, ranks2) and np.array_equal(rd, rd2)


# ------------------------------------------------------------ exact top-k ---
@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("edr", 0.25), ("lcss", 0.25),
                                        ("msm", 0.5)])
def test_exact_topk_is_the_kth_distance_with_ties(metric, arg):
    """FSMELD_exact with top_k = k (src/fsmeld.cpp:1097-1107, :1273-1290): the
    threshold is the k-th smallest of all exact distances (a plain sort of
    every distance the brute-force scan finds) and the result is every window
    at or below it -- more than k when distances tie (EDR / LCSS counts)."""
    import fsmeld_amd as F
    t = F.synth_target(23, 3000, 0.001, 5.0, 2, 150, 1)
    q, _ = F.synth_query(23, t, 32, 0.3)
    k = 7
    thr, ids, d, _ = O.exact_topk(t, q, 32, k, metric, 3, arg)
    all_ids, all_d, _ = O.exact(t, q, 32, 1e300, metric, 3, arg)
    assert thr == np.sort(all_d)[k - 1]
    assert np.array_equal(ids, all_ids[all_d <= thr]) and np.array_equal(d, all_d[all_d <= thr])
    assert ids.size >= k
    cthr, ranks, _, _ = O.cutoff(t, q, 32, k, metric, 3, arg)
    assert cthr == thr and np.all(np.isin(ranks, ids))
    if metric in ("edr", "lcss"):
        assert ids.size > k   # integer-valued distances tie at the threshold
    with pytest.raises(LookupError):
        O.exact_topk(t[:80], q, 32, 60, metric, 3, arg)


# --------------------------------------------------------------- raw mode ---
@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("erp", 0.5), ("edr", 0.3),
                                        ("lcss", 0.3), ("msm", 0.5)])
def test_raw_mode_is_the_cascade_on_unnormalised_values(metric, arg):
    """Raw matching (NormalizationMode::Raw, paper App. D): every distance is
    the verifier's on the raw query and raw window, a window is invalid only
    when it holds a non-finite sample (a constant run stays valid), search
    returns exactly the exact scan's windows, cutoff is the k-th of them, and
    the LB_IP bounds stay below the distances."""
    import fsmeld_amd as F
    t = F.synth_target(29, 4000, 0.001, 5.0, 2, 120, 2)
    q = t[700:740] + np.random.default_rng(1).normal(size=40) * 0.05
    m, w = 40, 3
    with O.raw_mode():
        mu, sg, inv = O.window_stats(t, m)
        ids, d, st = O.exact(t, q, m, 1e300, metric, w, arg, threads=4)
        thr, ranks, _, _ = O.cutoff(t, q, m, 6, metric, w, arg)
        s_ids, _ = O.search(t, q, m, thr, metric, w, arg)
        e_ids, _, _ = O.exact(t, q, m, thr, metric, w, arg)
        lb, _, _ = O.lbip_all(t, q, m, metric, w, arg)
    win = np.lib.stride_tricks.sliding_window_view(t, m)
    assert np.all(mu == 0.0) and np.all(sg == 1.0)
    assert np.array_equal(inv, ~np.isfinite(win).all(axis=1))
    assert not inv[np.flatnonzero(np.all(np.diff(win, axis=1) == 0, axis=1))].any()
    assert np.array_equal(ids, np.flatnonzero(~inv))
    want = np.array([O.eld(metric, q, t[g:g + m], w, arg) for g in ids])
    assert np.array_equal(d, want)
    assert thr == np.sort(d)[5]
    assert np.array_equal(s_ids, e_ids) and np.all(np.isin(ranks, s_ids))
    assert np.all(~(lb[ids] > d + 1e-9 * np.maximum(1.0, d)))
    # the mode matters: z-normalised distances are not these
    zd = O.exact(t, q, m, 1e300, metric, w, arg)[1]
    assert not np.array_equal(zd[: d.size], d[: zd.size])


# ------------------------------------------------------------------ SWALE ---
def test_swale_closed_forms():
    """SWALE (src/ElasticMeasures.cpp:200-231; match within eps 1, mismatch or
    gap 5): identical series cost m (all diagonal matches); with w = 0 the
    path is the diagonal, sum of 1 per matching and 5 per other pair; a
    brute-force full DP from the definition equals the banded one at
    w = m - 1; and no path beats 1 per step of the longer series."""
    rng = np.random.default_rng(31)
    for m in (2, 5, 17, 40):
        q = rng.normal(size=m)
        t = q + rng.normal(size=m) * 0.3
        assert O.eld("swale", q, q, int(rng.integers(0, m)), 0.1) == m
        diag = np.where(np.abs(q - t) <= 0.2, 1.0, 5.0).sum()
        assert O.eld("swale", q, t, 0, 0.2) == diag
        D = np.full((m + 1, m + 1), np.inf)
        D[0, 0] = 0.0
        for i in range(m + 1):
            for j in range(m + 1):
                if i == j == 0:
                    continue
                best = np.inf
                if i and j:
                    best = D[i - 1, j - 1] + (1.0 if abs(q[i - 1] - t[j - 1]) <= 0.2 else 5.0)
                if i:
                    best = min(best, D[i - 1, j] + 5.0)
                if j:
                    best = min(best, D[i, j - 1] + 5.0)
                D[i, j] = best
        assert O.eld("swale", q, t, m - 1, 0.2) == D[m, m]
        assert O.eld("swale", q, t, m // 3, 0.2) >= m


def test_swale_glb_known_answers_and_bound():
    # Case A, eps = 0.5: boundary min(sub(1, 0), 5) = 5 (|1 - 0| > .5); T -> Q at
    # T2 = 3 above U_Q2 = 2 by 1 > .5: 5; Q -> T nothing -> 5 + 5
    assert oracle_glb("swale", A[0], A[1], 0.5, 1) == 10.0
    # eps = 1: the boundary and the T -> Q term match within eps: 1 + 1
    assert oracle_glb("swale", A[0], A[1], 1.0, 1) == 2.0
    # Case B, eps = .5: boundary 5; Q -> T at Q1 = 3 above U_T1 = 1 by 2: 5 -> 10
    assert oracle_glb("swale", B[0], B[1], 0.5, 1) == 10.0
    rng = np.random.default_rng(32)
    for _ in range(300):
        m = int(rng.integers(3, 14))
        w = int(rng.integers(0, m))
        q, t = rng.normal(size=m), rng.normal(size=m)
        assert oracle_glb("swale", q, t, 0.3, w) <= O.eld("swale", q, t, w, 0.3)
