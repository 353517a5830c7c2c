"""Many queries over one target: fsm_query_batch and the serving loop.

Bar: every query's result in a batch or in the server's mix is identical to
the same query run alone (indices, distance bits, counters), whatever else
runs beside it; cancellation and the memory budget behave as documented."""
import time

import numpy as np
import pytest

import fsmeld_amd as F

pytestmark = pytest.mark.gpu

METRICS = [("dtw", 0.0), ("erp", 0.0), ("edr", 0.25), ("lcss", 0.25), ("msm", 0.5)]


def bits(x):
    return np.asarray(x, np.float64).view(np.uint64)


def mixed_specs(t, n, seed):
    """Queries of different m, w, metric and mode over target t; cutoffs from
    a cutoff query so that searches return matches."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        m = int(np.exp(rng.uniform(np.log(16), np.log(200))))
        w = max(0, int(0.05 * m) + int(rng.integers(0, 3)))
        metric, arg = METRICS[int(rng.integers(0, 5))]
        q, _ = F.synth_query(seed * 100 + i, t, m, 0.1)
        mode = ["search", "cutoff", "exact"][i % 3]
        spec = dict(mode=mode, query=q, m=m, metric=metric, w=w, arg=arg, k=5)
        specs.append(spec)
    return specs


def fill_cutoffs(T, specs):
    for sp in specs:
        if sp["mode"] != "cutoff":
            r = T.cutoff(sp["query"], sp["m"], 5, sp["metric"], sp["w"], sp["arg"])
            sp["cutoff"] = r.threshold


def same_result(a, b):
    assert np.array_equal(a.indices, b.indices)
    assert np.array_equal(bits(a.distances), bits(b.distances))
    for key in ("valid", "invalid", "after_lbip", "after_glb", "matches", "threshold"):
        x, y = a.stats[key], b.stats[key]
        assert (x == y) or (x != x and y != y), key


@pytest.fixture(scope="module")
def target():
    t = F.synth_target(41, 1 << 16, 0.001, 5.0, 4, 200, 3)
    return t, F.Target(t)


def test_batch_equals_single(gpu, target):
    t, T = target
    specs = mixed_specs(t, 18, 7)
    fill_cutoffs(T, specs)
    alone = [T.query(sp["mode"], sp["query"], sp["m"], sp["metric"], sp["w"], sp["arg"],
                     cutoff_value=sp.get("cutoff", 0.0), k=sp["k"]) for sp in specs]
    for conc in (1, 4):
        got = T.batch(specs, concurrency=conc)
        for g, a in zip(got, alone):
            assert not isinstance(g, Exception), g
            same_result(g, a)


def test_batch_reports_per_query_errors(gpu, target):
    t, T = target
    good = dict(mode="search", query=F.synth_query(3, t, 32, 0.1)[0], m=32, metric="dtw", w=2,
                arg=0.0, cutoff=1.0)
    bad = dict(good, w=40)  # w >= m
    out = T.batch([good, bad, good], concurrency=2)
    assert not isinstance(out[0], Exception) and not isinstance(out[2], Exception)
    assert isinstance(out[1], F.FsmeldError) and out[1].code == -1


def test_shared_window_stats_across_handles(gpu):
    """Window statistics and envelopes are kept per target and shared by its
    handles once a length / radius comes back: three handles alternating
    lengths and radii, with trims in between, give exactly the results of a
    target that shares nothing (a new one per query)."""
    t = F.synth_target(43, 1 << 15, 0.001, 5.0, 3, 150, 2)
    T = F.Target(t)
    H = [T, T.share(), T.share()]
    ms = [40, 57, 40, 40, 57, 90, 40, 90, 57, 40]
    for rnd in range(2):
        for i, m in enumerate(ms):
            metric, arg = METRICS[(i + rnd) % 5]
            w = max(1, m // 20) + (i % 2)
            q, _ = F.synth_query(500 + 37 * rnd + i, t, m, 0.1)
            got = H[i % 3].query("cutoff", q, m, metric, w, arg, k=5)
            with F.Target(t) as R:
                ref = R.query("cutoff", q, m, metric, w, arg, k=5)
            same_result(got, ref)
            c = ref.stats["threshold"]
            got = H[(i + 1) % 3].search(q, m, c, metric, w, arg)
            with F.Target(t) as R:
                ref = R.search(q, m, c, metric, w, arg)
            same_result(got, ref)
        H[1].trim()
        H[2].trim()
    for h in H[1:]:
        h.close()


def test_server_mix_equals_single(gpu, target):
    t, T = target
    specs = mixed_specs(t, 24, 11)
    fill_cutoffs(T, specs)
    alone = [T.query(sp["mode"], sp["query"], sp["m"], sp["metric"], sp["w"], sp["arg"],
                     cutoff_value=sp.get("cutoff", 0.0), k=sp["k"]) for sp in specs]
    biggest = max(F.query_bytes(len(t), sp["metric"], sp["m"], sp["w"]) for sp in specs)
    # a budget of ~2.5 queries' working sets: admission must hold some back
    # and idle workers must release cached buffers
    with F.Server(T, mem_budget=int(biggest * 2.5), max_workers=4) as S:
        tickets = [S.submit(sp) for sp in specs]
        for tk, a in zip(tickets, alone):
            state, res, lat = S.poll(tk, wait_ms=-1)
            assert state == "done" and lat >= 0.0
            assert not isinstance(res, Exception), res
            same_result(res, a)
        assert S.info()["queued"] == 0


def test_server_cancel_and_budget(gpu, target):
    t, T = target
    sp = dict(mode="exact", query=F.synth_query(5, t, 128, 0.1)[0], m=128, metric="msm", w=6,
              arg=0.5, cutoff=1e9)
    need = F.query_bytes(len(t), "msm", 128, 6)
    # the workers get three quarters of the budget (the rest is the shared
    # statistics cache): room for exactly one such query at a time
    with F.Server(T, mem_budget=need * 4 // 3 + 4, max_workers=2) as S:
        tickets = [S.submit(sp) for _ in range(6)]
        S.cancel(tickets[-1])   # still queued behind the others: dropped
        states = [S.poll(tk, wait_ms=-1)[0] for tk in tickets]
        assert states[-1] == "cancelled"
        assert all(st == "done" for st in states[:-1])
        big = dict(sp, m=1024, w=200)   # row-band path: its scratch exceeds the budget
        assert F.query_bytes(len(t), "msm", 1024, 200) > need
        with pytest.raises(F.FsmeldError) as e:
            S.submit(big)
        assert e.value.code == -2
        # pending poll returns at once
        tk = S.submit(sp)
        t0 = time.perf_counter()
        S.poll(tk, wait_ms=0)
        assert time.perf_counter() - t0 < 1.0
        assert S.poll(tk, wait_ms=-1)[0] == "done"
