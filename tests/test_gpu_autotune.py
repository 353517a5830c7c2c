"""The autotuner and the launch shapes it sets (fsm_autotune, fsm_tuning_*).

Bar: a launch shape changes speed only -- every search under every shape
returns the matches, distances and counters of the built-in shapes, which
equal the oracle's; the autotuner returns values from its candidate sets and
leaves them applied."""
import numpy as np
import pytest

import fsmeld_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

METRICS = [("dtw", 0.0), ("erp", 0.0), ("edr", 0.25), ("lcss", 0.25), ("msm", 0.5)]


def bits(x):
    return np.asarray(x, np.float64).view(np.uint64)


def same(a, b):
    assert np.array_equal(a.indices, b.indices)
    assert np.array_equal(bits(a.distances), bits(b.distances))
    for key in ("valid", "after_lbip", "after_glb", "matches"):
        assert a.stats[key] == b.stats[key], key


@pytest.fixture(scope="module")
def setup():
    t = F.synth_target(83, 1 << 18, 0.001, 5.0, 4, 200, 3)
    T = F.Target(t)
    specs = []
    # one length per LB_IP segment class (m_k 512 .. 8192) and beyond
    for i, m in enumerate([40, 100, 200, 480, 700, 1100]):
        metric, arg = METRICS[i % 5]
        w = max(1, m // 20)
        q, _ = F.synth_query(900 + i, t, m, 0.1)
        c = T.cutoff(q, m, 5, metric, w, arg)
        specs.append(dict(mode="search", query=q, m=m, metric=metric, w=w, arg=arg,
                          cutoff=c.threshold))
    base = [T.query("search", sp["query"], sp["m"], sp["metric"], sp["w"], sp["arg"],
                    cutoff_value=sp["cutoff"]) for sp in specs]
    return t, T, specs, base


def run_all(T, specs):
    return [T.query("search", sp["query"], sp["m"], sp["metric"], sp["w"], sp["arg"],
                    cutoff_value=sp["cutoff"]) for sp in specs]


@pytest.mark.parametrize("shape", [
    dict(lbip_threads=[128, 1024, 1024, 512], eld_refill=[0, 4, 24, 4, 24, 4, 4],
         glb_blocks=1024),
    dict(lbip_threads=[1024, 256, 512, 1024], eld_refill=[0, 64, 1, 16, 12, 8, 8],
         glb_blocks=8192),
])
def test_explicit_shapes_keep_results(gpu, setup, shape):
    t, T, specs, base = setup
    T.set_tuning(shape)
    try:
        assert T.tuning() == shape
        for r, b in zip(run_all(T, specs), base):
            same(r, b)
    finally:
        T.set_tuning({})
    assert T.tuning() == dict(lbip_threads=[0] * 4, eld_refill=[0] * 7, glb_blocks=0)
    # the built-in shapes agree with the oracle
    sp = specs[1]
    want, _ = O.search(t, sp["query"], sp["m"], sp["cutoff"], sp["metric"], sp["w"], sp["arg"])
    assert np.array_equal(base[1].indices, want)


def test_autotune_picks_candidates_and_keeps_results(gpu, setup):
    t, T, specs, base = setup
    try:
        tu = T.autotune(specs, repeats=1)
        assert tu == T.tuning()
        assert all(v in (0, 128, 256, 512, 1024) for v in tu["lbip_threads"])
        assert all(v in (0, 4, 8, 12, 16, 24) for v in tu["eld_refill"])
        assert tu["glb_blocks"] in (0, 1024, 2048, 4096, 8192)
        for r, b in zip(run_all(T, specs), base):
            same(r, b)
        # a shared handle sees the same shapes
        with T.share() as H:
            assert H.tuning() == tu
    finally:
        T.set_tuning({})


def test_tuning_rejects_bad_values(gpu, setup):
    _, T, specs, _ = setup
    for bad in (dict(lbip_threads=[100, 0, 0, 0]), dict(eld_refill=[0, 65]),
                dict(glb_blocks=-1)):
        with pytest.raises(F.FsmeldError):
            T.set_tuning(bad)
    with pytest.raises(F.FsmeldError):
        T.autotune([dict(specs[0], mode="cutoff", k=3)])
