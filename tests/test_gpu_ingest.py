"""Out-of-core queries (fsm_query_file): a target file streamed through the
GPU in chunks of whole macro-segments under a device budget smaller than
the target (the reference's file streaming, src/fsmeld.cpp:740-746,
read_tail_segment :159-168).  Every command's result -- indices, distance
bits, threshold bits, counters of valid / invalid windows -- equals the
whole-target query's and the oracle's, with an exclusion across chunk
borders, the EOF-aligned tail, chunks without k valid windows and raw mode."""
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
    for key in ("valid", "invalid", "matches"):
        assert a.stats[key] == b.stats[key], key
    assert bits(a.threshold) == bits(b.threshold)


@pytest.fixture(scope="module")
def target_file(tmp_path_factory):
    t = F.synth_target(77, 200_003, 0.001, 5.0, 6, 300, 4)
    t[90_000:91_000] = np.nan   # one region without valid windows
    p = tmp_path_factory.mktemp("ingest") / "target.bin"
    t.astype("<f8").tofile(p)
    return t, str(p)


def budget_for(n, metric, m, w, parts):
    nc = n // parts
    return 2 * (8 * nc + F.query_bytes(nc, metric, m, w))


@pytest.mark.parametrize("metric,arg", METRICS)
def test_streamed_commands_equal_whole_target(gpu, target_file, metric, arg):
    t, path = target_file
    T = F.Target(t)
    m, w, k = 64, 4, 9
    q, off = F.synth_query(77, t, m, 0.1)
    excl = (off - 3000, off + 40_000)           # crosses chunk borders
    B = budget_for(t.size, metric, m, w, 7)     # about seven chunks
    assert B < 8 * t.size + F.query_bytes(t.size, metric, m, w)
    for raw in (False, True):
        c = T.cutoff(q, m, k, metric, w, arg, exclude=excl, raw=raw)
        cf = F.query_file(path, "cutoff", q, m, metric, w, arg, k=k, exclude=excl, raw=raw,
                          device_budget=B)
        same(cf, c)
        for mode in ("search", "exact"):
            a = T.query(mode, q, m, metric, w, arg, cutoff_value=c.threshold, exclude=excl,
                        raw=raw)
            b = F.query_file(path, mode, q, m, metric, w, arg, cutoff_value=c.threshold,
                             exclude=excl, raw=raw, device_budget=B)
            same(b, a)
        tk = T.exact_topk(q, m, k, metric, w, arg, exclude=excl, raw=raw)
        tkf = F.query_file(path, "exact_topk", q, m, metric, w, arg, k=k, exclude=excl, raw=raw,
                           device_budget=B)
        same(tkf, tk)
    # and against the oracle directly
    thr, ranks, rd, _ = O.cutoff(t, q, m, k, metric, w, arg, exclude=excl, threads=16)
    cf = F.query_file(path, "cutoff", q, m, metric, w, arg, k=k, exclude=excl, device_budget=B)
    assert bits(cf.threshold) == bits(thr) and np.array_equal(cf.indices, ranks)
    ids, _ = O.search(t, q, m, thr, metric, w, arg, exclude=excl, threads=16)
    sf = F.query_file(path, "search", q, m, metric, w, arg, cutoff_value=thr, exclude=excl,
                      device_budget=B)
    assert np.array_equal(sf.indices, ids)


def test_streamed_query_edge_cases(gpu, target_file):
    t, path = target_file
    q, _ = F.synth_query(78, t, 64, 0.1)
    # a budget that holds no chunk of one macro-segment
    with pytest.raises(F.FsmeldError) as e:
        F.query_file(path, "search", q, 64, "dtw", 4, 0.0, cutoff_value=1.0, device_budget=1000)
    assert e.value.code == -2
    # one chunk (a budget above the whole target) equals the in-memory query
    c = F.Target(t).cutoff(q, 64, 5, "msm", 4, 0.5)
    cf = F.query_file(path, "cutoff", q, 64, "msm", 4, 0.5, k=5, device_budget=64 << 30)
    same(cf, c)
    # a short file: a single segment (N < m_k), streamed in one chunk
    short = path + ".short"
    t[:700].astype("<f8").tofile(short)
    q128, _ = F.synth_query(79, t[:700], 128, 0.1)
    c = F.Target(t[:700]).cutoff(q128, 128, 5, "erp", 6, 0.0)
    cf = F.query_file(short, "cutoff", q128, 128, "erp", 6, 0.0, k=5, device_budget=1 << 30)
    same(cf, c)
