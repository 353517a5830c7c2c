"""Pins the oracle's five exact verifiers to the reference's own
src/ElasticMeasures.cpp: live against oracle/_ref/libref_eld.so where the
reference tree was available to build it, and always against the golden
vectors it produced (tests/golden/eld_ref_vectors.json)."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "eld_ref_vectors.json")


def test_oracle_eld_matches_reference_golden_vectors():
    cases = json.load(open(GOLDEN))["cases"]
    assert len(cases) >= 100
    for c in cases:
        got = O.eld(c["metric"], np.array(c["q"]), np.array(c["t"]), c["w"], c["arg"])
        assert float(got).hex() == c["d"], c


# existence only: the library is loaded inside the test, so processes that
# merely collect this file (the -m gpu run) never map it
@pytest.mark.skipif(not os.path.exists(O.REF_ELD),
                    reason="reference tree not present to build oracle/_ref")
def test_oracle_eld_matches_reference_build_random():
    ref = O.ref_eld_lib()
    P = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(7)
    for trial in range(1500):
        m = int(rng.integers(2, 90))
        w = int(rng.integers(0, m))
        metric = int(rng.integers(1, 7))   # 6: SWALE
        arg = {1: 0.0, 2: float(rng.uniform(0, 1)), 3: float(rng.normal()),
               4: float(rng.uniform(0, 1)), 5: float(rng.uniform(0, 2)),
               6: float(rng.uniform(0, 1))}[metric]
        q = rng.normal(size=m)
        t = rng.normal(size=m) if trial % 3 else q + rng.normal(size=m) * 0.02
        want = ref.ref_eld(metric, q.ctypes.data_as(P), t.ctypes.data_as(P), m, w, arg)
        got = O.eld(metric, q, t, w, arg)
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (metric, m, w, arg)
