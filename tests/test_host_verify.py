"""The engine's host re-verification DP (fsmeld_amd/csrc/host_verify.cpp,
which decides windows within 1e-9 of a decision boundary) gives the oracle
verifier's bits -- itself pinned to the reference's ElasticMeasures.cpp --
on random pairs of every metric, radius 0 .. m - 1 and ties."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SHIM = r'''
#include <cstdint>
namespace fsm { double host_eld(int, const double*, const double*, int, int, double); }
extern "C" double shim_host_eld(int metric, const double* x, const double* y, int m, int w,
                                double arg) { return fsm::host_eld(metric, x, y, m, w, arg); }
'''


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    d = tmp_path_factory.mktemp("hostv")
    (d / "shim.cpp").write_text(SHIM)
    so = d / "libhostv.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "include"), "-o", str(so), str(d / "shim.cpp"),
                    os.path.join(ROOT, "fsmeld_amd", "csrc", "host_verify.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    P = ctypes.POINTER(ctypes.c_double)
    lib.shim_host_eld.restype = ctypes.c_double
    lib.shim_host_eld.argtypes = [ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    return lib, P


def test_host_eld_equals_oracle(host):
    lib, P = host
    rng = np.random.default_rng(3)
    for trial in range(2500):
        m = int(rng.integers(2, 70))
        w = int(rng.integers(0, m))
        metric = int(rng.integers(1, 7))
        arg = {1: 0.0, 2: float(rng.uniform(0, 1)), 3: float(rng.normal()),
               4: float(rng.uniform(0, 1)), 5: float(rng.uniform(0, 2)),
               6: float(rng.uniform(0, 1))}[metric]
        q = np.round(rng.normal(size=m), 1 if trial % 4 == 0 else 12)
        t = np.round(rng.normal(size=m), 1) if trial % 3 else q + rng.normal(size=m) * 0.02
        t = np.ascontiguousarray(t)
        got = lib.shim_host_eld(metric, q.ctypes.data_as(P), t.ctypes.data_as(P), m, w, arg)
        want = O.eld(metric, q, t, w, arg)
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (metric, m, w, arg)


def test_host_eld_equals_oracle_long(host):
    """Longer rows (m up to 700), so the band's interior, its edges and both
    grid borders all occur in one pair."""
    lib, P = host
    rng = np.random.default_rng(17)
    for trial in range(90):
        m = int(rng.integers(100, 700))
        w = int(rng.choice([0, 1, int(rng.integers(2, m)), m - 1]))
        metric = 1 + trial % 6
        arg = {1: 0.0, 2: 0.3, 3: 0.5, 4: 0.3, 5: 0.7, 6: 0.3}[metric]
        q = np.cumsum(rng.normal(size=m))
        t = np.ascontiguousarray(q + rng.normal(size=m) * (0.05 if trial % 2 else 1.0))
        got = lib.shim_host_eld(metric, q.ctypes.data_as(P), t.ctypes.data_as(P), m, w, arg)
        want = O.eld(metric, q, t, w, arg)
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (metric, m, w, arg)
