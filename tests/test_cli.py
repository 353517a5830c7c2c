"""The fsmeld CLI keeps the reference's command-line contract
(src/main.cpp:12-110): usage, validation messages and exit codes.  Every
case below is rejected before any GPU work."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fsmeld_amd", "bin", "fsmeld")


def run(*args, cwd=None):
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd)
    return p.returncode, p.stdout, p.stderr


@pytest.fixture()
def files(tmp_path):
    t = tmp_path / "target.bin"
    np.cumsum(np.random.default_rng(0).normal(size=1000)).astype("<f8").tofile(t)
    q = tmp_path / "query.txt"
    q.write_text(" ".join(f"{x:.17g}" for x in np.random.default_rng(1).normal(size=64)))
    return str(t), str(q), tmp_path


def test_help_and_usage():
    for args in ([], ["--help"], ["-h"]):
        rc, out, _ = run(*args)
        assert rc == 0 and out.startswith("FSMELD: exact z-normalized subsequence search")
    rc, out, err = run("search", "a")
    assert rc == 1 and "Usage:" in out
    assert err.strip() == "Error: Expected eight arguments, plus optional exclusion bounds."


@pytest.mark.parametrize("argv,msg", [
    (["find", "T", "Q", "64", "1", "dtw", "3", "0"], "Unknown command: find"),
    (["search", "T", "Q", "1", "1", "dtw", "0", "0"], "m must be at least 2 and fit the FFT indexing range."),
    (["search", "T", "Q", "64x", "1", "dtw", "3", "0"], "Invalid m"),
    (["search", "T", "Q", "64", "1", "swale", "3", "0"], "Unsupported metric: swale"),
    (["search", "T", "Q", "64", "1", "dtw", "64", "0"], "w must satisfy 0 <= w < m."),
    (["search", "T", "Q", "64", "1", "dtw", "3", "nan"], "Invalid finite metric parameter"),
    (["search", "T", "Q", "64", "1", "edr", "3", "-0.1"], "Matching tolerance and MSM cost must be non-negative."),
])
def test_argument_errors(files, argv, msg):
    t, q, _ = files
    argv = [t if a == "T" else q if a == "Q" else a for a in argv]
    rc, _, err = run(*argv)
    assert rc == 1 and err.strip() == "Error: " + msg


def test_target_and_query_errors(files):
    t, q, tmp = files
    short = tmp / "short.bin"
    np.zeros(100).tofile(short)
    rc, _, err = run("search", short, q, 64, 1, "dtw", 3, 0)
    assert rc == 1 and "at least 2*m-1 complete double values" in err
    odd = tmp / "odd.bin"
    odd.write_bytes(b"\0" * 8001)
    rc, _, err = run("search", odd, q, 64, 1, "dtw", 3, 0)
    assert rc == 1 and "at least 2*m-1 complete double values" in err
    bad_q = tmp / "bad.txt"
    bad_q.write_text("1 2 nan 4")
    rc, _, err = run("search", t, bad_q, 4, 1, "dtw", 1, 0)
    assert rc == 1 and err.strip() == "Error: Query must begin with m finite numeric values."
    flat = tmp / "flat.txt"
    flat.write_text(" ".join(["3.5"] * 64))
    rc, _, err = run("search", t, flat, 64, 1, "dtw", 3, 0)
    assert rc == 1 and "cannot be z-normalized" in err
    rc, _, err = run("search", t, q, 64, -1, "dtw", 3, 0)
    assert rc == 1 and err.strip() == "Error: cutoff must be non-negative."
    rc, _, err = run("cutoff", t, q, 64, 5000, "dtw", 3, 0)
    assert rc == 1 and "k exceeds the candidate count" in err
    rc, _, err = run("search", t, q, 64, 1, "dtw", 3, 0, 10, 5)
    assert rc == 1 and err.strip() == "Error: Exclusion requires 0 <= start < end."
