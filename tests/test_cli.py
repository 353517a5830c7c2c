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


DRIVER = r'''
#include "search_result.h"
#include <iostream>
int main(int argc, char** argv) {
    std::cout << dataset_name_of(argv[1]) << "|" << dataset_name_of("bare.bin") << "|"
              << dataset_name_of("") << std::endl;
    ExperimentRow r;
    r.dataset = "ecg,\"lead\"";
    r.metric = "dtw";
    r.m = 128;
    r.w = 6;
    r.xi = 0.1;
    r.method = "FSMELD";
    r.lb_time_seconds = 1.0 / 3.0;
    r.total_time_seconds = 2.5;
    r.valid_candidates = 8;
    r.invalid_candidates = 2;
    r.pruned_candidates = 3;
    r.survivors = 5;
    r.matches = 4;
    r.lbip_time_seconds = 0.25;
    r.lbip_pruning_rate = 0.125;
    emit_experiment_row(r);
    ExperimentRow e;   // exact-style row: no lbip cells; then one with no total / matches
    e.dataset = "d";
    e.metric = "msm";
    e.method = "ELD";
    e.total_time_seconds = 1.0;
    e.matches = 0;
    emit_experiment_row(e);
    e.total_time_seconds.reset();
    e.matches.reset();
    emit_experiment_row(e);
    return 0;
}
'''


def test_experiment_csv_format(tmp_path):
    """Columns, 17-digit reals, quoting, empty optional cells and the header
    only on a new file (format of src/search_result.cpp:39-101)."""
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    cli = os.path.join(ROOT, "fsmeld_amd", "cli")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", cli, "-o", str(exe), str(src),
                    os.path.join(cli, "search_result.cpp")], check=True)
    csv = tmp_path / "out.csv"
    env = dict(os.environ, FSMELD_RESULTS_CSV=str(csv))
    p = subprocess.run([str(exe), "/data/ecg/target.bin"], capture_output=True, text=True,
                       env=env, check=True)
    lines = p.stdout.splitlines()
    assert lines[0] == "ecg|bare|unknown"
    header = ("dataset,normalization,metric,m,w,xi,method,lb_time,total_time,valid_candidates,"
              "invalid_candidates,pruned_candidates,pruning_rate,survivors,matches,"
              "lbip_time,lbip_pruning_rate,timestamp")
    assert lines[1] == "EXPERIMENT_CSV_HEADER=" + header
    row = lines[2][len("EXPERIMENT_CSV="):]
    assert lines[2].startswith("EXPERIMENT_CSV=")
    stamp = row.rsplit(",", 1)[1]
    import re
    assert re.fullmatch(r"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\d", stamp)
    assert row.rsplit(",", 1)[0] == (
        '"ecg,""lead""",zscore,dtw,128,6,0.10000000000000001,FSMELD,0.33333333333333331,2.5,'
        "8,2,3,0.375,5,4,0.25,0.125")
    assert lines[3] == "Experiment CSV saved to: " + str(csv)
    eld = [l[len("EXPERIMENT_CSV="):].rsplit(",", 1)[0] for l in lines
           if l.startswith("EXPERIMENT_CSV=")][1:]
    assert eld == ["d,zscore,msm,0,0,0,ELD,0,1,0,0,0,0,0,0,,",
                   "d,zscore,msm,0,0,0,ELD,0,,0,0,0,0,0,,,"]
    file_lines = csv.read_text().splitlines()
    assert file_lines[0] == header and len(file_lines) == 4   # header once, three rows
    assert [l.rsplit(",", 1)[0] for l in file_lines[1:]] == [row.rsplit(",", 1)[0]] + eld
