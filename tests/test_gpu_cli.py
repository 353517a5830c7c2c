"""The fsmeld command line end to end on the GPU: cutoff -> search -> exact
on a small target file, every report line in the reference's format
(src/fsmeld.cpp:521-554 cutoff, :972-1026 search, :1219 / :1292-1334
exact; src/search_result.cpp:39-101 CSV), the indices and threshold equal
to the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

import fsmeld_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fsmeld_amd", "bin", "fsmeld")


def run(args, env):
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout


@pytest.mark.parametrize("metric,arg", [("dtw", 0.0), ("msm", 0.5)])
def test_cli_cutoff_search_exact(gpu, tmp_path, metric, arg):
    t = F.synth_target(71, 1 << 15, 0.001, 5.0, 3, 200, 2)
    m, w, k = 64, 3, 6
    q, _ = F.synth_query(71, t, m, 0.1)
    folder = tmp_path / "walk"
    folder.mkdir()
    tf, qf = folder / "target.bin", tmp_path / "query.txt"
    t.astype("<f8").tofile(tf)
    qf.write_text("\n".join(f"{x:.17g}" for x in q) + "\n")
    csv = tmp_path / "results.csv"
    env = dict(os.environ, FSMELD_RESULTS_CSV=str(csv))

    thr, ranks, _, _ = O.cutoff(t, q, m, k, metric, w, arg)
    out = run(["cutoff", tf, qf, m, k, metric, w, arg], env)
    got = re.search(r"^  - Final Top-K Threshold:      (\S+)$", out, re.M)
    assert got and float(got.group(1)) == thr
    rank_lines = re.findall(r"^Rank (\d+): Index A\[(\d+)\]=(\d+)$", out, re.M)
    assert [int(r[0]) for r in rank_lines] == list(range(1, k + 1))
    assert [int(r[1]) for r in rank_lines] == list(range(1, k + 1))
    assert np.array_equal([int(r[2]) for r in rank_lines], ranks)
    for key in ("Candidates checked by LB", "Pruned by LBIP", "Pruned by GLB",
                "Exact distances computed"):
        assert re.search(rf"^  - {key}:\s+\d+$", out, re.M), key
    assert re.search(r"^ts_n=32705$", out, re.M) and re.search(r"^m_k=512$", out, re.M)

    cut = f"{thr:.17g}"
    want, ost = O.search(t, q, m, thr, metric, w, arg)
    out = run(["search", tf, qf, m, cut, metric, w, arg], env)
    idx = re.findall(r"^A\[(\d+)\]=(\d+)$", out, re.M)
    assert [int(i) for i, _ in idx] == list(range(len(want)))
    assert np.array_equal([int(v) for _, v in idx], want)
    assert re.search(rf"^Final Matches: {len(want)}$", out, re.M)
    assert re.search(rf"^Total Valid Candidates: {ost['valid']}$", out, re.M)
    assert re.search(rf"^Invalid candidates: {ost['invalid']}$", out, re.M)
    assert re.search(r"^kk1=\d+$", out, re.M) and re.search(r"^T_total: \S+ s$", out, re.M)
    row = re.search(r"^EXPERIMENT_CSV=(.*)$", out, re.M).group(1).split(",")
    assert row[:7] == ["walk", "zscore", metric, str(m), str(w), cut, "FSMELD"]
    assert int(row[9]) == ost["valid"] and int(row[10]) == ost["invalid"]
    assert int(row[14]) == len(want) and row[15] != "" and row[16] != ""

    ids, dist, _ = O.exact(t, q, m, thr, metric, w, arg)
    out = run(["exact", tf, qf, m, cut, metric, w, arg], env)
    gets = re.findall(rf"^GET {metric}: (-?\d+\.\d{{6}}) \(index: (\d+)\)$", out, re.M)
    assert np.array_equal([int(g[1]) for g in gets], ids)
    assert [g[0] for g in gets] == ["%f" % d for d in dist]
    idx = re.findall(r"^A\[(\d+)\] = (\d+)$", out, re.M)
    assert np.array_equal([int(v) for _, v in idx], ids)
    assert re.search(rf"^Final Matches: {len(ids)}$", out, re.M)
    row = re.search(r"^EXPERIMENT_CSV=(.*)$", out, re.M).group(1).split(",")
    assert row[6] == "ELD" and int(row[14]) == len(ids) and row[15] == "" and row[16] == ""
    lines = csv.read_text().splitlines()
    assert lines[0].startswith("dataset,normalization,metric,") and len(lines) == 3
    # streamed through a device budget smaller than the target: the same report
    env_b = dict(env, FSMELD_DEVICE_BUDGET=str(2 * (8 * 8000 + F.query_bytes(8000, metric, m, w))))
    out_b = run(["search", tf, qf, m, cut, metric, w, arg], env_b)
    idx_b = re.findall(r"^A\[(\d+)\]=(\d+)$", out_b, re.M)
    assert np.array_equal([int(v) for _, v in idx_b], want)
    # LB_IP on the FP32 matrix cores (FSMELD_LBIP=f32): the same matches
    out_f = run(["search", tf, qf, m, cut, metric, w, arg], dict(env, FSMELD_LBIP="f32"))
    idx_f = re.findall(r"^A\[(\d+)\]=(\d+)$", out_f, re.M)
    assert np.array_equal([int(v) for _, v in idx_f], want)
