#!/usr/bin/env python3
"""Benchmark of the FSMELD cascade on one MI355X.

Workload (plain single query, large, all metrics): a 2^26-sample FP64 random
walk with 0.1% N(0,5) spikes and 1000 constant runs of 300 samples, held in
HBM; one m=512 query (copied from a seeded offset plus N(0,0.1) noise),
w=25, metrics DTW, ERP g=0, EDR eps=0.25, LCSS eps=0.25, MSM c=0.5, each at
the cutoff `cutoff` returns for k=100 (computed before timing).

One step = one `search` per metric (5 queries), each a full pass of the
cascade over all 67,108,353 windows -- window statistics, target envelope,
LB_IP, GLB, ELD -- with results copied back to the host.  The statistics and
envelope are recomputed for every query, as the reference computes them per
query (fsm_target_reuse_stats(0)); the serving and batching workloads keep
them resident instead.  value = queries per second over the timed steps.

Printed: one JSON line with the driver's fields plus
  candidates_verified_per_s  valid windows decided per second,
  p50_ms / p99_ms            per-query latency,
  kernels                    per stage kernel: device time from HIP events on
                             the engine's stream and its algorithmic work,
  roofline                   the kernel with the most device time,
  cpu_baseline               the CPU restatement (oracle/) on the host cores,
                             on a sample of macro-segments, scaled.

Algorithmic work per launch (unit: one candidate window; S segments of m_k
samples, k_m_l windows each):
  k_window_sums    8 S m_k bytes in (segments with their overlap), 16 B per
                   window slot out (the two running sums)
  k_window_finish  16 B in, 17 B out per window slot (mean, std, flag)
  k_envelope       8 N bytes in, 16 N out (lower and upper envelope)
  k_lbip           17 B in per window slot (mean, std, flag), 16 B in per
                   segment sample (lower and upper envelope), one FP64
                   bound out per window
  k_glb<M>         LB_IP survivors x 3 m FP64 operations (two directional
                   sums over the window), against the FP64 VALU roof
  k_eld<M>         band cells the kernel evaluates for live candidates x the
                   metric's FP64 operations per cell (DTW 5, ERP 9, EDR 7,
                   LCSS 6, MSM 12); the full band of every GLB survivor
                   (m (2w + 1) - w (w + 1) cells) is reported beside it
Peaks: HBM 8 TB/s; FP64 78.6 TFLOP/s (vector and matrix alike on MI355X,
an FMA counting 2, so adds / mins / maxes alone reach at most half of it).

--cutoff-cache PATH stores the cutoffs (and reads them on later runs), so a
profiler run can launch the kernels in the timed configuration only.

--workload batch: 2^24 random walk, 256 queries (m log-uniform in [64, 2048],
w = 5% of m, metric uniform over the five, cutoffs from k=10): the plain path
(one fsm_query at a time) against fsm_query_batch at several concurrencies;
every batched result is compared with the plain one.
--workload small: the 2^20 plain small configuration (m=128, w=6, DTW, k=10
cutoff): GPU search against the CPU restatement on 16 host threads.
--workload precision: the 2^26 target at m = 128 and 1024 (w = 5% of m):
LB_IP's pruning rate at the k=10 cutoff, match-set identity of search and
exact, and LB <= ELD on a sample of windows (the error against an FP64 oracle
is measured by tests/test_gpu_precision.py).
--workload serve: the 2^26 target, Poisson arrivals at --rates queries/s of
queries with m log-uniform in [64, 1024]: latency from submission to result
through the serving loop (fsm_server_*) against the same arrivals served one
query at a time (one worker).
--copy-calibration runs a plain 1 GiB device copy (the scale for HBM
counters) and exits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
ELD_OPS_PER_CELL = {"dtw": 5, "erp": 9, "edr": 7, "lcss": 6, "msm": 12}
METRIC_CODE = {"dtw": 1, "edr": 2, "erp": 3, "lcss": 4, "msm": 5}


def eld_kernel_name(code, m, w):
    """The ELD kernel the engine launches for (m, w): the lane-per-candidate
    register band for w <= 25, one wave per candidate up to w = 127, the row
    band beyond (or where the query staging exceeds 64 KiB of LDS)."""
    if w <= 25 and 24 * (m + 80) <= 65536:
        return f"k_eld<{code},{w}>"
    if w <= 127 and 16 * (m + 2) <= 65536:
        return f"k_eld_wave<{code},{1 if w <= 63 else 2}>"
    return f"k_eld_rows<{code}>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration(steps):
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(max(1, steps)):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                      "launches": max(1, steps)}), flush=True)


def load_cutoffs(path, key):
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("cutoffs") if d.get("key") == key else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--cpu-segments", type=int, default=640,
                    help="macro-segments per metric in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cutoff-cache", default=None)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "small", "batch", "serve", "precision"])
    ap.add_argument("--concurrency", default="4,8")
    ap.add_argument("--rates", default="50,200,1000")
    ap.add_argument("--requests", type=int, default=60, help="arrivals per rate (serve)")
    ap.add_argument("--workers", type=int, default=8, help="server workers (serve)")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="sample searches for the autotuner before timing (0: built-in shapes; "
                         "default: 32 for batch, 0 for plain, where it picks the built-in "
                         "shapes within noise)")
    ap.add_argument("--phase-gaps", action="store_true",
                    help="batch: leave the device idle 0.25 s between phases (warm-up, plain, "
                         "autotune, tuned plain, each concurrency), so that a kernel trace "
                         "splits into them")
    ap.add_argument("--prec-m", default="128,1024", help="query lengths (precision)")
    ap.add_argument("--prec-metrics", default="dtw,erp,edr,lcss,msm",
                    help="metrics (precision)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the engine is single-GPU; a multi-GPU search would give each GPU a range of "
              f"macro-segments).",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.steps)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "precision":
        run_precision(args)
        return
    if args.workload == "small":
        run_small(args)
        return

    import fsmeld_amd as F
    from fsmeld_amd import workloads

    n = 1 << args.log2n
    t0 = time.perf_counter()
    cfg = workloads.plain_large(n=n)
    t, q, m, w = cfg["target"], cfg["query"], cfg["m"], cfg["w"]
    log(f"workload generated in {time.perf_counter() - t0:.1f}s (n={n}, m={m}, w={w})")

    T = F.Target(t)
    T.reuse_stats(False)   # statistics and envelope inside every query
    key = f"plain_large n={n} m={m} w={w} k={cfg['k']}"
    cutoffs = load_cutoffs(args.cutoff_cache, key)
    if cutoffs is None:
        cutoffs = {}
        for metric, arg in cfg["metrics"]:
            r = T.cutoff(q, m, cfg["k"], metric, w, arg)
            cutoffs[metric] = r.threshold
            log(f"cutoff {metric}: k={cfg['k']} threshold={r.threshold:.17g} "
                f"({r.stats['ms_total']:.1f} ms)")
        if args.cutoff_cache:
            with open(args.cutoff_cache, "w") as f:
                json.dump({"key": key, "cutoffs": cutoffs}, f)
    else:
        log(f"cutoffs read from {args.cutoff_cache}")
    tuning = None
    if args.autotune > 0:  # (default off here)
        # launch shapes from the autotuner on the step's own searches (before
        # timing; shapes change speed only, the tuner checks the matches)
        t0 = time.perf_counter()
        tuning = T.autotune([dict(mode="search", query=q, m=m, metric=metric, w=w, arg=arg,
                                  cutoff=cutoffs[metric]) for metric, arg in cfg["metrics"]],
                            repeats=2)
        log(f"autotune {time.perf_counter() - t0:.1f}s: {tuning}")

    def step():
        out = []
        for metric, arg in cfg["metrics"]:
            s0 = time.perf_counter()
            r = T.search(q, m, cutoffs[metric], metric, w, arg)
            out.append((metric, time.perf_counter() - s0, r))
        return out

    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")

    stages = ("stats", "stats_finish", "env", "lbip", "glb", "eld")
    lat, stage, results = [], {k: 0.0 for k in stages}, []
    per_metric = {metric: {k: 0.0 for k in stages + ("eld_cells",)} for metric, _ in cfg["metrics"]}
    t_start = time.perf_counter()
    for i in range(args.steps):
        res = step()
        for metric, dt, r in res:
            lat.append(dt * 1000.0)
            for k in stages:
                stage[k] += r.stats["ms_" + k]
                per_metric[metric][k] += r.stats["ms_" + k]
            per_metric[metric]["eld_cells"] += r.stats["eld_cells"]
        results = res
        log(f"timed step {i + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start

    nq = args.steps * len(cfg["metrics"])
    ms_per_step = elapsed * 1000.0 / args.steps
    qps = nq / elapsed
    G = F.geometry(n, m)
    valid = results[0][2].stats["valid"]
    per_query = {metric: dict(matches=int(r.stats["matches"]), after_lbip=int(r.stats["after_lbip"]),
                              after_glb=int(r.stats["after_glb"]),
                              cutoff=cutoffs[metric]) for metric, _, r in results}

    # Per-kernel figures over the timed region (device time from HIP events
    # recorded on the engine's stream around each stage).
    steps = args.steps
    kernels = []
    slots = G["S"] * G["k_m_l"]

    def hbm(name, ms_total, nbytes):
        ms = ms_total / nq
        kernels.append(dict(kernel=name, bound="hbm", launches=nq, avg_launch_ms=round(ms, 4),
                            total_ms=round(ms_total, 3), work_per_launch=nbytes,
                            achieved=nbytes / (ms / 1e3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s"))

    hbm("k_window_sums", stage["stats"] - stage["stats_finish"],
        8.0 * G["S"] * G["m_k"] + 16.0 * slots)
    hbm("k_window_finish", stage["stats_finish"], 33.0 * slots)
    hbm("k_envelope", stage["env"], 24.0 * n)
    hbm("k_lbip", stage["lbip"], 17.0 * slots + 16.0 * G["S"] * G["m_k"] + 8.0 * G["ts_n"])
    band = m * (2 * w + 1) - w * (w + 1)
    for metric, _ in cfg["metrics"]:
        code = METRIC_CODE[metric]
        pq = per_query[metric]
        g_ms = per_metric[metric]["glb"] / steps
        g_flops = pq["after_lbip"] * 3.0 * m
        kernels.append(dict(kernel=f"k_glb<{code}>", metric=metric, bound="fp64-valu",
                            launches=steps, avg_launch_ms=round(g_ms, 4),
                            total_ms=round(per_metric[metric]["glb"], 3), work_per_launch=g_flops,
                            achieved=g_flops / (g_ms / 1e3) / 1e12, peak=FP64_PEAK_TFLOPS,
                            unit="TFLOP/s"))
        e_ms = per_metric[metric]["eld"] / steps
        cells = pq["after_glb"] * float(band)
        done = per_metric[metric]["eld_cells"] / steps
        flops = done * ELD_OPS_PER_CELL[metric]
        kernels.append(dict(kernel=eld_kernel_name(code, m, w), metric=metric, bound="fp64-valu",
                            launches=steps, avg_launch_ms=round(e_ms, 4),
                            total_ms=round(per_metric[metric]["eld"], 3), work_per_launch=flops,
                            achieved=flops / (e_ms / 1e3) / 1e12, peak=FP64_PEAK_TFLOPS,
                            unit="TFLOP/s", cells_evaluated_per_launch=done,
                            band_cells_per_launch=cells,
                            evaluated_frac=(done / cells) if cells else None))
    for k in kernels:
        k["achieved"] = round(k["achieved"], 3)
        k["frac"] = round(k["achieved"] / k["peak"], 5)
        if k.get("evaluated_frac") is not None:
            k["evaluated_frac"] = round(k["evaluated_frac"], 4)
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    # this workload (scaled by a copy kernel's readings), when recorded
    recorded = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            recorded = json.load(f)
    for k in kernels:
        tr = recorded.get(k["kernel"] + " " + key)
        k["traffic"] = tr["hbm_bytes_per_launch"] if tr else None
    top = max(kernels, key=lambda k: k["total_ms"])
    roofline = dict(kernel=top["kernel"], bound=top["bound"], achieved=top["achieved"],
                    peak=top["peak"], unit=top["unit"], frac=top["frac"], traffic=top["traffic"],
                    work_per_launch=top["work_per_launch"], avg_launch_ms=top["avg_launch_ms"])
    if top["kernel"].startswith("k_eld"):
        roofline["note"] = ("FP64 vector ALU bound, no MFMA (MI355X FP64 vector and matrix peaks "
                            "are both 78.6 TFLOP/s, an FMA counting 2; the DP's adds, mins and "
                            "maxes cap at 39.3); work = the band cells the kernel evaluates for "
                            "live candidates x the metric's FP64 operations per cell; early "
                            "abandoning leaves evaluated_frac of the full bands")
        roofline["evaluated_frac"] = top["evaluated_frac"]

    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        seg = min(args.cpu_segments, G["S"])
        per_metric_s = []
        c0 = time.perf_counter()
        gpu_ids = {metric: r.indices for metric, _, r in results}
        # the sample: seg consecutive full segments around the query's source
        # offset, where the matches are, so the comparison below has windows
        # to compare (segments [s0, s0 + seg) hold windows [s0 k_m_l, (s0 + seg) k_m_l))
        s0 = int(min(max(0, cfg["offset"] // G["k_m_l"] - seg // 2), max(0, G["forsize"] - seg)))
        lo, hi = s0 * G["k_m_l"], (s0 + seg) * G["k_m_l"]
        same, compared = True, 0
        for metric, arg in cfg["metrics"]:
            c1 = time.perf_counter()
            ids, _ = O.search(t, q, m, cutoffs[metric], metric, w, arg, segments=(s0, s0 + seg),
                              threads=threads)
            dt = time.perf_counter() - c1
            per_metric_s.append(dt * G["S"] / seg)
            g = gpu_ids[metric]
            ok = bool(np.array_equal(g[(g >= lo) & (g < hi)], ids))
            same = same and ok
            compared += int(ids.size)
            log(f"cpu baseline {metric}: segments [{s0},{s0 + seg}) in {dt:.2f}s; same windows "
                f"as the GPU there: {ok} ({ids.size} matches)")
        est = float(np.mean(per_metric_s))
        # one thread (the reference is single-threaded) on a smaller sample
        seg1 = max(1, min(seg, seg // 20))
        c1 = time.perf_counter()
        for metric, arg in cfg["metrics"]:
            O.search(t, q, m, cutoffs[metric], metric, w, arg, segments=(s0, s0 + seg1),
                     threads=1)
        est1 = (time.perf_counter() - c1) / len(cfg["metrics"]) * G["S"] / seg1
        log(f"cpu baseline, one thread: {seg1} segments per metric, {est1:.1f} s per query")
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f
                              if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = dict(value=round(1.0 / est, 6), unit="queries/s", cores=threads, kind="port",
                   cpu_model=model, one_thread_value=round(1.0 / est1, 7),
                   one_thread_sample=f"macro-segments [{s0},{s0 + seg1}) per metric, scaled",
                   sample=f"oracle search on macro-segments [{s0},{s0 + seg}) of {G['S']} (around "
                          f"the query's source offset) per metric (5 metrics, k=100 cutoffs), "
                          f"OpenMP over segments, scaled to the full target; "
                          f"{time.perf_counter() - c0:.1f}s of CPU work",
                   est_seconds_per_query=round(est, 3), same_windows=same,
                   windows_compared=compared)

    line = {
        "metric": "queries/sec + candidates verified/sec at exact match parity; p50/p99 latency",
        "value": round(qps, 4),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain single query, large, all metrics: 2^{args.log2n} FP64 "
                               f"random walk + 0.1% N(0,5) spikes + 1000 constant runs; m={m}, "
                               f"w={w}; dtw, erp g=0, edr 0.25, lcss 0.25, msm 0.5; cutoffs "
                               f"from k={cfg['k']}; one step = 5 searches, each recomputing "
                               f"its window statistics and envelope",
                   "n": n, "m": m, "w": w, "k": cfg["k"], "parallelism": "single GPU",
                   "launch_shapes": ("autotuned on the step's searches before timing: "
                                     f"{tuning}" if tuning else "built-in")},
        "candidates_verified_per_s": round(valid * nq / elapsed, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "stage_ms_per_query": {k: round(v / nq, 4) for k, v in stage.items()},
        "per_query": per_query,
        "same_windows": cpu["same_windows"] if cpu else None,
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def _specs_with_cutoffs(T, queries, k, cache, cache_key):
    """Search specs whose cutoffs come from k-NN queries (batched, untimed)."""
    cut = load_cutoffs(cache, cache_key)
    if cut is None:
        cspecs = [dict(mode="cutoff", query=q["query"], m=q["m"], metric=q["metric"], w=q["w"],
                       arg=q["arg"], k=k) for q in queries]
        t0 = time.perf_counter()
        cut = []
        for i in range(0, len(cspecs), 32):
            res = T.batch(cspecs[i:i + 32], concurrency=8)
            cut += [r.threshold if not isinstance(r, Exception) else None for r in res]
            log(f"cutoffs {len(cut)}/{len(cspecs)} (k={k}) after {time.perf_counter() - t0:.1f}s")
        if cache:
            with open(cache, "w") as f:
                json.dump({"key": cache_key, "cutoffs": cut}, f)
    return [dict(mode="search", query=q["query"], m=q["m"], metric=q["metric"], w=q["w"],
                 arg=q["arg"], cutoff=c) for q, c in zip(queries, cut) if c is not None]


def _pct(v, p):
    return round(float(np.percentile(v, p)), 3) if len(v) else None


def run_batch(args):
    import fsmeld_amd as F
    from fsmeld_amd import workloads
    cfg = workloads.batching()
    T = F.Target(cfg["target"])
    n = len(cfg["target"])
    specs = _specs_with_cutoffs(T, cfg["queries"], cfg["k"], args.cutoff_cache,
                                f"batching n={n} q={len(cfg['queries'])} k={cfg['k']}")
    def gap():
        if args.phase_gaps:
            time.sleep(0.25)

    # warm every window length's statistics path once (not timed)
    T.batch(specs[:16], concurrency=4)

    def run_plain(tag):
        gap()
        t0 = time.perf_counter()
        out = []
        for i, sp in enumerate(specs):
            out.append(T.query("search", sp["query"], sp["m"], sp["metric"], sp["w"],
                               sp["arg"], cutoff_value=sp["cutoff"]))
            if i % 32 == 31:
                log(f"plain ({tag}) {i + 1}/{len(specs)} after {time.perf_counter() - t0:.1f}s")
        el = time.perf_counter() - t0
        log(f"plain ({tag}): {len(specs)} queries in {el:.2f}s")
        return out, el

    plain0, plain0_s = run_plain("built-in shapes")
    tuning, tune_s = None, 0.0
    if args.autotune < 0:
        args.autotune = 32
    if args.autotune > 0:
        # the autotuner on an evenly spread sample of the batch (not timed
        # with the batch); plain and batched runs below both use its shapes
        step = max(1, len(specs) // args.autotune)
        gap()
        t0 = time.perf_counter()
        tuning = T.autotune(specs[::step][:args.autotune], repeats=2)
        tune_s = time.perf_counter() - t0
        log(f"autotune ({len(specs[::step][:args.autotune])} searches) {tune_s:.1f}s: {tuning}")
        plain, plain_s = run_plain("tuned shapes")
    else:
        plain, plain_s = plain0, plain0_s
    for a, b in zip(plain, plain0):
        if not np.array_equal(a.indices, b.indices):
            raise SystemExit("tuned shapes changed a search's matches")
    plain_lat = [r.stats["ms_total"] for r in plain]
    modes = []
    identical = True
    for c in [int(x) for x in args.concurrency.split(",")]:
        gap()
        t0 = time.perf_counter()
        got = T.batch(specs, concurrency=c)
        el = time.perf_counter() - t0
        for g, p in zip(got, plain):
            if isinstance(g, Exception) or not np.array_equal(g.indices, p.indices) or \
                    not np.array_equal(g.distances.view(np.uint64), p.distances.view(np.uint64)):
                identical = False
        lat = [g.stats["ms_total"] for g in got if not isinstance(g, Exception)]
        modes.append(dict(concurrency=c, queries_per_s=round(len(specs) / el, 3),
                          speedup=round(plain_s / el, 3), p50_ms=_pct(lat, 50),
                          p99_ms=_pct(lat, 99)))
        log(f"batch concurrency {c}: {el:.2f}s")
    best = max(modes, key=lambda x: x["queries_per_s"])
    print(json.dumps({
        "metric": "queries/sec, batched vs plain one-at-a-time (config 4); p50/p99 latency",
        "value": best["queries_per_s"], "unit": "queries/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batching: 2^24 random walk, {len(specs)} searches, m log-uniform "
                               f"[64, 2048], w = 5% m, metric uniform over five, cutoffs k=10"},
        "plain": dict(queries_per_s=round(len(specs) / plain_s, 3), p50_ms=_pct(plain_lat, 50),
                      p99_ms=_pct(plain_lat, 99)),
        "plain_builtin_shapes": dict(
            queries_per_s=round(len(specs) / plain0_s, 3),
            p50_ms=_pct([r.stats["ms_total"] for r in plain0], 50),
            p99_ms=_pct([r.stats["ms_total"] for r in plain0], 99)),
        "autotune": dict(samples=args.autotune, seconds=round(tune_s, 2), shapes=tuning),
        "batched": modes, "results_identical_to_plain": identical}), flush=True)


def run_serve(args):
    import threading
    import fsmeld_amd as F
    from fsmeld_amd import workloads
    cfg = workloads.serving()
    T = F.Target(cfg["target"])
    n = len(cfg["target"])
    specs = _specs_with_cutoffs(T, cfg["queries"], cfg["k"], args.cutoff_cache,
                                f"serving n={n} q={len(cfg['queries'])} k={cfg['k']}")
    budget = 96 << 30
    rng = np.random.default_rng(7)
    biggest = max(specs, key=lambda sp: sp["m"])

    def serve(rate, workers):
        gaps = rng.exponential(1.0 / rate, args.requests)
        order = rng.integers(0, len(specs), args.requests)
        lat, done_t = [], []
        with F.Server(T, mem_budget=budget, max_workers=workers) as S:
            # warm-up (not recorded): one of the largest queries per worker,
            # so every worker's buffers exist before the arrivals start, as in
            # a server that has been up for a while
            for tk in [S.submit(biggest) for _ in range(workers)]:
                S.poll(tk, wait_ms=-1)
            tickets = []
            lock = threading.Lock()

            def collect():
                i = 0
                while True:
                    with lock:
                        if i >= len(tickets) and len(tickets) == args.requests:
                            return
                        tk = tickets[i] if i < len(tickets) else None
                    if tk is None:
                        time.sleep(0.0005)
                        continue
                    state, res, l = S.poll(tk, wait_ms=-1)
                    lat.append(l)
                    done_t.append(time.perf_counter())
                    i += 1

            th = threading.Thread(target=collect)
            th.start()
            t0 = time.perf_counter()
            due = t0
            for g, j in zip(gaps, order):
                due += g
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tk = S.submit(specs[int(j)])
                with lock:
                    tickets.append(tk)
            th.join()
            span = max(done_t) - t0
        return dict(offered_rate=rate, completed=len(lat), queries_per_s=round(len(lat) / span, 3),
                    p50_ms=_pct(lat, 50), p99_ms=_pct(lat, 99))

    rows = []
    for rate in [float(x) for x in args.rates.split(",")]:
        a = serve(rate, args.workers)
        b = serve(rate, 1)
        rows.append(dict(rate=rate, served=a, one_at_a_time=b))
        log(f"rate {rate}: served {a['queries_per_s']} q/s p99 {a['p99_ms']} ms; "
            f"one at a time {b['queries_per_s']} q/s p99 {b['p99_ms']} ms")
    print(json.dumps({
        "metric": "served queries/sec and p50/p99 latency under Poisson arrivals (config 5)",
        "value": max(r["served"]["queries_per_s"] for r in rows), "unit": "queries/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving: 2^26 target, {len(specs)} distinct searches, m "
                               f"log-uniform [64, 1024], w = 5% m, mixed metrics, cutoffs k=10; "
                               f"{args.requests} Poisson arrivals per rate; {args.workers} workers, "
                               f"{budget >> 30} GiB budget"},
        "rates": rows}), flush=True)


def run_small(args):
    """The plain small configuration end to end: a k = 10 cutoff query, then
    `search` at its threshold, on the GPU (steps timed after warmup) against
    the CPU restatement of the same search on 16 host threads (the whole
    target, no sampling); the two must return the same windows."""
    import fsmeld_amd as F
    import oracle as O
    from fsmeld_amd import workloads
    cfg = workloads.plain_small()
    t, q, m, w, metric, k = cfg["target"], cfg["query"], cfg["m"], cfg["w"], cfg["metric"], cfg["k"]
    T = F.Target(t)
    c = T.cutoff(q, m, k, metric, w, 0.0)
    thr = c.threshold
    for _ in range(max(args.warmup, 1)):
        T.search(q, m, thr, metric, w, 0.0)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps * 20):
        s0 = time.perf_counter()
        r = T.search(q, m, thr, metric, w, 0.0)
        lat.append((time.perf_counter() - s0) * 1000.0)
    gpu_s = (time.perf_counter() - t0) / (args.steps * 20)
    threads = 16
    c0 = time.perf_counter()
    cpu_idx, cst = O.search(t, q, m, thr, metric, w, 0.0, threads=threads)
    cpu_s = time.perf_counter() - c0
    same = bool(np.array_equal(np.asarray(cpu_idx), r.indices))
    log(f"small: gpu {gpu_s * 1e3:.2f} ms/search, cpu ({threads} threads) {cpu_s * 1e3:.1f} ms, "
        f"same windows: {same}")
    print(json.dumps({
        "metric": "plain small configuration: GPU search vs the CPU restatement on the host cores",
        "value": round(1.0 / gpu_s, 3), "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain small: 2^20 random walk, 10 constant runs of 300, 10 NaNs; "
                               f"m={m}, w={w}, {metric}, cutoff from k={k}", "n": int(t.size)},
        "p50_ms": round(_pct(lat, 50), 3), "p99_ms": round(_pct(lat, 99), 3),
        "cutoff": thr, "matches": int(r.indices.size),
        "cpu_baseline": {"value": round(1.0 / cpu_s, 4), "unit": "queries/s", "cores": threads,
                         "kind": "port", "sample": "the whole search on the host"},
        "speedup_vs_cpu": round(cpu_s / gpu_s, 1), "same_windows_as_cpu": same}), flush=True)


def _mfma_flops(n, m):
    """FP32 matrix-core flops one LB_IP launch issues in its contraction: per
    segment ceil(windows / 256) tiles of 2 sides x ceil((m + 15) / 4) steps of
    a 16x16x4 product (2048 flops)."""
    import fsmeld_amd as F
    G = F.geometry(n, m)
    tiles = G["forsize"] * -(-G["k_m_l"] // 256)
    if G["tail"]:
        tiles += -(-G["tail"] // 256)
    return tiles * 2 * ((m + 18) // 4) * 2048


def run_precision(args):
    """Config 3: LB_IP's sliding sums exact (64-bit fixed point) against the
    FP32 matrix-core contraction with certified slack, on the plain large
    target at m in {128, 1024}: LB error, pruning, LB_IP time, matrix-core
    rate, and identity of cutoff / search answers with each other and with
    the exact scan."""
    import fsmeld_amd as F
    from fsmeld_amd import workloads
    t = workloads.plain_large()["target"]
    T = F.Target(t)
    rows = []
    lengths = [int(x) for x in args.prec_m.split(",")]
    chosen = args.prec_metrics.split(",")
    for m in lengths:
        w = m // 20
        flops = _mfma_flops(t.size, m)
        for metric, arg in workloads.CONFIG2_METRICS:
            if metric not in chosen:
                continue
            q, _ = F.synth_query(3 * m + 1, t, m, 0.1)
            c = T.cutoff(q, m, 10, metric, w, arg)
            c32 = T.cutoff(q, m, 10, metric, w, arg, lbip_f32=True)
            s = T.search(q, m, c.threshold, metric, w, arg)
            e = T.exact(q, m, c.threshold, metric, w, arg)
            extra = np.setdiff1d(s.indices, e.indices).size
            missing = np.setdiff1d(e.indices, s.indices)
            at_cut = bool(np.all(e.distances[np.isin(e.indices, missing)] == c.threshold))
            ms = {False: [], True: []}
            runs = {}
            for rep in range(3):
                for f32 in (False, True):
                    r = T.search(q, m, c.threshold, metric, w, arg, lbip_f32=f32)
                    ms[f32].append(r.stats["ms_lbip"])
                    runs[f32] = r
            s32 = runs[True]
            lb, _, _ = T.lbip(q, m, metric, w, arg)
            lb32, _, _ = T.lbip(q, m, metric, w, arg, lbip_f32=True)
            fin = np.isfinite(lb)
            za, zb = lb32[fin], lb[fin]
            if metric in ("dtw", "erp"):
                za, zb = za * za, zb * zb
            rel = (zb - za) / np.maximum(zb, 1e-300)
            win = np.flatnonzero(fin)[::65537]
            d = T.eld(q, m, metric, w, arg, win)
            viol = int(np.sum(lb[win] > d + 1e-9 * np.maximum(1.0, d)))
            viol32 = int(np.sum(lb32[win] > d + 1e-9 * np.maximum(1.0, d)))
            st, st32 = s.stats, s32.stats
            t64, t32 = float(np.median(ms[False])), float(np.median(ms[True]))
            rows.append(dict(
                m=m, w=w, metric=metric, cutoff=c.threshold,
                lbip_pruned_frac=round(1.0 - st["after_lbip"] / st["valid"], 6),
                lbip_f32_pruned_frac=round(1.0 - st32["after_lbip"] / st32["valid"], 6),
                glb_pruned_frac=round(1.0 - st["after_glb"] / st["valid"], 6),
                matches=int(st["matches"]),
                search_equals_exact=(extra == 0 and (missing.size == 0 or at_cut)),
                f32_equals_exact_path=bool(
                    np.array_equal(s32.indices, s.indices)
                    and np.array_equal(s32.distances, s.distances)
                    and c32.threshold == c.threshold and np.array_equal(c32.indices, c.indices)),
                f32_lb_rel_below_exact_max=float(rel.max()),
                f32_lb_rel_below_exact_mean=float(rel.mean()),
                f32_lb_above_exact=int(np.sum(rel < -1e-12)),
                lb_above_eld=viol, f32_lb_above_eld=viol32, lb_checked=int(win.size),
                lbip_ms=round(t64, 3), lbip_f32_ms=round(t32, 3),
                mfma_tflops=round(flops / t32 / 1e9, 2),
                mfma_frac=round(flops / t32 / 1e9 / 157.3, 4),
                search_ms=round(st["ms_total"], 2), search_f32_ms=round(st32["ms_total"], 2)))
            log(f"m={m} {metric}: {rows[-1]}")
    print(json.dumps({
        "metric": "LB_IP precision at long queries (config 3): exact fixed-point sums against "
                  "the FP32 matrix-core contraction with certified slack",
        "value": float(np.mean([r["lbip_f32_pruned_frac"] for r in rows])),
        "unit": "fraction pruned by the FP32 bound",
        "n_gpus": 1, "higher_is_better": True,
        "dtype": "LB_IP sums: exact 64-bit fixed point (f64 penalties) and FP32 MFMA "
                 "(v_mfma_f32_16x16x4_f32); GLB / ELD f64",
        "data": "synthetic",
        "config": {"workload": "precision: plain large 2^26 target, m in {128, 1024}, w = 5% m, "
                               "five metrics, cutoffs k=10"},
        "mfma_peak_tflops": 157.3,
        "rows": rows}), flush=True)


if __name__ == "__main__":
    main()
