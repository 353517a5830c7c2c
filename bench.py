#!/usr/bin/env python3
"""Benchmark of the FSMELD cascade on one MI355X.

Workload (plain single query, large, all metrics): a 2^26-sample FP64 random
walk with 0.1% N(0,5) spikes and 1000 constant runs of 300 samples, held in
HBM; one m=512 query (copied from a seeded offset plus N(0,0.1) noise),
w=25, metrics DTW, ERP g=0, EDR eps=0.25, LCSS eps=0.25, MSM c=0.5, each at
the cutoff `cutoff` returns for k=100 (computed before timing).

One step = one `search` per metric (5 queries), each a full pass of the
LB_IP -> GLB -> ELD cascade over all 67,108,353 windows with results copied
back to the host.  value = queries per second over the timed steps.

Printed: one JSON line with the driver's fields plus
  candidates_verified_per_s  valid windows decided per second,
  p50_ms / p99_ms            per-query latency,
  roofline                   the LB_IP kernel (the cascade's streaming stage)
                             against HBM bandwidth, from HIP events,
  cpu_baseline               the CPU restatement (oracle/) on the host cores,
                             on a sample of macro-segments, scaled.
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--cpu-segments", type=int, default=256,
                    help="macro-segments per metric in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the engine is single-GPU; see DESIGN.md for how the path shards).",
              file=sys.stderr)
        sys.exit(2)

    import fsmeld_amd as F
    from fsmeld_amd import workloads

    n = 1 << args.log2n
    t0 = time.perf_counter()
    cfg = workloads.plain_large(n=n)
    t, q, m, w = cfg["target"], cfg["query"], cfg["m"], cfg["w"]
    log(f"workload generated in {time.perf_counter() - t0:.1f}s (n={n}, m={m}, w={w})")

    T = F.Target(t)
    cutoffs = {}
    for metric, arg in cfg["metrics"]:
        r = T.cutoff(q, m, cfg["k"], metric, w, arg)
        cutoffs[metric] = r.threshold
        log(f"cutoff {metric}: k={cfg['k']} threshold={r.threshold:.17g} "
            f"({r.stats['ms_total']:.1f} ms)")

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

    lat, stage, results = [], {"stats": 0.0, "lbip": 0.0, "glb": 0.0, "eld": 0.0}, []
    t_start = time.perf_counter()
    for i in range(args.steps):
        res = step()
        for metric, dt, r in res:
            lat.append(dt * 1000.0)
            for k in stage:
                stage[k] += r.stats["ms_" + k]
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

    # LB_IP kernel roofline: algorithmic bytes per launch = ts_n windows x
    # (8 * m_k / k_m_l + 8) B (target read with segment overlap + one FP64 bound out)
    lbip_ms = stage["lbip"] / nq
    bytes_per_launch = G["ts_n"] * (8.0 * G["m_k"] / G["k_m_l"] + 8.0)
    achieved = bytes_per_launch / (lbip_ms / 1000.0) / 1e9
    roofline = dict(kernel="k_lbip", bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None,
                    bytes_per_launch=int(bytes_per_launch), avg_launch_ms=round(lbip_ms, 4))

    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        seg = min(args.cpu_segments, G["S"])
        per_metric_s = []
        c0 = time.perf_counter()
        for metric, arg in cfg["metrics"]:
            s0 = time.perf_counter()
            O.search(t, q, m, cutoffs[metric], metric, w, arg, segments=(0, seg), threads=threads)
            dt = time.perf_counter() - s0
            per_metric_s.append(dt * G["S"] / seg)
            log(f"cpu baseline {metric}: {seg} segments in {dt:.2f}s")
        est = float(np.mean(per_metric_s))
        cpu = dict(value=round(1.0 / est, 6), unit="queries/s", cores=threads, kind="port",
                   sample=f"oracle search on macro-segments [0,{seg}) of {G['S']} per metric "
                          f"(5 metrics, k=100 cutoffs), OpenMP over segments, scaled to the "
                          f"full target; {time.perf_counter() - c0:.1f}s of CPU work",
                   est_seconds_per_query=round(est, 3))

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
                               f"from k={cfg['k']}; one step = 5 searches",
                   "n": n, "m": m, "w": w, "k": cfg["k"], "parallelism": "single GPU"},
        "candidates_verified_per_s": round(valid * nq / elapsed, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "stage_ms_per_query": {k: round(v / nq, 4) for k, v in stage.items()},
        "per_query": per_query,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
