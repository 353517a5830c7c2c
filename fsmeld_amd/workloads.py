"""Seeded synthetic workloads of the five benchmark configurations.

Each builder returns plain numpy arrays; the generator is the engine's own
``fsm_synth_target`` / ``fsm_synth_query`` (SplitMix64-seeded xoshiro256**,
Box-Muller), so the same seed gives the same data on every machine.
"""
from __future__ import annotations

import math

import numpy as np

from . import synth_query, synth_target

# metric -> arg for the plain large configuration (ERP g=0, EDR/LCSS eps=0.25, MSM c=0.5)
CONFIG2_METRICS = [("dtw", 0.0), ("erp", 0.0), ("edr", 0.25), ("lcss", 0.25), ("msm", 0.5)]


def plain_small(seed: int = 1, n: int = 1 << 20):
    """2^20 random walk, 10 constant runs of 300 samples, 10 single NaNs;
    m=128 query copied from a seeded offset plus N(0, 0.1) noise; DTW w=6."""
    t = synth_target(seed, n, 0.0, 0.0, n_const_runs=10, const_len=300, n_nans=10)
    q, off = synth_query(seed, t, 128, 0.1)
    return dict(target=t, query=q, offset=off, m=128, w=6, metric="dtw", arg=0.0, k=10)


def plain_large(seed: int = 2, n: int = 1 << 26):
    """2^26 random walk with 0.1% N(0,5) spikes and 1000 constant runs of 300
    samples; m=512, w=25, all five metrics; cutoffs from k=100."""
    t = synth_target(seed, n, 0.001, 5.0, n_const_runs=1000, const_len=300, n_nans=0)
    q, off = synth_query(seed, t, 512, 0.1)
    return dict(target=t, query=q, offset=off, m=512, w=25, metrics=CONFIG2_METRICS, k=100)


def batch_queries(target: np.ndarray, count: int = 256, seed: int = 4, m_lo: int = 64,
                  m_hi: int = 2048):
    """Queries with m log-uniform in [m_lo, m_hi], w = floor(0.05 m), metric
    uniform over the five (args of the plain large configuration)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m = int(round(math.exp(rng.uniform(math.log(m_lo), math.log(m_hi)))))
        metric, arg = CONFIG2_METRICS[int(rng.integers(0, 5))]
        q, off = synth_query(seed * 100003 + i, target, m, 0.1)
        out.append(dict(query=q, offset=off, m=m, w=int(0.05 * m), metric=metric, arg=arg))
    return out


def batching(seed: int = 4, n: int = 1 << 24, count: int = 256):
    """2^24 random walk; `count` queries with m log-uniform in [64, 2048],
    w = floor(0.05 m), metric uniform over the five; cutoffs from k=10."""
    t = synth_target(seed, n, 0.0, 0.0, n_const_runs=0, const_len=0, n_nans=0)
    return dict(target=t, queries=batch_queries(t, count, seed, 64, 2048), k=10)


def serving(seed: int = 5, n: int = 1 << 26, distinct: int = 64):
    """The plain large configuration's 2^26 target; `distinct` queries with m
    log-uniform in [64, 1024], w = floor(0.05 m), mixed metrics; cutoffs from
    k=10.  Arrivals are drawn by the caller (Poisson, rate lambda)."""
    big = plain_large(n=n)
    t = big["target"]
    return dict(target=t, queries=batch_queries(t, distinct, seed, 64, 1024), k=10)
