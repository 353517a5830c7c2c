#!/usr/bin/env python3
"""Writes tests/golden/eld_ref_vectors.json: exact elastic distances computed
by the reference's own src/ElasticMeasures.cpp (compiled from the reference
tree into oracle/_ref/libref_eld.so by oracle/Makefile) on seeded inputs.

Inputs are stored with 6 decimals (exactly reproducible by parsing), outputs
as float.hex so tests compare bits.  Run from the repository root after
`make -C oracle`."""
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

ARGS = {1: [0.0], 3: [0.0, 0.7], 2: [0.0, 0.25, 1.5], 4: [0.0, 0.25, 1.5], 5: [0.0, 0.5, 3.0],
        6: [0.0, 0.25, 1.5]}


def main():
    ref = oracle.ref_eld_lib()
    if ref is None:
        sys.exit("oracle/_ref/libref_eld.so not built (needs the reference tree)")
    rng = np.random.default_rng(20261203)
    P = ctypes.POINTER(ctypes.c_double)
    cases = []
    for metric in (1, 2, 3, 4, 5, 6):   # 6: SWALE (appended: earlier cases unchanged)
        for arg in ARGS[metric]:
            for _ in range(12):
                m = int(rng.integers(2, 25))
                w = int(rng.integers(0, m))
                q = np.round(rng.normal(size=m), 6)
                kind = rng.integers(0, 3)
                if kind == 0:
                    t = np.round(rng.normal(size=m), 6)
                elif kind == 1:
                    t = np.round(q + rng.normal(size=m) * 0.05, 6)
                else:
                    t = np.round(rng.normal(size=m), 1)
                d = ref.ref_eld(metric, q.ctypes.data_as(P), t.ctypes.data_as(P), m, w, arg)
                cases.append(dict(metric=metric, m=m, w=w, arg=arg, q=q.tolist(), t=t.tolist(),
                                  d=float(d).hex()))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eld_ref_vectors.json")
    with open(out, "w") as f:
        json.dump({"source": "reference src/ElasticMeasures.cpp via oracle/_ref",
                   "cases": cases}, f, separators=(",", ":"))
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
