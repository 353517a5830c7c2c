"""The C ABI library loads on a CPU-only machine, exports every entry point
include/fsmeld.h declares, and its host-only parts behave without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np

import fsmeld_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fsmeld.h")).read()
    return sorted(set(re.findall(r"\b(fsm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 18
    lib = ctypes.CDLL(F.lib_path)
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", F.lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (fsm_\w+)", out))
    assert set(syms) <= exported


def test_no_device_paths_fail_cleanly():
    if F.device_count() > 0:
        return
    t = np.arange(100.0)
    try:
        F.Target(t)
        raise AssertionError("expected FsmeldError")
    except F.FsmeldError as e:
        assert e.code == -5


def test_synth_is_seeded_and_shaped():
    a = F.synth_target(5, 10000, 0.001, 5.0, 3, 300, 4)
    b = F.synth_target(5, 10000, 0.001, 5.0, 3, 300, 4)
    c = F.synth_target(6, 10000, 0.001, 5.0, 3, 300, 4)
    assert np.array_equal(a, b, equal_nan=True) and not np.array_equal(a, c, equal_nan=True)
    assert 1 <= np.isnan(a).sum() <= 4
    runs = np.sum(np.diff(a) == 0)
    assert runs >= 3 * 299 - 300
    # random walk of N(0,1) increments: step std close to 1 away from spikes/runs
    steps = np.diff(F.synth_target(1, 200000))
    assert abs(steps.std() - 1.0) < 0.01 and abs(steps.mean()) < 0.01
    q, off = F.synth_query(5, a, 64, 0.0)
    assert np.array_equal(q, a[off:off + 64])
    assert np.isfinite(a[off:off + 64]).all() and np.all(np.diff(a[off:off + 64]) != 0)


def test_ctypes_structs_match_the_header(tmp_path):
    """The binding's descriptor and statistics structs have the C header's
    size and field offsets (gcc on include/fsmeld.h)."""
    structs = {"fsm_query_desc": F._Desc, "fsm_stats": F._Stats, "fsm_tuning": F._Tuning}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "fsmeld.h"',
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    got = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                             check=True).stdout.splitlines():
        cname, fname, val = ln.split()
        got[(cname, fname)] = int(val)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
