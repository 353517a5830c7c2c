"""fsmeld_amd -- FSMELD's exact z-normalised subsequence search on MI355X.

The engine is the HIP library ``lib/libfsmeld_hip.so`` (C ABI in
``include/fsmeld.h``); this module is a thin ctypes layer over it that mirrors
the reference's query-level entry points (include/funcs.h:64-72):

* :func:`search`  -- ``FSMELD``: every window whose exact elastic distance to
  the query is <= cutoff (LB_IP -> GLB -> ELD cascade);
* :func:`cutoff`  -- ``FSMELD_find_threshold``: the k-th smallest exact
  distance and the k best windows in (distance, index) order;
* :func:`exact`   -- ``FSMELD_exact``: the brute-force scan.

Importing the package loads the library and raises if it is missing: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "METRICS", "Target", "QueryResult", "FsmeldError", "search", "cutoff", "exact",
    "synth_target", "synth_query", "device_count", "lib_path", "query_file",
]

METRICS = {"dtw": 1, "edr": 2, "erp": 3, "lcss": 4, "msm": 5, "swale": 6}
MODES = {"search": 0, "cutoff": 1, "exact": 2, "exact_topk": 3}

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "lib", "libfsmeld_hip.so")

if not os.path.exists(lib_path):
    raise ImportError(
        f"fsmeld_amd: engine library not built ({lib_path}); run __graft_entry__.build()")

_lib = ctypes.CDLL(lib_path)


class FsmeldError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class _Desc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("metric", ctypes.c_int), ("m", ctypes.c_int),
                ("w", ctypes.c_int), ("arg", ctypes.c_double), ("cutoff", ctypes.c_double),
                ("k", ctypes.c_int), ("exclude_begin", ctypes.c_int64),
                ("exclude_end", ctypes.c_int64), ("raw", ctypes.c_int),
                ("lbip_f32", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("valid", ctypes.c_uint64), ("invalid", ctypes.c_uint64),
                ("after_lbip", ctypes.c_uint64), ("after_glb", ctypes.c_uint64),
                ("exact_evals", ctypes.c_uint64), ("matches", ctypes.c_uint64),
                ("threshold", ctypes.c_double), ("ts_n", ctypes.c_int64),
                ("k_m_l", ctypes.c_int64), ("forsize", ctypes.c_int64), ("tail", ctypes.c_int64),
                ("m_k", ctypes.c_int), ("ms_stats", ctypes.c_double),
                ("ms_lbip", ctypes.c_double), ("ms_glb", ctypes.c_double),
                ("ms_eld", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("eld_cells", ctypes.c_uint64), ("ms_stats_finish", ctypes.c_double),
                ("ms_env", ctypes.c_double), ("near_rechecked", ctypes.c_uint64),
                ("near_changed", ctypes.c_uint64), ("lbip_f32", ctypes.c_int)]


class _Tuning(ctypes.Structure):
    _fields_ = [("lbip_threads", ctypes.c_int * 4), ("eld_refill", ctypes.c_int * 7),
                ("glb_blocks", ctypes.c_int)]


_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_VP = ctypes.c_void_p


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("fsm_device_count", ctypes.c_int, [])
_sig("fsm_last_error", ctypes.c_char_p, [])
_sig("fsm_target_open", ctypes.c_int, [_D, ctypes.c_int64, ctypes.POINTER(_VP)])
_sig("fsm_target_open_file", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP)])
_sig("fsm_target_length", ctypes.c_int64, [_VP])
_sig("fsm_target_close", None, [_VP])
_sig("fsm_query", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, ctypes.c_int64,
                                 ctypes.POINTER(_VP)])
_sig("fsm_result_stats", ctypes.c_int, [_VP, ctypes.POINTER(_Stats)])
_sig("fsm_result_count", ctypes.c_int64, [_VP])
_sig("fsm_result_indices", _I64, [_VP])
_sig("fsm_result_distances", _D, [_VP])
_sig("fsm_result_free", None, [_VP])
_sig("fsm_stage_window_stats", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int64,
                                              ctypes.c_int64, _D, _D, _U8])
_sig("fsm_stage_lbip", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, _D, _D, _D])
_sig("fsm_stage_glb", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, _I64, ctypes.c_int64, _D])
_sig("fsm_stage_glb_filter", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, _I64, ctypes.c_int64,
                                            _U8])
_sig("fsm_stage_eld", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, _I64, ctypes.c_int64, _D])
_sig("fsm_eld_pairs", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     _D, _D, ctypes.c_int64, _D])
# device-memory stage entries: pointers and the stream as integer addresses
_sig("fsm_stage_window_stats_dev", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int64,
                                                  ctypes.c_int64, _VP, _VP, _VP, _VP])
_sig("fsm_stage_lbip_dev", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, _VP, _VP, _VP, _VP])
for _n in ("fsm_stage_glb_dev", "fsm_stage_glb_filter_dev", "fsm_stage_eld_dev"):
    _sig(_n, ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, _VP, ctypes.c_int64, _VP, _VP])
_sig("fsm_target_share", ctypes.c_int, [_VP, ctypes.POINTER(_VP)])
_sig("fsm_target_trim", None, [_VP])
_sig("fsm_query_file", ctypes.c_int, [ctypes.c_char_p, _VP, _D, ctypes.c_int64, ctypes.c_uint64,
                                      ctypes.POINTER(_VP)])
_sig("fsm_target_reuse_stats", ctypes.c_int, [_VP, ctypes.c_int])
_sig("fsm_target_cache_limit", ctypes.c_int, [_VP, ctypes.c_uint64])
_sig("fsm_query_bytes", ctypes.c_uint64, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int])
_sig("fsm_query_batch", ctypes.c_int, [_VP, ctypes.c_int64, ctypes.POINTER(_Desc),
                                       ctypes.POINTER(_D), _I64, ctypes.c_int,
                                       ctypes.POINTER(_VP), ctypes.POINTER(ctypes.c_int)])
_sig("fsm_tuning_get", ctypes.c_int, [_VP, ctypes.POINTER(_Tuning)])
_sig("fsm_tuning_set", ctypes.c_int, [_VP, ctypes.POINTER(_Tuning)])
_sig("fsm_autotune", ctypes.c_int, [_VP, ctypes.c_int64, ctypes.POINTER(_Desc),
                                    ctypes.POINTER(_D), _I64, ctypes.c_int,
                                    ctypes.POINTER(_Tuning)])
_sig("fsm_server_create", ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_VP)])
_sig("fsm_server_submit", ctypes.c_int, [_VP, ctypes.POINTER(_Desc), _D, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_uint64)])
_sig("fsm_server_poll", ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_VP),
                                       ctypes.POINTER(ctypes.c_int), _D])
_sig("fsm_server_cancel", ctypes.c_int, [_VP, ctypes.c_uint64])
_sig("fsm_server_info", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), _I64, _I64])
_sig("fsm_server_destroy", None, [_VP])
_sig("fsm_synth_target", ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, _D])
_sig("fsm_synth_query", ctypes.c_int, [ctypes.c_uint64, _D, ctypes.c_int64, ctypes.c_int,
                                       ctypes.c_double, _D, _I64])


def _ptr(a: np.ndarray, t=_D):
    return a.ctypes.data_as(t)


def _check(rc: int):
    if rc != 0:
        raise FsmeldError(rc, (_lib.fsm_last_error() or b"").decode(errors="replace"))


def device_count() -> int:
    return int(_lib.fsm_device_count())


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


@dataclass
class QueryResult:
    """Indices (ascending for search/exact, rank order for cutoff), their exact
    distances and the engine's counters (kk1/kk2/kk3 of src/fsmeld.cpp:981-996)."""
    indices: np.ndarray
    distances: np.ndarray
    stats: dict = field(default_factory=dict)

    @property
    def threshold(self) -> float:
        return self.stats.get("threshold", float("nan"))


FSM_PENDING = 1
FSM_ERR_CANCELLED = -7


def _take_result(r) -> QueryResult:
    """Copies a library result object into a QueryResult and frees it."""
    try:
        st = _Stats()
        _check(_lib.fsm_result_stats(r, ctypes.byref(st)))
        n = int(_lib.fsm_result_count(r))
        idx = np.ctypeslib.as_array(_lib.fsm_result_indices(r), shape=(n,)).copy() if n else \
            np.zeros(0, np.int64)
        dist = np.ctypeslib.as_array(_lib.fsm_result_distances(r), shape=(n,)).copy() if n \
            else np.zeros(0, np.float64)
        stats = {name: getattr(st, name) for name, _ in _Stats._fields_}
    finally:
        _lib.fsm_result_free(r)
    return QueryResult(idx, dist, stats)


def _spec_desc(spec) -> _Desc:
    """Descriptor of a query spec: dict(mode, m, metric, w, arg, cutoff | k, exclude, raw,
    lbip_f32)."""
    return _desc(spec["mode"], spec["metric"], spec["m"], spec["w"], spec.get("arg", 0.0),
                 spec.get("cutoff", 0.0), spec.get("k", 0), spec.get("exclude"),
                 spec.get("raw", False), spec.get("lbip_f32", False))


def query_bytes(n, metric, m, w) -> int:
    """Device bytes one handle holds for a query of (metric, m, w) over n samples."""
    return int(_lib.fsm_query_bytes(int(n), METRICS[metric] if isinstance(metric, str)
                                    else int(metric), int(m), int(w)))


def _desc(mode, metric, m, w, arg, cutoff_value=0.0, k=0, exclude=None, raw=False,
          lbip_f32=False) -> _Desc:
    d = _Desc()
    d.raw = 1 if raw else 0
    d.lbip_f32 = 1 if lbip_f32 else 0
    d.mode = MODES[mode]
    d.metric = METRICS[metric] if isinstance(metric, str) else int(metric)
    d.m, d.w, d.arg = int(m), int(w), float(arg)
    d.cutoff, d.k = float(cutoff_value), int(k)
    if exclude is None:
        d.exclude_begin = d.exclude_end = -1
    else:
        d.exclude_begin, d.exclude_end = int(exclude[0]), int(exclude[1])
    return d


class Target:
    """A target series resident in GPU memory (HBM) for repeated queries."""

    def __init__(self, values=None, path: Optional[str] = None):
        self._h = _VP()
        if path is not None:
            _check(_lib.fsm_target_open_file(os.fsencode(path), ctypes.byref(self._h)))
            self.values = None
        else:
            v = _f64(values)
            self.values = v
            _check(_lib.fsm_target_open(_ptr(v), int(v.size), ctypes.byref(self._h)))

    def __len__(self):
        return int(_lib.fsm_target_length(self._h))

    def close(self):
        if self._h:
            _lib.fsm_target_close(self._h)
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- query-level ----
    def query(self, mode, query, m, metric, w, arg, cutoff_value=0.0, k=0,
              exclude=None, raw=False, lbip_f32=False) -> QueryResult:
        """raw=True: raw (non-normalised) matching, the reference's
        NormalizationMode::Raw (paper App. D).  lbip_f32=True: LB_IP's sliding
        sums from the FP32 matrix-core contraction with certified slack."""
        q = _f64(query)
        d = _desc(mode, metric, m, w, arg, cutoff_value, k, exclude, raw, lbip_f32)
        r = _VP()
        _check(_lib.fsm_query(self._h, ctypes.byref(d), _ptr(q), int(q.size), ctypes.byref(r)))
        return _take_result(r)

    def batch(self, specs, concurrency=4):
        """Runs every spec (dict(mode, query, m, metric, w, arg, cutoff | k,
        exclude)) with up to `concurrency` queries in flight; returns one
        QueryResult or FsmeldError per spec, in order."""
        n = len(specs)
        qs = [_f64(sp["query"]) for sp in specs]
        descs = (_Desc * max(n, 1))(*[_spec_desc(sp) for sp in specs])
        qptr = (_D * max(n, 1))(*[_ptr(q) for q in qs])
        lens = np.array([q.size for q in qs] or [0], dtype=np.int64)
        res = (_VP * max(n, 1))()
        status = (ctypes.c_int * max(n, 1))()
        _check(_lib.fsm_query_batch(self._h, n, descs, qptr, _ptr(lens, _I64), int(concurrency),
                                    res, status))
        out = []
        for i in range(n):
            if status[i] == 0:
                out.append(_take_result(res[i]))
            else:
                out.append(FsmeldError(status[i], "query failed"))
        return out

    @staticmethod
    def _tuning_dict(tu) -> dict:
        return dict(lbip_threads=list(tu.lbip_threads), eld_refill=list(tu.eld_refill),
                    glb_blocks=int(tu.glb_blocks))

    def tuning(self) -> dict:
        """The target's launch shapes (fsm_tuning_get); 0 = built-in rule."""
        tu = _Tuning()
        _check(_lib.fsm_tuning_get(self._h, ctypes.byref(tu)))
        return self._tuning_dict(tu)

    def set_tuning(self, tuning: dict):
        """Sets the target's launch shapes (fsm_tuning_set)."""
        tu = _Tuning()
        for i, v in enumerate(tuning.get("lbip_threads", [])):
            tu.lbip_threads[i] = int(v)
        for i, v in enumerate(tuning.get("eld_refill", [])):
            tu.eld_refill[i] = int(v)
        tu.glb_blocks = int(tuning.get("glb_blocks", 0))
        _check(_lib.fsm_tuning_set(self._h, ctypes.byref(tu)))

    def autotune(self, specs, repeats=2) -> dict:
        """Times the sample searches (dicts as for batch(), mode "search")
        under each candidate launch shape, applies the fastest and returns
        them (fsm_autotune)."""
        n = len(specs)
        qs = [_f64(sp["query"]) for sp in specs]
        descs = (_Desc * max(n, 1))(*[_spec_desc(sp) for sp in specs])
        qptr = (_D * max(n, 1))(*[_ptr(q) for q in qs])
        lens = np.array([q.size for q in qs] or [0], dtype=np.int64)
        tu = _Tuning()
        _check(_lib.fsm_autotune(self._h, n, descs, qptr, _ptr(lens, _I64), int(repeats),
                                 ctypes.byref(tu)))
        return self._tuning_dict(tu)

    def trim(self):
        """Frees this handle's working buffers and the target's cached window
        statistics / envelopes that no handle is using (fsm_target_trim)."""
        _lib.fsm_target_trim(self._h)

    def reuse_stats(self, enable: bool):
        """False: every query recomputes its window statistics and envelope,
        as the reference does per query (fsm_target_reuse_stats)."""
        _check(_lib.fsm_target_reuse_stats(self._h, 1 if enable else 0))

    def cache_limit(self, nbytes: int):
        """Caps the target's shared statistics / envelope cache (bytes)."""
        _check(_lib.fsm_target_cache_limit(self._h, int(nbytes)))

    def share(self) -> "Target":
        """Another handle to the same device samples (own stream and buffers)."""
        t = Target.__new__(Target)
        t._h = _VP()
        t.values = self.values
        _check(_lib.fsm_target_share(self._h, ctypes.byref(t._h)))
        return t

    def search(self, query, m, cutoff_value, metric, w, arg, exclude=None,
               raw=False, lbip_f32=False) -> QueryResult:
        return self.query("search", query, m, metric, w, arg, cutoff_value=cutoff_value,
                          exclude=exclude, raw=raw, lbip_f32=lbip_f32)

    def cutoff(self, query, m, k, metric, w, arg, exclude=None, raw=False,
               lbip_f32=False) -> QueryResult:
        return self.query("cutoff", query, m, metric, w, arg, k=k, exclude=exclude, raw=raw,
                          lbip_f32=lbip_f32)

    def exact_topk(self, query, m, k, metric, w, arg, exclude=None, raw=False) -> QueryResult:
        """FSMELD_exact with top_k = k: every window at or below the k-th
        smallest exact distance (stats["threshold"]), ascending, by brute force."""
        return self.query("exact_topk", query, m, metric, w, arg, 0.0, k, exclude, raw)

    def exact(self, query, m, cutoff_value, metric, w, arg, exclude=None,
              raw=False) -> QueryResult:
        return self.query("exact", query, m, metric, w, arg, cutoff_value=cutoff_value,
                          exclude=exclude, raw=raw)

    # ---- stage-level (parity tests) ----
    def window_stats(self, m, exclude=None):
        n = len(self) - m + 1
        mu, sg = np.empty(n), np.empty(n)
        inv = np.empty(n, np.uint8)
        xb, xe = (-1, -1) if exclude is None else exclude
        _check(_lib.fsm_stage_window_stats(self._h, int(m), int(xb), int(xe), _ptr(mu),
                                           _ptr(sg), _ptr(inv, _U8)))
        return mu, sg, inv.astype(bool)

    def lbip(self, query, m, metric, w, arg, raw=False, lbip_f32=False):
        q = _f64(query)
        d = _desc("search", metric, m, w, arg, raw=raw, lbip_f32=lbip_f32)
        n = len(self) - m + 1
        lb = np.empty(n)
        seg = _geometry(len(self), m)["S"]
        tu, tl = np.empty(seg), np.empty(seg)
        _check(_lib.fsm_stage_lbip(self._h, ctypes.byref(d), _ptr(q), _ptr(lb), _ptr(tu),
                                   _ptr(tl)))
        return lb, tu, tl

    def glb(self, query, m, metric, w, arg, windows, bsf):
        q = _f64(query)
        win = np.ascontiguousarray(np.asarray(windows, dtype=np.int64))
        d = _desc("search", metric, m, w, arg, cutoff_value=bsf)
        out = np.empty(win.size)
        _check(_lib.fsm_stage_glb(self._h, ctypes.byref(d), _ptr(q), _ptr(win, _I64),
                                  int(win.size), _ptr(out)))
        return out

    def glb_keep(self, query, m, metric, w, arg, windows, bsf):
        """GLB keep decision (``glb <= bsf``) per window from the search filter kernel."""
        q = _f64(query)
        win = np.ascontiguousarray(np.asarray(windows, dtype=np.int64))
        d = _desc("search", metric, m, w, arg, cutoff_value=bsf)
        out = np.empty(win.size, dtype=np.uint8)
        _check(_lib.fsm_stage_glb_filter(self._h, ctypes.byref(d), _ptr(q), _ptr(win, _I64),
                                         int(win.size), _ptr(out, _U8)))
        return out.astype(bool)

    def eld(self, query, m, metric, w, arg, windows):
        q = _f64(query)
        win = np.ascontiguousarray(np.asarray(windows, dtype=np.int64))
        d = _desc("search", metric, m, w, arg)
        out = np.empty(win.size)
        _check(_lib.fsm_stage_eld(self._h, ctypes.byref(d), _ptr(q), _ptr(win, _I64),
                                  int(win.size), _ptr(out)))
        return out


    # ---- the stages on device memory: device addresses (ints) in, outputs
    # written in place on the device, ordered on `stream` (a hipStream_t
    # address; 0: the null stream) ----
    def window_stats_dev(self, m, mu, sigma, invalid, stream=0, exclude=None):
        xb, xe = (-1, -1) if exclude is None else exclude
        _check(_lib.fsm_stage_window_stats_dev(self._h, int(m), int(xb), int(xe), mu, sigma,
                                               invalid, stream))

    def lbip_dev(self, query, m, metric, w, arg, lb, t_u=None, t_l=None, stream=0, raw=False):
        q = _f64(query)
        d = _desc("search", metric, m, w, arg, raw=raw)
        _check(_lib.fsm_stage_lbip_dev(self._h, ctypes.byref(d), _ptr(q), lb, t_u, t_l, stream))

    def _verify_dev(self, fn, query, m, metric, w, arg, windows, n, out, stream, bsf=None):
        q = _f64(query)
        d = _desc("search", metric, m, w, arg, cutoff_value=0.0 if bsf is None else bsf)
        _check(fn(self._h, ctypes.byref(d), _ptr(q), windows, int(n), out, stream))

    def glb_dev(self, query, m, metric, w, arg, windows, n, out, bsf, stream=0):
        self._verify_dev(_lib.fsm_stage_glb_dev, query, m, metric, w, arg, windows, n, out,
                         stream, bsf)

    def glb_keep_dev(self, query, m, metric, w, arg, windows, n, keep, bsf, stream=0):
        self._verify_dev(_lib.fsm_stage_glb_filter_dev, query, m, metric, w, arg, windows, n,
                         keep, stream, bsf)

    def eld_dev(self, query, m, metric, w, arg, windows, n, out, stream=0):
        self._verify_dev(_lib.fsm_stage_eld_dev, query, m, metric, w, arg, windows, n, out,
                         stream)


def eld_pairs(metric, q, cands, w, arg) -> np.ndarray:
    """Exact distance of each normalised row of ``cands`` to normalised ``q``."""
    q = _f64(q)
    c = _f64(cands).reshape(-1, q.size)
    out = np.empty(c.shape[0])
    m = METRICS[metric] if isinstance(metric, str) else int(metric)
    _check(_lib.fsm_eld_pairs(m, int(q.size), int(w), float(arg), _ptr(q), _ptr(c),
                              int(c.shape[0]), _ptr(out)))
    return out


def _geometry(N: int, m: int) -> dict:
    """Macro-segmentation of src/fsmeld.cpp:28-37, :647-661."""
    mk = 1
    while mk < 8 * m:
        mk *= 2
    mk = min(mk, N)
    kml = mk - m + 1
    ts_n = N - m + 1
    forsize, tail = divmod(ts_n, kml)
    return dict(m_k=mk, k_m_l=kml, ts_n=ts_n, forsize=forsize, tail=tail,
                tail_first=kml - tail, S=forsize + (1 if tail else 0))


geometry = _geometry


# ---- module-level wrappers named after the reference's entry points ----
def query_file(path, mode, query, m, metric, w, arg, cutoff_value=0.0, k=0, exclude=None,
               raw=False, device_budget=8 << 30) -> QueryResult:
    """A query over a target FILE with at most device_budget bytes of device
    memory (fsm_query_file): the file streams through the GPU in chunks of
    whole macro-segments; the result equals the whole-target query's."""
    q = _f64(query)
    d = _desc(mode, metric, m, w, arg, cutoff_value, k, exclude, raw)
    r = _VP()
    _check(_lib.fsm_query_file(os.fsencode(path), ctypes.byref(d), _ptr(q), int(q.size),
                               ctypes.c_uint64(int(device_budget)), ctypes.byref(r)))
    return _take_result(r)


def search(target, query, m, cutoff_value, metric, w, arg, exclude=None) -> QueryResult:
    """FSMELD (include/funcs.h:67-68)."""
    t = target if isinstance(target, Target) else Target(target)
    return t.search(query, m, cutoff_value, metric, w, arg, exclude)


def cutoff(target, query, m, k, metric, w, arg, exclude=None) -> QueryResult:
    """FSMELD_find_threshold (include/funcs.h:64-65)."""
    t = target if isinstance(target, Target) else Target(target)
    return t.cutoff(query, m, k, metric, w, arg, exclude)


def exact(target, query, m, cutoff_value, metric, w, arg, exclude=None) -> QueryResult:
    """FSMELD_exact with top_k = 0 (include/funcs.h:70-72)."""
    t = target if isinstance(target, Target) else Target(target)
    return t.exact(query, m, cutoff_value, metric, w, arg, exclude)


def synth_target(seed: int, n: int, spike_frac=0.0, spike_sd=0.0, n_const_runs=0,
                 const_len=300, n_nans=0) -> np.ndarray:
    out = np.empty(int(n))
    _check(_lib.fsm_synth_target(int(seed), int(n), float(spike_frac), float(spike_sd),
                                 int(n_const_runs), int(const_len), int(n_nans), _ptr(out)))
    return out


def synth_query(seed: int, target: np.ndarray, m: int, noise_sd=0.1):
    t = _f64(target)
    q = np.empty(int(m))
    off = ctypes.c_int64(0)
    _check(_lib.fsm_synth_query(int(seed), _ptr(t), int(t.size), int(m), float(noise_sd),
                                _ptr(q), ctypes.byref(off)))
    return q, int(off.value)


class Server:
    """Serving loop over a Target: submit / poll / cancel with memory-based
    admission and concurrent execution (fsm_server_* of include/fsmeld.h)."""

    def __init__(self, target: Target, mem_budget: int, max_workers: int = 4):
        self._target = target
        self._h = _VP()
        _check(_lib.fsm_server_create(target._h, int(mem_budget), int(max_workers),
                                      ctypes.byref(self._h)))
        self._queries = {}

    def submit(self, spec) -> int:
        q = _f64(spec["query"])
        d = _spec_desc(spec)
        ticket = ctypes.c_uint64()
        _check(_lib.fsm_server_submit(self._h, ctypes.byref(d), _ptr(q), int(q.size),
                                      ctypes.byref(ticket)))
        return int(ticket.value)

    def poll(self, ticket, wait_ms=0):
        """(state, result, latency_ms): state "pending", "done" (result is a
        QueryResult or FsmeldError), or "cancelled"."""
        r = _VP()
        status = ctypes.c_int()
        lat = ctypes.c_double()
        rc = _lib.fsm_server_poll(self._h, int(ticket), int(wait_ms), ctypes.byref(r),
                                  ctypes.byref(status), ctypes.byref(lat))
        if rc == FSM_PENDING:
            return "pending", None, None
        if rc == FSM_ERR_CANCELLED:
            return "cancelled", None, None
        _check(rc)
        if status.value != 0:
            return "done", FsmeldError(status.value, "query failed"), lat.value
        return "done", _take_result(r), lat.value

    def cancel(self, ticket):
        _check(_lib.fsm_server_cancel(self._h, int(ticket)))

    def info(self):
        w = ctypes.c_int()
        qd = np.zeros(1, np.int64)
        rn = np.zeros(1, np.int64)
        _check(_lib.fsm_server_info(self._h, ctypes.byref(w), _ptr(qd, _I64), _ptr(rn, _I64)))
        return dict(workers=w.value, queued=int(qd[0]), running=int(rn[0]))

    def close(self):
        if self._h:
            _lib.fsm_server_destroy(self._h)
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
