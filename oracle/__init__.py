"""CPU oracle for FSMELD's LB_IP -> GLB -> ELD cascade -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (built from ``fsmeld_oracle.cpp`` by
``oracle/Makefile``), a restatement of the reference's algorithm that follows
its arithmetic order.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg use it, always as the checker.

Parity status per function (how each is pinned; see tests/test_oracle_*.py):
  eld (5 metrics)        -- bit-identical to the reference's own
                            src/ElasticMeasures.cpp (oracle/_ref build) and to
                            tests/golden/eld_ref_vectors.json; closed forms.
  mvmean / invalid mask  -- closed forms, two-pass brute force, NaN cases.
  min_max_window, lemire -- brute force on NaN-free input.
  thresholds (find_t_lp) -- brute-force objective on small inputs; NaN inputs:
                            reference behaviour undefined ("parity unpinned").
  sip                    -- direct sum vs numpy.correlate; FFTW's own bits are
                            not reproduced ("parity unpinned" at ~1e-12).
  glb                    -- closed forms, GLB <= ELD property.
  search/cutoff/exact    -- search == exact, cutoff == k-th of exact, index
                            coverage of the segmentation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_ELD = os.path.join(_HERE, "_ref", "libref_eld.so")

METRICS = {"dtw": 1, "edr": 2, "erp": 3, "lcss": 4, "msm": 5, "swale": 6}

_lib = ctypes.CDLL(LIB)
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I = ctypes.c_int
_L = ctypes.c_int64


class Stats(ctypes.Structure):
    _fields_ = [("valid", ctypes.c_uint64), ("invalid", ctypes.c_uint64),
                ("after_lbip", ctypes.c_uint64), ("after_glb", ctypes.c_uint64),
                ("exact_evals", ctypes.c_uint64), ("matches", ctypes.c_uint64),
                ("threshold", ctypes.c_double), ("ts_n", _L), ("k_m_l", _L),
                ("forsize", _L), ("tail", _L), ("m_k", _I)]


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args


_sig("orc_mvmean", None, [_D, _I, _I, _D, _D])
_sig("orc_invalid_mask", None, [_D, _D, _I, _I, _U8])
_sig("orc_min_max_window", None, [_D, _I, _I, _D, _D])
_sig("orc_lower_upper_lemire", None, [_D, _I, _I, _D, _D])
_sig("orc_normalized_arrays", None, [_D, _D, _D, _D, _I, _D, _D, _D, _D])
_sig("orc_thresholds", None, [_I, _D, _D, _I, _D, _D, _I, ctypes.c_double, _D, _D])
_sig("orc_q_to_h", None, [_I, _D, _I, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D])
_sig("orc_s_to_g", None, [_D, _D, _I, ctypes.c_double, ctypes.c_double, _D, _D])
_sig("orc_sip_direct", None, [_D, _I, _D, _I, _D])
_sig("orc_set_normalization_mode", None, [_I])
_sig("orc_trans", ctypes.c_double, [_I, ctypes.c_double, ctypes.c_double, _I])
_sig("orc_lbip_exceeds_cutoff", _I, [ctypes.c_double, ctypes.c_double, _I])
_sig("orc_glb", ctypes.c_double, [_I, _D, _D, _I, ctypes.c_double, _D, _D, _I, ctypes.c_double])
_sig("orc_eld", ctypes.c_double, [_I, _D, _D, _I, _I, ctypes.c_double])
_sig("orc_normalize_query", None, [_D, _I, _D])
_sig("orc_search", _I, [_D, _L, _D, _I, ctypes.c_double, _I, _I, ctypes.c_double, _L, _L, _L,
                        _L, _I, ctypes.POINTER(Stats), _I64, _L, _I64])
_sig("orc_exact", _I, [_D, _L, _D, _I, ctypes.c_double, _I, _I, ctypes.c_double, _L, _L, _I,
                       ctypes.POINTER(Stats), _I64, _D, _L, _I64])
_sig("orc_cutoff", _I, [_D, _L, _D, _I, _I, _I, _I, ctypes.c_double, _L, _L, _I,
                        ctypes.POINTER(Stats), _I64, _D])
_sig("orc_exact_topk", _I, [_D, _L, _D, _I, _I, _I, _I, ctypes.c_double, _L, _L, _I,
                            ctypes.POINTER(Stats), _I64, _D, _L, _I64, _D])
_sig("orc_window_stats", _I, [_D, _L, _I, _L, _L, _D, _D, _U8])
_sig("orc_lbip_all", _I, [_D, _L, _D, _I, _I, _I, ctypes.c_double, _D, _D, _D])


def _a(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _metric(x):
    return METRICS[x] if isinstance(x, str) else int(x)


def _stats(st):
    return {n: getattr(st, n) for n, _ in Stats._fields_}


def _excl(exclude):
    return (-1, -1) if exclude is None else (int(exclude[0]), int(exclude[1]))


# ---- stage functions ----
def mvmean(a, l):
    a = _a(a)
    n = a.size - l + 1
    mu, si = np.empty(n), np.empty(n)
    _lib.orc_mvmean(_p(a), a.size, l, _p(mu), _p(si))
    return mu, si


def invalid_mask(t, si, m):
    t, si = _a(t), _a(si)
    out = np.zeros(t.size, np.uint8)
    _lib.orc_invalid_mask(_p(t), _p(si), t.size, m, _p(out, _U8))
    return out[: t.size - m + 1].astype(bool)


def min_max_window(a, n, m):
    a = _a(a)
    mn, mx = np.empty(n), np.empty(n)
    _lib.orc_min_max_window(_p(a), n, m, _p(mn), _p(mx))
    return mn, mx


def lemire(a, r):
    a = _a(a)
    lo, hi = np.empty(a.size), np.empty(a.size)
    _lib.orc_lower_upper_lemire(_p(a), a.size, r, _p(lo), _p(hi))
    return lo, hi


def normalized_arrays(U, L, max_mean, min_mean, max_std, min_std):
    U, L = _a(U), _a(L)
    un, ln = np.empty(U.size), np.empty(U.size)
    _lib.orc_normalized_arrays(_p(U), _p(L), _p(un), _p(ln), U.size, _p(_a(max_mean)),
                               _p(_a(min_mean)), _p(_a(max_std)), _p(_a(min_std)))
    return un, ln


def thresholds(metric, Un, Ln, qn, arg):
    Un, Ln, qn = _a(Un), _a(Ln), _a(qn)
    qs = np.sort(qn)
    tu, tl = ctypes.c_double(), ctypes.c_double()
    _lib.orc_thresholds(_metric(metric), _p(Un), _p(Ln), Un.size, _p(qn), _p(qs), qn.size,
                        float(arg), ctypes.byref(tu), ctypes.byref(tl))
    return tu.value, tl.value


def q_to_h(metric, q, tl, tu, arg):
    q = _a(q)
    hl, hu = np.empty(q.size), np.empty(q.size)
    _lib.orc_q_to_h(_metric(metric), _p(q), q.size, tl, tu, float(arg), _p(hl), _p(hu))
    return hl, hu


def sip_direct(g, h):
    g, h = _a(g), _a(h)
    z = np.zeros(g.size)
    _lib.orc_sip_direct(_p(g), g.size, _p(h), h.size, _p(z))
    return z


def trans(metric, zl, zu, m):
    return _lib.orc_trans(_metric(metric), float(zl), float(zu), int(m))


def s_to_g(Ln, Un, tl, tu):
    Ln, Un = _a(Ln), _a(Un)
    gl, gu = np.empty(Ln.size), np.empty(Un.size)
    _lib.orc_s_to_g(_p(Ln), _p(Un), Ln.size, float(tl), float(tu), _p(gl), _p(gu))
    return gl, gu


def lbip_exceeds_cutoff(bound, cutoff, metric):
    return bool(_lib.orc_lbip_exceeds_cutoff(float(bound), float(cutoff), _metric(metric)))


def glb(metric, Q, T, arg, LQ, UQ, w, bsf):
    Q, T, LQ, UQ = _a(Q), _a(T), _a(LQ), _a(UQ)
    return _lib.orc_glb(_metric(metric), _p(Q), _p(T), Q.size, float(arg), _p(LQ), _p(UQ), w,
                        float(bsf))


def eld(metric, q, t, w, arg):
    q, t = _a(q), _a(t)
    return _lib.orc_eld(_metric(metric), _p(q), _p(t), q.size, w, float(arg))


def normalize_query(q_raw, m):
    q = _a(q_raw)
    out = np.empty(m)
    _lib.orc_normalize_query(_p(q), m, _p(out))
    return out


class raw_mode:
    """Context manager: the drivers run raw (non-normalised) matching inside
    it (orc_set_normalization_mode(1)), z-normalised matching outside."""

    def __enter__(self):
        _lib.orc_set_normalization_mode(1)
        return self

    def __exit__(self, *exc):
        _lib.orc_set_normalization_mode(0)


# ---- drivers ----
def search(T, q_raw, m, cutoff, metric, w, arg, exclude=None, segments=None, threads=1):
    T, q = _a(T), _a(q_raw)
    st = Stats()
    xb, xe = _excl(exclude)
    sb, se = (0, -1) if segments is None else segments
    cap = 1 << 20
    out = np.empty(cap, np.int64)
    n = ctypes.c_int64()
    rc = _lib.orc_search(_p(T), T.size, _p(q), m, float(cutoff), _metric(metric), w, float(arg),
                         xb, xe, sb, se, threads, ctypes.byref(st), _p(out, _I64), cap,
                         ctypes.byref(n))
    if rc:
        raise ValueError(f"orc_search rc={rc}")
    if n.value > cap:
        raise ValueError("too many matches for the oracle buffer")
    return out[: n.value].copy(), _stats(st)


def exact(T, q_raw, m, cutoff, metric, w, arg, exclude=None, threads=1):
    T, q = _a(T), _a(q_raw)
    st = Stats()
    xb, xe = _excl(exclude)
    cap = 1 << 22
    out = np.empty(cap, np.int64)
    dist = np.empty(cap)
    n = ctypes.c_int64()
    rc = _lib.orc_exact(_p(T), T.size, _p(q), m, float(cutoff), _metric(metric), w, float(arg),
                        xb, xe, threads, ctypes.byref(st), _p(out, _I64), _p(dist), cap,
                        ctypes.byref(n))
    if rc:
        raise ValueError(f"orc_exact rc={rc}")
    if n.value > cap:
        raise ValueError("too many matches for the oracle buffer")
    return out[: n.value].copy(), dist[: n.value].copy(), _stats(st)


def cutoff(T, q_raw, m, k, metric, w, arg, exclude=None, threads=1):
    """Returns (threshold, ranks, rank distances, stats); raises LookupError
    when fewer than k candidates exist (src/fsmeld.cpp:519-520).  threads > 1
    merges the heaps of parallel segment ranges (same ranks and threshold)."""
    T, q = _a(T), _a(q_raw)
    st = Stats()
    xb, xe = _excl(exclude)
    ranks = np.empty(k, np.int64)
    rd = np.empty(k)
    rc = _lib.orc_cutoff(_p(T), T.size, _p(q), m, k, _metric(metric), w, float(arg), xb, xe,
                         int(threads), ctypes.byref(st), _p(ranks, _I64), _p(rd))
    if rc == -4:
        raise LookupError("Fewer valid, non-excluded candidates than requested k.")
    if rc:
        raise ValueError(f"orc_cutoff rc={rc}")
    return st.threshold, ranks, rd, _stats(st)


def exact_topk(T, q_raw, m, k, metric, w, arg, exclude=None, threads=1):
    """(threshold, indices, distances, stats) of FSMELD_exact with top_k = k;
    raises LookupError below k finite distances."""
    T, q = _a(T), _a(q_raw)
    st = Stats()
    xb, xe = _excl(exclude)
    cap = 1 << 22
    out = np.empty(cap, np.int64)
    dist = np.empty(cap)
    n = ctypes.c_int64()
    kth = ctypes.c_double()
    rc = _lib.orc_exact_topk(_p(T), T.size, _p(q), m, k, _metric(metric), w, float(arg), xb, xe,
                             threads, ctypes.byref(st), _p(out, _I64), _p(dist), cap,
                             ctypes.byref(n), ctypes.byref(kth))
    if rc == -4:
        raise LookupError("fewer finite exact distances than k")
    if rc:
        raise ValueError(f"orc_exact_topk rc={rc}")
    return kth.value, out[: n.value].copy(), dist[: n.value].copy(), _stats(st)


def window_stats(T, m, exclude=None):
    T = _a(T)
    n = T.size - m + 1
    mu, sg = np.empty(n), np.empty(n)
    inv = np.empty(n, np.uint8)
    xb, xe = _excl(exclude)
    rc = _lib.orc_window_stats(_p(T), T.size, m, xb, xe, _p(mu), _p(sg), _p(inv, _U8))
    if rc:
        raise ValueError(f"orc_window_stats rc={rc}")
    return mu, sg, inv.astype(bool)


def lbip_all(T, q_raw, m, metric, w, arg):
    T, q = _a(T), _a(q_raw)
    n = T.size - m + 1
    mk = 1
    while mk < 8 * m:
        mk *= 2
    mk = min(mk, T.size)
    kml = mk - m + 1
    S = n // kml + (1 if n % kml else 0)
    lb, tu, tl = np.empty(n), np.empty(S), np.empty(S)
    rc = _lib.orc_lbip_all(_p(T), T.size, _p(q), m, _metric(metric), w, float(arg), _p(lb),
                           _p(tu), _p(tl))
    if rc:
        raise ValueError(f"orc_lbip_all rc={rc}")
    return lb, tu, tl


def ref_eld_lib():
    """The reference's own ElasticMeasures.cpp (built by oracle/Makefile where
    the reference tree exists), or None."""
    if not os.path.exists(REF_ELD):
        return None
    lib = ctypes.CDLL(REF_ELD)
    lib.ref_eld.restype = ctypes.c_double
    lib.ref_eld.argtypes = [_I, _D, _D, _I, _I, ctypes.c_double]
    return lib
