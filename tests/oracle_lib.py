"""ctypes bindings to the oracle (oracle/gmk_oracle.cpp).  Test infrastructure:
only tests, smoke() and the CPU-baseline leg of bench.py import this module."""
from __future__ import annotations

import ctypes
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class Digest(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("fp_sum", ctypes.c_uint64), ("fp_xor", ctypes.c_uint64),
                ("canon_lo", ctypes.c_uint64), ("canon_hi", ctypes.c_uint64)]

    def as_tuple(self):
        return (self.count, self.fp_sum, self.fp_xor)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        u32p = P(ctypes.c_uint32)
        L.or_bppivot.restype = ctypes.c_longlong
        L.or_bppivot.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint64, ctypes.c_int, ctypes.c_int, P(u32p), P(ctypes.c_uint64),
                                 P(Digest), P(ctypes.c_uint64)]
        L.or_brute.restype = ctypes.c_longlong
        L.or_brute.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32,
                               P(u32p), P(ctypes.c_uint64), P(Digest)]
        L.or_canon_digest.restype = ctypes.c_int
        L.or_canon_digest.argtypes = [u32p, ctypes.c_uint64, P(Digest)]
        L.or_validate.restype = ctypes.c_longlong
        L.or_validate.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                  u32p, ctypes.c_uint64, P(ctypes.c_int)]
        L.or_fingerprint.restype = ctypes.c_uint64
        L.or_fingerprint.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32]
        L.or_free.argtypes = [ctypes.c_void_p]
        u64p = P(ctypes.c_uint64)
        L.or_fingerprints.restype = ctypes.c_longlong
        L.or_fingerprints.argtypes = [u32p, ctypes.c_uint64, u64p, ctypes.c_uint64]
        L.or_record_offsets.restype = ctypes.c_longlong
        L.or_record_offsets.argtypes = [u32p, ctypes.c_uint64, u64p, ctypes.c_uint64]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _take(words_p, nwords):
    n = int(nwords.value)
    arr = np.ctypeslib.as_array(words_p, shape=(max(n, 1),))[:n].copy()
    lib().or_free(ctypes.cast(words_p, ctypes.c_void_p))
    return arr


def bppivot(g, k, theta, limit=0, mode=2, keep=True):
    """Returns (records uint32 array or None, Digest, branches)."""
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    wp = ctypes.POINTER(ctypes.c_uint32)()
    nw = ctypes.c_uint64(0)
    dg = Digest()
    br = ctypes.c_uint64(0)
    lib().or_bppivot(g.n_left, g.n, _p(off), _p(adj), k, theta, limit, mode, int(keep),
                     ctypes.byref(wp), ctypes.byref(nw), ctypes.byref(dg), ctypes.byref(br))
    words = _take(wp, nw)
    return (words if keep else None), dg, int(br.value)


def brute(g, k, theta):
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    wp = ctypes.POINTER(ctypes.c_uint32)()
    nw = ctypes.c_uint64(0)
    dg = Digest()
    r = lib().or_brute(g.n_left, g.n, _p(off), _p(adj), k, theta, ctypes.byref(wp), ctypes.byref(nw), ctypes.byref(dg))
    if r < 0:
        raise ValueError("graph too large for brute force")
    return _take(wp, nw), dg


def digest(words):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    dg = Digest()
    if lib().or_canon_digest(_p(words), words.size, ctypes.byref(dg)) != 0:
        raise ValueError("malformed record stream")
    return dg


def validate(g, k, theta, words):
    """0 when every record is a distinct maximal k-biplex with sides >= theta;
    otherwise (record index, reason code)."""
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    words = np.ascontiguousarray(words, dtype=np.uint32)
    why = ctypes.c_int(0)
    r = lib().or_validate(g.n_left, g.n, _p(off), _p(adj), k, theta, _p(words), words.size, ctypes.byref(why))
    return 0 if r == 0 else (int(r), int(why.value))


def fingerprints(words):
    """Per-record fingerprints (same function as the device digest), uint64."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cap = max(1, words.size // 2)
    out = np.empty(cap, dtype=np.uint64)
    n = lib().or_fingerprints(_p(words), words.size, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap)
    assert n >= 0, "malformed record stream"
    return out[:n]


def sample_records(words, count, seed=0):
    """A record stream holding `count` records spread over the stream."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cap = max(1, words.size // 2)
    offs = np.empty(cap, dtype=np.uint64)
    n = lib().or_record_offsets(_p(words), words.size, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap)
    assert n >= 0, "malformed record stream"
    offs = offs[:n].astype(np.int64)
    pick = np.sort(np.random.default_rng(seed).choice(n, size=min(count, n), replace=False))
    parts = []
    for i in pick:
        o = offs[i]
        nl = int(words[o]); nr = int(words[o + 1 + nl])
        parts.append(words[o:o + 2 + nl + nr])
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint32)


def records(words):
    """Iterate (L tuple, R tuple) over a record stream."""
    i, n = 0, len(words)
    while i < n:
        nl = int(words[i]); L = tuple(int(x) for x in words[i + 1:i + 1 + nl])
        nr = int(words[i + 1 + nl]); R = tuple(int(x) for x in words[i + 2 + nl:i + 2 + nl + nr])
        yield L, R
        i += 2 + nl + nr


MODEL_PATH = os.path.join(ROOT, "tests", "model", "_build", "libthinmodel.so")
_model = None


def _model_init():
    global _model
    if _model is None:
        _model = ctypes.CDLL(MODEL_PATH)
        P = ctypes.POINTER
        u32p = P(ctypes.c_uint32)
        _model.thin_model.restype = ctypes.c_longlong
        _model.thin_model.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_int, ctypes.c_uint32, P(u32p), P(ctypes.c_uint64)]
        _model.thin_dominated.restype = ctypes.c_longlong
        _model.thin_dominated.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_int]
        _model.thin_model_set_hash_min.restype = None
        _model.thin_model_set_hash_min.argtypes = [ctypes.c_uint32]
        _model.thin_model_hash_check.restype = ctypes.c_longlong
        _model.thin_model_hash_check.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32]
        u64p = P(ctypes.c_uint64)
        _model.thin_scan_totals_model.restype = None
        _model.thin_scan_totals_model.argtypes = [u64p, ctypes.c_uint32, u64p]
        _model.thin_model_set_guards.restype = None
        _model.thin_model_set_guards.argtypes = [ctypes.c_int]
        _model.thin_model_guard_modes.restype = None
        _model.thin_model_guard_modes.argtypes = [u64p]
        _model.thin_model_binom.restype = ctypes.c_uint64
        _model.thin_model_binom.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        _model.thin_model_unrank.restype = None
        _model.thin_model_unrank.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, u32p]


def scan_totals_model(totals):
    """Exclusive saturating prefix sums of family totals by the device scan's
    own phases, run on the CPU (tests/model/thin_model.cpp); n + 1 entries."""
    _model_init()
    t = np.ascontiguousarray(totals, dtype=np.uint64)
    out = np.zeros(len(t) + 1, dtype=np.uint64)
    P = ctypes.POINTER(ctypes.c_uint64)
    _model.thin_scan_totals_model(t.ctypes.data_as(P), len(t), out.ctypes.data_as(P))
    return out


def model_binom(n, c):
    """C(n, c) as the kernels compute it (saturating at 2^62)."""
    _model_init()
    return int(_model.thin_model_binom(n, c))


def model_unrank(r, c, n):
    """The kernels' colex unrank of rank r among the c-subsets of range(n)."""
    _model_init()
    out = (ctypes.c_uint32 * 3)()
    _model.thin_model_unrank(r, c, n, out)
    return [int(out[i]) for i in range(c)]


def hash_check(g, nh_min):
    """Neighbour-hash membership against the sorted lists, all (hashed vertex,
    vertex) pairs: the number of disagreements (tests/model/thin_model.cpp)."""
    _model_init()
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    return int(_model.thin_model_hash_check(g.n, _p(off), _p(adj), nh_min))


def thin_dominated(g, k, side_a):
    """One-vertex tuples of side_a with a dominating vertex (tests/model/thin_model.cpp)."""
    _model_init()
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    return int(_model.thin_dominated(g.n, g.n_left, _p(off), _p(adj), k, side_a))


def thin_model(g, k, theta, side_a, min_b, hash_min=0, guards=True, modes=None):
    """CPU model of the thin-result kernels (tests/model/thin_model.cpp);
    hash_min > 0: adjacency tests through neighbour hashes of the vertices of
    at least that degree; guards: the families' guard tables (as the device
    builds them) or plain candidate lists only.  `modes` (a dict) receives the
    families per guard mode."""
    _model_init()
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    wp = ctypes.POINTER(ctypes.c_uint32)()
    nw = ctypes.c_uint64(0)
    cnt = np.zeros(4, dtype=np.uint64)
    _model.thin_model_guard_modes(cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    _model.thin_model_set_hash_min(hash_min)
    _model.thin_model_set_guards(1 if guards else 0)
    try:
        _model.thin_model(g.n, g.n_left, _p(off), _p(adj), k, theta, side_a, min_b, ctypes.byref(wp),
                          ctypes.byref(nw))
    finally:
        _model.thin_model_set_hash_min(0)
        _model.thin_model_set_guards(1)
    _model.thin_model_guard_modes(cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if modes is not None:
        modes.update(none=int(cnt[0]), table=int(cnt[1]), allmax=int(cnt[2]), dominated=int(cnt[3]))
    return _take(wp, nw)


PORT_PATH = os.path.join(ROOT, "oracle", "_build", "libcpuport.so")
_port = None


def cpu_thin_port(g, k, limit=0, threads=0, budget=0.0):
    """Multi-core CPU port of the thin-result method (oracle/cpu_port.cpp):
    (results, seconds, ranks tested) over the left side's one-vertex tuples,
    stopping after about `budget` seconds when budget > 0."""
    global _port
    if _port is None:
        _port = ctypes.CDLL(PORT_PATH)
        P = ctypes.POINTER
        u32p = P(ctypes.c_uint32)
        _port.cpu_thin_port.restype = ctypes.c_longlong
        _port.cpu_thin_port.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.c_int, ctypes.c_double, P(ctypes.c_double), P(ctypes.c_uint64)]
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    sec = ctypes.c_double(0)
    rk = ctypes.c_uint64(0)
    n = _port.cpu_thin_port(g.n, g.n_left, _p(off), _p(adj), k, limit, threads, float(budget), ctypes.byref(sec),
                            ctypes.byref(rk))
    return int(n), float(sec.value), int(rk.value)


def closure_bicliques(g, theta):
    """Maximal bicliques (k = 0) with both sides >= theta, as a record stream,
    by closure under intersection: R is the right side of a maximal biclique
    iff R = the intersection of N(u) over its common neighbours u, and each such
    R is reached from some N(u) by intersecting with further N(u') while its
    size stays >= theta.  Independent of the branch-and-bound (a textbook
    characterisation of maximal bicliques); pinned against brute force in
    tests/test_oracle.py."""
    nl = int(g.n_left)
    off = np.asarray(g.off, dtype=np.int64)
    adj = np.asarray(g.adj, dtype=np.int64)
    nbr = lambda v: adj[off[v]:off[v + 1]]
    base = {tuple(nbr(u)) for u in range(nl) if off[u + 1] - off[u] >= theta}
    found = set(base)
    frontier = list(base)
    cover = {}
    while frontier:
        new = []
        for s in frontier:
            # cnt[u] = |s & N(u)| over the left vertices
            cnt = np.bincount(np.concatenate([nbr(x) for x in s]), minlength=nl)
            cover[s] = np.nonzero(cnt == len(s))[0]
            for u in np.nonzero((cnt >= theta) & (cnt < len(s)))[0]:
                t = tuple(np.intersect1d(np.asarray(s, dtype=np.int64), nbr(u)))
                if t not in found:
                    found.add(t)
                    new.append(t)
        frontier = new
    parts = []
    for s in sorted(found):
        L = cover[s]
        if len(L) >= theta:
            parts.append(np.concatenate([[len(L)], L, [len(s)], np.asarray(s)]).astype(np.uint32))
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint32)
