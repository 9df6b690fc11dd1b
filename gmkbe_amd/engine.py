"""Python host interface over the C ABI of libgmk.so (include/gmk.h).

Mirrors the reference's process interface for the enumeration path: a graph in
.graph.bin CSR form, k (-k), a result limit (-r) and a device (-d), producing the
set of maximal k-biplexes with both sides at least theta.  The HIP library is
required: importing succeeds without it, but every call raises when the library
is missing or no GPU is present -- there is no CPU path.
"""
from __future__ import annotations

import ctypes
import json
import dataclasses
import os

import numpy as np

from .graphio import Graph

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_lib", "libgmk.so")

MODES = {"plain": 0, "batched": 1, "learned": 2}
PARTS = {"auto": 0, "global": 1, "seeded": 2}
STATUS = {0: "ok", -1: "invalid argument", -2: "io error", -3: "device error", -4: "out of device memory",
          -5: "no graph", -6: "overflow", -7: "unsupported"}


class Params(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("theta", ctypes.c_uint32), ("limit", ctypes.c_uint64),
                ("mode", ctypes.c_int32), ("partition", ctypes.c_int32), ("keep_results", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("result_capacity_words", ctypes.c_uint64),
                ("dedup", ctypes.c_int32), ("dedup_load_permille", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("fp_sum", ctypes.c_uint64), ("fp_xor", ctypes.c_uint64),
                ("states", ctypes.c_uint64), ("intersections", ctypes.c_uint64), ("algo_bytes", ctypes.c_uint64),
                ("partitions", ctypes.c_uint64), ("launches", ctypes.c_uint64), ("dedup_inserts", ctypes.c_uint64),
                ("dedup_probes", ctypes.c_uint64), ("dedup_dups", ctypes.c_uint64), ("ms", ctypes.c_double),
                ("ms_expand", ctypes.c_double), ("hit_limit", ctypes.c_int32), ("spills", ctypes.c_uint32),
                ("ms_kernel", ctypes.c_double * 4), ("launches_kernel", ctypes.c_uint64 * 4),
                ("bytes_kernel", ctypes.c_uint64 * 4), ("dedup_undecided", ctypes.c_uint64),
                ("ms_schedule", ctypes.c_double), ("ms_reduce", ctypes.c_double),
                ("bytes_reduce", ctypes.c_uint64), ("reduce_removed", ctypes.c_uint64),
                ("widest_side", ctypes.c_uint64), ("reduce_edges", ctypes.c_uint64), ("ms_build", ctypes.c_double)]

    KERNELS = ("bb_expand", "thin_build_families", "thin_check", "partition_build")

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            if name == "pad":
                continue
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


MLP_IN, MLP_HID = 8, 64
STATIC_ITEMS = 65536  # GMK_STATIC_ITEMS


class GraphDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("n_left", ctypes.c_uint32), ("off", ctypes.POINTER(ctypes.c_uint32)),
                ("adj", ctypes.POINTER(ctypes.c_uint32)), ("k", ctypes.c_uint32), ("theta", ctypes.c_uint32)]


class GraphStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("fp_sum", ctypes.c_uint64), ("fp_xor", ctypes.c_uint64),
                ("ms", ctypes.c_double), ("hit_cap", ctypes.c_int32), ("pad", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    """gmk_tuning: expansion launch parameters per bitset width (2^l words)."""
    _fields_ = [("version", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32 * 7),
                ("batch_states", ctypes.c_uint32 * 7), ("dedup_load_permille", ctypes.c_uint32),
                ("check_form", ctypes.c_uint32 * 4)]

    def as_dict(self) -> dict:
        return {"version": self.version, "waves_per_cu": list(self.waves_per_cu),
                "batch_states": list(self.batch_states), "dedup_load_permille": self.dedup_load_permille,
                "check_form": list(self.check_form)}

    @classmethod
    def from_dict(cls, d: dict) -> "Tuning":
        t = cls()
        t.version = int(d.get("version", TUNING_VERSION))
        for i, v in enumerate(d["waves_per_cu"]):
            t.waves_per_cu[i] = int(v)
        for i, v in enumerate(d["batch_states"]):
            t.batch_states[i] = int(v)
        t.dedup_load_permille = int(d.get("dedup_load_permille", 500))
        for i, v in enumerate(d.get("check_form", [0, 0, 0, 0])):
            t.check_form[i] = int(v)
        return t


TUNING_VERSION = 2
# Table written by gmkbe_amd/autotune.py for this device family; loaded by
# every Engine when present.
TUNE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune", "mi355x.json")


class Mlp(ctypes.Structure):
    _fields_ = [("w1", ctypes.c_float * (MLP_IN * MLP_HID)), ("b1", ctypes.c_float * MLP_HID),
                ("w2", ctypes.c_float * (MLP_HID * MLP_HID)), ("b2", ctypes.c_float * MLP_HID),
                ("w3", ctypes.c_float * MLP_HID), ("b3", ctypes.c_float),
                ("mu", ctypes.c_float * MLP_IN), ("sd", ctypes.c_float * MLP_IN)]

    def arrays(self) -> dict:
        """Parameters as numpy arrays (w1: in x hid, w2: hid x hid)."""
        a = lambda f, shape=None: np.ctypeslib.as_array(f).copy().reshape(shape) if shape else np.ctypeslib.as_array(f).copy()
        return {"w1": a(self.w1, (MLP_IN, MLP_HID)), "b1": a(self.b1), "w2": a(self.w2, (MLP_HID, MLP_HID)),
                "b2": a(self.b2), "w3": a(self.w3), "b3": np.float32(self.b3), "mu": a(self.mu), "sd": a(self.sd)}

    @staticmethod
    def from_arrays(d: dict) -> "Mlp":
        m = Mlp()
        for name in ("w1", "b1", "w2", "b2", "w3", "mu", "sd"):
            src = np.ascontiguousarray(np.asarray(d[name], dtype=np.float32).ravel())
            ctypes.memmove(getattr(m, name), src.ctypes.data, src.nbytes)
        m.b3 = float(d["b3"])
        return m


class GmkError(RuntimeError):
    pass


_lib = None


def load_library():
    """Loads libgmk.so; raises GmkError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GmkError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    u32p = P(ctypes.c_uint32)
    L.gmk_create.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
    L.gmk_create.restype = ctypes.c_int
    L.gmk_destroy.argtypes = [ctypes.c_void_p]
    L.gmk_destroy.restype = None
    L.gmk_last_error.argtypes = [ctypes.c_void_p]
    L.gmk_last_error.restype = ctypes.c_char_p
    L.gmk_parse_edge_list.argtypes = [ctypes.c_char_p, ctypes.c_int, P(u32p), P(u32p), P(ctypes.c_uint64)]
    L.gmk_parse_edge_list.restype = ctypes.c_int
    L.gmk_free.argtypes = [ctypes.c_void_p]
    L.gmk_free.restype = None
    L.gmk_device_name.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32]
    L.gmk_device_name.restype = ctypes.c_int
    L.gmk_load_graph_bin.argtypes = [ctypes.c_void_p, ctypes.c_char_p, u32p, u32p, P(ctypes.c_uint64)]
    L.gmk_load_graph_bin.restype = ctypes.c_int
    L.gmk_set_graph_csr.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
    L.gmk_set_graph_csr.restype = ctypes.c_int
    L.gmk_enumerate.argtypes = [ctypes.c_void_p, P(Params), P(Stats)]
    L.gmk_enumerate.restype = ctypes.c_int
    L.gmk_fetch_results.argtypes = [ctypes.c_void_p, u32p, ctypes.c_uint64, P(ctypes.c_uint64)]
    L.gmk_fetch_results.restype = ctypes.c_int
    L.gmk_fetch_progress.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_double), ctypes.c_uint32,
                                     P(ctypes.c_uint32)]
    L.gmk_fetch_progress.restype = ctypes.c_int
    L.gmk_enumerate_batch.argtypes = [ctypes.c_void_p, P(GraphDesc), ctypes.c_uint32, P(Params), u32p,
                                      P(GraphStats), P(Stats)]
    L.gmk_enumerate_batch.restype = ctypes.c_int
    fp = P(ctypes.c_float)
    L.gmk_graph_features.argtypes = [P(GraphDesc), fp]
    L.gmk_graph_features.restype = ctypes.c_int
    L.gmk_mlp_train.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                ctypes.c_uint32, P(Mlp), fp]
    L.gmk_mlp_train.restype = ctypes.c_int
    L.gmk_mlp_predict.argtypes = [ctypes.c_void_p, P(Mlp), fp, ctypes.c_uint32, fp]
    L.gmk_mlp_predict.restype = ctypes.c_int
    L.gmk_mlp_grad.argtypes = [ctypes.c_void_p, P(Mlp), fp, fp, ctypes.c_uint32, P(Mlp), fp]
    L.gmk_mlp_grad.restype = ctypes.c_int
    L.gmk_set_cost_model.argtypes = [ctypes.c_void_p, P(Mlp)]
    L.gmk_set_cost_model.restype = ctypes.c_int
    L.gmk_set_item_model.argtypes = [ctypes.c_void_p, P(Mlp)]
    L.gmk_set_item_model.restype = ctypes.c_int
    L.gmk_item_features.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, fp]
    L.gmk_item_features.restype = ctypes.c_int
    L.gmk_item_predict.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, fp]
    L.gmk_item_predict.restype = ctypes.c_int
    L.gmk_item_predict_static.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, fp]
    L.gmk_item_predict_static.restype = ctypes.c_int
    L.gmk_probe_items.argtypes = [ctypes.c_void_p, P(Params), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                  P(ctypes.c_uint64), P(ctypes.c_double)]
    L.gmk_probe_items.restype = ctypes.c_int
    L.gmk_set_tuning.argtypes = [ctypes.c_void_p, P(Tuning)]
    L.gmk_set_tuning.restype = ctypes.c_int
    L.gmk_get_tuning.argtypes = [ctypes.c_void_p, P(Tuning)]
    L.gmk_get_tuning.restype = ctypes.c_int
    _lib = L
    return L


@dataclasses.dataclass
class Result:
    stats: dict
    words: np.ndarray | None  # record stream [|L|, L.., |R|, R..] per result

    @property
    def count(self) -> int:
        return int(self.stats["count"])


class Engine:
    """One device context (gmk_create).  Not thread-safe; one run at a time."""

    def __init__(self, device: int = 0, tuning: str | dict | None = TUNE_TABLE):
        """tuning: a tuning table (path or dict from gmkbe_amd/autotune.py);
        the committed table for this device family by default, library
        defaults when it is absent or None."""
        self._L = load_library()
        h = ctypes.c_void_p()
        rc = self._L.gmk_create(device, ctypes.byref(h))
        if rc != 0:
            raise GmkError(f"gmk_create(device={device}) failed: {STATUS.get(rc, rc)}")
        self._h = h
        self.graph: Graph | None = None
        if isinstance(tuning, str):
            if os.path.exists(tuning):
                with open(tuning) as f:
                    self.set_tuning(json.load(f)["tuning"])
        elif tuning is not None:
            self.set_tuning(tuning)

    def set_tuning(self, t) -> None:
        t = t if isinstance(t, Tuning) else Tuning.from_dict(t)
        self._check(self._L.gmk_set_tuning(self._h, ctypes.byref(t)), "gmk_set_tuning")

    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        self._check(self._L.gmk_device_name(self._h, buf, 256), "gmk_device_name")
        return buf.value.decode()

    def get_tuning(self) -> dict:
        t = Tuning()
        self._check(self._L.gmk_get_tuning(self._h, ctypes.byref(t)), "gmk_get_tuning")
        return t.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.gmk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self._L.gmk_last_error(self._h)
            raise GmkError(f"{what}: {STATUS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def set_graph(self, g: Graph) -> None:
        off = np.ascontiguousarray(g.off, dtype=np.uint32)
        adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
        p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self._check(self._L.gmk_set_graph_csr(self._h, g.n, g.n_left, p(off), p(adj)), "gmk_set_graph_csr")
        self.graph = g

    def load_graph_bin(self, path: str) -> tuple[int, int, int]:
        n, nl, nnz = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        self._check(self._L.gmk_load_graph_bin(self._h, path.encode(), ctypes.byref(n), ctypes.byref(nl),
                                               ctypes.byref(nnz)), f"gmk_load_graph_bin({path})")
        return n.value, nl.value, nnz.value

    def enumerate(self, k: int = 1, theta: int = 1, limit: int = 0, mode: str = "plain",
                  partition: str = "auto", keep: bool = True, capacity_words: int = 0, dedup: bool = False,
                  dedup_load: float = 0.5, verbose: bool = False, fetch: bool = True) -> Result:
        """keep: write result records (in device memory); fetch: copy them to
        the host (Result.words, None otherwise).  verbose: the library prints a
        progress line to stderr every 10 s."""
        prm = Params(k, theta, limit, MODES[mode], PARTS[partition], int(keep), int(verbose), capacity_words,
                     int(dedup), int(round(dedup_load * 1000)))
        st = Stats()
        self._check(self._L.gmk_enumerate(self._h, ctypes.byref(prm), ctypes.byref(st)), "gmk_enumerate")
        words = None
        if keep and fetch:
            nw = ctypes.c_uint64()
            self._check(self._L.gmk_fetch_results(self._h, None, 0, ctypes.byref(nw)), "gmk_fetch_results")
            words = np.empty(int(nw.value), dtype=np.uint32)
            if nw.value:
                self._check(self._L.gmk_fetch_results(self._h, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                      nw.value, ctypes.byref(nw)), "gmk_fetch_results")
        return Result(st.as_dict(), words)

    def progress(self) -> list[tuple[int, float]]:
        """(results so far, elapsed ms) after each launch of the last run."""
        n = ctypes.c_uint32()
        self._check(self._L.gmk_fetch_progress(self._h, None, None, 0, ctypes.byref(n)), "gmk_fetch_progress")
        cnt = (ctypes.c_uint64 * max(1, n.value))()
        ms = (ctypes.c_double * max(1, n.value))()
        self._check(self._L.gmk_fetch_progress(self._h, cnt, ms, n.value, ctypes.byref(n)), "gmk_fetch_progress")
        return [(int(cnt[i]), float(ms[i])) for i in range(n.value)]


def _graph_desc(g: Graph, k: int, theta: int, keep: list) -> GraphDesc:
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    keep += [off, adj]
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    return GraphDesc(g.n, g.n_left, p(off), p(adj), k, theta)


def graph_features(g: Graph, k: int, theta: int = 1) -> np.ndarray:
    """The cost model's input features of one graph (gmk_graph_features)."""
    L = load_library()
    keep: list = []
    d = _graph_desc(g, k, theta, keep)
    out = np.zeros(MLP_IN, dtype=np.float32)
    rc = L.gmk_graph_features(ctypes.byref(d), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc:
        raise GmkError(f"gmk_graph_features: {STATUS.get(rc, rc)}")
    return out


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _engine_batch_methods():
    def enumerate_batch(self, graphs, limit: int = 0, mode: str = "batched", order=None, verbose: bool = False):
        """graphs: list of (Graph, k, theta).  Returns (per-graph stats dicts, total stats dict).
        verbose: the library prints a progress line to stderr every 10 s."""
        keep: list = []
        arr = (GraphDesc * len(graphs))(*[_graph_desc(g, k, t, keep) for g, k, t in graphs])
        prm = Params(0, 1, limit, MODES[mode], PARTS["global"], 0, int(verbose), 0, 0, 0)
        per = (GraphStats * len(graphs))()
        tot = Stats()
        op = None
        if order is not None:
            o = np.ascontiguousarray(order, dtype=np.uint32)
            keep.append(o)
            op = o.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self._check(self._L.gmk_enumerate_batch(self._h, arr, len(graphs), ctypes.byref(prm), op, per,
                                                ctypes.byref(tot)), "gmk_enumerate_batch")
        out = [{"count": p.count, "fp_sum": p.fp_sum, "fp_xor": p.fp_xor, "ms": p.ms, "hit_cap": bool(p.hit_cap)}
               for p in per]
        return out, tot.as_dict()

    def mlp_train(self, X, y, epochs: int = 500, lr: float = 1e-2, seed: int = 1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        m = Mlp()
        loss = ctypes.c_float()
        self._check(self._L.gmk_mlp_train(self._h, _fp(X), _fp(y), len(y), epochs, lr, seed, ctypes.byref(m),
                                          ctypes.byref(loss)), "gmk_mlp_train")
        return m, float(loss.value)

    def mlp_predict(self, model: Mlp, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(len(X), dtype=np.float32)
        self._check(self._L.gmk_mlp_predict(self._h, ctypes.byref(model), _fp(X), len(X), _fp(y)), "gmk_mlp_predict")
        return y

    def mlp_grad(self, model: Mlp, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = Mlp()
        loss = ctypes.c_float()
        self._check(self._L.gmk_mlp_grad(self._h, ctypes.byref(model), _fp(X), _fp(y), len(y), ctypes.byref(g),
                                         ctypes.byref(loss)), "gmk_mlp_grad")
        return g, float(loss.value)

    def set_cost_model(self, model):
        self._check(self._L.gmk_set_cost_model(self._h, ctypes.byref(model) if model is not None else None),
                    "gmk_set_cost_model")

    def item_features(self, first: int, count: int, k: int):
        """Features of the one-vertex work items first .. first+count-1 (gmk_item_features)."""
        out = np.zeros((count, MLP_IN), dtype=np.float32)
        self._check(self._L.gmk_item_features(self._h, first, count, k, _fp(out)), "gmk_item_features")
        return out

    def item_predict(self, first: int, count: int, k: int):
        """The set item model's scores of items first .. first+count-1 (gmk_item_predict)."""
        out = np.zeros(count, dtype=np.float32)
        self._check(self._L.gmk_item_predict(self._h, first, count, k, _fp(out)), "gmk_item_predict")
        return out

    def item_predict_static(self, first: int, count: int, k: int):
        """As item_predict, through the entry point that allocates nothing
        (gmk_item_predict_static; at most STATIC_ITEMS items per call)."""
        out = np.zeros(count, dtype=np.float32)
        self._check(self._L.gmk_item_predict_static(self._h, first, count, k, _fp(out)), "gmk_item_predict_static")
        return out

    def probe_items(self, first: int, count: int, k: int, ranks: int = 1024):
        """Results and device lane ms of the first `ranks` combinations of each
        family of items first .. first+count-1 (gmk_probe_items)."""
        prm = Params(k, 1, 0, MODES["batched"], PARTS["seeded"], 0, 0, 0, 0, 0)
        res = np.zeros(count, dtype=np.uint64)
        ms = np.zeros(count, dtype=np.float64)
        self._check(self._L.gmk_probe_items(self._h, ctypes.byref(prm), first, count, ranks,
                                            res.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                            ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "gmk_probe_items")
        return res, ms

    def set_item_model(self, model):
        self._check(self._L.gmk_set_item_model(self._h, ctypes.byref(model) if model is not None else None),
                    "gmk_set_item_model")

    Engine.item_features = item_features
    Engine.item_predict = item_predict
    Engine.item_predict_static = item_predict_static
    Engine.probe_items = probe_items
    Engine.set_item_model = set_item_model
    Engine.enumerate_batch = enumerate_batch
    Engine.mlp_train = mlp_train
    Engine.mlp_predict = mlp_predict
    Engine.mlp_grad = mlp_grad
    Engine.set_cost_model = set_cost_model


_engine_batch_methods()
