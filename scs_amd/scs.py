"""Python host API over the C ABI in include/scs.h (ctypes).

Mirrors the reference's process surface (main.cu): a graph loaded from an
edge list or a CSR, an engine configured by the reference's twelve
command-line parameters plus named extensions, queries submitted as
``<l> <h> <q> <heuristic_flag> <limit_dominating_set>`` and results with the
reference's columns (N1, N2, QID, time, degree, overtime, heu).

The library must be built (scs_amd/build.py); nothing here falls back to a
CPU path.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libscs.so")
_LIB = None

MAX_H = 1000
MAX_DOMS = 32
STATUS = {0: "OK", 1: "HEU_OPT", 2: "CAPPED", 3: "INFEASIBLE", 4: "REJECTED"}
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EAGAIN", -4: "EHIP", -5: "EIO", -6: "EBIG"}


class SCSError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: {ERRORS.get(code, code)}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("partition_size", C.c_uint32), ("buffer_size", C.c_uint32),
                ("copy_limit", C.c_double), ("read_limit", C.c_uint32),
                ("limit_queries", C.c_uint32), ("factor", C.c_uint32),
                ("red1", C.c_uint32), ("red2", C.c_uint32), ("red3", C.c_uint32),
                ("prun1", C.c_uint32), ("prun2", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("limit_tasks", C.c_uint32),
                ("node_budget", C.c_uint64), ("mem_budget", C.c_uint64),
                ("max_kept", C.c_uint32), ("time_slice_us", C.c_uint32),
                ("deterministic", C.c_uint32), ("learned", C.c_uint32),
                ("model", C.c_void_p), ("heuristic_threads", C.c_uint32),
                ("admit_per_step", C.c_uint32), ("no_lds_bitmaps", C.c_uint32),
                ("levels", C.c_uint32), ("pair", C.c_uint32), ("deadline_ms", C.c_uint32),
                ("whole_graph", C.c_uint32), ("gpu_heuristic", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("l", C.c_uint32), ("h", C.c_uint32), ("q", C.c_uint32),
                ("kl", C.c_uint32), ("heu_kl", C.c_uint32), ("ku", C.c_uint32),
                ("status", C.c_int32), ("kept", C.c_uint32), ("nodes", C.c_uint64),
                ("t_submit_us", C.c_double), ("t_done_us", C.c_double),
                ("wsize", C.c_uint32), ("witness", C.c_uint32 * MAX_H)]


class Stats(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("donations", C.c_uint64), ("rebuilds", C.c_uint64),
                ("engine_us", C.c_double), ("prep_us", C.c_double), ("heu_us", C.c_double),
                ("prep_bytes", C.c_uint64), ("epochs", C.c_uint64), ("waves", C.c_uint32),
                ("kcap", C.c_uint32), ("edges", C.c_uint64), ("core_us", C.c_double),
                ("prep_batches", C.c_uint64), ("phase_ticks", C.c_uint64 * 13),
                ("dom_front", C.c_uint64), ("dom_cand", C.c_uint64), ("dom_found", C.c_uint64),
                ("full_scans", C.c_uint64), ("full_rows", C.c_uint64)]


class Model(C.Structure):
    """scs_model: FP32 master weights of the branching MLP (include/scs.h)."""
    _fields_ = [("w1", (C.c_float * 64) * 8), ("b1", C.c_float * 64),
                ("w2", (C.c_float * 64) * 64), ("b2", C.c_float * 64), ("w3", C.c_float * 64),
                ("b3", C.c_float), ("mean", C.c_float * 8), ("scale", C.c_float * 8),
                ("pad", C.c_float * 7)]

    def to_arrays(self):
        return {f: np.ctypeslib.as_array(getattr(self, f)).copy() if f != "b3" else np.float32(self.b3)
                for f, _ in self._fields_ if f != "pad"}

    @classmethod
    def from_arrays(cls, d):
        m = cls()
        for f, _ in cls._fields_:
            if f == "pad":
                continue
            if f == "b3":
                m.b3 = float(d["b3"])
            else:
                np.ctypeslib.as_array(getattr(m, f))[...] = np.asarray(d[f], np.float32)
        return m

    def save(self, path):
        np.savez(path, **self.to_arrays())

    def save_raw(self, path):
        """Raw scs_model bytes, the format scs_server --learned reads."""
        with open(path, "wb") as f:
            f.write(bytes(self))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls.from_arrays({k: z[k] for k in z.files})


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError(f"{_LIBPATH} is missing: run scs_amd/build.py (no CPU fallback)")
        L = C.CDLL(_LIBPATH)
        P = C.c_void_p
        L.scs_graph_create.argtypes = [np.ctypeslib.ndpointer(np.int64, flags="C"),
                                       np.ctypeslib.ndpointer(np.int32, flags="C"),
                                       C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(P)]
        L.scs_graph_load.argtypes = [C.c_char_p, C.POINTER(P)]
        L.scs_graph_info.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint32)]
        L.scs_graph_core.argtypes = [P, np.ctypeslib.ndpointer(np.uint32, flags="C")]
        L.scs_graph_csr.argtypes = [P, np.ctypeslib.ndpointer(np.int64, flags="C"),
                                    np.ctypeslib.ndpointer(np.int32, flags="C")]
        L.scs_graph_core_time_us.restype = C.c_double
        L.scs_graph_core_time_us.argtypes = [P]
        L.scs_graph_destroy.argtypes = [P]
        L.scs_params_default.argtypes = [C.POINTER(Params)]
        L.scs_engine_create.argtypes = [P, C.POINTER(Params), C.POINTER(P)]
        L.scs_submit.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint64]
        L.scs_submit_text.argtypes = [P, C.c_char_p, C.c_uint64]
        L.scs_step.argtypes = [P]
        L.scs_poll.argtypes = [P, C.POINTER(Result), C.c_int]
        L.scs_drain.argtypes = [P]
        L.scs_engine_destroy.argtypes = [P]
        L.scs_engine_stats.argtypes = [P, C.POINTER(Stats)]
        L.scs_strerror.restype = C.c_char_p
        L.scs_strerror.argtypes = [C.c_int]
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
        L.scs_model_init.argtypes = [C.POINTER(Model), C.c_uint64]
        L.scs_mlp_grad.argtypes = [C.POINTER(Model), f32p, f32p, C.c_uint32, C.c_float,
                                   C.POINTER(Model), C.POINTER(C.c_float)]
        L.scs_mlp_train.argtypes = [C.POINTER(Model), f32p, f32p, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_float, C.c_float, C.c_uint64,
                                    C.POINTER(C.c_float)]
        L.scs_mlp_infer.argtypes = [C.POINTER(Model), f32p, C.c_uint32, f32p]
        L.scs_engine_set_model.argtypes = [P, C.POINTER(Model), C.c_uint32]
        L.scs_engine_trace.argtypes = [P, C.c_uint32, C.c_uint32]
        L.scs_engine_trace_read.argtypes = [P, np.ctypeslib.ndpointer(np.uint32, flags="C"),
                                            C.c_uint32, C.POINTER(C.c_uint32)]
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc < 0:
        raise SCSError(rc, what)
    return rc


class Graph:
    """A graph resident on the GPU with its core numbers (computed there)."""

    def __init__(self, handle):
        self._h = handle
        n, nnz, dmax = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check(lib().scs_graph_info(self._h, C.byref(n), C.byref(nnz), C.byref(dmax)), "info")
        self.n, self.nnz, self.dmax = n.value, nnz.value, dmax.value

    @classmethod
    def from_csr(cls, off, adj):
        off = np.ascontiguousarray(off, dtype=np.int64)
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        h = C.c_void_p()
        _check(lib().scs_graph_create(off, adj, len(off) - 1, adj.size, 0, C.byref(h)),
               "scs_graph_create")
        return cls(h)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        _check(lib().scs_graph_load(path.encode(), C.byref(h)), f"scs_graph_load({path})")
        return cls(h)

    def core(self):
        c = np.zeros(self.n, np.uint32)
        _check(lib().scs_graph_core(self._h, c), "core")
        return c

    def csr(self):
        off = np.zeros(self.n + 1, np.int64)
        adj = np.zeros(max(self.nnz, 1), np.int32)
        _check(lib().scs_graph_csr(self._h, off, adj), "csr")
        return off, adj[:self.nnz]

    @property
    def core_time_us(self):
        return lib().scs_graph_core_time_us(self._h)

    def close(self):
        if self._h:
            lib().scs_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class QueryResult:
    tag: int
    l: int
    h: int
    q: int
    kl: int
    heu_kl: int
    ku: int
    status: str
    kept: int
    nodes: int
    latency_us: float
    witness: np.ndarray

    def line(self):
        """Reference results-file row N1|N2|QID|Time|Degree|Overtime|Heu|0
        (main.cu:123-128, 328-333); Overtime 1 = node budget reached."""
        over = 1 if self.status == "CAPPED" else 0
        heu = 1 if self.status == "HEU_OPT" else 0
        return f"{self.l}|{self.h}|{self.q}|{_commas(int(self.latency_us))}|{self.kl}|{over}|{heu}|0"


def _commas(x: int) -> str:
    return f"{x:,}"


def heuristic_gpu(graph: "Graph", queries, cap=MAX_H):
    """CSSC_heu on the GPU for a list of (l, h, q): [(kl, H)] per query, H as
    a uint32 array (cut at cap ids)."""
    n = len(queries)
    if n == 0:
        return []
    l = np.ascontiguousarray([x[0] for x in queries], np.uint32)
    h = np.ascontiguousarray([x[1] for x in queries], np.uint32)
    q = np.ascontiguousarray([x[2] for x in queries], np.uint32)
    kl = np.zeros(n, np.uint32)
    hs = np.zeros(n, np.uint32)
    H = np.zeros(n * cap, np.uint32)
    f = lib().scs_heuristic_gpu
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    f.argtypes = [C.c_void_p, C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p, C.c_uint32]
    _check(f(graph._h, n, l, h, q, kl, hs, H, cap), "scs_heuristic_gpu")
    return [(int(kl[i]), H[i * cap:i * cap + min(int(hs[i]), cap)].copy()) for i in range(n)]


def model_init(seed=0):
    m = Model()
    _check(lib().scs_model_init(C.byref(m), seed), "scs_model_init")
    return m


def mlp_grad(model, X, y, pos_weight=1.0):
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = Model()
    loss = C.c_float()
    _check(lib().scs_mlp_grad(C.byref(model), X, y, len(y), pos_weight, C.byref(g),
                              C.byref(loss)), "scs_mlp_grad")
    return g, loss.value


def mlp_train(model, X, y, epochs=10, batch=1024, lr=1e-3, pos_weight=1.0, seed=0):
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    loss = C.c_float()
    _check(lib().scs_mlp_train(C.byref(model), X, y, len(y), epochs, batch, lr, pos_weight,
                               seed, C.byref(loss)), "scs_mlp_train")
    return loss.value


def mlp_infer(model, X):
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros(len(X), np.float32)
    _check(lib().scs_mlp_infer(C.byref(model), X, len(X), out), "scs_mlp_infer")
    return out


class Engine:
    """Search engine over one graph (see scs_params in include/scs.h)."""

    def __init__(self, graph: Graph, **kw):
        p = Params()
        lib().scs_params_default(C.byref(p))
        model = kw.pop("model", None)
        if model is not None:
            self._model_ref = model
            p.model = C.cast(C.pointer(model), C.c_void_p)
        for k, v in kw.items():
            if not hasattr(p, k):
                raise TypeError(f"unknown engine parameter {k}")
            setattr(p, k, v)
        self.graph = graph
        self.params = p
        h = C.c_void_p()
        _check(lib().scs_engine_create(graph._h, C.byref(p), C.byref(h)), "scs_engine_create")
        self._h = h
        self._buf = (Result * 256)()

    def submit(self, l, h, q, heu=1, limit_doms=8, tag=0):
        return _check(lib().scs_submit(self._h, l, h, q, heu, limit_doms, tag), "scs_submit")

    def submit_text(self, text, tag=0):
        return _check(lib().scs_submit_text(self._h, text.encode(), tag), "scs_submit_text")

    def step(self):
        return _check(lib().scs_step(self._h), "scs_step")

    def poll(self):
        out = []
        while True:
            k = _check(lib().scs_poll(self._h, self._buf, 256), "scs_poll")
            for i in range(k):
                r = self._buf[i]
                out.append(QueryResult(
                    tag=r.tag, l=r.l, h=r.h, q=r.q, kl=r.kl, heu_kl=r.heu_kl, ku=r.ku,
                    status=STATUS.get(r.status, str(r.status)), kept=r.kept, nodes=r.nodes,
                    latency_us=r.t_done_us - r.t_submit_us,
                    witness=np.ctypeslib.as_array(r.witness)[:r.wsize].copy()))
            if k < 256:
                return out

    def drain(self):
        return _check(lib().scs_drain(self._h), "scs_drain")

    def set_model(self, model, learned=True):
        _check(lib().scs_engine_set_model(self._h, C.byref(model) if model is not None else None,
                                          1 if learned else 0), "scs_engine_set_model")

    def trace(self, capacity_rows, stride=1):
        _check(lib().scs_engine_trace(self._h, capacity_rows, stride), "scs_engine_trace")
        self._trace_cap = capacity_rows

    def heuristic(self, l, h, q):
        """(kl, ubD): the admission lower bound and BFS depth bound of a query."""
        kl, ubd = C.c_uint32(), C.c_uint32()
        _check(lib().scs_engine_heuristic(self._h, l, h, q, C.byref(kl), C.byref(ubd)),
               "heuristic")
        return kl.value, ubd.value

    def prep_many(self, q, kl0, ubd):
        """Batched preprocessing of up to 64 queries (local graphs built on the
        device, then released); returns (K, nnz) per query."""
        n = len(q)
        q = np.ascontiguousarray(q, np.uint32)
        kl0 = np.ascontiguousarray(kl0, np.uint32)
        ubd = np.ascontiguousarray(ubd, np.uint32)
        K = np.zeros(n, np.uint32)
        nnz = np.zeros(n, np.uint64)
        f = lib().scs_prep_probe_many
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        f.argtypes = [C.c_void_p, C.c_uint32, u32p, u32p, u32p, u32p,
                      np.ctypeslib.ndpointer(np.uint64, flags="C"), C.c_void_p, C.c_void_p,
                      C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]
        _check(f(self._h, n, q, kl0, ubd, K, nnz, None, None, None, 0, None, 0), "prep_many")
        return K, nnz

    def trace_read(self):
        cap = getattr(self, "_trace_cap", 0)
        rows = np.zeros((max(cap, 1), 8), np.uint32)
        n = C.c_uint32()
        _check(lib().scs_engine_trace_read(self._h, rows, cap, C.byref(n)), "trace_read")
        return rows[:n.value].copy()

    def stats(self):
        s = Stats()
        _check(lib().scs_engine_stats(self._h, C.byref(s)), "stats")
        out = {f: getattr(s, f) for f, _ in Stats._fields_}
        out["phase_ticks"] = list(out["phase_ticks"])
        return out

    def close(self):
        if self._h:
            lib().scs_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, queries, tags=None, progress=None):
        """Submits all queries, drains the engine and returns results by tag.
        progress: seconds between progress lines on stderr (None = quiet)."""
        import sys
        import time
        for i, qq in enumerate(queries):
            l, h, q, heu, ld = qq
            self.submit(l, h, q, heu, ld, i if tags is None else tags[i])
        res = []
        last = time.time()
        while True:
            left = self.step()
            res.extend(self.poll())
            if progress is not None and time.time() - last > progress:
                last = time.time()
                print(f"[scs] {len(res)}/{len(queries)} queries done", file=sys.stderr, flush=True)
            if left == 0:
                break
        res.extend(self.poll())
        res.sort(key=lambda r: r.tag)
        return res
