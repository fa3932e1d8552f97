"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU-baseline leg import this
module.  It loads oracle/liboracle.so, built from oracle/scs_oracle.cpp by
oracle/Makefile.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FLAGS_ALL = 0x3F  # red1 red2 red3 prun1 prun2 doms
FLAG_RECOMPUTE = 0x40  # use the recompute solver (same nodes as the trail solver)
FLAG_LEVELS = 0x80  # descending level floors from ku-1 (same optimum, different search)
FLAG_PAIR = 0x100  # common-neighbour pair rule (trail solver)

STATUS = {0: "OK", 1: "HEU_OPT", 2: "CAPPED", 3: "INFEASIBLE", -1: "INVALID"}


class OrcResult(C.Structure):
    _fields_ = [("kl", C.c_uint32), ("heu_kl", C.c_uint32), ("ku", C.c_uint32),
                ("status", C.c_int32), ("nodes", C.c_uint64), ("K", C.c_uint32),
                ("ubD", C.c_uint32), ("wsize", C.c_uint32), ("levels", C.c_uint32)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        L.orc_graph_from_csr.restype = P
        L.orc_graph_from_csr.argtypes = [C.c_uint32, np.ctypeslib.ndpointer(np.int64, flags="C"),
                                         np.ctypeslib.ndpointer(np.int32, flags="C")]
        L.orc_graph_load.restype = P
        L.orc_graph_load.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.orc_graph_free.argtypes = [P]
        L.orc_graph_n.restype = C.c_uint32
        L.orc_graph_n.argtypes = [P]
        L.orc_graph_nnz.restype = C.c_int64
        L.orc_graph_nnz.argtypes = [P]
        L.orc_graph_dmax.restype = C.c_uint32
        L.orc_graph_dmax.argtypes = [P]
        L.orc_graph_csr.argtypes = [P, np.ctypeslib.ndpointer(np.int64, flags="C"),
                                    np.ctypeslib.ndpointer(np.int32, flags="C")]
        L.orc_graph_core.argtypes = [P, u32p]
        L.orc_bfs.argtypes = [P, C.c_uint32, u32p]
        L.orc_ubd_host.restype = C.c_uint32
        L.orc_ubd_host.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_ubd_node.restype = C.c_uint32
        L.orc_ubd_node.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_unr_bound.restype = C.c_uint32
        L.orc_unr_bound.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32]
        L.orc_heuristic.restype = C.c_uint32
        L.orc_heuristic.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p,
                                    C.c_uint32, C.POINTER(C.c_uint32)]
        L.orc_preprocess.restype = C.c_uint32
        L.orc_preprocess.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_uint32,
                                     u32p, u32p, C.c_uint64, u32p, C.POINTER(C.c_uint64)]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_uint64, C.POINTER(OrcResult), u32p, C.c_uint32]
        L.orc_solve_timed.restype = C.c_int
        L.orc_solve_timed.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_uint64, C.c_double,
                                      C.POINTER(OrcResult), u32p, C.c_uint32]
        L.orc_solve_plain.restype = C.c_uint32
        L.orc_solve_plain.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.POINTER(C.c_uint64)]
        _LIB = L
    return _LIB


class OracleGraph:
    """A graph held by the oracle, with its core numbers."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_csr(cls, n, off, adj):
        return cls(lib().orc_graph_from_csr(n, np.ascontiguousarray(off, np.int64),
                                            np.ascontiguousarray(adj, np.int32)))

    @classmethod
    def load(cls, path):
        err = C.c_int(0)
        h = lib().orc_graph_load(path.encode(), C.byref(err))
        if not h:
            raise IOError(f"oracle loader failed ({err.value}) on {path}")
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_graph_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().orc_graph_n(self.h)

    @property
    def dmax(self):
        return lib().orc_graph_dmax(self.h)

    def csr(self):
        n = self.n
        off = np.zeros(n + 1, np.int64)
        adj = np.zeros(lib().orc_graph_nnz(self.h), np.int32)
        lib().orc_graph_csr(self.h, off, adj)
        return off, adj

    def core(self):
        c = np.zeros(self.n, np.uint32)
        lib().orc_graph_core(self.h, c)
        return c

    def bfs(self, q):
        d = np.zeros(self.n, np.uint32)
        lib().orc_bfs(self.h, q, d)
        return d

    def heuristic(self, which, l, h, q):
        cap = max(h, 1) + 1 + self.n if which == 4 else max(h, 1) + 1
        w = np.zeros(cap, np.uint32)
        ws = C.c_uint32(0)
        kl = lib().orc_heuristic(self.h, which, l, h, q, w, cap, C.byref(ws))
        return int(kl), w[:min(ws.value, cap)].copy()

    def preprocess(self, q, kl0, ubd):
        cap = self.n
        kept = np.zeros(cap, np.uint32)
        loff = np.zeros(cap + 1, np.uint32)
        ldist = np.zeros(cap, np.uint32)
        nnz = C.c_uint64(0)
        adj_cap = int(lib().orc_graph_nnz(self.h))
        ladj = np.zeros(max(adj_cap, 1), np.uint32)
        K = lib().orc_preprocess(self.h, q, kl0, ubd, kept, cap, loff, ladj, adj_cap, ldist,
                                 C.byref(nnz))
        return dict(K=K, kept=kept[:K].copy(), off=loff[:K + 1].copy(),
                    adj=ladj[:nnz.value].copy(), dist=ldist[:K].copy())

    def solve(self, l, h, q, heu=1, limit_doms=8, flags=FLAGS_ALL, node_cap=1 << 40,
              time_limit=0.0):
        """time_limit: wall-clock seconds after which the search ends CAPPED
        (0 = none)."""
        r = OrcResult()
        wcap = max(h, 1) + self.n
        w = np.zeros(wcap, np.uint32)
        lib().orc_solve_timed(self.h, l, h, q, heu, limit_doms, flags, node_cap, float(time_limit),
                              C.byref(r), w, wcap)
        return dict(kl=r.kl, heu_kl=r.heu_kl, ku=r.ku, status=STATUS[r.status], nodes=r.nodes,
                    K=r.K, ubD=r.ubD, levels=r.levels, witness=w[:min(r.wsize, wcap)].copy())

    def solve_plain(self, l, h, q, node_cap=1 << 40):
        nodes = C.c_uint64(0)
        v = lib().orc_solve_plain(self.h, l, h, q, node_cap, C.byref(nodes))
        return (None if v == 0xFFFFFFFF else int(v)), nodes.value


def ubd_host(kl, h):
    return int(lib().orc_ubd_host(kl, h))


def unr_bound(cvals, rvals, k):
    c = np.ascontiguousarray(cvals, np.uint32)
    r = np.ascontiguousarray(rvals, np.uint32)
    if r.size == 0:
        r = np.zeros(1, np.uint32)
        return int(lib().orc_unr_bound(c, c.size, r, 0, k))
    return int(lib().orc_unr_bound(c, c.size, r, r.size, k))


def ubd_node(kl, h):
    return int(lib().orc_ubd_node(kl, h))
