"""Deterministic synthetic graphs and query streams for the SCS workloads.

The reference ships only three small graphs (data/graph.txt, data/100.txt,
data/hepPH_SCS) and evaluates on SNAP graphs that are not shipped.  The
workloads used here are therefore generated: Chung-Lu power-law graphs and
Graph500-style R-MAT graphs, both from numpy's PCG64 so that a seed fixes
the graph bit for bit.  Every graph is returned as a simple undirected CSR
(self-loops and duplicates removed, adjacency sorted ascending), the same
shape ``load_graph`` builds in the reference (src/Graph.h:5-67).

Query streams use a separate PCG64 stream seeded with ``graph_seed + 1000``.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np


@dataclass
class CSR:
    """Undirected simple graph in CSR form (both directions stored)."""

    n: int
    off: np.ndarray  # int64[n+1]
    adj: np.ndarray  # int32[off[n]]

    @property
    def m(self) -> int:
        return int(self.off[-1]) // 2

    @property
    def degree(self) -> np.ndarray:
        return np.diff(self.off).astype(np.int64)

    def sha256(self) -> str:
        h = hashlib.sha256()
        h.update(np.int64(self.n).tobytes())
        h.update(np.ascontiguousarray(self.off, dtype=np.int64).tobytes())
        h.update(np.ascontiguousarray(self.adj, dtype=np.int32).tobytes())
        return h.hexdigest()


def csr_from_pairs(n: int, u: np.ndarray, v: np.ndarray) -> CSR:
    """Builds the simple undirected CSR of the pairs (u, v).

    Self-loops are dropped and duplicates merged, as ``load_graph`` does with
    its ``map<ui, set<ui>>`` (src/Graph.h:19-23); adjacency lists come out
    sorted ascending like the reference's ``set`` iteration order.
    """
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    keep = u != v
    u, v = u[keep], v[keep]
    if u.size and (min(u.min(), v.min()) < 0 or max(u.max(), v.max()) >= n):
        raise ValueError("vertex id out of range")
    a = np.minimum(u, v)
    b = np.maximum(u, v)
    key = np.unique(a * n + b)
    a = key // n
    b = key % n
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    order = np.argsort(src * n + dst, kind="stable")
    src = src[order]
    dst = dst[order]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.bincount(src, minlength=n)
    np.cumsum(off, out=off)
    return CSR(n=n, off=off, adj=dst.astype(np.int32))


def _unique_pairs_in_order(n: int, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Keys min*n+max of the distinct non-loop pairs, in first-draw order."""
    keep = u != v
    u, v = u[keep], v[keep]
    key = np.minimum(u, v) * n + np.maximum(u, v)
    _, first = np.unique(key, return_index=True)
    first.sort()
    return key[first]


def chung_lu(n: int, m: int, beta: float, dmax: float, seed: int) -> CSR:
    """Chung-Lu graph with power-law expected degrees.

    Weights w_i = c * (i + i0) ** (-1/(beta-1)); i0 and c are fixed so that
    w_0 == dmax and sum(w) == 2*m.  Endpoints are drawn independently with
    probability proportional to w from PCG64(seed).  Draws continue in
    rounds until m distinct undirected non-loop edges exist; the first m
    distinct edges in draw order are kept, so the edge count is exactly m.
    """
    alpha = 1.0 / (beta - 1.0)
    idx = np.arange(n, dtype=np.float64)

    def total(i0: float) -> float:
        return float(np.sum(dmax * (i0 / (idx + i0)) ** alpha))

    lo, hi = 1e-9, 1e12
    if total(hi) < 2 * m:
        raise ValueError("dmax too small for the requested edge count")
    for _ in range(200):
        mid = np.sqrt(lo * hi)
        if total(mid) < 2 * m:
            lo = mid
        else:
            hi = mid
    i0 = hi
    w = dmax * (i0 / (idx + i0)) ** alpha
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    rng = np.random.Generator(np.random.PCG64(seed))
    keys = np.zeros(0, dtype=np.int64)
    draw = m
    while True:
        u = np.searchsorted(cdf, rng.random(draw), side="right")
        v = np.searchsorted(cdf, rng.random(draw), side="right")
        u = np.minimum(u, n - 1).astype(np.int64)
        v = np.minimum(v, n - 1).astype(np.int64)
        new = _unique_pairs_in_order(n, u, v)
        # merge preserving first-seen order
        allk = np.concatenate([keys, new])
        _, first = np.unique(allk, return_index=True)
        first.sort()
        keys = allk[first]
        if keys.size >= m:
            keys = keys[:m]
            break
        draw = int((m - keys.size) * 1.5) + 1024
    return csr_from_pairs(n, keys // n, keys % n)


def rmat(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int,
         chunk: int = 1 << 22) -> CSR:
    """Graph500-style R-MAT graph, symmetrised and deduplicated.

    edge_factor * 2**scale directed edges are drawn; at each of the ``scale``
    levels one uniform draw picks the quadrant (a | b | c | d=1-a-b-c).
    Vertex labels are not permuted.
    """
    n = 1 << scale
    m_raw = edge_factor * n
    rng = np.random.Generator(np.random.PCG64(seed))
    us, vs = [], []
    ab = a + b
    abc = a + b + c
    for s in range(0, m_raw, chunk):
        cnt = min(chunk, m_raw - s)
        u = np.zeros(cnt, dtype=np.int64)
        v = np.zeros(cnt, dtype=np.int64)
        for lvl in range(scale):
            r = rng.random(cnt)
            bit = np.int64(1) << (scale - 1 - lvl)
            down = r >= ab  # quadrants c, d: row bit set
            right = ((r >= a) & (r < ab)) | (r >= abc)  # quadrants b, d
            u += down * bit
            v += right * bit
        us.append(u)
        vs.append(v)
    return csr_from_pairs(n, np.concatenate(us), np.concatenate(vs))


def core_numbers(g: CSR) -> np.ndarray:
    """Core number of every vertex by level-synchronous peeling.

    Core numbers of a graph are unique, so this agrees with the reference's
    linear-heap peeling ``core_decomposition_linear_list``
    (src/Graph.h:82-109).  At level k every vertex whose remaining degree is
    <= k is removed in one vectorised step and its neighbours' degrees are
    decremented, until none is left at that level.
    """
    n = g.n
    deg = g.degree.astype(np.int64).copy()
    core = np.zeros(n, dtype=np.int32)
    alive = np.ones(n, dtype=bool)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(g.off))
    adj = g.adj.astype(np.int64)
    left = n
    k = 0
    while left:
        k = max(k, int(deg[alive].min()))
        while True:
            rem = np.flatnonzero(alive & (deg <= k))
            if rem.size == 0:
                break
            core[rem] = k
            alive[rem] = False
            left -= rem.size
            mask = np.zeros(n, dtype=bool)
            mask[rem] = True
            sel = mask[src]
            tgt = adj[sel]
            tgt = tgt[alive[tgt]]
            deg -= np.bincount(tgt, minlength=n)
        k += 1
    return core


@dataclass
class Query:
    l: int
    h: int
    q: int
    heu: int
    limit_doms: int

    def text(self) -> str:
        return f"{self.l} {self.h} {self.q} {self.heu} {self.limit_doms}"


def make_queries(core: np.ndarray, count: int, seed: int, *, core_min: int,
                 l_lo: int, l_hi: int, dh_lo: int, dh_hi: int,
                 heu=1, limit_doms=8, skew_core: int | None = None,
                 skew_p: float = 0.0) -> list[Query]:
    """Query stream: q uniform over {core >= core_min}; l ~ U[l_lo, l_hi];
    h = l + U[dh_lo, dh_hi] (inclusive bounds).  ``heu`` and ``limit_doms``
    are either constants or sequences to draw uniformly from.  With
    ``skew_core`` set, q is drawn from {core >= skew_core} with probability
    ``skew_p`` and from {core >= core_min} otherwise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pool = np.flatnonzero(core >= core_min)
    if pool.size == 0:
        raise ValueError("no vertex satisfies the core bound")
    hi_pool = np.flatnonzero(core >= skew_core) if skew_core is not None else pool
    out = []
    for _ in range(count):
        if skew_core is not None and hi_pool.size and rng.random() < skew_p:
            q = int(hi_pool[rng.integers(0, hi_pool.size)])
        else:
            q = int(pool[rng.integers(0, pool.size)])
        l = int(rng.integers(l_lo, l_hi, endpoint=True))
        h = l + int(rng.integers(dh_lo, dh_hi, endpoint=True))
        hh = int(rng.choice(heu)) if np.ndim(heu) else int(heu)
        ld = int(rng.choice(limit_doms)) if np.ndim(limit_doms) else int(limit_doms)
        out.append(Query(l, h, q, hh, ld))
    return out


# Binary CSR cache.  One directory entry per parameter set: <key>.npz holds
# n, off (int64), adj (int32) and the CSR's SHA-256; the key is the SHA-256
# of the generator parameters (sorted JSON) and CACHE_VERSION, so a change of
# generator code that changes graphs only needs a new version string.
CACHE_VERSION = "scs-csr-1"


def cache_key(params: dict) -> str:
    import json
    blob = json.dumps({"v": CACHE_VERSION, **params}, sort_keys=True).encode()
    return hashlib.sha256(blob).hexdigest()[:32]


def cache_store(directory: str, params: dict, g: CSR) -> str:
    """Writes g under the parameters' key (atomically: temp file, rename)."""
    import os
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, cache_key(params) + ".npz")
    tmp = path + f".tmp{os.getpid()}.npz"
    np.savez(tmp, n=np.int64(g.n), off=np.ascontiguousarray(g.off, np.int64),
             adj=np.ascontiguousarray(g.adj, np.int32),
             sha256=np.frombuffer(bytes.fromhex(g.sha256()), np.uint8))
    os.replace(tmp, path)
    return path


def cache_load(directory: str, params: dict, verify: bool = True) -> CSR | None:
    """The cached CSR of these parameters, or None when absent or when its
    stored SHA-256 does not match its contents (the entry is then ignored)."""
    import os
    path = os.path.join(directory, cache_key(params) + ".npz")
    if not os.path.exists(path):
        return None
    try:
        with np.load(path, allow_pickle=False) as z:
            g = CSR(n=int(z["n"]), off=z["off"], adj=z["adj"])
            want = bytes(z["sha256"]).hex()
    except (OSError, ValueError, KeyError):
        return None
    if verify and g.sha256() != want:
        return None
    return g


def write_edge_list(g: CSR, path: str) -> None:
    """Writes the reference's edge-list format: header ``n m`` then one
    tab-separated undirected edge per line (src/Graph.h:16-23)."""
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.off))
    dst = g.adj.astype(np.int64)
    keep = src < dst
    with open(path, "w") as f:
        f.write(f"{g.n}\t{int(keep.sum())}\n")
        np.savetxt(f, np.stack([src[keep], dst[keep]], axis=1), fmt="%d", delimiter="\t")
