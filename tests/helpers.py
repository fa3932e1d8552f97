"""Independent checkers used by the tests: brute force SCS on small graphs,
witness validation and small graph builders.  None of this calls the oracle
or the library."""

from __future__ import annotations

import itertools
from collections import deque

import numpy as np


def csr_from_edges(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    off = np.zeros(n + 1, np.int64)
    lst = []
    for v in range(n):
        s = sorted(adj[v])
        lst.extend(s)
        off[v + 1] = off[v] + len(s)
    return off, np.array(lst, dtype=np.int32)


def neighbours(off, adj, v):
    return adj[off[v]:off[v + 1]]


def min_degree(off, adj, H):
    Hs = set(int(x) for x in H)
    return min(sum(1 for w in neighbours(off, adj, v) if int(w) in Hs) for v in Hs)


def connected(off, adj, H):
    Hs = set(int(x) for x in H)
    if not Hs:
        return False
    start = next(iter(Hs))
    seen = {start}
    dq = deque([start])
    while dq:
        v = dq.popleft()
        for w in neighbours(off, adj, v):
            w = int(w)
            if w in Hs and w not in seen:
                seen.add(w)
                dq.append(w)
    return len(seen) == len(Hs)


def check_witness(off, adj, q, l, h, kl, H):
    """A witness H must contain q, be connected, satisfy l <= |H| <= h and
    have minimum degree exactly kl."""
    H = [int(x) for x in H]
    assert len(set(H)) == len(H), "duplicate vertex in witness"
    assert q in H, "witness misses q"
    assert l <= len(H) <= h, f"|H|={len(H)} outside [{l},{h}]"
    assert connected(off, adj, H), "witness not connected"
    assert min_degree(off, adj, H) == kl, "witness min degree differs from kl"


def brute_force(off, adj, l, h, q):
    """Max min degree over connected H containing q with l <= |H| <= h by
    enumerating every subset (small n only).  None when infeasible."""
    n = len(off) - 1
    others = [v for v in range(n) if v != q]
    best = None
    for size in range(max(l, 1), min(h, n) + 1):
        for comb in itertools.combinations(others, size - 1):
            H = (q,) + comb
            if not connected(off, adj, H):
                continue
            d = min_degree(off, adj, H)
            if best is None or d > best:
                best = d
    return best


def random_graph(rng, n, p):
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
    return csr_from_edges(n, edges)
