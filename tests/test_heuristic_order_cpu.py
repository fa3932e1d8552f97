"""The selection rule the GPU heuristic (scs_heu.hip) is built on, checked
on the CPU.  The reference's heu2/heu3/heu4 (inc/heuristic.h:8-317) pop a
std::priority_queue that keeps stale duplicates; the GPU kernel replaces
each pop by one deterministic choice:
  heu2  the frontier vertex with the largest (H-neighbour count, id);
  heu3  the frontier vertex with the largest (largest score pushed, id);
  heu4  the member of S with the smallest (current degree in S, id).
The restatement below makes exactly those choices, with no heap, and must
give the oracle's kl and H (the oracle keeps the reference's heaps) on
random graphs whose many equal keys exercise the tie rules."""

import numpy as np
import pytest

from oracle.oracle import OracleGraph
from tests.helpers import random_graph


def nbrs(off, adj, v):
    return [int(x) for x in adj[off[v]:off[v + 1]]]


def heu2(off, adj, N1, N2, q):
    sta, cnt, deg = {q: 1}, {}, {}
    cand, tH, tkl, tsize = [q], [], 0, 0
    while True:
        F = [x for x in cand if sta.get(x) == 1]
        if not F:
            break
        v = max(F, key=lambda x: (cnt.get(x, 0), x))
        tH.append(v)
        sta[v] = 2
        for nb in nbrs(off, adj, v):
            cnt[nb] = cnt.get(nb, 0) + 1
            if sta.get(nb, 0) == 2:
                deg[nb] = deg.get(nb, 0) + 1
                deg[v] = deg.get(v, 0) + 1
            elif sta.get(nb, 0) == 0:
                sta[nb] = 1
                cand.append(nb)
        if len(tH) >= N1:
            mn = min(deg.get(x, 0) for x in tH)
            if mn >= tkl:
                tkl, tsize = mn, len(tH)
        if len(tH) == N2:
            break
    return tkl, tH[:tsize]


def heu3(off, adj, N1, N2, q, dmax):
    sta, key, deg = {q: 1}, {q: 0.0}, {}
    tH, tkl, tsize = [], 0, 0
    while True:
        F = [x for x in key if sta.get(x) == 1]
        if not F:
            break
        v = max(F, key=lambda x: (key[x], x))
        tH.append(v)
        sta[v] = 2
        N = nbrs(off, adj, v)
        for nb in N:
            if sta.get(nb, 0) == 2:
                deg[nb] = deg.get(nb, 0) + 1
                deg[v] = deg.get(v, 0) + 1
        for nb in N:
            if sta.get(nb, 0) == 2:
                continue
            sc = 0.0
            for w in nbrs(off, adj, nb):  # ascending ids
                if sta.get(w, 0) == 2 and deg.get(w, 0) != 0:
                    sc += 1.0 / deg[w]
            sc += (off[nb + 1] - off[nb]) / dmax
            key[nb] = max(key[nb], sc) if nb in key else sc
            if sta.get(nb, 0) == 0:
                sta[nb] = 1
        if len(tH) >= N1:
            mn = min(deg.get(x, 0) for x in tH)
            if mn >= tkl:
                tkl, tsize = mn, len(tH)
        if len(tH) == N2:
            break
    return tkl, tH[:tsize]


def heu4(off, adj, N1, N2, q):
    Nq = nbrs(off, adj, q)
    if len(Nq) < N1 - 1:
        return 0, []
    S = [q] + Nq
    inS = set(S)
    dg = {x: sum(1 for y in nbrs(off, adj, x) if y in inS) for x in S}
    dg[q] += 1
    rmv, rmv_idx, mindeg = [], 0, 0
    while True:
        live = [x for x in S if x in inS]
        if not live:
            break
        v = min(live, key=lambda x: (dg[x], x))
        d = dg[v] - (1 if v == q else 0)
        remain = len(S) - len(rmv)
        if N1 <= remain <= N2 and d > mindeg:
            mindeg, rmv_idx = d, len(rmv)
        if remain == N1:
            break
        inS.discard(v)
        rmv.append(v)
        for x in nbrs(off, adj, v):
            if x in inS:
                dg[x] -= 1
    gone = set(rmv[:rmv_idx])
    return mindeg, [x for x in S if x not in gone]


@pytest.mark.parametrize("seed", range(6))
def test_choice_rules_give_the_reference_heaps_results(seed):
    rng = np.random.default_rng(500 + seed)
    checked = 0
    for trial in range(25):
        n = int(rng.integers(6, 40))
        off, adj = random_graph(rng, n, float(rng.uniform(0.08, 0.6)))
        if len(adj) == 0:
            continue
        og = OracleGraph.from_csr(n, off, adj)
        dmax = float(max(off[v + 1] - off[v] for v in range(n)))
        for _ in range(4):
            q = int(rng.integers(0, n))
            l = int(rng.integers(1, min(n, 12) + 1))
            h = l + int(rng.integers(0, 8))
            got = {2: heu2(off, adj, l, h, q), 3: heu3(off, adj, l, h, q, dmax),
                   4: heu4(off, adj, l, h, q)}
            for which in (2, 3, 4):
                kl, H = og.heuristic(which, l, h, q)
                assert got[which][0] == kl, (seed, trial, which, l, h, q)
                assert got[which][1] == [int(x) for x in H], (seed, trial, which, l, h, q)
            checked += 1
    assert checked >= 50
