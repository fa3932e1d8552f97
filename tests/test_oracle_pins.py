"""Pins for the oracle functions that the optimum checks leave open.

- Per-query preprocessing (kept set, local CSR, distances) against a
  construction made with networkx alone: the subgraph induced by
  {core > kl0} (plus q), the vertices within ubD hops of q in it, the
  induced subgraph on those vertices relabelled by rank of original id
  (reference semantics: src/helpers.cc:232-233 keeps v iff core > kl and
  dist <= ubD; src/helpers.cc:317-368 builds the kept vertices' lists).
- The heuristics heu2/heu3/heu4 and CSSC_heu (inc/heuristic.h:8-348):
  every returned H is a valid witness of its kl and kl never exceeds the
  brute-force optimum; on hand-built graphs the greedy order, including
  std::priority_queue's tie rule, is stepped through by hand in comments and
  the expected H written down from that.
- The diameter bound ubD (main.cu:154-172) against its definition: the
  largest diameter of a connected graph with at most h vertices and minimum
  degree at least kl, taken over every graph of networkx's atlas (all graphs
  with up to 7 vertices).
"""

import networkx as nx
import numpy as np
import pytest

from oracle.oracle import OracleGraph, ubd_host, ubd_node
from scs_amd import configs
from tests.helpers import brute_force, check_witness, csr_from_edges, random_graph

import os

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def nx_graph(off, adj):
    G = nx.Graph()
    n = len(off) - 1
    G.add_nodes_from(range(n))
    for v in range(n):
        for w in adj[off[v]:off[v + 1]]:
            if v < int(w):
                G.add_edge(v, int(w))
    return G


def nx_preprocess(G, core, q, kl0, ubd):
    S = [v for v in G.nodes if core[v] > kl0 or v == q]
    GS = G.subgraph(S)
    dist = nx.single_source_shortest_path_length(GS, q, cutoff=ubd)
    kept = sorted(dist)
    rank = {v: i for i, v in enumerate(kept)}
    H = G.subgraph(kept)
    off = [0]
    adj = []
    for v in kept:
        adj.extend(sorted(rank[w] for w in H.neighbors(v)))
        off.append(len(adj))
    return kept, off, adj, [dist[v] for v in kept]


def graph_cases():
    out = []
    for name in ("100.txt", "hepPH_SCS"):
        og = OracleGraph.load(os.path.join(GOLD, name))
        out.append((name, og))
    g = configs.graph("C1")
    out.append(("C1", OracleGraph.from_csr(g.n, g.off, g.adj)))
    return out


@pytest.mark.parametrize("idx", range(3))
def test_preprocess_matches_networkx(idx):
    name, og = graph_cases()[idx]
    off, adj = og.csr()
    G = nx_graph(off, adj)
    core = og.core()
    rng = np.random.default_rng(11 + idx)
    cand = np.flatnonzero(core >= 2)
    for t in range(12):
        q = int(cand[rng.integers(0, cand.size)])
        kl0 = int(rng.integers(0, int(core[q])))
        h = int(rng.integers(kl0 + 2, kl0 + 40))
        ubd = ubd_host(kl0, h) if t % 3 else int(rng.integers(1, 4))
        got = og.preprocess(q, kl0, ubd)
        kept, loff, ladj, dist = nx_preprocess(G, core, q, kl0, ubd)
        assert got["K"] == len(kept), (name, q, kl0, ubd)
        assert list(got["kept"]) == kept
        assert list(got["off"]) == loff
        assert list(got["adj"]) == ladj
        assert list(got["dist"]) == dist


def test_heuristics_give_valid_witnesses_below_the_optimum():
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(150):
        n = int(rng.integers(5, 12))
        off, adj = random_graph(rng, n, float(rng.uniform(0.25, 0.75)))
        og = OracleGraph.from_csr(n, off, adj)
        q = int(rng.integers(0, n))
        l = int(rng.integers(1, n + 1))
        h = int(rng.integers(l, n + 2))
        bf = brute_force(off, adj, l, h, q)
        opt = 0 if bf is None else bf
        for which in (2, 3, 4, 0):
            kl, H = og.heuristic(which, l, h, q)
            assert kl <= opt, (trial, which, kl, opt)
            if kl > 0:
                check_witness(off, adj, q, l, h, kl, H)
                checked += 1
    assert checked > 100


def test_heu2_tie_goes_to_the_larger_vertex():
    # Edges 0-1 0-2 1-3 2-3 2-4, q = 0, N1 = N2 = 3.
    # pop 0: push (1, 1) and (1, 2): each has one neighbour (0) in H.
    # pop (1, 2): key tie, std::priority_queue<pair> takes the larger id.
    #   pushes (1, 3) (neighbour 2 in H) and (1, 4).
    # pop (1, 4): tie among (1,1) (1,3) (1,4), larger id again.
    # |H| = 3: degrees inside {0, 2, 4} are 1, 2, 1 -> kl 1.
    off, adj = csr_from_edges(5, [(0, 1), (0, 2), (1, 3), (2, 3), (2, 4)])
    og = OracleGraph.from_csr(5, off, adj)
    kl, H = og.heuristic(2, 3, 3, 0)
    assert (kl, list(H)) == (1, [0, 2, 4])


def test_heu3_tie_goes_to_the_larger_vertex():
    # 5-cycle 0-1-3-4-2-0, every degree 2 so dMAX = 2; q = 0, N1 = N2 = 3.
    # pop 0 (deg 0 inside H, contributes nothing): push 1 and 2 with score
    #   degree/dMAX = 1.0 each; tie -> 2 first.
    # pop 2: edge 0-2 gives deg[0] = deg[2] = 1; push 4 with score
    #   1/deg[2] + 2/2 = 2.0; re-push 0? no (in H); 1 keeps 1.0.
    # pop (2.0, 4): edge 2-4 -> deg[2] = 2, deg[4] = 1.
    # |H| = 3: min(1, 2, 1) = 1 -> kl 1, H = [0, 2, 4].
    off, adj = csr_from_edges(5, [(0, 1), (1, 3), (3, 4), (4, 2), (2, 0)])
    og = OracleGraph.from_csr(5, off, adj)
    kl, H = og.heuristic(3, 3, 3, 0)
    assert (kl, list(H)) == (1, [0, 2, 4])


def test_heu4_peels_the_smallest_degree_then_smallest_id():
    # q = 0 with N(q) = {1, 2, 3, 4}; among them edges 1-2, 2-3, 3-4, 1-4
    # (a 4-cycle), plus 1-3.  Vertex 5 hangs off 1 and stays outside S.
    # S = q then N(q) ascending = [0, 1, 2, 3, 4]; degrees inside S are
    # 0:4 (key 5 with the +1 on q), 1:4, 2:3, 3:4, 4:3.  N1 = 3, N2 = 5.
    # Min-heap on (key, id), so ties go to the smaller id:
    #   pop (3, 2): remain 5 in [3, 5] and 3 > 0 -> mindeg 3 with nothing
    #     removed; remove 2 -> keys 0:4, 1:3, 3:3.
    #   pop (3, 1): remain 4, 3 > 3 is false; remove 1 -> 0:3, 3:2, 4:2.
    #   pop (2, 3): remain 3 == N1 -> stop.
    # The best set is all of S, min degree 3.
    edges = [(0, v) for v in (1, 2, 3, 4)] + [(1, 2), (2, 3), (3, 4), (1, 4), (1, 3), (1, 5)]
    off, adj = csr_from_edges(6, edges)
    og = OracleGraph.from_csr(6, off, adj)
    kl, H = og.heuristic(4, 3, 5, 0)
    assert (kl, sorted(int(x) for x in H)) == (3, [0, 1, 2, 3, 4])
    # N2 = 4: pop (3, 2) sees remain 5 > N2 and only removes 2 (keys 0:4,
    # 1:3, 3:3, 4:3); pop (3, 1) sees remain 4 in [3, 4] and 3 > 0 -> mindeg
    # 3 with {2} removed; removing 1 leaves 3:2, 4:2 and pop (2, 3) stops at
    # remain 3.  H = {0, 1, 3, 4}, every vertex of degree 3 inside it.
    kl, H = og.heuristic(4, 3, 4, 0)
    assert (kl, sorted(int(x) for x in H)) == (3, [0, 1, 3, 4])


def test_cssc_heu_prefers_heu4_then_heu3_on_ties():
    # inc/heuristic.h:321-348 keeps heu4's H when its kl is not below the
    # others', then heu3's when not below heu2's.  On K5 every heuristic
    # reaches 4 with the whole clique, so CSSC returns heu4's witness, whose
    # order is q followed by its neighbours ascending.
    off, adj = csr_from_edges(5, [(a, b) for a in range(5) for b in range(a + 1, 5)])
    og = OracleGraph.from_csr(5, off, adj)
    k4, H4 = og.heuristic(4, 5, 5, 2)
    k0, H0 = og.heuristic(0, 5, 5, 2)
    assert k4 == 4 and k0 == 4
    assert list(H0) == list(H4) == [2, 0, 1, 3, 4]


def _atlas_max_diameter():
    """best[(k, n)] = largest diameter of a connected graph with exactly n
    vertices and minimum degree >= k, over all graphs with <= 7 vertices."""
    best = {}
    for G in nx.graph_atlas_g():
        n = G.number_of_nodes()
        if n < 2 or not nx.is_connected(G):
            continue
        dmin = min(d for _, d in G.degree())
        diam = nx.diameter(G)
        for k in range(1, dmin + 1):
            best[(k, n)] = max(best.get((k, n), 0), diam)
    return best


def test_ubd_is_the_largest_feasible_diameter():
    best = _atlas_max_diameter()
    for k in range(2, 7):
        for h in range(k + 1, 8):
            exp = max([best[(k, n)] for n in range(2, h + 1) if (k, n) in best], default=0)
            assert ubd_host(k, h) >= exp, (k, h)  # sound: never cuts a feasible H
            assert ubd_host(k, h) == exp, (k, h, ubd_host(k, h), exp)
            assert ubd_node(k, h) == ubd_host(k, h)
    # kl <= 1 gives h - 1 (a path)
    for h in range(2, 8):
        assert ubd_host(1, h) == h - 1 == max(best[(1, n)] for n in range(2, h + 1))
