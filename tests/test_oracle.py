"""Pins the CPU oracle before it is trusted as the parity reference.

Each check is independent of the oracle's own formulas: networkx for the
loader, core numbers and BFS; brute-force enumeration for the optimum on
small graphs; closed forms (cliques, cycles, trees); the known answer the
reference ships for data/graph.txt (pics/problem_code.png: mindeg 3,
H = {2, 3, 0, 1}); two independent exact solvers that must agree; and a
brute-force relaxation for the U_nr bound.
"""

import json
import os

import networkx as nx
import numpy as np
import pytest

from oracle.oracle import (FLAG_LEVELS, FLAG_PAIR, FLAG_RECOMPUTE, FLAGS_ALL, OracleGraph,
                           ubd_host, ubd_node, unr_bound)
from tests.helpers import brute_force, check_witness, csr_from_edges, random_graph

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def nx_from_file(path):
    G = nx.Graph()
    with open(path) as f:
        n, m = map(int, f.readline().split())
        G.add_nodes_from(range(n))
        for line in f:
            p = line.split()
            if len(p) >= 2:
                a, b = int(p[0]), int(p[1])
                if a != b:
                    G.add_edge(a, b)
    return G


@pytest.mark.parametrize("name", ["graph.txt", "100.txt", "hepPH_SCS"])
def test_loader_core_bfs_match_networkx(name):
    path = os.path.join(GOLD, name)
    og = OracleGraph.load(path)
    G = nx_from_file(path)
    off, adj = og.csr()
    assert og.n == G.number_of_nodes()
    for v in range(og.n):
        assert list(adj[off[v]:off[v + 1]]) == sorted(G.neighbors(v))
    core = og.core()
    nxc = nx.core_number(G)
    assert all(int(core[v]) == nxc[v] for v in range(og.n))
    for q in (0, og.n // 2, og.n - 1):
        d = og.bfs(q)
        ref = nx.single_source_shortest_path_length(G, q)
        for v in range(og.n):
            assert int(d[v]) == ref.get(v, 1000000000)


def test_hepph_edge_count():
    # 224,034 lines, asymmetric; 64,437 distinct undirected edges
    off, adj = OracleGraph.load(os.path.join(GOLD, "hepPH_SCS")).csr()
    assert int(off[-1]) == 2 * 64437


def test_known_answer_graph_txt():
    og = OracleGraph.load(os.path.join(GOLD, "graph.txt"))
    off, adj = og.csr()
    for heu in (1, 0):
        r = og.solve(1, 10, 2, heu, 1)
        assert r["kl"] == 3
        assert sorted(int(x) for x in r["witness"]) == [0, 1, 2, 3]
        check_witness(off, adj, 2, 1, 10, 3, r["witness"])


def test_closed_forms():
    # K6 glued to a path: the clique gives 5 whenever 6 fits in [l, h]
    edges = [(a, b) for a in range(6) for b in range(a + 1, 6)] + [(5, 6), (6, 7), (7, 8)]
    off, adj = csr_from_edges(9, edges)
    og = OracleGraph.from_csr(9, off, adj)
    assert og.solve(3, 6, 0, 0, 4)["kl"] == 5
    assert og.solve(2, 9, 0, 1, 4)["kl"] == 5
    assert og.solve(7, 9, 0, 0, 4)["kl"] == 1  # must take path vertices
    assert og.solve(1, 1, 8, 0, 4)["kl"] == 0  # single vertex
    # cycle C8: any proper connected subset is a path (min degree 1), all 8 give 2
    off, adj = csr_from_edges(8, [(i, (i + 1) % 8) for i in range(8)])
    og = OracleGraph.from_csr(8, off, adj)
    assert og.solve(3, 7, 0, 0, 4)["kl"] == 1
    assert og.solve(3, 8, 0, 0, 4)["kl"] == 2
    # tree: core 1 everywhere, optimum 1
    off, adj = csr_from_edges(7, [(0, 1), (0, 2), (1, 3), (1, 4), (2, 5), (2, 6)])
    og = OracleGraph.from_csr(7, off, adj)
    assert list(og.core()) == [1] * 7
    assert og.solve(2, 5, 3, 0, 4)["kl"] == 1
    # infeasible: component smaller than l
    off, adj = csr_from_edges(6, [(0, 1), (1, 2), (3, 4), (4, 5)])
    og = OracleGraph.from_csr(6, off, adj)
    r = og.solve(4, 6, 0, 0, 4)
    assert r["kl"] == 0 and r["status"] == "INFEASIBLE"


@pytest.mark.parametrize("seed", range(6))
def test_bruteforce_random_graphs(seed):
    rng = np.random.default_rng(seed)
    for trial in range(25):
        n = int(rng.integers(5, 12))
        off, adj = random_graph(rng, n, float(rng.uniform(0.25, 0.7)))
        og = OracleGraph.from_csr(n, off, adj)
        q = int(rng.integers(0, n))
        l = int(rng.integers(1, n + 1))
        h = int(rng.integers(l, n + 2))
        bf = brute_force(off, adj, l, h, q)
        for heu in (0, 1):
            for flags in (FLAGS_ALL, 0, FLAGS_ALL & ~0x20, FLAGS_ALL | FLAG_LEVELS,
                          FLAG_LEVELS, FLAGS_ALL | FLAG_LEVELS | FLAG_PAIR, FLAG_PAIR):
                r = og.solve(l, h, q, heu, int(rng.integers(1, 5)), flags)
                exp = 0 if bf is None else bf
                assert r["kl"] == exp, (seed, trial, n, l, h, q, heu, flags, r, bf)
                if bf is not None and (bf > 0 or l <= 1):
                    check_witness(off, adj, q, l, h, r["kl"], r["witness"])
                if bf is None or (bf == 0 and l >= 2):
                    assert r["status"] == "INFEASIBLE"
        plain, _ = og.solve_plain(l, h, q)
        assert plain == (0 if bf is None else bf)


@pytest.mark.parametrize("seed", range(4))
def test_pair_rule_on_dense_small_windows(seed):
    """The pair rule only acts when 2(kl+1) > h - 2: dense graphs with a
    narrow size window [l, h] put most nodes in that regime.  Every optimum
    and witness agrees with brute force."""
    rng = np.random.default_rng(100 + seed)
    active = 0
    for trial in range(30):
        n = int(rng.integers(7, 12))
        off, adj = random_graph(rng, n, float(rng.uniform(0.5, 0.95)))
        og = OracleGraph.from_csr(n, off, adj)
        q = int(rng.integers(0, n))
        h = int(rng.integers(3, min(n, 8) + 1))
        l = int(rng.integers(1, h + 1))
        bf = brute_force(off, adj, l, h, q)
        exp = 0 if bf is None else bf
        active += 2 * (exp + 1) > h - 2
        for flags in (FLAGS_ALL | FLAG_PAIR, FLAGS_ALL | FLAG_LEVELS | FLAG_PAIR, FLAG_PAIR):
            for heu in (0, 1):
                r = og.solve(l, h, q, heu, 4, flags)
                assert r["kl"] == exp, (seed, trial, n, l, h, q, flags, heu, r, bf)
                if bf is not None and (bf > 0 or l <= 1):
                    check_witness(off, adj, q, l, h, r["kl"], r["witness"])
    assert active >= 10


@pytest.mark.parametrize("extra", [0, FLAG_LEVELS])
def test_recompute_and_trail_solvers_expand_the_same_nodes(extra):
    og = OracleGraph.load(os.path.join(GOLD, "hepPH_SCS"))
    for q in [(3, 6, 2975, 1, 1), (6, 9, 8145, 1, 1), (9, 12, 6792, 1, 1), (12, 15, 5721, 1, 1),
              (9, 12, 6792, 0, 4)]:
        a = og.solve(*q, flags=FLAGS_ALL | extra, node_cap=4000)
        b = og.solve(*q, flags=FLAGS_ALL | FLAG_RECOMPUTE | extra, node_cap=4000)
        assert (a["kl"], a["nodes"], a["status"]) == (b["kl"], b["nodes"], b["status"])
        assert (a["levels"] > 1) == (extra != 0 and a["ku"] > a["heu_kl"] + 1 and a["levels"] > 1)


def test_levels_keep_every_golden_optimum():
    """Descending level floors change the search, never the optimum: every
    golden query of the shipped graphs gets the same kl with them."""
    gold = json.load(open(os.path.join(GOLD, "golden.json")))
    n = 0
    for name in ("graph.txt", "100.txt", "hepPH_SCS"):
        og = OracleGraph.load(os.path.join(GOLD, name))
        off, adj = og.csr()
        for r in gold[name]:
            for extra in (FLAG_LEVELS, FLAG_LEVELS | FLAG_PAIR, FLAG_PAIR):
                o = og.solve(*r["query"], flags=FLAGS_ALL | extra, node_cap=2_000_000)
                assert (o["kl"], o["status"]) == (r["kl"], r["status"]), (name, r["query"], extra)
                q = r["query"]
                if o["kl"] > 0:
                    check_witness(off, adj, q[2], q[0], q[1], o["kl"], o["witness"])
            n += 1
    assert n >= 20


def test_monotone_in_h_and_l():
    og = OracleGraph.load(os.path.join(GOLD, "100.txt"))
    for q in (10, 34, 54):
        prev = -1
        for h in range(4, 14):
            k = og.solve(3, h, q, 1, 4)["kl"]
            assert k >= prev
            prev = k
        prev = 10**9
        for l in range(1, 12):
            k = og.solve(l, 12, q, 1, 4)["kl"]
            assert k <= prev
            prev = k


def _relaxation_opt(cvals, rvals, k):
    """max over ways to give each of the k largest R values r_i increments to
    r_i distinct C entries, of the resulting minimum (brute force)."""
    import itertools
    m = len(cvals)
    rs = sorted(rvals, reverse=True)[:k]
    best = -1

    def rec(i, cur):
        nonlocal best
        if i == len(rs):
            best = max(best, min(cur))
            return
        for S in itertools.combinations(range(m), min(rs[i], m)):
            nxt = list(cur)
            for s in S:
                nxt[s] += 1
            rec(i + 1, nxt)

    rec(0, list(cvals))
    return best


def test_unr_bound_equals_relaxation_optimum():
    rng = np.random.default_rng(7)
    for _ in range(300):
        m = int(rng.integers(1, 5))
        cv = [int(x) for x in rng.integers(0, 4, m)]
        nr = int(rng.integers(0, 5))
        rv = [int(x) for x in rng.integers(0, 5, nr)]
        k = int(rng.integers(0, nr + 1))
        got = unr_bound(cv, rv, k)
        assert got == _relaxation_opt(cv, rv, k), (cv, rv, k)


def test_diameter_bound_small_cases():
    # the smallest graph with minimum degree 2 and diameter 3 is C6: with
    # h = 5 no such graph exists (bound 2), with h = 6 it does (bound >= 3)
    assert ubd_host(2, 5) == 2 and ubd_host(2, 6) >= 3
    # a diameter-1 graph with minimum degree kl already needs kl+1 > h
    # vertices: bound 0
    assert ubd_host(9, 9) == 0 and ubd_host(9, 10) == 1
    for kl in range(2, 40):
        for h in range(1, 200):
            assert ubd_host(kl, h) == ubd_node(kl, h)
    # node form never tighter than the trivial bound h-1 for kl < 2
    for h in range(1, 100):
        assert ubd_node(0, h) <= h - 1 and ubd_node(1, h) <= h - 1


def test_golden_vectors_reproduce():
    gold = json.load(open(os.path.join(GOLD, "golden.json")))
    loaders = {"graph.txt": lambda: OracleGraph.load(os.path.join(GOLD, "graph.txt")),
               "100.txt": lambda: OracleGraph.load(os.path.join(GOLD, "100.txt")),
               "hepPH_SCS": lambda: OracleGraph.load(os.path.join(GOLD, "hepPH_SCS"))}
    for name, mk in loaders.items():
        og = mk()
        off, adj = og.csr()
        for r in gold[name]:
            if r["nodes"] > 20000:
                continue
            o = og.solve(*r["query"], node_cap=400_000)
            assert (o["kl"], o["status"], o["nodes"]) == (r["kl"], r["status"], r["nodes"])
            q = r["query"]
            if r["kl"] > 0:
                check_witness(off, adj, q[2], q[0], q[1], r["kl"], r["witness"])


def test_c1_golden_reproduces_and_witnesses_hold():
    """The C1 golden file: every proven entry's witness is a valid community
    of its kl, every entry's kl is within [heu_kl, ku], and the entries the
    oracle proved in at most 5,000 nodes reproduce node for node with the
    recorded flags and cap."""
    from scs_amd import configs
    gold = json.load(open(os.path.join(GOLD, "golden.json")))
    g = configs.graph("C1")
    assert g.sha256() == gold["C1_sha256"]
    og = OracleGraph.from_csr(g.n, g.off, g.adj)
    srch = gold["C1_search"]
    rows = gold["C1"]
    assert len(rows) == 100 and srch["proven"] == sum(r["status"] != "CAPPED" for r in rows)
    redo = 0
    for r in rows:
        q = r["query"]
        assert r["heu_kl"] <= r["kl"] <= r["ku"]
        if r["kl"] > 0:
            check_witness(g.off, g.adj, q[2], q[0], q[1], r["kl"], r["witness"])
        if r["status"] != "CAPPED" and r["nodes"] <= 5000:
            o = og.solve(*q, flags=srch["flags"], node_cap=srch["node_cap"])
            assert (o["kl"], o["status"], o["nodes"]) == (r["kl"], r["status"], r["nodes"]), q
            redo += 1
    assert redo >= 20


def test_time_limit_ends_a_search_capped():
    """The oracle's wall-clock limit ends a long search CAPPED with a valid
    lower bound (a C1 query the golden file records as unproven)."""
    import time
    from scs_amd import configs
    gold = json.load(open(os.path.join(GOLD, "golden.json")))
    r = next(x for x in gold["C1"] if x["status"] == "CAPPED")
    g = configs.graph("C1")
    og = OracleGraph.from_csr(g.n, g.off, g.adj)
    t = time.time()
    o = og.solve(*r["query"], flags=gold["C1_search"]["flags"], time_limit=0.5)
    assert time.time() - t < 5.0
    assert o["status"] == "CAPPED" and o["heu_kl"] <= o["kl"] <= o["ku"]
    q = r["query"]
    if o["kl"] > 0:
        check_witness(g.off, g.adj, q[2], q[0], q[1], o["kl"], o["witness"])
