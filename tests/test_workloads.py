"""Generated workloads are reproducible bit for bit (graph SHA-256 and the
first queries of each stream) and their core numbers match networkx."""

import json
import os

import networkx as nx
import numpy as np
import pytest

from scs_amd import configs, graphgen as gg

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workloads.json")))


@pytest.mark.parametrize("name", ["C1", "C2"])
def test_graph_and_queries_reproduce(name):
    g = configs.graph(name)
    assert g.sha256() == GOLD[name]["sha256"]
    assert (g.n, g.m) == (GOLD[name]["n"], GOLD[name]["m"])
    # simple undirected CSR: sorted, symmetric, no loops or duplicates
    for v in range(0, g.n, max(1, g.n // 500)):
        nb = g.adj[g.off[v]:g.off[v + 1]]
        assert np.all(np.diff(nb) > 0) and not np.any(nb == v)
    core = gg.core_numbers(g)
    qs = configs.queries(name, core, count=len(GOLD[name]["first_queries"]))
    assert [q.text() for q in qs] == GOLD[name]["first_queries"]


def test_core_numbers_match_networkx():
    g = configs.graph("C1")
    G = nx.Graph()
    G.add_nodes_from(range(g.n))
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    G.add_edges_from(zip(src.tolist(), g.adj.tolist()))
    ref = nx.core_number(G)
    core = gg.core_numbers(g)
    assert all(int(core[v]) == ref[v] for v in range(g.n))


def test_query_bounds():
    g = configs.graph("C1")
    core = gg.core_numbers(g)
    for q in configs.queries("C1", core):
        assert 5 <= q.l <= 15 and q.l + 5 <= q.h <= q.l + 15
        assert core[q.q] >= 10 and q.heu == 1 and q.limit_doms == 8


def test_edge_list_round_trip(tmp_path):
    from oracle.oracle import OracleGraph
    g = configs.graph("C1")
    p = tmp_path / "c1.txt"
    gg.write_edge_list(g, str(p))
    off, adj = OracleGraph.load(str(p)).csr()
    assert np.array_equal(off, g.off) and np.array_equal(adj, g.adj)
