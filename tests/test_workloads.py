"""Generated workloads are reproducible bit for bit (graph SHA-256 and the
first queries of each stream) and their core numbers match networkx."""

import json
import os

import networkx as nx
import numpy as np
import pytest

from scs_amd import configs, graphgen as gg

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workloads.json")))


@pytest.mark.parametrize("name", ["C1", "C2", "C3", "C4"])
def test_graph_and_queries_reproduce(name):
    g = configs.graph(name, cache_dir="off")
    assert g.sha256() == GOLD[name]["sha256"]
    assert (g.n, g.m) == (GOLD[name]["n"], GOLD[name]["m"])
    # simple undirected CSR: sorted, symmetric, no loops or duplicates
    for v in range(0, g.n, max(1, g.n // 500)):
        nb = g.adj[g.off[v]:g.off[v + 1]]
        assert np.all(np.diff(nb) > 0) and not np.any(nb == v)
    core = gg.core_numbers(g) if g.n <= 100_000 else _oracle_core(g)
    qs = configs.queries(name, core, count=len(GOLD[name]["first_queries"]))
    assert [q.text() for q in qs] == GOLD[name]["first_queries"]


def _oracle_core(g):
    from oracle.oracle import OracleGraph
    return OracleGraph.from_csr(g.n, g.off, g.adj).core().astype(np.int64)


def test_c5_graph_reproduces():
    """R-MAT scale 22, edge factor 16 (the int64-offset workload): 64M
    undirected edges generated from the seed, pinned by SHA-256."""
    g = configs.graph("C5", cache_dir="off")
    assert g.sha256() == GOLD["C5"]["sha256"]
    assert (g.n, g.m) == (GOLD["C5"]["n"], GOLD["C5"]["m"])
    assert g.off.dtype == np.int64
    qs = configs.queries("C5", _oracle_core(g), count=len(GOLD["C5"]["first_queries"]))
    assert [q.text() for q in qs] == GOLD["C5"]["first_queries"]


def test_csr_cache_round_trip(tmp_path):
    """The binary CSR cache returns the generated graph bit for bit, keys
    entries by the generator parameters, and ignores an entry whose
    contents no longer match its stored SHA-256."""
    d = str(tmp_path)
    params = configs.graph_params("C3")
    assert gg.cache_load(d, params) is None
    g = configs.generate("C3")
    path = gg.cache_store(d, params, g)
    h = gg.cache_load(d, params)
    assert h is not None and h.sha256() == g.sha256() == GOLD["C3"]["sha256"]
    assert np.array_equal(h.off, g.off) and np.array_equal(h.adj, g.adj)
    other = dict(params, seed=params["seed"] + 1)
    assert gg.cache_key(other) != gg.cache_key(params) and gg.cache_load(d, other) is None
    # a corrupted entry (one adjacency id changed) is rejected
    import numpy as _np
    with _np.load(path, allow_pickle=False) as z:
        parts = {k: z[k].copy() for k in z.files}
    parts["adj"][0] ^= 1
    _np.savez(path, **parts)
    assert gg.cache_load(d, params) is None
    # configs.graph goes through the cache directory it is given
    configs.graph.cache_clear()
    assert configs.graph("C3", cache_dir=d).sha256() == g.sha256()
    assert gg.cache_load(d, params) is not None


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
