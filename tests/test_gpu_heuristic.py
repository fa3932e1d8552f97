"""The greedy lower bound on the GPU (scs_heuristic_gpu, one wave per
query) against the oracle's restatement of the reference's CSSC_heu
(inc/heuristic.h:8-348): the same kl and the same H, vertex for vertex, on
the shipped graphs, C1, a hub graph and the C2 and C4 streams; and the
engine's admission with the GPU heuristic expands exactly the oracle's
nodes in deterministic mode."""

import os

import numpy as np
import pytest

from oracle.oracle import FLAGS_ALL, OracleGraph
from scs_amd import configs, scs
from tests.test_gpu_parity import GOLD, hub_graph

pytestmark = pytest.mark.gpu


def random_queries(n, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        l = int(rng.integers(1, 30))
        out.append((l, l + int(rng.integers(0, 40)), int(rng.integers(0, n))))
    return out


def cases(name):
    if name == "hubs":
        n, off, adj = hub_graph()
        G, og = scs.Graph.from_csr(off, adj), OracleGraph.from_csr(n, off, adj)
        qs = random_queries(n, 40, 5) + [(5, 30, 0), (10, 60, 1), (3, 12, 2)]
    elif name in ("C1", "C2", "C4"):
        g = configs.graph(name)
        G, og = scs.Graph.from_csr(g.off, g.adj), OracleGraph.from_csr(g.n, g.off, g.adj)
        core = og.core().astype(np.int64)
        count = {"C1": 100, "C2": 24, "C4": 8}[name]
        qs = [(x.l, x.h, x.q) for x in configs.queries(name, core, count=count)]
    else:
        path = os.path.join(GOLD, name)
        G, og = scs.Graph.load(path), OracleGraph.load(path)
        qs = random_queries(og.n, 60, 9)
    return G, og, qs


@pytest.mark.parametrize("name", ["100.txt", "hepPH_SCS", "C1", "hubs", "C2", "C4"])
def test_gpu_heuristic_equals_oracle(name):
    G, og, qs = cases(name)
    print(name, len(qs), "queries", flush=True)
    got = scs.heuristic_gpu(G, qs)
    G.close()
    for (l, h, q), (kl, H) in zip(qs, got):
        ref_kl, ref_H = og.heuristic(0, l, h, q)
        assert kl == ref_kl, (name, l, h, q, kl, ref_kl)
        if kl > 0:
            assert list(H) == list(ref_H), (name, l, h, q)


def test_engine_with_gpu_heuristic_expands_the_oracles_nodes():
    g = configs.graph("C1")
    G, og = scs.Graph.from_csr(g.off, g.adj), OracleGraph.from_csr(g.n, g.off, g.adj)
    core = og.core().astype(np.int64)
    qs = [(x.l, x.h, x.q, x.heu, x.limit_doms) for x in configs.queries("C1", core)[:12]]
    E = scs.Engine(G, deterministic=1, node_budget=2000, limit_queries=16, levels=0, pair=0,
                   gpu_heuristic=1)
    res = E.run(qs)
    E.close()
    G.close()
    for r, q in zip(res, qs):
        o = og.solve(*q, flags=FLAGS_ALL, node_cap=2000)
        assert (r.kl, r.status, r.nodes) == (o["kl"], o["status"], o["nodes"]), q
