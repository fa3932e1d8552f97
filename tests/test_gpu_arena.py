"""Admission under device-memory pressure.  The per-query local graphs live
in one device arena; when it is full a query waits in the queue and is
admitted once resident queries finish, and a query whose local graph does
not fit even the empty arena is rejected.  With an arena holding about two
local graphs, every golden C1 query the oracle proved still returns its
optimum.  These engines build every query's local graph (whole_graph=0); a
whole-graph query needs only its 4n-byte distance array, which the small
arena still holds."""

import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleGraph, ubd_host
from scs_amd import configs, scs

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def granule(b):
    return max(256, (b + 255) // 256 * 256)


def test_arena_pressure_requeues_and_matches_golden():
    rows = [r for r in json.load(open(os.path.join(GOLD, "golden.json")))["C1"]
            if r["status"] != "CAPPED" and r["nodes"] <= 20000]
    assert len(rows) >= 20
    g = configs.graph("C1")
    og = OracleGraph.from_csr(g.n, g.off, g.adj)
    need = []
    for r in rows:
        l, h, q, heu, ld = r["query"]
        p = og.preprocess(q, r["heu_kl"], ubd_host(r["heu_kl"], h))
        need.append(granule(4 * (4 * p["K"] + 1)) + granule(4 * len(p["adj"])))
    budget = 2 * max(need) + 4096
    print("arena bytes", budget, "per-query", need, flush=True)
    G = scs.Graph.from_csr(g.off, g.adj)
    # proven golden queries only; the budget only bounds a regression
    E = scs.Engine(G, limit_queries=16, mem_budget=budget, node_budget=5_000_000,
                   whole_graph=0)
    res = E.run([tuple(r["query"]) for r in rows], progress=10.0)
    E.close()
    for r, row in zip(res, rows):
        assert r.status in ("OK", "HEU_OPT"), (row["query"], r.status)
        assert r.kl == row["kl"], row["query"]

    # an arena smaller than the smallest local graph rejects every query
    # that needs one, and the engine stays usable
    E = scs.Engine(G, limit_queries=16, mem_budget=min(need) // 2, whole_graph=0)
    res = E.run([tuple(r["query"]) for r in rows[:3]])
    E.close()
    for r, row in zip(res, rows[:3]):
        if row["heu_kl"] == row["ku"]:
            assert r.status == "HEU_OPT"
        else:
            assert r.status == "REJECTED" and r.heu_kl <= r.kl <= row["kl"]
    # the same arena admits every query as a whole-graph query
    assert min(need) // 2 >= granule(4 * g.n)
    E = scs.Engine(G, limit_queries=16, mem_budget=min(need) // 2, whole_graph=2,
                   node_budget=5_000_000)
    res = E.run([tuple(r["query"]) for r in rows[:3]])
    E.close()
    for r, row in zip(res, rows[:3]):
        assert r.status in ("OK", "HEU_OPT") and r.kl == row["kl"], row["query"]
    G.close()
