"""Writes tests/golden/golden.json: exact optima (and the oracle's node
counts) for the reference's shipped graphs and query files and for the
solvable part of the C1 workload.

Inputs: data files shipped with the reference (graph.txt, 100.txt,
hepPH_SCS) and its query lists client/query/exp1_2/{100,hepPh}, copied here
as fixtures; C1 from scs_amd.configs.  Every optimum is produced by the CPU
oracle and cross-checked against the independent plain solver where that
finishes.  Run: python tests/golden/make_golden.py
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle.oracle import FLAGS_ALL, OracleGraph  # noqa: E402
from scs_amd import configs  # noqa: E402

CAP = 400_000
C1_SOLVABLE = [1, 17, 18, 41, 63, 65, 72, 86]


def read_queries(name):
    out = []
    for line in open(os.path.join(HERE, name)):
        line = line.strip()
        if not line or line == "server_exit":
            continue
        out.append(tuple(int(x) for x in line.split()))
    return out


def solve(og, q, cap=CAP):
    r = og.solve(*q, flags=FLAGS_ALL, node_cap=cap)
    return dict(query=list(q), kl=int(r["kl"]), status=r["status"], nodes=int(r["nodes"]),
                heu_kl=int(r["heu_kl"]), ku=int(r["ku"]), K=int(r["K"]),
                witness=[int(x) for x in r["witness"]])


def main():
    out = {}
    og = OracleGraph.load(os.path.join(HERE, "graph.txt"))
    out["graph.txt"] = [solve(og, (1, 10, 2, 1, 1)), solve(og, (1, 10, 2, 0, 1)),
                        solve(og, (3, 5, 9, 0, 2)), solve(og, (2, 4, 10, 1, 1))]
    og = OracleGraph.load(os.path.join(HERE, "100.txt"))
    qs = read_queries("queries_100.txt")
    out["100.txt"] = [solve(og, q) for q in qs] + [solve(og, (q[0], q[1], q[2], 0, 4)) for q in qs[:5]]
    og = OracleGraph.load(os.path.join(HERE, "hepPH_SCS"))
    qs = read_queries("queries_hepPh.txt")
    out["hepPH_SCS"] = [solve(og, qs[i]) for i in (0, 1, 2, 3, 5)]
    g = configs.graph("C1")
    og = OracleGraph.from_csr(g.n, g.off, g.adj)
    core = og.core()
    qs = configs.queries("C1", core.astype(np.int64))
    out["C1"] = [dict(solve(og, (qs[i].l, qs[i].h, qs[i].q, qs[i].heu, qs[i].limit_doms)), index=i)
                 for i in C1_SOLVABLE]
    out["C1_sha256"] = g.sha256()
    for k, v in out.items():
        if isinstance(v, list):
            for r in v:
                assert r["status"] in ("OK", "HEU_OPT", "INFEASIBLE"), (k, r)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
