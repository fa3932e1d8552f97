"""Writes tests/golden/golden.json: exact optima (and the oracle's node
counts) for the reference's shipped graphs and query files, and for every
query of the C1 workload the oracle proves within its node cap.

Inputs: data files shipped with the reference (graph.txt, 100.txt,
hepPH_SCS) and its query lists client/query/exp1_2/{100,hepPh}, copied here
as fixtures; C1 from scs_amd.configs.  The shipped graphs' optima come from
the oracle with the reference's rules alone and are cross-checked against
the independent plain solver where that finishes.  C1 is solved with the
reference's rules plus the descending level search and the pair rule (the
same optimum, far fewer nodes), all 100 queries in parallel processes, each
with node cap C1_CAP; a query the cap stops stays in the file as CAPPED with
its lower bound.  A second pass (--retry-cap N) re-solves only those with a
larger cap.  Run: python tests/golden/make_golden.py [--jobs N], then
python tests/golden/make_golden.py --retry-cap 1500000 (the file is rewritten after every
answer, so a pass cut short keeps what it proved)
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle.oracle import FLAG_LEVELS, FLAG_PAIR, FLAGS_ALL, OracleGraph  # noqa: E402
from scs_amd import configs  # noqa: E402

CAP = 400_000
C1_CAP = 300_000
C1_FLAGS = FLAGS_ALL | FLAG_LEVELS | FLAG_PAIR


def read_queries(name):
    out = []
    for line in open(os.path.join(HERE, name)):
        line = line.strip()
        if not line or line == "server_exit":
            continue
        out.append(tuple(int(x) for x in line.split()))
    return out


def solve(og, q, cap=CAP, flags=FLAGS_ALL):
    import time
    t = time.time()
    r = og.solve(*q, flags=flags, node_cap=cap)
    return dict(query=list(q), kl=int(r["kl"]), status=r["status"], nodes=int(r["nodes"]),
                heu_kl=int(r["heu_kl"]), ku=int(r["ku"]), K=int(r["K"]),
                witness=[int(x) for x in r["witness"]], seconds=round(time.time() - t, 3))


_C1 = {}


def _c1_init():
    g = configs.graph("C1")
    og = OracleGraph.from_csr(g.n, g.off, g.adj)
    _C1["og"] = og
    _C1["qs"] = configs.queries("C1", og.core().astype(np.int64))


def _c1_one(i):
    q = _C1["qs"][i]
    r = solve(_C1["og"], (q.l, q.h, q.q, q.heu, q.limit_doms), cap=C1_CAP, flags=C1_FLAGS)
    r["index"] = i
    return r


def solve_c1(jobs):
    from multiprocessing import Pool
    with Pool(jobs, initializer=_c1_init) as p:
        rows = list(p.imap(_c1_one, range(100)))
    return rows


_RETRY_CAP = [0]


def _c1_retry(i):
    q = _C1["qs"][i]
    r = solve(_C1["og"], (q.l, q.h, q.q, q.heu, q.limit_doms), cap=_RETRY_CAP[0], flags=C1_FLAGS)
    r["index"] = i
    return r


def retry_capped(jobs, cap, only=None):
    """Second pass: the C1 entries still CAPPED in golden.json (or those of
    them listed in `only`) are solved again with node cap `cap`; the proven
    ones replace their entries (each entry records the cap it was solved
    with), and the pass appends its record to C1_search.retries."""
    import time
    from multiprocessing import Pool
    path = os.path.join(HERE, "golden.json")
    out = json.load(open(path))
    todo = [r["index"] for r in out["C1"] if r["status"] == "CAPPED"
            and (only is None or r["index"] in only)]
    _RETRY_CAP[0] = cap
    t0 = time.time()
    done = proven = 0
    rec = {"node_cap": cap, "queries": len(todo), "answered": 0, "jobs": jobs,
           "wall_seconds": 0.0, "proven": 0}
    if only is not None:
        rec["only"] = sorted(only)
    out["C1_search"].setdefault("retries", []).append(rec)
    with Pool(jobs, initializer=_c1_init) as p:
        # the file is rewritten after every answer, so a pass cut short keeps
        # what it proved
        for r in p.imap_unordered(_c1_retry, todo):
            done += 1
            r["node_cap"] = cap
            if r["status"] != "CAPPED":
                out["C1"][r["index"]] = r
                proven += 1
            rec.update(answered=done, wall_seconds=round(time.time() - t0, 1), proven=proven)
            out["C1_search"]["proven"] = sum(x["status"] != "CAPPED" for x in out["C1"])
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(out, f, indent=1)
            os.replace(tmp, path)
            print(f"retry {done}/{len(todo)}: index {r['index']} {r['status']} kl {r['kl']} "
                  f"nodes {r['nodes']}", flush=True)
    print("retried", len(todo), "proven", proven)


def main(jobs=8):
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
    import time
    t0 = time.time()
    out["C1"] = solve_c1(jobs)
    g = configs.graph("C1")
    out["C1_search"] = {"flags": C1_FLAGS, "node_cap": C1_CAP, "jobs": jobs,
                        "wall_seconds": round(time.time() - t0, 1),
                        "proven": sum(r["status"] != "CAPPED" for r in out["C1"])}
    out["C1_sha256"] = g.sha256()
    for k, v in out.items():
        if isinstance(v, list) and k != "C1":
            for r in v:
                assert r["status"] in ("OK", "HEU_OPT", "INFEASIBLE"), (k, r)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden.json")


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--retry-cap", type=int, default=0,
                    help="only re-solve the C1 entries still CAPPED, with this node cap")
    ap.add_argument("--only", default="",
                    help="with --retry-cap: comma-separated C1 indices to retry")
    a = ap.parse_args()
    if a.retry_cap:
        only = {int(x) for x in a.only.split(",") if x} or None
        retry_capped(a.jobs, a.retry_cap, only)
    else:
        main(a.jobs)
