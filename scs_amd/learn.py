"""Learned branching order: traces, training and evaluation.

1. Traces.  The engine runs a stream of training queries with tracing on;
   every stride-th search node records up to 64 of its branching
   candidates with six features (dC, dR, dist(q, v), core(v),
   slack = dC + min(dR, h-|C|-1) - kl, and the plain connection score).
2. Labels.  Only queries whose optimum the search proved give rows: a
   candidate is positive when it belongs to the optimal community H the
   search returned for its query (the witness), i.e. choosing it leads to
   the optimum.  Rows of budget-capped queries are dropped, since their
   witness marks only a lower bound.
3. Training.  The 6-64-64-1 MLP is trained on the GPU by the library's own
   forward/backward/Adam kernels (scs_mlp_train) with class-balanced binary
   cross-entropy.
4. Evaluation.  Unseen queries (the workload's own stream) are run with the
   plain connection-score rule and with the learned order at the same node
   budget; every optimum both prove is the same (the order never prunes),
   and on the queries both prove the nodes expanded and the time to the
   optimum (submission to result) are compared, beside how many each
   proves.

Usage:
  python -m scs_amd.learn --config C1 --out models/branching_c1.npz
  python -m scs_amd.learn --config C4 --per-seed 8 --deadline-ms 2000 \
      --out models/branching_c4.npz   # stops: C4 proves none of its training queries
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import configs, graphgen as gg, scs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def training_queries(name: str, core: np.ndarray, seeds=range(100, 200), per_seed=4):
    """Query streams drawn with seeds 100..199 (disjoint from the evaluation
    stream, whose seed is graph seed + 1000)."""
    c = configs.CONFIGS[name]
    qp = dict(c["queries"])
    qp["count"] = per_seed
    out = []
    for s in seeds:
        out.extend(gg.make_queries(core, seed=s, **qp))
    return [(q.l, q.h, q.q, q.heu, q.limit_doms) for q in out]


PROVEN = ("OK", "HEU_OPT")


def collect(G, queries, budget, stride=2, cap=2_000_000, **engine_kw):
    E = scs.Engine(G, node_budget=budget, **engine_kw)
    E.trace(cap, stride)
    res = E.run(queries, tags=list(range(len(queries))), progress=10.0)
    rows = E.trace_read()
    E.close()
    wit = {r.tag: set(int(x) for x in r.witness) for r in res
           if r.status in PROVEN and r.kl > 0}
    rows = rows[np.isin(rows[:, 0], np.array(sorted(wit), dtype=rows.dtype))] if wit else rows[:0]
    X = np.zeros((len(rows), 6), np.float32)
    X[:, 0] = rows[:, 2]
    X[:, 1] = rows[:, 3]
    X[:, 2] = rows[:, 4]
    X[:, 3] = rows[:, 5]
    X[:, 4] = rows[:, 6].view(np.int32)
    X[:, 5] = rows[:, 7].view(np.float32)
    y = np.array([1.0 if int(v) in wit.get(int(t), ()) else 0.0 for t, v in rows[:, :2]],
                 np.float32)
    return X, y, res


def train(X, y, epochs=20, batch=2048, lr=2e-3, seed=0):
    m = scs.model_init(seed)
    pos = max(1.0, float(y.sum()))
    pw = float(min(50.0, (len(y) - pos) / pos))
    loss = scs.mlp_train(m, X, y, epochs=epochs, batch=min(batch, len(y)), lr=lr,
                         pos_weight=pw, seed=seed)
    return m, loss, pw


def evaluate(G, queries, model, budget, **engine_kw):
    """Runs the stream with the plain rule and with the learned order."""
    out = {}
    for mode in ("plain", "learned"):
        E = scs.Engine(G, node_budget=budget, **engine_kw)
        if mode == "learned":
            E.set_model(model, True)
        t0 = time.perf_counter()
        res = E.run(queries, progress=10.0)
        dt = time.perf_counter() - t0
        print(f"[learn] {mode}: {dt:.1f} s", file=sys.stderr, flush=True)
        st = E.stats()
        E.close()
        proven = sum(r.status in ("OK", "HEU_OPT", "INFEASIBLE") for r in res)
        out[mode] = {"seconds": dt, "nodes": int(st["nodes"]), "proven": proven,
                     "kl": [r.kl for r in res], "status": [r.status for r in res],
                     "nodes_per_query": [int(r.nodes) for r in res],
                     "latency_ms": [r.latency_us / 1000.0 for r in res]}
    p, l = out["plain"], out["learned"]
    both = [i for i, (a, b) in enumerate(zip(p["status"], l["status"]))
            if a != "CAPPED" and b != "CAPPED"]
    out["both_proven"] = {
        "queries": len(both),
        "optimum_equal": all(p["kl"][i] == l["kl"][i] for i in both),
        "nodes_plain": int(sum(p["nodes_per_query"][i] for i in both)),
        "nodes_learned": int(sum(l["nodes_per_query"][i] for i in both)),
        "time_to_optimum_ms_plain": round(sum(p["latency_ms"][i] for i in both), 1),
        "time_to_optimum_ms_learned": round(sum(l["latency_ms"][i] for i in both), 1)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C1")
    ap.add_argument("--per-seed", type=int, default=4)
    ap.add_argument("--budget", type=int, default=20000)
    ap.add_argument("--stride", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--eval-queries", type=int, default=100)
    ap.add_argument("--out", default=os.path.join(ROOT, "models", "branching_c1.npz"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--limit-queries", type=int, default=1024)
    ap.add_argument("--levels", type=int, default=1)
    ap.add_argument("--pair", type=int, default=1)
    ap.add_argument("--deadline-ms", type=int, default=0,
                    help="per-query deadline from submission (0 = none), in collection and evaluation")
    args = ap.parse_args()
    ekw = dict(limit_queries=args.limit_queries, levels=args.levels, pair=args.pair,
               deadline_ms=args.deadline_ms, mem_budget=32 << 30)

    print(f"[learn] building {args.config}", file=sys.stderr, flush=True)
    g = configs.graph(args.config)
    G = scs.Graph.from_csr(g.off, g.adj)
    core = G.core().astype(np.int64)
    tq = training_queries(args.config, core, per_seed=args.per_seed)
    t0 = time.perf_counter()
    X, y, res = collect(G, tq, args.budget, stride=args.stride, **ekw)
    n_proven = sum(r.status in PROVEN for r in res)
    t_col = time.perf_counter() - t0
    print(f"[learn] {len(tq)} training queries ({n_proven} proven) -> {len(y)} rows, "
          f"{int(y.sum())} positive ({t_col:.1f} s)", file=sys.stderr, flush=True)
    if len(y) == 0:
        sys.exit(f"[learn] no training query of {args.config} was proven within the budget and "
                 "deadline, so there are no labels from proven optima to train on")
    t0 = time.perf_counter()
    m, loss, pw = train(X, y, epochs=args.epochs)
    t_tr = time.perf_counter() - t0
    print(f"[learn] trained: loss {loss:.4f} pos_weight {pw:.1f} ({t_tr:.1f} s)", file=sys.stderr,
          flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    m.save(args.out)
    eq = [(q.l, q.h, q.q, q.heu, q.limit_doms)
          for q in configs.queries(args.config, core, count=args.eval_queries)]
    ev = evaluate(G, eq, m, args.budget, **ekw)
    same = all(a == b for a, b, s1, s2 in zip(ev["plain"]["kl"], ev["learned"]["kl"],
                                                 ev["plain"]["status"], ev["learned"]["status"])
               if s1 != "CAPPED" and s2 != "CAPPED")
    rep = {"config": args.config, "train_queries": len(tq), "train_proven": n_proven,
           "levels": args.levels, "pair": args.pair, "deadline_ms": args.deadline_ms,
           "rows": int(len(y)),
           "positives": int(y.sum()), "loss": loss, "collect_s": t_col, "train_s": t_tr,
           "eval_queries": len(eq), "budget": args.budget,
           "plain": {k: ev["plain"][k] for k in ("seconds", "nodes", "proven")},
           "learned": {k: ev["learned"][k] for k in ("seconds", "nodes", "proven")},
           "optimum_equal_where_both_complete": same,
           "both_proven": ev["both_proven"],
           "kl_sum_plain": int(sum(ev["plain"]["kl"])), "kl_sum_learned": int(sum(ev["learned"]["kl"]))}
    print(json.dumps(rep), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    G.close()


if __name__ == "__main__":
    main()
