"""Benchmark: size-bounded community search queries per second on one MI355X.

Workload (default): the batched GPU configuration, a Chung-Lu graph (seed 2,
n = 100,000, 1.5M edges, exponent 2.3) with 2,000 queries (l in [5,40],
h = l + [5,60], q over core >= 8, heuristic on, limit_dominating_set 8),
all submitted at once to the persistent engine.  A step is one pass of the
whole query batch through the engine (admission: heuristic + GPU
preprocessing; search; results).  Each query runs until its optimum is
proven or its node budget is spent (status CAPPED, kl is then a lower
bound); the fraction proven optimal is reported beside the rate.

Prints one JSON line (rank 0).  The CPU baseline is the project's serial
exact solver (the oracle) on a bounded sample of the same queries with the
same node budget, timed on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per adjacency entry scanned by the search: the 4-byte
# neighbour id plus the neighbour's state read and written (status 1 B,
# dC 2 B, dR 4 B read; dC 2 B and dR 4 B written)
BYTES_PER_EDGE = 4 + 7 + 6


PHASES = ("rebuild", "rules", "bounds", "score", "doms", "move", "donate", "idle")


def phase_shares(s0, s1):
    """Share of the search waves' wall-clock time per phase in the timed region."""
    d = [b - a for a, b in zip(s0["phase_ticks"], s1["phase_ticks"])]
    tot = sum(d) or 1
    return {k: round(v / tot, 4) for k, v in zip(PHASES, d)}


def percentile(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if len(xs) else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--queries", type=int, default=2000)
    ap.add_argument("--budget", type=int, default=300, help="node budget per query")
    # defaults: the autotuner's pick for this workload (scs_amd/tuned_c2.json)
    ap.add_argument("--waves-per-cu", type=int, default=4)
    ap.add_argument("--limit-queries", type=int, default=512)
    ap.add_argument("--slice-us", type=int, default=5000)
    ap.add_argument("--copy-limit", type=float, default=0.5)
    ap.add_argument("--learned", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); queries are independent "
              "and would be sharded across GPUs with no collective on the hot path", file=sys.stderr)
        sys.exit(2)

    from scs_amd import configs, graphgen as gg, scs

    g = configs.graph(args.config)
    G = scs.Graph.from_csr(g.off, g.adj)
    core = G.core()
    qs = configs.queries(args.config, core.astype(np.int64), count=args.queries)
    queries = [(q.l, q.h, q.q, q.heu, q.limit_doms) for q in qs]
    E = scs.Engine(G, node_budget=args.budget, waves_per_cu=args.waves_per_cu,
                   limit_queries=args.limit_queries, time_slice_us=args.slice_us,
                   copy_limit=args.copy_limit, learned=args.learned, mem_budget=32 << 30)
    # load code objects outside the timed region
    E.run([queries[0]])

    for _ in range(args.warmup):
        E.run(queries)

    s0 = E.stats()
    t0 = time.perf_counter()
    all_res = []
    for _ in range(args.steps):
        all_res = E.run(queries, progress=10.0)
    t1 = time.perf_counter()
    s1 = E.stats()

    wall = t1 - t0
    nq = len(queries) * args.steps
    qps = nq / wall
    lat_ms = [r.latency_us / 1000.0 for r in all_res]
    status = {}
    for r in all_res:
        status[r.status] = status.get(r.status, 0) + 1
    proven = sum(status.get(k, 0) for k in ("OK", "HEU_OPT", "INFEASIBLE")) / max(1, len(all_res))
    nodes = s1["nodes"] - s0["nodes"]
    # roofline of the dominant kernel (scs_epoch) over every launch of the
    # process, so the average launch duration matches a kernel-trace profile
    # of the same command
    engine_us = s1["engine_us"]
    epochs = s1["epochs"]
    edges = s1["edges"]
    ach = (edges * BYTES_PER_EDGE) / (engine_us * 1e-6) / 1e9 if engine_us > 0 else 0.0

    line = {
        "metric": "queries/sec (p99 latency, proven-optimal fraction beside it)",
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (deterministic Chung-Lu graph and query stream)",
        "config": {"workload": f"{args.config}: Chung-Lu n={g.n} m={g.m}, {len(queries)} queries",
                   "node_budget": args.budget, "learned": bool(args.learned),
                   "waves": s1["waves"], "limit_queries": args.limit_queries},
        "p50_ms": round(percentile(lat_ms, 50), 3),
        "p99_ms": round(percentile(lat_ms, 99), 3),
        "proven_optimal_fraction": round(proven, 4),
        "status_counts": status,
        "nodes": int(nodes),
        "nodes_per_s": round(nodes / wall, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": None,
                     "kernel": "scs_epoch", "epoch_avg_us": round(engine_us / max(1, epochs), 2),
                     "bytes_per_edge": BYTES_PER_EDGE, "edges": int(edges)},
        "prep": {"device_us": round(s1["prep_us"] - s0["prep_us"], 1),
                 "algorithmic_bytes": int(s1["prep_bytes"] - s0["prep_bytes"])},
        "search_phases": phase_shares(s0, s1),
        "domination": {k: int(s1[k] - s0[k]) for k in ("dom_front", "dom_cand", "dom_found")},
        "heuristic_host_us": round(s1["heu_us"] - s0["heu_us"], 1),
        "core_decomposition_us": round(s1["core_us"], 1),
    }
    pb = line["prep"]["algorithmic_bytes"]
    pu = line["prep"]["device_us"]
    line["prep"]["achieved_GBs"] = round(pb / (pu * 1e-6) / 1e9, 2) if pu > 0 else 0.0

    if not args.no_cpu:
        from oracle.oracle import OracleGraph
        og = OracleGraph.from_csr(g.n, g.off, g.adj)
        done = 0
        tc0 = time.perf_counter()
        for q in queries:
            og.solve(*q, node_cap=args.budget)
            done += 1
            if done % 10 == 0:
                print(f"[cpu] {done} queries, {time.perf_counter() - tc0:.1f} s", file=sys.stderr,
                      flush=True)
            if time.perf_counter() - tc0 > args.cpu_seconds:
                break
        tc = time.perf_counter() - tc0
        line["cpu_baseline"] = {"value": round(done / tc, 4), "unit": "queries/s", "cores": 1,
                                "kind": "port",
                                "sample": f"first {done} queries of the same stream, node budget "
                                          f"{args.budget}, serial trail solver, {tc:.1f} s"}
    E.close()
    G.close()
    out = json.dumps(line)
    print(out, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")


if __name__ == "__main__":
    main()
