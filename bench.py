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


PHASES = ("rebuild", "rules", "bounds", "score", "doms", "move", "donate", "idle", "closure_scan",
          "closure_remove", "argmax", "clear", "pair")


def phase_shares(s0, s1):
    """Share of the search waves' wall-clock time per phase in the timed region."""
    d = [b - a for a, b in zip(s0["phase_ticks"], s1["phase_ticks"])]
    tot = sum(d) or 1
    return {k: round(v / tot, 4) for k, v in zip(PHASES, d)}


def percentile(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if len(xs) else 0.0


def prep_leg(scs, G, core, queries):
    """Batched preprocessing of the queries alone (64 per batch, local graphs
    built on the device, then released), with no search kernel on the GPU:
    device time from HIP events and the kernels' algorithmic-byte count."""
    print("[bench] standalone preprocessing leg", file=sys.stderr, flush=True)
    EP = scs.Engine(G, limit_queries=64, waves_per_cu=1, mem_budget=32 << 30)
    core64 = core.astype(np.int64)
    todo = []
    for (l, h, q, heu, ld) in queries:
        kl0, ubd = EP.heuristic(l, h, q)
        if kl0 < min(int(core64[q]), h - 1):
            todo.append((q, kl0, ubd))
    EP.prep_many(*zip(*todo[:1])) if todo else None  # code objects
    p0 = EP.stats()
    Ks = []
    for c0 in range(0, len(todo), 64):
        qq, kk, uu = zip(*todo[c0:c0 + 64])
        K, _ = EP.prep_many(qq, kk, uu)
        Ks.extend(int(x) for x in K)
    p1 = EP.stats()
    EP.close()
    pus = p1["prep_us"] - p0["prep_us"]
    pby = p1["prep_bytes"] - p0["prep_bytes"]
    gbs = pby / (pus * 1e-6) / 1e9 if pus > 0 else 0.0
    return {"queries": len(todo), "batches": int(p1["prep_batches"] - p0["prep_batches"]),
            "device_us": round(pus, 1), "us_per_query": round(pus / max(1, len(todo)), 2),
            "algorithmic_bytes": int(pby), "achieved_GBs": round(gbs, 2),
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "mean_kept": round(float(np.mean(Ks)), 1) if Ks else 0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--queries", type=int, default=2000)
    ap.add_argument("--budget", type=int, default=300, help="node budget per query")
    # defaults: the autotuner's pick for this workload (scs_amd/tuned_c2.json)
    ap.add_argument("--heuristic-threads", type=int, default=12)
    ap.add_argument("--waves-per-cu", type=int, default=4)
    ap.add_argument("--limit-queries", type=int, default=2048)
    ap.add_argument("--slice-us", type=int, default=200000)
    ap.add_argument("--copy-limit", type=float, default=0.5)
    ap.add_argument("--admit-per-step", type=int, default=256,
                    help="queries admitted (heuristic + GPU preprocessing) per epoch")
    ap.add_argument("--learned", type=int, default=0,
                    help="1 = the main leg itself uses the learned order")
    ap.add_argument("--model", default=os.path.join(ROOT, "models", "branching_c1.npz"),
                    help="branching model for the learned leg (skipped if the file is absent)")
    ap.add_argument("--prep-sample", type=int, default=512,
                    help="queries preprocessed alone (batched) for the preprocessing roofline")
    ap.add_argument("--single-sample", type=int, default=24,
                    help="queries run one at a time (plain order) for the one-query GPU form")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prep-only", action="store_true",
                    help="only run the standalone batched-preprocessing leg (--prep-sample queries)")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    ap.add_argument("--copy-calibrate-mib", type=int, default=0,
                    help="only run a plain elementwise pass over this many MiB (read once, "
                         "written once) for scaling hardware traffic counters, then exit")
    args = ap.parse_args()

    if args.copy_calibrate_mib:
        import torch
        n = args.copy_calibrate_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "elementwise mul", "bytes_read": n * 4,
                          "bytes_written": n * 4, "dispatches": 3}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); queries are independent "
              "and would be sharded across GPUs with no collective on the hot path", file=sys.stderr)
        sys.exit(2)

    from scs_amd import configs, graphgen as gg, scs

    print(f"[bench] building {args.config}", file=sys.stderr, flush=True)
    g = configs.graph(args.config)
    print(f"[bench] graph n={g.n} m={g.m}", file=sys.stderr, flush=True)
    G = scs.Graph.from_csr(g.off, g.adj)
    core = G.core()
    qs = configs.queries(args.config, core.astype(np.int64), count=args.queries)
    queries = [(q.l, q.h, q.q, q.heu, q.limit_doms) for q in qs]
    model = scs.Model.load(args.model) if args.model and os.path.exists(args.model) else None
    eng_kw = dict(node_budget=args.budget, waves_per_cu=args.waves_per_cu,
                  limit_queries=args.limit_queries, time_slice_us=args.slice_us,
                  copy_limit=args.copy_limit, mem_budget=32 << 30,
                  heuristic_threads=args.heuristic_threads, admit_per_step=args.admit_per_step)
    if args.prep_only:
        print(json.dumps({"workload": args.config, "prep_standalone": prep_leg(
            scs, G, core, queries[:args.prep_sample or len(queries)])}), flush=True)
        return
    E = scs.Engine(G, **eng_kw)
    if args.learned:
        if model is None:
            sys.exit("--learned needs --model")
        E.set_model(model, True)
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
        "rule12_full_rescans": {k: int(s1[k] - s0[k]) for k in ("full_scans", "full_rows")},
        "heuristic_host_us": round(s1["heu_us"] - s0["heu_us"], 1),
        "core_decomposition_us": round(s1["core_us"], 1),
    }
    # HBM traffic per launch from the committed hardware-counter profile of
    # this workload and engine setting (two rocprofv3 --pmc runs, FETCH_SIZE
    # and WRITE_SIZE, scaled by a copy pass of known size)
    prof = os.path.join(ROOT, "profiles", "scs_epoch_c2_traffic.json")
    if os.path.exists(prof):
        t = json.load(open(prof))
        if (t["workload"] == args.config and t["waves_per_cu"] == args.waves_per_cu and
                t["time_slice_us"] == args.slice_us and t["node_budget"] == args.budget and
                not args.learned):
            line["roofline"]["traffic"] = t["traffic_bytes_per_launch"]
            line["roofline"]["traffic_source"] = "profiles/scs_epoch_c2_traffic.json"
            line["roofline"]["algorithmic_bytes_per_launch"] = int(
                edges * BYTES_PER_EDGE / max(1, epochs))
    line["kl_sum"] = int(sum(r.kl for r in all_res))
    line["kcap"] = int(s1["kcap"])
    E.close()

    # the same queries in the learned order (batched), and the plain order
    # one query at a time
    if model is not None and not args.learned:
        print("[bench] learned leg", file=sys.stderr, flush=True)
        EL = scs.Engine(G, **eng_kw)
        EL.set_model(model, True)
        EL.run([queries[0]])
        t0 = time.perf_counter()
        lres = EL.run(queries, progress=10.0)
        lw = time.perf_counter() - t0
        EL.close()
        same = all(a.kl == b.kl for a, b in zip(all_res, lres)
                   if a.status != "CAPPED" and b.status != "CAPPED")
        llat = [r.latency_us / 1000.0 for r in lres]
        line["learned_batched"] = {
            "value": round(len(queries) / lw, 3), "unit": "queries/s",
            "p50_ms": round(percentile(llat, 50), 3), "p99_ms": round(percentile(llat, 99), 3),
            "kl_sum": int(sum(r.kl for r in lres)),
            "proven": int(sum(r.status in ("OK", "HEU_OPT", "INFEASIBLE") for r in lres)),
            "optimum_equal_where_both_complete": same, "model": os.path.basename(args.model)}
    if args.single_sample > 0:
        print("[bench] one-query-at-a-time leg", file=sys.stderr, flush=True)
        sq = queries[:args.single_sample]
        ES = scs.Engine(G, **dict(eng_kw, limit_queries=1))
        ES.run([queries[0]])
        t0 = time.perf_counter()
        sres = ES.run(sq, progress=10.0)
        sw = time.perf_counter() - t0
        ES.close()
        slat = [r.latency_us / 1000.0 for r in sres]
        line["plain_one_at_a_time"] = {
            "value": round(len(sq) / sw, 3), "unit": "queries/s", "sample": len(sq),
            "p99_ms": round(percentile(slat, 99), 3)}
        line["batched_vs_one_at_a_time"] = round(qps / (len(sq) / sw), 2)
        if "learned_batched" in line:
            line["learned_batched_vs_one_at_a_time"] = round(
                line["learned_batched"]["value"] / (len(sq) / sw), 2)

    # batched preprocessing alone (no search kernel sharing the GPU): the
    # roofline of the preprocessing kernels
    if args.prep_sample > 0:
        line["prep_standalone"] = prep_leg(scs, G, core, queries[:args.prep_sample])

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
    G.close()
    out = json.dumps(line)
    print(out, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")


if __name__ == "__main__":
    main()
