"""Benchmark: proven-optimal size-bounded community search queries per second
on one MI355X.

Workload (default): the serving configuration, a Chung-Lu graph (seed 4,
n = 500,000, 8M edges, exponent 2.2) and its query stream (l in [5,50],
h = l + [5,150], q over every vertex, heuristic 0/1 at random,
limit_dominating_set in {4,8,16}).  A step is one batch of consecutive
queries of the stream (64 by default) submitted together to the persistent
engine (admission: heuristic + batched GPU preprocessing; search; results).
Batches are fed continuously: the next one goes in as soon as at most
inflight - 1 batches' worth of queries remain in the engine (4 batches = the
engine's 256 query slots by default).  Each query is searched until its
optimum is proven, its node budget is spent or its deadline (from
submission) passes (status CAPPED: kl is then only a lower bound).  Warm-up
steps take the stream's first batches and the timed steps the batches after
them, so no query is timed twice.

value = proven-optimal queries per second over the timed steps, in the
serving form: batched, with the learned branching order (the workload's
model, else the one trained on C1's proven optima).  CAPPED answers are
excluded from the numerator and counted beside it, with the proven
fraction.  Beside it the line reports, on the same timed queries: the plain
order (batched), the plain order one query at a time, the CPU oracle (same
rules, same budget, one core, on a bounded sample) and whether every
optimum two forms proved is equal; the preprocessing kernels alone; and the
roofline of the search kernel.

Prints one JSON line (rank 0).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
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
PROVEN = ("OK", "HEU_OPT", "INFEASIBLE")
PHASES = ("rebuild", "rules", "bounds", "score", "doms", "move", "donate", "idle", "closure_scan",
          "closure_remove", "argmax", "clear", "pair")


def phase_shares(s0, s1):
    """Share of the search waves' wall-clock time per phase in the timed region."""
    d = [b - a for a, b in zip(s0["phase_ticks"], s1["phase_ticks"])]
    tot = sum(d) or 1
    return {k: round(v / tot, 4) for k, v in zip(PHASES, d)}


def percentile(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if len(xs) else 0.0


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def summary(res, wall):
    proven = [r for r in res if r.status in PROVEN]
    lat = [r.latency_us / 1000.0 for r in proven]
    st = {}
    for r in res:
        st[r.status] = st.get(r.status, 0) + 1
    return {"value": round(len(proven) / wall, 3) if wall > 0 else 0.0, "unit": "queries/s",
            "proven": len(proven), "queries": len(res),
            "proven_optimal_fraction": round(len(proven) / max(1, len(res)), 4),
            "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
            "status_counts": st, "nodes": int(sum(r.nodes for r in res)), "wall_s": round(wall, 3)}


def run_batches(E, batches, progress=None, inflight=1):
    """Submits the batches in order, each batch at once; tags number the
    queries across batches.  inflight = 1 drains each batch before the next
    goes in.  inflight = d submits the next batch as soon as at most d - 1
    batches' worth of queries are still in the engine, so new queries are
    admitted while the previous ones search (continuous batching)."""
    total = sum(len(b) for b in batches)
    res, i, tag = [], 0, 0
    last = time.time()
    while True:
        while i < len(batches) and tag - len(res) <= (inflight - 1) * len(batches[i]):
            for (l, h, q, heu, ld) in batches[i]:
                E.submit(l, h, q, heu, ld, tag)
                tag += 1
            i += 1
        left = E.step()
        res.extend(E.poll())
        if progress is not None and time.time() - last > progress:
            last = time.time()
            print(f"[bench] {len(res)}/{total} queries done", file=sys.stderr, flush=True)
        if left == 0 and i == len(batches):
            break
    res.extend(E.poll())
    res.sort(key=lambda r: r.tag)
    return res


def run_one_at_a_time(E, queries, progress=None):
    """Each query submitted only after the previous one has returned, so its
    deadline runs while it alone is on the GPU."""
    out = []
    last = time.time()
    for i, qq in enumerate(queries):
        out.extend(E.run([qq], tags=[i]))
        if progress is not None and time.time() - last > progress:
            last = time.time()
            print(f"[bench] {len(out)}/{len(queries)} queries done", file=sys.stderr, flush=True)
    return out


def prep_leg(scs, G, core, queries, whole=1):
    """Batched preprocessing of the queries alone (64 per batch, built on the
    device, then released), with no search kernel on the GPU: device time
    from HIP events and the kernels' algorithmic-byte count.  whole=1 is the
    engine's own admission (queries keeping at least half the graph get a
    distance array over the resident graph); whole=0 builds every query's
    compacted local CSR."""
    print(f"[bench] standalone preprocessing leg (whole_graph={whole})", file=sys.stderr,
          flush=True)
    EP = scs.Engine(G, limit_queries=64, waves_per_cu=1, mem_budget=32 << 30, whole_graph=whole)
    core64 = core.astype(np.int64)
    todo = []
    for (l, h, q, heu, ld) in queries:
        kl0, ubd = EP.heuristic(l, h, q)
        if kl0 < min(int(core64[q]), h - 1):
            todo.append((q, kl0, ubd))
    if todo:
        EP.prep_many(*zip(*todo[:1]))  # code objects
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


def extra_leg(scs, configs, name, args, eng_kw):
    """A further workload beside the headline.  C1: all 100 queries of its
    stream in one batch at the golden file's node cap and rules, every
    optimum the GPU proves compared with the golden optimum where the file
    has one.  Other workloads: batches of the stream like the headline's."""
    print(f"[bench] extra workload {name}", file=sys.stderr, flush=True)
    g = configs.graph(name)
    G = scs.Graph.from_csr(g.off, g.adj)
    core = G.core().astype(np.int64)
    out = {}
    if name == "C1":
        gold = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
        rows = gold["C1"]
        qs = [tuple(r["query"]) for r in rows]
        kw = dict(eng_kw, node_budget=gold["C1_search"]["node_cap"], deadline_ms=0)
        E = scs.Engine(G, **kw)
        E.run([qs[0]])
        t0 = time.perf_counter()
        res = E.run(qs, tags=list(range(len(qs))), progress=20.0)
        wall = time.perf_counter() - t0
        E.close()
        out = summary(res, wall)

        def golden_check(rr):
            both = [(r, row) for r, row in zip(rr, rows)
                    if r.status in PROVEN and row["status"] != "CAPPED"]
            return {"golden_both_proven": len(both),
                    "equal_to_golden": sum(r.kl == row["kl"] for r, row in both),
                    "at_or_above_golden_lower_bound": sum(
                        r.kl >= row["kl"] for r, row in zip(rr, rows)
                        if row["status"] == "CAPPED" and r.status in PROVEN)}
        out.update({"node_budget": kw["node_budget"], "golden_proven": gold["C1_search"]["proven"]})
        out.update(golden_check(res))
        # the learned order on the same 100 queries, with the C1 model
        mpath = os.path.join(ROOT, "models", "branching_c1.npz")
        if os.path.exists(mpath):
            EL = scs.Engine(G, **kw)
            EL.set_model(scs.Model.load(mpath), True)
            EL.run([qs[0]])
            t0 = time.perf_counter()
            lres = EL.run(qs, tags=list(range(len(qs))), progress=20.0)
            lw = time.perf_counter() - t0
            EL.close()
            ls = summary(lres, lw)
            ls.update(golden_check(lres))
            both = [(a, b) for a, b in zip(res, lres) if a.status in PROVEN and b.status in PROVEN]
            ls.update({"model": os.path.relpath(mpath, ROOT), "both_proven": len(both),
                       "optimum_equal": all(a.kl == b.kl for a, b in both),
                       "nodes_on_both_proven": {"plain": int(sum(a.nodes for a, _ in both)),
                                                "learned": int(sum(b.nodes for _, b in both))}})
            out["learned"] = ls
            out["learned_vs_plain"] = round(ls["value"] / max(1e-9, out["value"]), 3)
        # the plain order one query at a time on the first queries
        sq = qs[:args.c1_single]
        if sq:
            ES = scs.Engine(G, **dict(kw, limit_queries=1))
            ES.run([qs[0]])
            t0 = time.perf_counter()
            sres = run_one_at_a_time(ES, sq, progress=20.0)
            sw = time.perf_counter() - t0
            ES.close()
            ss = summary(sres, sw)
            ss["sample"] = len(sq)
            out["plain_one_at_a_time"] = ss
            if ss["value"] > 0:
                out["batched_vs_one_at_a_time"] = round(out["value"] / ss["value"], 2)
                if "learned" in out:
                    out["learned_batched_vs_one_at_a_time"] = round(
                        out["learned"]["value"] / ss["value"], 2)
    else:
        B = args.batch
        qs = configs.queries(name, core, count=args.extra_batches * B)
        batches = [[(q.l, q.h, q.q, q.heu, q.limit_doms) for q in qs[i * B:(i + 1) * B]]
                   for i in range(args.extra_batches)]
        E = scs.Engine(G, **eng_kw)
        E.run([batches[0][0]])
        t0 = time.perf_counter()
        res = run_batches(E, batches, progress=20.0, inflight=args.inflight)
        wall = time.perf_counter() - t0
        E.close()
        out = summary(res, wall)
        out["batches"] = f"{args.extra_batches} x {B}"
    G.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="C4")
    ap.add_argument("--batch", type=int, default=64, help="queries per step")
    ap.add_argument("--inflight", type=int, default=4,
                    help="batches in the engine at once (1 = each batch drained before the next)")
    ap.add_argument("--budget", type=int, default=20000, help="node budget per query")
    ap.add_argument("--deadline-ms", type=int, default=2000,
                    help="per-query wall-clock bound from submission (CAPPED past it)")
    ap.add_argument("--heuristic-threads", type=int, default=12)
    ap.add_argument("--gpu-heuristic", type=int, default=0,
                    help="1 = admission computes the greedy heuristic on the GPU")
    ap.add_argument("--waves-per-cu", type=int, default=4)
    ap.add_argument("--limit-queries", type=int, default=256)
    ap.add_argument("--slice-us", type=int, default=50000)
    ap.add_argument("--copy-limit", type=float, default=0.5)
    ap.add_argument("--admit-per-step", type=int, default=256)
    ap.add_argument("--levels", type=int, default=1, help="descending level floors")
    ap.add_argument("--pair", type=int, default=1, help="common-neighbour pair rule")
    ap.add_argument("--learned", type=int, default=1,
                    help="1 = the main leg uses the learned order (the serving form; the plain "
                         "order runs beside it), 0 = the main leg is the plain order")
    ap.add_argument("--model", default=None,
                    help="branching model (default models/branching_<config>.npz when present, "
                         "else the C1 model)")
    ap.add_argument("--prep-sample", type=int, default=256,
                    help="queries preprocessed alone (batched) for the preprocessing roofline")
    ap.add_argument("--single-sample", type=int, default=16,
                    help="timed queries also run one at a time (plain order)")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prep-modes", default="1,0",
                    help="--prep-only legs: 1 = as admitted (whole-graph queries), 0 = local CSRs")
    ap.add_argument("--prep-only", action="store_true",
                    help="only run the standalone batched-preprocessing leg (--prep-sample queries)")
    ap.add_argument("--extra", default="C1,C3",
                    help="further workloads reported beside the headline: C1 (all 100 queries at "
                         "the golden file's node cap, optima checked against it) and/or C3 "
                         "(--extra-batches batches like the headline's)")
    ap.add_argument("--extra-batches", type=int, default=5)
    ap.add_argument("--c1-single", type=int, default=100,
                    help="C1 queries also run one at a time (plain order)")
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

    from scs_amd import configs, scs

    print(f"[bench] building {args.config}", file=sys.stderr, flush=True)
    g = configs.graph(args.config)
    cfg = configs.CONFIGS[args.config]
    print(f"[bench] graph n={g.n} m={g.m}", file=sys.stderr, flush=True)
    G = scs.Graph.from_csr(g.off, g.adj)
    core = G.core()
    B = args.batch
    need = (args.warmup + args.steps) * B
    qs = configs.queries(args.config, core.astype(np.int64), count=max(need, args.prep_sample))
    allq = [(q.l, q.h, q.q, q.heu, q.limit_doms) for q in qs]
    warm = [allq[i * B:(i + 1) * B] for i in range(args.warmup)]
    timed = [allq[(args.warmup + i) * B:(args.warmup + i + 1) * B] for i in range(args.steps)]
    timed_flat = [q for b in timed for q in b]
    # a workload's own model when it has one; otherwise the model trained on
    # the proven optima of C1 (C4 gives no proven training queries: none of
    # 800 within its budget and deadline)
    mpath = args.model or os.path.join(ROOT, "models", f"branching_{args.config.lower()}.npz")
    if not args.model and not os.path.exists(mpath):
        mpath = os.path.join(ROOT, "models", "branching_c1.npz")
    model = scs.Model.load(mpath) if os.path.exists(mpath) else None
    eng_kw = dict(node_budget=args.budget, waves_per_cu=args.waves_per_cu,
                  limit_queries=args.limit_queries, time_slice_us=args.slice_us,
                  copy_limit=args.copy_limit, mem_budget=32 << 30,
                  heuristic_threads=args.heuristic_threads, admit_per_step=args.admit_per_step,
                  gpu_heuristic=args.gpu_heuristic,
                  levels=args.levels, pair=args.pair, deadline_ms=args.deadline_ms)
    if args.prep_only:
        sample = allq[:args.prep_sample or len(allq)]
        out = {"workload": args.config}
        modes = [int(x) for x in args.prep_modes.split(",") if x]
        if 1 in modes:
            out["prep_standalone"] = prep_leg(scs, G, core, sample)
        if 0 in modes:
            out["prep_standalone_local"] = prep_leg(scs, G, core, sample, whole=0)
        print(json.dumps(out), flush=True)
        return
    # the main leg runs the learned order (the serving form) when a model is
    # present; the other order runs beside it on the same queries
    learned = bool(args.learned) and model is not None
    E = scs.Engine(G, **eng_kw)
    if learned:
        E.set_model(model, True)
    # code objects load outside the timed region
    E.run([allq[0]])
    if warm:
        print(f"[bench] warmup: {args.warmup} batches of {B}", file=sys.stderr, flush=True)
        run_batches(E, warm, inflight=args.inflight)

    s0 = E.stats()
    print(f"[bench] timed: {args.steps} batches of {B}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res = run_batches(E, timed, progress=20.0, inflight=args.inflight)
    t1 = time.perf_counter()
    s1 = E.stats()
    E.close()
    wall = t1 - t0
    main_sum = summary(res, wall)
    nodes = s1["nodes"] - s0["nodes"]
    # roofline of the dominant kernel (scs_epoch) over every launch of the
    # process, so the average launch duration matches a kernel-trace profile
    # of the same command
    engine_us = s1["engine_us"]
    epochs = s1["epochs"]
    edges = s1["edges"]
    ach = (edges * BYTES_PER_EDGE) / (engine_us * 1e-6) / 1e9 if engine_us > 0 else 0.0
    gdesc = (f"Chung-Lu n={g.n} m={g.m}" if cfg["kind"] == "chung_lu" else
             f"R-MAT scale {cfg['scale']} n={g.n} m={g.m}")
    line = {
        "metric": "proven-optimal queries/sec (p99 latency beside it), learned vs plain branching",
        "value": main_sum["value"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (deterministic generated graph and query stream)",
        "config": {"workload": f"{args.config}: {gdesc}, {args.steps} batches of {B} queries "
                               f"from the stream",
                   "node_budget": args.budget, "deadline_ms": args.deadline_ms,
                   "learned": learned,
                   "levels": args.levels, "pair": args.pair, "waves": s1["waves"],
                   "limit_queries": args.limit_queries, "inflight": args.inflight,
                   "gpu_heuristic": args.gpu_heuristic},
        "proven": main_sum["proven"],
        "queries": main_sum["queries"],
        "proven_optimal_fraction": main_sum["proven_optimal_fraction"],
        "p50_ms": main_sum["p50_ms"],
        "p99_ms": main_sum["p99_ms"],
        "p99_ms_all": round(percentile([r.latency_us / 1000.0 for r in res], 99), 3),
        "status_counts": main_sum["status_counts"],
        "nodes": int(nodes),
        "nodes_per_s": round(nodes / wall, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": None,
                     "kernel": "scs_epoch", "epoch_avg_us": round(engine_us / max(1, epochs), 2),
                     "epochs": int(epochs), "bytes_per_edge": BYTES_PER_EDGE, "edges": int(edges),
                     "algorithmic_bytes_per_launch": int(edges * BYTES_PER_EDGE / max(1, epochs))},
        "prep": {"device_us": round(s1["prep_us"] - s0["prep_us"], 1),
                 "algorithmic_bytes": int(s1["prep_bytes"] - s0["prep_bytes"])},
        "search_phases": phase_shares(s0, s1),
        "heuristic_host_us": round(s1["heu_us"] - s0["heu_us"], 1),
        "core_decomposition_us": round(s1["core_us"], 1),
        "kcap": int(s1["kcap"]),
    }
    pb, pu = line["prep"]["algorithmic_bytes"], line["prep"]["device_us"]
    line["prep"]["achieved_GBs"] = round(pb / (pu * 1e-6) / 1e9, 2) if pu > 0 else 0.0
    # HBM traffic per launch from the committed hardware-counter profile of
    # this workload and engine setting (two rocprofv3 --pmc runs, FETCH_SIZE
    # and WRITE_SIZE, scaled by a copy pass of known size)
    prof = os.path.join(ROOT, "profiles", f"scs_epoch_{args.config.lower()}_traffic.json")
    if os.path.exists(prof):
        t = json.load(open(prof))
        same = all(t.get(k) == v for k, v in (("workload", args.config),
                                              ("waves_per_cu", args.waves_per_cu),
                                              ("time_slice_us", args.slice_us),
                                              ("node_budget", args.budget),
                                              ("batch", B), ("levels", args.levels),
                                              ("pair", args.pair),
                                              ("deadline_ms", args.deadline_ms),
                                              ("inflight", args.inflight),
                                              ("learned", int(learned))))
        if same:
            line["roofline"]["traffic"] = t["traffic_bytes_per_launch"]
            line["roofline"]["traffic_source"] = os.path.relpath(prof, ROOT)
    by_tag = {r.tag: r for r in res}

    def agree(other):
        both = [(by_tag[r.tag], r) for r in other
                if r.status in PROVEN and r.tag in by_tag and by_tag[r.tag].status in PROVEN]
        return len(both), sum(a.kl == b.kl for a, b in both)

    # the same timed queries in the other order (batched)
    if model is not None:
        other = "plain" if learned else "learned"
        print(f"[bench] {other} leg", file=sys.stderr, flush=True)
        EL = scs.Engine(G, **eng_kw)
        if not learned:
            EL.set_model(model, True)
        EL.run([allq[0]])
        t0 = time.perf_counter()
        lres = run_batches(EL, timed, progress=20.0, inflight=args.inflight)
        lw = time.perf_counter() - t0
        EL.close()
        ls = summary(lres, lw)
        nb, ne = agree(lres)
        ls.update({"both_proven": nb, "optimum_equal": ne == nb})
        if not learned:
            ls["model"] = os.path.relpath(mpath, ROOT)
        line[f"{other}_batched"] = ls
        lv, pv = (main_sum["value"], ls["value"]) if learned else (ls["value"], main_sum["value"])
        line["learned_vs_plain_batched"] = round(lv / max(1e-9, pv), 3)
    if learned:
        line["config"]["model"] = os.path.relpath(mpath, ROOT)
    # the plain order one query at a time (one resident query), on a sample
    # of the timed queries, against the batched rate
    if args.single_sample > 0:
        print("[bench] one-query-at-a-time leg", file=sys.stderr, flush=True)
        sq = timed_flat[:args.single_sample]
        ES = scs.Engine(G, **dict(eng_kw, limit_queries=1))
        ES.run([allq[0]])
        t0 = time.perf_counter()
        sres = run_one_at_a_time(ES, sq, progress=20.0)
        sw = time.perf_counter() - t0
        ES.close()
        ss = summary(sres, sw)
        ss["sample"] = len(sq)
        nb, ne = agree(sres)
        ss.update({"both_proven": nb, "optimum_equal": ne == nb})
        line["plain_one_at_a_time"] = ss
        if ss["value"] > 0:
            line["batched_vs_one_at_a_time"] = round(main_sum["value"] / ss["value"], 2)
            for k in ("learned_batched", "plain_batched"):
                if k in line:
                    line[f"{k}_vs_one_at_a_time"] = round(line[k]["value"] / ss["value"], 2)

    # batched preprocessing alone (no search kernel sharing the GPU): the
    # roofline of the preprocessing kernels
    if args.prep_sample > 0:
        line["prep_standalone"] = prep_leg(scs, G, core, allq[:args.prep_sample])
        line["prep_standalone_local"] = prep_leg(scs, G, core, allq[:args.prep_sample], whole=0)

    for name in [x for x in args.extra.split(",") if x]:
        line[f"extra_{name.lower()}"] = extra_leg(scs, configs, name, args, eng_kw)

    if not args.no_cpu:
        # the oracle (same rules and level order, same budget) on the timed
        # queries in order, one core, until the time bound; its proven
        # answers must equal the GPU's where both proved
        from oracle.oracle import FLAG_LEVELS, FLAG_PAIR, FLAGS_ALL, OracleGraph
        flags = FLAGS_ALL | (FLAG_LEVELS if args.levels else 0) | (FLAG_PAIR if args.pair else 0)
        og = OracleGraph.from_csr(g.n, g.off, g.adj)
        done = proven = both = equal = 0
        tc0 = time.perf_counter()
        for i, q in enumerate(timed_flat):
            o = og.solve(*q, flags=flags, node_cap=args.budget,
                         time_limit=args.deadline_ms / 1000.0)
            done += 1
            if o["status"] in PROVEN:
                proven += 1
                r = by_tag.get(i)
                if r is not None and r.status in PROVEN:
                    both += 1
                    equal += int(r.kl == o["kl"])
            if done % 10 == 0:
                print(f"[cpu] {done} queries, {time.perf_counter() - tc0:.1f} s", file=sys.stderr,
                      flush=True)
            if time.perf_counter() - tc0 > args.cpu_seconds:
                break
        tc = time.perf_counter() - tc0
        line["cpu_baseline"] = {
            "value": round(proven / tc, 4), "unit": "queries/s", "cores": 1, "kind": "port",
            "sample": (f"first {done} timed queries in stream order, serial oracle (same rules, "
                       f"node budget {args.budget}, {args.deadline_ms} ms per query), "
                       f"{proven} proven in {tc:.1f} s"),
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}
        line["oracle_check"] = {"both_proven": both, "equal": equal}
    G.close()
    out = json.dumps(line)
    print(out, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")


if __name__ == "__main__":
    main()
