"""Continuous-batching serving driver: Poisson query arrivals.

Queries arrive at exponential gaps (rate in queries/s) and are submitted
when their arrival time has passed; every scs_step admits what fits (slots
and device memory) while the previous admissions search on the GPU, so new
arrivals overlap running work.  Reports achieved queries/s and the
arrival-to-result latency percentiles.
"""

from __future__ import annotations

import time

import numpy as np


def poisson_run(E, queries, rate_qps, seed=0, timeout_s=600.0):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_qps, len(queries)))
    res = []
    i = 0
    t0 = time.perf_counter()
    while len(res) < len(queries):
        now = time.perf_counter() - t0
        if now > timeout_s:
            raise TimeoutError("serving run exceeded its time limit")
        while i < len(queries) and arrivals[i] <= now:
            l, h, q, heu, ld = queries[i]
            E.submit(l, h, q, heu, ld, i)
            i += 1
        left = E.step()
        res.extend(E.poll())
        if left == 0 and i < len(queries):
            time.sleep(max(0.0, min(0.002, arrivals[i] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    lat = np.array([r.latency_us for r in res]) / 1000.0
    res.sort(key=lambda r: r.tag)
    return {"offered_qps": rate_qps, "achieved_qps": len(res) / wall, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "results": res}


def main():
    import argparse
    import json
    import sys

    from . import configs, scs

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C2")
    ap.add_argument("--queries", type=int, default=300)
    ap.add_argument("--rates", default="50,200,1000")
    ap.add_argument("--budget", type=int, default=300)
    ap.add_argument("--model", default=None, help="learned-order model (.npz)")
    ap.add_argument("--limit-queries", type=int, default=256)
    ap.add_argument("--slice-us", type=int, default=50000)
    ap.add_argument("--mem-gb", type=float, default=16.0, help="device arena for local graphs")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    print(f"[serve] building {args.config}", file=sys.stderr, flush=True)
    g = configs.graph(args.config)
    G = scs.Graph.from_csr(g.off, g.adj)
    core = G.core().astype(np.int64)
    qs = [(q.l, q.h, q.q, q.heu, q.limit_doms)
          for q in configs.queries(args.config, core, count=args.queries)]
    model = scs.Model.load(args.model) if args.model else None
    report = []
    for mode in (["plain", "learned"] if model is not None else ["plain"]):
        for rate in [float(x) for x in args.rates.split(",")]:
            E = scs.Engine(G, node_budget=args.budget, limit_queries=args.limit_queries,
                           time_slice_us=args.slice_us, mem_budget=int(args.mem_gb * (1 << 30)))
            if mode == "learned":
                E.set_model(model, True)
            E.run([qs[0]])
            r = poisson_run(E, qs, rate)
            E.close()
            kl_sum = int(sum(x.kl for x in r["results"]))
            row = {"mode": mode, "offered_qps": rate, "achieved_qps": round(r["achieved_qps"], 2),
                   "p50_ms": round(r["p50_ms"], 2), "p99_ms": round(r["p99_ms"], 2),
                   "kl_sum": kl_sum}
            print(json.dumps(row), flush=True)
            report.append(row)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(report, f, indent=1)
    G.close()


if __name__ == "__main__":
    main()
