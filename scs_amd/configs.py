"""The five SCS workloads (graph + query stream) with their exact parameters.

C1  plain method, CPU-runnable: Chung-Lu seed 1, n=4,000, 88,000 edges,
    exponent 2.5, max degree ~1,000; 100 queries, q over core>=10,
    l in [5,15], h = l + [5,15], heuristic on, limit_dominating_set 8.
C2  plain method, batched on GPU: Chung-Lu seed 2, n=100,000, 1.5M edges,
    exponent 2.3; 2,000 queries, l in [5,40], h = l + [5,60], q over core>=8.
C3  learned branching: R-MAT seed 3, scale 20, edge factor 10,
    (a,b,c) = (.57,.19,.19); 5,000 queries, l in [10,60], h = l + [10,100].
C4  serving: Chung-Lu seed 4, n=500,000, 8M edges, exponent 2.2; l in [5,50],
    h = l + [5,150], heuristic 0/1, limit_dominating_set in {4,8,16}.
C5  large combined: R-MAT seed 5, scale 22, edge factor 16; 20,000 queries,
    l in [10,100], h = l + [20,200], q over core>=20 w.p. 0.7.
"""

from __future__ import annotations

import os
from functools import lru_cache

import numpy as np

from . import graphgen as gg

# max degree targets for the Chung-Lu graphs (only C1 states one; the others
# use sqrt-scaled caps so the largest weight stays well below 2m)
CONFIGS = {
    "C1": dict(kind="chung_lu", n=4_000, m=88_000, beta=2.5, dmax=1_000, seed=1,
               queries=dict(count=100, core_min=10, l_lo=5, l_hi=15, dh_lo=5, dh_hi=15,
                            heu=1, limit_doms=8)),
    "C2": dict(kind="chung_lu", n=100_000, m=1_500_000, beta=2.3, dmax=5_000, seed=2,
               queries=dict(count=2_000, core_min=8, l_lo=5, l_hi=40, dh_lo=5, dh_hi=60,
                            heu=1, limit_doms=8)),
    "C3": dict(kind="rmat", scale=20, edge_factor=10, a=0.57, b=0.19, c=0.19, seed=3,
               queries=dict(count=5_000, core_min=1, l_lo=10, l_hi=60, dh_lo=10, dh_hi=100,
                            heu=1, limit_doms=8)),
    "C4": dict(kind="chung_lu", n=500_000, m=8_000_000, beta=2.2, dmax=20_000, seed=4,
               queries=dict(count=2_000, core_min=1, l_lo=5, l_hi=50, dh_lo=5, dh_hi=150,
                            heu=(0, 1), limit_doms=(4, 8, 16))),
    "C5": dict(kind="rmat", scale=22, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=5,
               queries=dict(count=20_000, core_min=1, l_lo=10, l_hi=100, dh_lo=20, dh_hi=200,
                            heu=1, limit_doms=8, skew_core=20, skew_p=0.7)),
}


def generate(name: str) -> gg.CSR:
    """Builds the graph of a workload from its generator parameters."""
    c = CONFIGS[name]
    if c["kind"] == "chung_lu":
        return gg.chung_lu(c["n"], c["m"], c["beta"], c["dmax"], c["seed"])
    return gg.rmat(c["scale"], c["edge_factor"], c["a"], c["b"], c["c"], c["seed"])


def graph_params(name: str) -> dict:
    """The generator parameters that fix a workload's graph bit for bit."""
    return {k: v for k, v in CONFIGS[name].items() if k != "queries"}


@lru_cache(maxsize=4)
def graph(name: str, cache_dir: str | None = None) -> gg.CSR:
    """The workload's graph, from the binary CSR cache when it holds one for
    these generator parameters, else generated (and then cached).  The cache
    directory is ``cache_dir``, or $SCS_CSR_CACHE, or _cache/csr beside the
    package; SCS_CSR_CACHE=off disables it."""
    d = cache_dir or os.environ.get("SCS_CSR_CACHE") or os.path.join(
        os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_cache", "csr")
    if d == "off":
        return generate(name)
    g = gg.cache_load(d, graph_params(name))
    if g is None:
        g = generate(name)
        try:
            gg.cache_store(d, graph_params(name), g)
        except OSError:
            pass  # read-only home: generation still works
    return g


def queries(name: str, core: np.ndarray, count: int | None = None) -> list[gg.Query]:
    c = CONFIGS[name]
    qp = dict(c["queries"])
    if count is not None:
        qp["count"] = count
    return gg.make_queries(core, seed=c["seed"] + 1000, **qp)
