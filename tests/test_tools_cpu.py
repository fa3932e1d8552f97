"""Host-side pieces of the bench, autotuner and learning driver that run
without a GPU: argument contracts, grid parsing and the deterministic
training-query stream."""

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_autotune_grid_parsing():
    from scs_amd.autotune import GRID, parse_grid
    g = parse_grid("waves_per_cu=2,4;time_slice_us=200000")
    assert g["waves_per_cu"] == [2, 4]
    assert g["time_slice_us"] == [200000]
    assert g["limit_queries"] == GRID["limit_queries"]
    assert all(isinstance(x, float) for x in parse_grid("copy_limit=0.1,1")["copy_limit"])


def test_training_queries_are_disjoint_from_evaluation_stream():
    from scs_amd import configs, graphgen as gg, learn
    g = configs.graph("C1")
    core = gg.core_numbers(g).astype(np.int64)
    a = learn.training_queries("C1", core, seeds=range(100, 104), per_seed=4)
    b = learn.training_queries("C1", core, seeds=range(100, 104), per_seed=4)
    assert a == b and len(a) == 16
    ev = [(q.l, q.h, q.q, q.heu, q.limit_doms) for q in configs.queries("C1", core)]
    # seeds 100..199 never equal the evaluation seed (graph seed + 1000)
    assert configs.CONFIGS["C1"]["seed"] + 1000 not in range(100, 200)
    for l, h, q, heu, ld in a:
        assert 5 <= l <= 15 and l + 5 <= h <= l + 15 and core[q] >= 10
