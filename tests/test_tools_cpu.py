"""Host-side pieces of the bench, autotuner and learning driver that run
without a GPU: argument contracts, grid parsing, the deterministic
training-query stream, and the continuous-batching feeds (against a host
stand-in for the engine's submit/step/poll)."""

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


class FakeEngine:
    """The submit/step/poll surface of scs.Engine on the host: a submitted
    query finishes `delay` steps later.  Records how many queries were in the
    engine after each submission."""

    def __init__(self, delay=3):
        self.delay = delay
        self.pending = []  # [steps left, tag, query]
        self.done = []
        self.max_inside = 0

    def submit(self, l, h, q, heu, ld, tag):
        self.pending.append([self.delay, tag, (l, h, q, heu, ld)])
        self.max_inside = max(self.max_inside, len(self.pending))

    def step(self):
        for p in self.pending:
            p[0] -= 1
        fin = [p for p in self.pending if p[0] <= 0]
        self.pending = [p for p in self.pending if p[0] > 0]

        class R:
            def __init__(self, tag, qq):
                self.tag, self.query = tag, qq

        self.done.extend(R(t, qq) for _, t, qq in fin)
        return len(self.pending)

    def poll(self):
        out, self.done = self.done, []
        return out


def _batches(nb, b):
    return [[(i * b + j, 1, 2, 1, 8) for j in range(b)] for i in range(nb)]


def test_bench_feed_keeps_at_most_inflight_batches_and_returns_every_query():
    sys.path.insert(0, ROOT)
    import bench
    for inflight in (1, 2, 4):
        E = FakeEngine()
        batches = _batches(7, 5)
        res = bench.run_batches(E, batches, inflight=inflight)
        assert [r.tag for r in res] == list(range(35))
        assert [r.query[0] for r in res] == list(range(35))
        assert E.max_inside <= inflight * 5
        assert E.max_inside > (inflight - 1) * 5


def test_bench_one_at_a_time_has_a_single_query_inside():
    sys.path.insert(0, ROOT)
    import bench

    class E1(FakeEngine):
        def run(self, qs, tags):
            for q, t in zip(qs, tags):
                self.submit(*q, t)
            out = []
            while self.step():
                out.extend(self.poll())
            return out + self.poll()

    E = E1()
    res = bench.run_one_at_a_time(E, [q for b in _batches(2, 3) for q in b])
    assert [r.tag for r in res] == list(range(6)) and E.max_inside == 1


def test_autotune_feed_matches_the_bench_feed():
    from scs_amd.autotune import feed
    E = FakeEngine(delay=2)
    res = feed(E, _batches(5, 4), 3)
    assert sorted(r.tag for r in res) == list(range(20))
    assert 8 < E.max_inside <= 12
