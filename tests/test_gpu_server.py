"""End to end through the process surface: scs_server (12 positional
arguments) receives the reference's query file over the SysV queue from
scs_client and writes the pipe-separated results file; every Degree equals
the golden optimum."""

import json
import os
import subprocess
import time

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVER = os.path.join(ROOT, "scs_amd", "server", "scs_server")
CLIENT = os.path.join(ROOT, "scs_amd", "server", "scs_client")
GOLD = os.path.join(ROOT, "tests", "golden")


# the 11 arguments after the graph: this build's defaults, and the fixed
# arguments of the reference's own runtime script (scripts/GPU_runtime.sh:22:
# partitionSize 300000, bufferSize 100000 words, copyLimit 0.5, readLimit 1,
# limitQueries 8, factor 100, all rules and bounds on)
ARGS = {"defaults": ["0", "0", "0.5", "1", "8", "1", "1", "1", "1", "1", "1"],
        "reference_script": ["300000", "100000", "0.5", "1", "8", "100", "1", "1", "1", "1", "1"]}


@pytest.mark.parametrize("args", list(ARGS))
def test_server_round_trip(tmp_path, args):
    gold = {tuple(r["query"]): r["kl"] for r in json.load(open(os.path.join(GOLD, "golden.json")))["100.txt"]}
    res_dir = tmp_path / "results"
    srv = subprocess.Popen([SERVER, os.path.join(GOLD, "100.txt"), *ARGS[args],
                            "--results", str(res_dir), "--queue-key", "x"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    try:
        time.sleep(1.0)
        subprocess.run([CLIENT, os.path.join(GOLD, "queries_100.txt"), "x"], check=True, timeout=60)
        out, _ = srv.communicate(timeout=120)
    finally:
        if srv.poll() is None:
            srv.kill()
    assert srv.returncode == 0, out
    lines = (res_dir / "results_100.txt").read_text().strip().split("\n")
    assert lines[0] == "N1|N2|QID|Time|Degree|Overtime|Heu|TotalTime"
    assert lines[-1].startswith("-1|-1|-1|-1|-1|-1|-1|")
    rows = [l.split("|") for l in lines[1:-1]]
    assert len(rows) == 7
    for r in rows:
        q = (int(r[0]), int(r[1]), int(r[2]), 1, 1)
        assert int(r[4]) == gold[q], r
        assert r[5] == "0"
