"""Host code under the sanitizers (CPU only, g++): the admission path's
heuristic threads (several threads running the greedy heuristic on the one
shared graph, each with its own scratch) under ThreadSanitizer and under
AddressSanitizer with UndefinedBehaviorSanitizer, and the device-arena
allocator's randomised check under AddressSanitizer.  Every run must be
clean and give the single-threaded answers."""

import os
import subprocess

import numpy as np
import pytest

from scs_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "scs_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")


def build(tmp_path, name, sources, san):
    exe = str(tmp_path / name)
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-Wall", "-fno-omit-frame-pointer", "-pthread",
           f"-fsanitize={san}", "-I", CSRC] + sources + ["-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def write_csr(path, g):
    with open(path, "wb") as f:
        f.write(np.uint32(g.n).tobytes())
        f.write(np.uint64(len(g.adj)).tobytes())
        f.write(np.ascontiguousarray(g.off, np.int64).tobytes())
        f.write(np.ascontiguousarray(g.adj, np.int32).tobytes())


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_heuristic_threads(tmp_path, san):
    exe = build(tmp_path, "heu_threads", [os.path.join(NATIVE, "heu_threads_check.cpp"),
                                          os.path.join(CSRC, "scs_heuristic.cpp")], san)
    csr = str(tmp_path / "c1.bin")
    write_csr(csr, configs.graph("C1"))
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    out = subprocess.run([exe, csr, "200", "8"], capture_output=True, text=True, timeout=600,
                         env=env)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "threads ok" in out.stdout
    assert "WARNING" not in out.stderr and "runtime error" not in out.stderr, out.stderr[-4000:]


def test_arena_under_asan(tmp_path):
    exe = build(tmp_path, "arena_asan", [os.path.join(NATIVE, "arena_check.cpp")],
                "address,undefined")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, ASAN_OPTIONS="halt_on_error=1"))
    assert out.returncode == 0, out.stderr[-4000:]
    assert "arena ok" in out.stdout
