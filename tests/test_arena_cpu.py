"""Host-side allocator of the per-query device arena, built with g++ and
exercised on the CPU: disjoint in-bounds allocations, coalescing, and
refusal of releases of ranges that were never handed out (a stale offset
returned to the free list would let two resident queries share memory)."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_allocator(tmp_path):
    exe = str(tmp_path / "arena_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "scs_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "arena_check.cpp"), "-o", exe],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "arena ok" in out.stdout
