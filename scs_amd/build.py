"""Builds the SCS shared library (HIP kernels + C ABI) for gfx950, in-tree.

Output: scs_amd/libscs.so.  The host server binary (scs_amd/server) is built
by build_server().  Both are plain hipcc / g++ invocations; no JIT caches.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libscs.so")
SERVER = os.path.join(HERE, "server", "scs_server")
CLIENT = os.path.join(HERE, "server", "scs_client")

HIP_SOURCES = ["scs_graph.hip", "scs_prep.hip", "scs_engine.hip", "scs_mlp.hip", "scs_heu.hip"]
CPP_SOURCES = ["scs_capi.cpp", "scs_heuristic.cpp"]
ARCH = os.environ.get("SCS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for every HIP translation unit.
DEV_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build_lib(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", *DEV_FLAGS,
              "-c", path, "-o", obj])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_server(verbose: bool = False) -> str:
    sdir = os.path.join(HERE, "server")
    for src, out in (("scs_server.cpp", SERVER), ("scs_client.cpp", CLIENT)):
        path = os.path.join(sdir, src)
        if not os.path.exists(path):
            continue
        libs = ["-L" + HERE, "-lscs", "-Wl,-rpath,$ORIGIN/.."] if src == "scs_server.cpp" else []
        _run(["g++", "-O2", "-std=c++17", path, "-o", out, "-I" + os.path.join(HERE, ".."),
              *libs, "-lpthread"])
        if verbose:
            print("built", out)
    return SERVER


if __name__ == "__main__":
    build_lib(verbose=True)
    build_server(verbose=True)
