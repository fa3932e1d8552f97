"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/scs.h, and its host heuristic reproduces the oracle's
restatement of CSSC_heu (kl and H) exactly."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle.oracle import OracleGraph
from scs_amd import configs, graphgen as gg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "scs_amd", "libscs.so")
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "scs.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[a-z_ ]+[ \*]+(scs_[a-z_]+)\(", text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libscs.so not built: run python scs_amd/build.py")
    return C.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n


def host_heuristic(lib, off, adj, l, h, q):
    kl = C.c_uint32()
    hs = C.c_uint32()
    H = np.zeros(len(off) + h + 1, np.uint32)
    f = lib.scs_heuristic_csr
    f.argtypes = [np.ctypeslib.ndpointer(np.int64, flags="C"), np.ctypeslib.ndpointer(np.int32, flags="C"),
                  C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                  np.ctypeslib.ndpointer(np.uint32, flags="C"), C.c_uint32, C.POINTER(C.c_uint32)]
    rc = f(off, adj, len(off) - 1, l, h, q, C.byref(kl), H, H.size, C.byref(hs))
    assert rc == 0
    return kl.value, H[:hs.value]


def dense_graph(n, p, seed):
    """G(n, p) as a CSR (sorted rows, symmetric, no self-loops)."""
    rng = np.random.default_rng(seed)
    a = np.triu(rng.random((n, n)) < p, 1)
    a = a | a.T
    off = np.concatenate([[0], np.cumsum(a.sum(1))]).astype(np.int64)
    adj = np.nonzero(a)[1].astype(np.int32)
    return OracleGraph.from_csr(n, off, adj)


@pytest.mark.parametrize("gname", ["100.txt", "hepPH_SCS", "C1", "C3", "dense"])
def test_host_heuristic_equals_oracle(lib, gname):
    if gname in ("C1", "C3"):
        # C3 (R-MAT, 1M vertices): every query's greedy walks into the hubs,
        # touching about a million vertices per construction
        g = configs.graph(gname)
        og = OracleGraph.from_csr(g.n, g.off, g.adj)
        core = og.core()
        qs = [(x.l, x.h, x.q) for x in
              configs.queries(gname, core.astype(np.int64), count=None if gname == "C1" else 8)]
        qs = qs[:40]
    elif gname == "dense":
        # vertices with up to h neighbours in H: heu3's neighbour lists grow
        # past their in-place room and through several doublings
        og = dense_graph(300, 0.4, 5)
        rng = np.random.default_rng(4)
        qs = []
        for _ in range(20):
            l = int(rng.integers(2, 60))
            qs.append((l, l + int(rng.integers(0, 80)), int(rng.integers(0, og.n))))
    else:
        og = OracleGraph.load(os.path.join(GOLD, gname))
        rng = np.random.default_rng(3)
        qs = []
        for _ in range(40):
            l = int(rng.integers(1, 20))
            qs.append((l, l + int(rng.integers(0, 20)), int(rng.integers(0, og.n))))
    off, adj = og.csr()
    for l, h, q in qs:
        a = host_heuristic(lib, off, adj, l, h, q)
        b = og.heuristic(0, l, h, q)
        assert a[0] == b[0], (gname, l, h, q)
        assert list(a[1]) == list(b[1]), (gname, l, h, q)


def test_host_heuristic_rejects_bad_input(lib):
    """scs_heuristic_csr returns SCS_EINVAL (-1) instead of reading outside
    the caller's arrays."""
    f = lib.scs_heuristic_csr
    f.argtypes = [np.ctypeslib.ndpointer(np.int64, flags="C"), np.ctypeslib.ndpointer(np.int32, flags="C"),
                  C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                  C.c_void_p, C.c_uint32, C.c_void_p]
    kl = C.c_uint32()
    off = np.array([0, 1, 2], np.int64)
    adj = np.array([1, 0], np.int32)
    assert f(off, adj, 2, 1, 2, 0, C.byref(kl), None, 0, None) == 0 and kl.value == 1
    assert f(off, adj, 2, 1, 2, 2, C.byref(kl), None, 0, None) == -1      # q >= n
    assert f(off, adj, 2, 3, 2, 0, C.byref(kl), None, 0, None) == -1      # l > h
    assert f(off, adj, 2, 0, 0, 0, C.byref(kl), None, 0, None) == -1      # h == 0
    assert f(off, np.array([1, 2], np.int32), 2, 1, 2, 0, C.byref(kl), None, 0, None) == -1
    assert f(off, np.array([1, -1], np.int32), 2, 1, 2, 0, C.byref(kl), None, 0, None) == -1
    assert f(np.array([0, 2, 1], np.int64), adj, 2, 1, 2, 0, C.byref(kl), None, 0, None) == -1
    assert f(np.array([1, 1, 2], np.int64), adj, 2, 1, 2, 0, C.byref(kl), None, 0, None) == -1


def test_ctypes_structs_match_the_header(tmp_path):
    """The Python binding's structures have the size and field offsets the C
    header gives (a program built with g++ against include/scs.h prints
    them), so a field added on one side only fails here."""
    import subprocess
    from scs_amd import scs
    structs = {"scs_params": scs.Params, "scs_result": scs.Result, "scs_stats": scs.Stats,
               "scs_model": scs.Model}
    lines = ['#include <cstddef>', '#include <cstdio>', '#include "scs.h"', "int main() {"]
    for cname, py in structs.items():
        lines.append(f'  std::printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  std::printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    exe = str(tmp_path / "layout")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, c = ln.split()
            got[(a, b)] = int(c)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def _csr(n, rows):
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    adj = np.array([x for r in rows for x in r], np.int32)
    return off, adj


@pytest.mark.parametrize("case,rows", [
    ("asymmetric", [[1], [], [0]]),            # 0-1 without 1-0; 2-0 without 0-2
    ("self_loop", [[0, 1], [0], []]),
    ("duplicate", [[1, 1], [0, 0], []]),
    ("unsorted", [[2, 1], [0], [0]]),
    ("missing_reverse", [[1, 2], [0], []]),
    ("out_of_range", [[3], [], []]),
])
def test_graph_create_rejects_non_simple_csr(case, rows):
    """scs_graph_create validates the CSR on the host before any device call:
    sorted unique rows, no self-loops, symmetric adjacency.  The search
    kernels' per-neighbour updates assume exactly that."""
    from scs_amd import scs
    off, adj = _csr(3, rows)
    with pytest.raises(scs.SCSError) as e:
        scs.Graph.from_csr(off, adj)
    assert e.value.code == -1, case


def test_reference_loader_output_passes_validation(lib):
    """A CSR that the reference-semantics loader produces (here the oracle's
    restatement of it) is simple and symmetric, so validation accepts its
    structure: the call goes on to the device and fails there only for lack
    of a GPU, never with SCS_EINVAL."""
    from scs_amd import scs
    og = OracleGraph.load(os.path.join(GOLD, "hepPH_SCS"))
    off, adj = og.csr()
    try:
        scs.Graph.from_csr(off, adj).close()
    except scs.SCSError as e:
        assert e.code != -1
