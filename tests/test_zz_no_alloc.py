"""scs_core_numbers_noalloc against scs_graph_create's core numbers.

The allocation-free entry point works in a fixed device array inside the
library; the caller passes the CSR and the output in its own device memory.
Here the caller's buffers come from hipMalloc of the HIP runtime the library
itself links (torch bundles a second HIP runtime, which cannot initialise
once the library's has).  Two calls in a row, the second on a smaller graph,
must each give exactly the core numbers the allocating entry point computes
for the same graph (the array keeps the first call's contents, so the second
call shows that everything it reads is reset).  A graph above the array's
limit is refused with SCS_EBIG before any device work.
"""

import ctypes as C

import numpy as np
import pytest

from scs_amd import configs, scs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    scs.lib()  # loads the library and, with it, its HIP runtime
    for path in ("/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so.7"):
        try:
            h = C.CDLL(path)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 4)) == 0

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, arr):
        assert self.hip.hipMemcpy(arr.ctypes.data, self.p, arr.nbytes, D2H) == 0
        return arr

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc_fn():
    f = scs.lib().scs_core_numbers_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _noalloc(hip, off, adj, n):
    off = np.ascontiguousarray(off, np.int64)
    adj = np.ascontiguousarray(adj, np.int32)
    d_off, d_adj, d_core = DevBuf(hip, off.nbytes), DevBuf(hip, adj.nbytes), DevBuf(hip, 4 * n)
    d_off.put(off)
    d_adj.put(adj)
    d_core.put(np.full(n, 0xFFFFFFFF, np.uint32))
    rc = _noalloc_fn()(d_off.p, d_adj.p, n, d_core.p, None)
    assert hip.hipDeviceSynchronize() == 0
    core = d_core.get(np.zeros(n, np.uint32))
    for b in (d_off, d_adj, d_core):
        b.free()
    return rc, core


def _reference(off, adj):
    G = scs.Graph.from_csr(off, adj)
    core = G.core()
    G.close()
    return core


def test_noalloc_core_numbers_match_the_allocating_entry_point():
    hip = _hip()
    big = configs.graph("C2")      # 100,000 vertices, 3M adjacency entries
    small = configs.graph("C1")    # 4,000 vertices
    for g in (big, small):
        print("no-alloc core numbers, n =", g.n, flush=True)
        rc, got = _noalloc(hip, g.off, g.adj, g.n)
        assert rc == 0
        assert np.array_equal(got, _reference(g.off, g.adj)), g.n


def test_noalloc_rejects_graphs_above_its_limit():
    hip = _hip()
    d = DevBuf(hip, 64)
    n = (1 << 20) + 1
    # refused on the size alone, before any pointer is read
    assert _noalloc_fn()(d.p, d.p, n, d.p, None) == -6
    d.free()
