"""tc_csr_count_device_noalloc (scratch in a static device array, no device
allocation) against tc_csr_count_device, the entry point it is a variant of,
through the C-ABI: two calls in a row on one GPU, the second on a smaller
graph (the scratch keeps the first call's contents and must be cleared), then
malformed inputs, which must be rejected with the library's error status."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import custingeralg_amd as T
import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc():
    f = T.lib().tc_csr_count_device_noalloc
    f.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_uint64), C.c_void_p]
    f.restype = C.c_int
    return f


def _both(off, adj):
    """(per-vertex, global) of the allocating entry and of the no-alloc one on
    the same device CSR."""
    nv = len(off) - 1
    doff, dadj = T.DeviceArray.from_host(off), T.DeviceArray.from_host(adj)
    out_a, out_b = T.DeviceArray(nv, np.uint32), T.DeviceArray(nv, np.uint32)
    ga, gb = C.c_uint64(0), C.c_uint64(0)
    T.check(T.lib().tc_csr_count_device(nv, len(adj), doff.ptr, dadj.ptr, out_a.ptr,
                                        C.byref(ga), None))
    # fill the output with garbage first: the entry point overwrites all of it
    T.check(T.lib().tc_memset(out_b.ptr, 0xA5, out_b.nbytes))
    T.check(_noalloc()(nv, len(adj), doff.ptr, dadj.ptr, out_b.ptr, C.byref(gb), None))
    res = (out_a.to_host(), int(ga.value), out_b.to_host(), int(gb.value))
    for a in (doff, dadj, out_a, out_b):
        a.free()
    return res


@pytest.mark.parametrize("graphs", [
    [("rmat", 16, 16, 3), ("rmat", 12, 8, 5)],          # larger, then smaller
    [("rmat", 14, 24, 7), ("complete", 130, 0, 0)],     # skewed, then dense (merge path)
])
def test_noalloc_matches_allocating_entry(graphs):
    for kind, a, b, seed in graphs:
        if kind == "rmat":
            off, adj = O.rmat_graph(a, b, seed)
        else:
            n = a
            rows = np.repeat(np.arange(n), n - 1)
            cols = np.array([c for r in range(n) for c in range(n) if c != r], np.int32)
            off = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.int64)
            adj = cols
            assert len(rows) == len(adj)
        pa, ga, pb, gb = _both(off, adj)
        assert np.array_equal(pa, pb), (kind, a, b)
        assert ga == gb, (kind, a, b)
        ref = O.csr_count(off, adj)
        assert np.array_equal(pb, ref)
        assert gb == int(ref.astype(np.uint64).sum())


def test_noalloc_hub_and_empty_rows():
    """A star with a few chords (one hub row far longer than the rest: the
    binary-search path) plus isolated vertices (empty rows share offsets)."""
    n = 5000
    edges = [(0, i) for i in range(1, 4000)] + [(i, i + 1) for i in range(1, 3999, 2)]
    off, adj = O.csr_from_edges(n, edges)
    pa, ga, pb, gb = _both(off, adj)
    assert np.array_equal(pa, pb) and ga == gb
    assert gb == 3 * 1999  # one triangle per chord


def test_noalloc_rejects_malformed():
    f = _noalloc()
    off = np.array([0, 2, 3, 3], np.int64)
    for adj in (np.array([1, 7, 0], np.int32),      # id out of range
                np.array([2, 1, 0], np.int32)):     # row 0 not increasing
        doff, dadj = T.DeviceArray.from_host(off), T.DeviceArray.from_host(adj)
        out = T.DeviceArray(3, np.uint32)
        g = C.c_uint64(0)
        assert f(3, 3, doff.ptr, dadj.ptr, out.ptr, C.byref(g), None) == T.TC_ERR_INVALID
        for a in (doff, dadj, out):
            a.free()
    assert f(-1, 0, None, None, None, None, None) == T.TC_ERR_INVALID
