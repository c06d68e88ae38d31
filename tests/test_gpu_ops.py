"""The operator layer (include/custinger_ops.hpp) on the GPU: each traversal
runs a small operator from tests/cpp/ops_test.hip over a store and its result
is checked against numpy on the same CSR (the reference's operators,
include/operators.cuh, have no tests of their own; these are the properties
each operator defines)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

import custingeralg_amd as T
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_ops = None


def ops():
    global _ops
    if _ops is None:
        path = os.path.join(HERE, "cpp", "libops_test.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
        T.lib()
        _ops = C.CDLL(path)
        P, I64, I32 = C.c_void_p, C.c_int64, C.c_int32
        for name, args in {"ops_degrees": [P, P], "ops_subset_degrees": [P, P, I64, P],
                           "ops_edge_sums": [P, P, I64, P, P],
                           "ops_edge_sums_unbalanced_name": [P, P, I64, P, P],
                           "ops_indegree": [P, P], "ops_batch_found": [P, P, P, I64, P],
                           "ops_bfs": [P, I32, P, C.POINTER(C.c_int32)]}.items():
            f = getattr(_ops, name)
            f.argtypes = args
            f.restype = C.c_int
    return _ops


def _graph(kind):
    if kind == "rmat":
        return O.rmat_graph(14, 12, 7)
    if kind == "star":  # one hub row spread over many wavefronts, many empty rows
        n = 50000
        leaves = np.arange(1, 30001, dtype=np.int32)
        return O.canon_csr(n, np.zeros(len(leaves), np.int32), leaves)
    if kind == "empty":
        return np.zeros(11, np.int64), np.zeros(0, np.int32)
    raise ValueError(kind)


def _ok(st):
    assert st == 0, T.lib().tc_last_error().decode(errors="replace")


@pytest.mark.parametrize("kind", ["rmat", "star", "empty"])
def test_vertex_and_edge_traversals(kind):
    off, adj = _graph(kind)
    n = len(off) - 1
    deg = np.diff(off)
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    g = T.cuStinger(off, adj)
    lib = ops()
    try:
        out = T.DeviceArray(n, np.uint32).zero()
        _ok(lib.ops_degrees(g.h, out.ptr))
        assert np.array_equal(out.to_host(), deg.astype(np.uint32))

        rng = np.random.default_rng(3)
        verts = rng.integers(0, max(n, 1), 3000).astype(np.int32) if n else np.zeros(0, np.int32)
        dv = T.DeviceArray.from_host(verts)
        out.zero()
        _ok(lib.ops_subset_degrees(g.h, dv.ptr, len(verts), out.ptr))
        want = np.zeros(n, np.uint32)
        want[verts] = deg[verts]
        assert np.array_equal(out.to_host(), want)

        # every edge of the chosen vertices, once per occurrence in the array
        mult = np.bincount(verts, minlength=n).astype(np.uint64)
        want_cnt = (mult * deg).astype(np.uint32)
        want_sum = mult * np.bincount(rows, weights=adj, minlength=n).astype(np.uint64)
        for fn in ("ops_edge_sums", "ops_edge_sums_unbalanced_name"):
            s = T.DeviceArray(n, np.uint64).zero()
            c = T.DeviceArray(n, np.uint32).zero()
            _ok(getattr(lib, fn)(g.h, dv.ptr, len(verts), s.ptr, c.ptr))
            assert np.array_equal(c.to_host(), want_cnt), fn
            assert np.array_equal(s.to_host(), want_sum), fn

        out.zero()
        _ok(lib.ops_indegree(g.h, out.ptr))
        assert np.array_equal(out.to_host(), np.bincount(adj, minlength=n).astype(np.uint32))
    finally:
        g.close()


def test_batch_edges_and_updates():
    """allEinA over a batch, before and after the store takes the batch."""
    off, adj = _graph("rmat")
    n = len(off) - 1
    rng = np.random.default_rng(5)
    u = rng.integers(0, n, 4000).astype(np.int32)
    v = rng.integers(0, n, 4000).astype(np.int32)
    keep = u != v
    u, v = u[keep], v[keep]
    rows = np.repeat(np.arange(n), np.diff(off))
    present = set(zip(rows.tolist(), adj.tolist()))
    g = T.cuStinger(off, adj)
    lib = ops()
    try:
        du, dv = T.DeviceArray.from_host(u), T.DeviceArray.from_host(v)
        found = T.DeviceArray(1, np.uint64).zero()
        _ok(lib.ops_batch_found(g.h, du.ptr, dv.ptr, len(u), found.ptr))
        assert int(found.to_host()[0]) == sum((a, b) in present for a, b in zip(u.tolist(), v.tolist()))
        delta = T.DeviceArray(n, np.uint32).zero()
        T.batch_apply(g, u, v, np.zeros(len(u), np.uint8), delta)
        found.zero()
        _ok(lib.ops_batch_found(g.h, du.ptr, dv.ptr, len(u), found.ptr))
        assert int(found.to_host()[0]) == len(u)  # every pair is an edge now
    finally:
        g.close()


@pytest.mark.parametrize("kind,root", [("rmat", 0), ("rmat", 77), ("star", 5), ("star", 0)])
def test_bfs_levels(kind, root):
    """A top-down BFS built on allVinA_TraverseEdges_LB and the device queue
    gives scipy's unweighted shortest-path lengths."""
    off, adj = _graph(kind)
    n = len(off) - 1
    g = T.cuStinger(off, adj)
    lib = ops()
    try:
        level = T.DeviceArray(n, np.int32)
        depth = C.c_int32(0)
        _ok(lib.ops_bfs(g.h, root, level.ptr, C.byref(depth)))
        got = level.to_host()
    finally:
        g.close()
    A = csr_matrix((np.ones(len(adj)), adj, off), shape=(n, n))
    d = shortest_path(A, unweighted=True, indices=root)
    want = np.where(np.isinf(d), np.iinfo(np.int32).max, d).astype(np.int32)
    assert np.array_equal(got, want)
    assert depth.value == int(d[np.isfinite(d)].max()) + 1
