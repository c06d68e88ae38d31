"""The store's API beyond triangle counting, as other cuStingerAlg consumers
use it: directed edgeInsertions / edgeDeletions (tests/katz-tester.cu:211-252),
the update allocator's growth policy (tests/cct.cpp:261 passes
defaultUpdateAllocater), compareCUS (src/streaming_triangle_counting/
cct_upd.cu:442-484) and the int32-offset CSR count (tests/cct_csr.cu:279-285).
Expected stores are built on the host with set operations on the entries."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import custingeralg_amd as T
import oracle as O

pytestmark = pytest.mark.gpu


def directed_csr(n, keys):
    """CSR of the distinct directed entries src * n + dst (rows sorted)."""
    keys = np.unique(np.asarray(keys, np.int64))
    src, dst = keys // n, keys % n
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), dst.astype(np.int32)


def keys_of(off, adj):
    n = len(off) - 1
    return np.repeat(np.arange(n, dtype=np.int64), np.diff(off)) * n + adj


def test_directed_insert_then_delete():
    """katz-tester's flow: a directed batch (uniform entries, repeats, some
    already present, some self-loops) inserted, then the same batch deleted."""
    n = 3000
    rng = np.random.default_rng(5)
    base = rng.integers(0, n, size=(40000, 2))
    base = base[base[:, 0] != base[:, 1]]
    off, adj = directed_csr(n, base[:, 0] * n + base[:, 1])
    g = T.cuStinger(off, adj)
    try:
        assert g.info()["symmetric"] == 0
        bs = rng.integers(0, n, size=(5000, 2)).astype(np.int32)
        bs[::7] = bs[1::7][: len(bs[::7])]                 # in-batch repeats
        k0 = keys_of(off, adj)
        pick = rng.choice(len(k0), 500, replace=False)      # entries already present
        bs[-500:, 0], bs[-500:, 1] = k0[pick] // n, k0[pick] % n
        bs[:20, 1] = bs[:20, 0]                              # self-loops are entries too
        bk = bs[:, 0].astype(np.int64) * n + bs[:, 1]
        st = T.batch_update_directed(g, bs[:, 0], bs[:, 1])
        want = np.union1d(k0, bk)
        o1, a1 = g.to_csr()
        assert np.array_equal(keys_of(o1, a1), want)
        assert st["inserted"] == len(want) - len(k0)
        assert st["dup_inserts"] == len(bs) - st["inserted"]
        assert g.info()["nnz"] == len(want)
        # counting the directed store (with loops) follows the reference formula
        pv, glob, stc = T.static_count(g)
        assert stc["method_used"] == T.METHOD_DIRECT
        assert np.array_equal(pv, O.custinger_count(o1, a1))
        st2 = T.batch_update_directed(g, bs[:, 0], bs[:, 1], np.ones(len(bs), np.uint8))
        left = np.setdiff1d(k0, bk)
        o2, a2 = g.to_csr()
        assert np.array_equal(keys_of(o2, a2), left)
        assert st2["deleted"] == len(np.unique(bk))
        assert st2["missing_deletes"] == len(bs) - st2["deleted"]
    finally:
        g.close()


def test_directed_batches_on_a_symmetric_store():
    """Both directions of every pair keep the store symmetric (the triangle
    count stays on the oriented method and tc_batch_apply still runs); one
    direction makes it asymmetric: the count falls back to the reference
    formula verbatim and undirected updates are refused."""
    off, adj = O.rmat_graph(12, 8, 3)
    n = len(off) - 1
    g = T.cuStinger(off, adj)
    try:
        rng = np.random.default_rng(9)
        u = rng.integers(0, n, 2000).astype(np.int32)
        v = rng.integers(0, n, 2000).astype(np.int32)
        keep = u != v
        u, v = u[keep], v[keep]
        T.batch_update_directed(g, np.concatenate([u, v]), np.concatenate([v, u]))
        assert g.info()["symmetric"] == 1
        o1, a1 = g.to_csr()
        pv, glob, st = T.static_count(g)
        assert st["method_used"] == T.METHOD_ORIENTED
        assert np.array_equal(pv, O.custinger_count(o1, a1))
        # deleting both directions keeps it symmetric as well
        T.batch_update_directed(g, np.concatenate([u[:500], v[:500]]),
                                np.concatenate([v[:500], u[:500]]), np.ones(1000, np.uint8))
        assert g.info()["symmetric"] == 1
        T.batch_apply(g, u[500:600], v[500:600], np.ones(100, np.uint8))  # undirected still fine
        # one direction only
        T.batch_update_directed(g, u[:50], v[:50])
        o2, a2 = g.to_csr()
        pv, glob, st = T.static_count(g)
        assert st["method_used"] == T.METHOD_DIRECT
        assert np.array_equal(pv, O.custinger_count(o2, a2))
        assert g.info()["symmetric"] == 0
        with pytest.raises(T.TCError):
            T.batch_apply(g, u[:10], v[:10])
    finally:
        g.close()


@pytest.mark.parametrize("policy,expect", [((0, 0), lambda nl: 1 << (nl - 1).bit_length()),
                                           ((32, 0), lambda nl: 2 * nl),
                                           ((16, 10), lambda nl: nl + 10)])
def test_growth_policy(policy, expect):
    """A row that outgrows its block moves to a block of the policy's size:
    the pool's high-water mark (getNumberEdgesAllocated) grows by exactly that."""
    n = 64
    edges = [(0, i) for i in range(1, 6)] + [(i, i + 1) for i in range(10, 40)]
    off, adj = O.csr_from_edges(n, edges)
    g = T.cuStinger(off, adj)
    try:
        g.set_growth(*policy)
        a0 = g.info()["allocated"]
        T.batch_update_directed(g, np.zeros(3, np.int32), np.array([50, 51, 52], np.int32))
        info = g.info()
        assert info["reallocations"] == 1
        assert info["allocated"] - a0 == expect(5 + 3)
        o1, a1 = g.to_csr()
        assert a1[o1[0]:o1[1]].tolist() == [1, 2, 3, 4, 5, 50, 51, 52]
    finally:
        g.close()


def test_compare_stores():
    off, adj = O.rmat_graph(11, 8, 4)
    n = len(off) - 1
    a, b = T.cuStinger(off, adj), T.cuStinger(off, adj)
    try:
        assert a.compare(b) == {"size_mismatch": 0, "entry_mismatch": 0, "first_vertex": -1}
        # row 7 gains an entry; row 3 swaps one entry for another (same length)
        r3 = adj[off[3]:off[4]]
        new3 = next(x for x in range(n) if x != 3 and x not in set(r3.tolist()))
        new7 = next(x for x in range(n) if x != 7 and x not in set(adj[off[7]:off[8]].tolist()))
        T.batch_update_directed(b, np.array([3, 7], np.int32), np.array([new3, new7], np.int32))
        T.batch_update_directed(b, np.array([3], np.int32), np.array([r3[0]], np.int32),
                                np.ones(1, np.uint8))
        assert a.compare(b) == {"size_mismatch": 1, "entry_mismatch": 1, "first_vertex": 3}
        assert b.compare(a)["first_vertex"] == 3
    finally:
        a.close()
        b.close()


def test_csr_count_int32_offsets():
    """callDeviceAllTrianglesCSR's int32 offsets, widened on the device."""
    off, adj = O.rmat_graph(13, 12, 6)
    nv = len(off) - 1
    d32 = T.DeviceArray.from_host(off.astype(np.int32))
    dadj = T.DeviceArray.from_host(adj)
    out = T.DeviceArray(nv, np.uint32)
    gl = C.c_uint64(0)
    try:
        T.check(T.lib().tc_csr_count_device32(nv, d32.ptr, dadj.ptr, out.ptr, C.byref(gl), None))
        ref = O.csr_count(off, adj)
        assert np.array_equal(out.to_host(), ref)
        assert gl.value == int(ref.astype(np.uint64).sum())
    finally:
        for x in (d32, dadj, out):
            x.free()


@pytest.mark.parametrize("method", [T.METHOD_AUTO, T.METHOD_DIRECT])
def test_recorded_counts_replay(method):
    """tc_schedule.capture: the second count of an unchanged graph with the
    same schedule records the sequence as a hipGraph, later ones replay it;
    every count equals the oracle's, and an update makes the next count a
    plain one again (the recorded graph belongs to the old store)."""
    off, adj = O.rmat_graph(14, 16, 21)
    ref = O.custinger_count(off, adj)
    g = T.cuStinger(off, adj)
    out = T.DeviceArray(g.nv, np.uint32)
    try:
        sched = T.default_schedule(method=method, capture=1)
        replays = []
        for k in range(5):
            T.check(T.lib().tc_memset(out.ptr, 0xFF, out.nbytes))  # every replay rewrites it
            _, glob, st = T.static_count(g, sched, out=out)
            assert np.array_equal(out.to_host(), ref), k
            assert glob == int(ref.astype(np.uint64).sum())
            replays.append(st["graph_replay"])
        assert replays == [0, 1, 1, 1, 1]
        # another schedule is another recording
        _, glob, st = T.static_count(g, T.default_schedule(method=method, capture=1, job_items=2048),
                                     out=out)
        assert st["graph_replay"] == 0 and glob == int(ref.astype(np.uint64).sum())
        # an update: the next count runs plainly and sees the new store
        u, v = np.array([1, 2], np.int32), np.array([9000, 9001], np.int32)
        T.batch_apply(g, u, v)
        o2, a2 = g.to_csr()
        ref2 = O.custinger_count(o2, a2)
        for k in range(3):
            _, glob, st = T.static_count(g, sched, out=out)
            assert np.array_equal(out.to_host(), ref2)
            assert st["graph_replay"] == (0 if k == 0 else 1)
    finally:
        out.free()
        g.close()


def test_batched_handle_replay():
    """The batched handle counted repeatedly with capture on: per-graph
    counts equal every time, replays from the third count."""
    n, ef, seeds = O.batched_spec(G=64, nmin=256, nmax=4096)
    voff_h, off_h, adj_h = O.batched_graph(n, ef, seeds)
    want_pv = O.custinger_count(off_h, adj_h)
    want = np.array([int(want_pv[voff_h[i]:voff_h[i + 1]].astype(np.uint64).sum())
                     for i in range(len(n))], np.uint64)
    dvoff, doff, dadj = T.batched_graph_device(n, ef, seeds)
    bg = T.BatchedGraph(dvoff, doff, dadj)
    pg = T.DeviceArray(len(n), np.uint64)
    try:
        sched = T.default_schedule(capture=1)
        flags = []
        for _ in range(4):
            st = bg.count(pg, None, sched)
            assert np.array_equal(pg.to_host(), want)
            flags.append(st["graph_replay"])
        assert flags == [0, 1, 1, 1]
    finally:
        bg.close()
        pg.free()
        for a in (dvoff, doff, dadj):
            a.free()
