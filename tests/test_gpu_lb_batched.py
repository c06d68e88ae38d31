"""Load-balance entry points and batched multi-graph counting on the GPU."""
from __future__ import annotations

import numpy as np
import pytest

import custingeralg_amd as T
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,ef,seed", [(10, 2, 1), (12, 5, 2), (14, 16, 3), (7, 1, 4)])
def test_lb_partition_matches_reference_restatement(scale, ef, seed):
    off, adj = O.rmat_graph(scale, ef, seed)
    n = len(off) - 1
    g = T.cuStinger(off, adj)
    deg = np.diff(off)
    rng = np.random.default_rng(seed)
    for verts in (np.arange(n, dtype=np.int32), rng.permutation(n)[: max(3, n // 3)].astype(np.int32)):
        parts = T.lb_partition(g, verts)
        _, want = O.lb_partition(deg[verts])
        assert np.array_equal(parts, want)
    g.close()


def test_lb_degree_bins():
    off, adj = O.rmat_graph(12, 16, 5)
    g = T.cuStinger(off, adj)
    deg = np.diff(off)
    verts = np.arange(len(deg), dtype=np.int32)
    got = T.lb_degree_bins(g, verts)
    assert list(got) == [int((deg >= 64).sum()), int(((deg >= 16) & (deg < 64)).sum()),
                         int((deg < 16).sum())]
    g.close()


def test_batched_generation_and_counts():
    n, ef, seeds = O.batched_spec(G=40, nmin=64, nmax=4096)
    voff_h, off_h, adj_h = O.batched_graph(n, ef, seeds)
    dvoff, doff, dadj = T.batched_graph_device(n, ef, seeds)
    assert np.array_equal(dvoff.to_host(), voff_h)
    assert np.array_equal(doff.to_host(), off_h)
    assert np.array_equal(dadj.to_host(), adj_h)
    pg = T.DeviceArray(len(n), np.uint64)
    st = T.batched_count_device(dvoff, doff, dadj, pg)
    got = pg.to_host()
    want = []
    for i in range(len(n)):
        a, b = voff_h[i], voff_h[i + 1]
        o = off_h[a:b + 1] - off_h[a]
        adj = adj_h[off_h[a]:off_h[b]] - a
        want.append(6 * O.forward_count(o, adj.astype(np.int32))[0])
    assert np.array_equal(got, np.array(want, np.uint64))
    assert st["method_used"] == T.METHOD_ORIENTED
    # per-vertex output of the batched call equals the union graph's static count
    pg2, pv, _ = T.batched_count(voff_h, off_h, adj_h, per_vertex=True)
    assert np.array_equal(pg2, got)
    assert np.array_equal(pv, O.custinger_count(off_h, adj_h))


def test_batched_handle_repeated_counts():
    """tc_batched_graph_create / _count: one validation, several counts (also
    with per-vertex output and another schedule), each equal to the oracle."""
    n, ef, seeds = O.batched_spec(G=24, nmin=256, nmax=8192)
    voff_h, off_h, adj_h = O.batched_graph(n, ef, seeds)
    dvoff, doff, dadj = T.batched_graph_device(n, ef, seeds)
    want_pv = O.custinger_count(off_h, adj_h)
    want = np.array([int(want_pv[voff_h[i]:voff_h[i + 1]].astype(np.uint64).sum())
                     for i in range(len(n))], np.uint64)
    bg = T.BatchedGraph(dvoff, doff, dadj)
    pg = T.DeviceArray(len(n), np.uint64)
    pv = T.DeviceArray(int(voff_h[-1]), np.uint32)
    for sched in (None, T.default_schedule(small_max=32, mid_max=256, job_items=1024),
                  T.default_schedule(policy=T.POLICY_MERGE),
                  T.default_schedule(policy=T.POLICY_MERGEPATH)):
        bg.count(pg, pv, sched)
        assert np.array_equal(pg.to_host(), want)
        assert np.array_equal(pv.to_host(), want_pv)
    bg.close()
