"""Streaming updates on the GPU: the store after each batch and the per-vertex
delta are checked against the oracle (bit-exact)."""
from __future__ import annotations

import numpy as np
import pytest

import custingeralg_amd as T
import oracle as O

pytestmark = pytest.mark.gpu


def _batch(rng, n, off, adj, n_ins, n_del, rmat_scale=None):
    """Random undirected pairs: inserts (uniform or R-MAT) plus deletions of present edges."""
    if rmat_scale:
        u, v = O.rmat_pairs(rmat_scale, n_ins, int(rng.integers(1 << 30)))
    else:
        u = rng.integers(0, n, n_ins).astype(np.int32)
        v = rng.integers(0, n, n_ins).astype(np.int32)
    keep = u != v
    u, v = u[keep], v[keep]
    deg = np.diff(off)
    rows = np.repeat(np.arange(n, dtype=np.int32), deg)
    if len(adj) and n_del:
        e = rng.integers(0, len(adj), n_del)
        du, dv = rows[e], adj[e]
    else:
        du = dv = np.zeros(0, np.int32)
    bu = np.concatenate([u, du]).astype(np.int32)
    bv = np.concatenate([v, dv]).astype(np.int32)
    dl = np.concatenate([np.zeros(len(u), np.uint8), np.ones(len(du), np.uint8)])
    return bu, bv, dl


def _host_apply(n, off, adj, bu, bv, dl):
    d = dl.astype(bool)
    return O.apply_batch_csr(n, off, adj, ins_u=bu[~d], ins_v=bv[~d], del_u=bu[d], del_v=bv[d])


def _run_batches(off, adj, batches, slack, seed, check_every=1):
    n = len(off) - 1
    rng = np.random.default_rng(seed)
    g = T.cuStinger(off, adj, slack_pct=slack)
    delta = T.DeviceArray(n, np.uint32).zero()
    base = O.custinger_count(off, adj)
    total = int(base.astype(np.uint64).sum())
    cur_off, cur_adj = off, adj
    for bi, (ni, nd, rm) in enumerate(batches):
        bu, bv, dl = _batch(rng, n, cur_off, cur_adj, ni, nd, rm)
        dg, st = T.batch_apply(g, bu, bv, dl, delta)
        total += dg
        want_off, want_adj = _host_apply(n, cur_off, cur_adj, bu, bv, dl)
        got_off, got_adj = g.to_csr()
        assert np.array_equal(got_off, want_off), f"batch {bi}: offsets"
        assert np.array_equal(got_adj, want_adj), f"batch {bi}: adjacency"
        cur_off, cur_adj = want_off, want_adj
        if bi % check_every == 0 or bi == len(batches) - 1:
            now = O.custinger_count(cur_off, cur_adj)
            assert np.array_equal(base + delta.to_host(), now), f"batch {bi}: per-vertex"
            assert total == int(now.astype(np.uint64).sum()), f"batch {bi}: global"
            pv, glob, _ = T.static_count(g)
            assert np.array_equal(pv, now) and glob == total
    g.close()
    return st


@pytest.mark.parametrize("slack", [0, 100])
@pytest.mark.parametrize("seed", [1, 2])
def test_small_mixed_batches(slack, seed):
    off, adj = O.rmat_graph(8, 6, seed)
    _run_batches(off, adj, [(200, 0, None), (0, 150, None), (300, 100, None), (300, 100, 8),
                            (50, 300, None), (1000, 20, 8)], slack, seed)


def test_insert_into_empty_and_delete_all():
    n = 60
    off = np.zeros(n + 1, np.int64)
    adj = np.zeros(0, np.int32)
    g = T.cuStinger(off, adj)
    delta = T.DeviceArray(n, np.uint32).zero()
    pairs = [(a, b) for a in range(12) for b in range(a + 1, 12)]  # K_12
    u = np.array([p[0] for p in pairs], np.int32)
    v = np.array([p[1] for p in pairs], np.int32)
    dg, st = T.batch_apply(g, u, v, None, delta)
    assert st["inserted"] == len(pairs) and dg == 6 * 220
    assert np.array_equal(delta.to_host()[:12], np.full(12, 2 * 55, np.uint32))
    dg2, st2 = T.batch_apply(g, u, v, np.ones(len(u), np.uint8), delta)
    assert st2["deleted"] == len(pairs) and dg2 == -6 * 220
    assert not delta.to_host().any()
    assert g.info()["nnz"] == 0
    g.close()


def test_duplicates_and_noops():
    off, adj = O.rmat_graph(7, 4, 3)
    n = len(off) - 1
    g = T.cuStinger(off, adj)
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    present_u, present_v = rows[:5], adj[:5]
    u = np.array([1, 1, 2] + list(present_u) + [3], np.int32)
    v = np.array([9, 9, 9] + list(present_v) + [4], np.int32)
    dg, st = T.batch_apply(g, u, v)
    want_off, want_adj = O.apply_batch_csr(n, off, adj, ins_u=u, ins_v=v)
    got = g.to_csr()
    assert np.array_equal(got[0], want_off) and np.array_equal(got[1], want_adj)
    # every request that changed nothing is a no-op, in-batch repeats included
    added = (len(want_adj) - len(adj)) // 2
    assert st["inserted"] == added
    assert st["dup_inserts"] == len(u) - added
    assert st["deleted"] == 0 and st["missing_deletes"] == 0
    # deletions: an absent pair twice, a present pair three times
    cur_off, cur_adj = got
    assert not np.isin(n - 1, cur_adj[cur_off[0]:cur_off[1]])
    pu, pv = int(present_u[0]), int(present_v[0])
    du = np.array([0, 0, pu, pv, pu], np.int32)
    dv = np.array([n - 1, n - 1, pv, pu, pv], np.int32)
    dg, st = T.batch_apply(g, du, dv, np.ones(5, np.uint8))
    assert st["deleted"] == 1 and st["missing_deletes"] == 4
    assert st["inserted"] == 0 and st["dup_inserts"] == 0
    g.close()


def test_invalid_batch_leaves_store_untouched():
    off, adj = O.rmat_graph(7, 4, 3)
    g = T.cuStinger(off, adj)
    with pytest.raises(T.TCError):
        T.batch_apply(g, np.array([1, 2], np.int32), np.array([5, 2], np.int32))  # self-loop
    with pytest.raises(T.TCError):
        T.batch_apply(g, np.array([1], np.int32), np.array([1 << 20], np.int32))  # out of range
    # a batch above the one-block size (1,024 pairs) takes the multi-block
    # front kernels: one bad id among 3,000 valid pairs, then one self-loop
    rng = np.random.default_rng(4)
    n = len(off) - 1
    bu = rng.integers(0, n, 3000).astype(np.int32)
    bv = ((bu + 1 + rng.integers(0, n - 1, 3000)) % n).astype(np.int32)
    for k, bad in ((1500, n), (2999, None)):
        xu, xv = bu.copy(), bv.copy()
        if bad is None:
            xv[k] = xu[k]
        else:
            xv[k] = bad
        with pytest.raises(T.TCError):
            T.batch_apply(g, xu, xv)
    got = g.to_csr()
    assert np.array_equal(got[0], off) and np.array_equal(got[1], adj)
    g.close()


def test_reference_passes_match_oracle():
    """tc_stream_delta_passes runs the reference's three kernels; each pass is
    compared with the oracle's restatement on a batch with planted repeats and
    already-present edges (tests/cct_upd.cpp:313-332)."""
    off, adj = O.rmat_graph(9, 8, 5)
    n = len(off) - 1
    rng = np.random.default_rng(3)
    u = rng.integers(0, n, 400).astype(np.int32)
    v = rng.integers(0, n, 400).astype(np.int32)
    k = u != v
    u, v = u[k], v[k]
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    u = np.concatenate([u, u[:30], rows[:20]])
    v = np.concatenate([v, v[:30], adj[:20]])
    src = np.concatenate([u, v]).astype(np.int32)
    dst = np.concatenate([v, u]).astype(np.int32)
    s_src, s_dst, dup, boff = O.batch_prepare(n, src, dst, off, adj)
    goff, gadj = O.apply_batch_csr(n, off, adj, ins_u=u, ins_v=v)
    g = T.cuStinger(goff, gadj)
    d = [T.DeviceArray.from_host(x) for x in (s_src, s_dst, dup, boff)]
    for mask in (1, 2, 4, 7):
        out = T.DeviceArray(n, np.uint32).zero()
        T.check(T.lib().tc_stream_delta_passes(g.h, len(s_src), d[0].ptr, d[1].ptr, d[2].ptr,
                                               d[3].ptr, out.ptr, mask))
        want = O.stream_delta(goff, gadj, s_src, s_dst, dup, boff, pass_mask=mask)
        assert np.array_equal(out.to_host(), want), mask
    # all three passes together give the recount difference
    assert np.array_equal(O.custinger_count(off, adj) + want, O.custinger_count(goff, gadj))
    g.close()


def test_streaming_class_surface():
    off, adj = O.rmat_graph(8, 8, 4)
    n = len(off) - 1
    g = T.cuStinger(off, adj, slack_pct=50)
    alg = T.StreamingTriangleCounting()
    alg.Init(g)
    alg.runStatic(g)
    base = alg.copyPerVertexToHost()
    assert np.array_equal(base, O.custinger_count(off, adj))
    rng = np.random.default_rng(9)
    u = rng.integers(0, n, 300).astype(np.int32)
    v = rng.integers(0, n, 300).astype(np.int32)
    k = u != v
    u, v = u[k], v[k]
    alg.batchUpdateInserted(g, u, v)
    off2, adj2 = O.apply_batch_csr(n, off, adj, ins_u=u, ins_v=v)
    assert np.array_equal(alg.copyPerVertexToHost(), O.custinger_count(off2, adj2))
    alg.batchUpdateDeleted(g, u[:100], v[:100])
    off3, adj3 = O.apply_batch_csr(n, off2, adj2, del_u=u[:100], del_v=v[:100])
    assert np.array_equal(alg.copyPerVertexToHost(), O.custinger_count(off3, adj3))
    assert alg.getGlobal() == int(O.custinger_count(off3, adj3).astype(np.uint64).sum())
    alg.Release()
    g.close()


def test_rmat16_twenty_batches_recount():
    """Scaled-down streaming config: R-MAT base, 20 mixed batches (80% inserts,
    half uniform / half R-MAT, 20% deletions of present edges); after every
    batch the running total equals a GPU static recount, and at the end the
    per-vertex vector equals the oracle."""
    off, adj = O.rmat_graph(14, 16, 3)
    sizes = [100, 1000, 4000, 10000] * 5
    batches = []
    for i, b in enumerate(sizes):
        ni = int(0.8 * b)
        batches.append((ni, b - ni, 14 if i % 2 else None))
    _run_batches(off, adj, batches, 0, 11, check_every=5)


@pytest.mark.parametrize("seed", [0, 1])
def test_batch_prepare_matches_oracle(seed):
    """tc_batch_prepare (sortDeviceBUD + duplicate flags) against the oracle on a
    batch holding in-batch repeats and entries already in the store."""
    off, adj = O.rmat_graph(9, 8, 6 + seed)
    n = len(off) - 1
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, 500).astype(np.int32)
    v = rng.integers(0, n, 500).astype(np.int32)
    k = u != v
    u, v = u[k], v[k]
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    u = np.concatenate([u, u[:50], rows[::37]])
    v = np.concatenate([v, v[:50], adj[::37]])
    src = np.concatenate([u, v]).astype(np.int32)
    dst = np.concatenate([v, u]).astype(np.int32)
    g = T.cuStinger(off, adj)
    got = T.batch_prepare(g, src, dst)
    want = O.batch_prepare(n, src, dst, off, adj)
    for a, b, name in zip(got, want, ("s_src", "s_dst", "dup", "boff")):
        assert np.array_equal(a, b), name
    assert got[2].sum() >= 50  # the planted repeats are flagged
    # empty batch: all offsets zero
    e = T.batch_prepare(g, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert len(e[0]) == 0 and not e[3].any()
    g.close()


def _hub_graph(n, hubs, hub_deg, seed):
    """Background R-MAT pairs plus a few hub rows of hub_deg neighbours each."""
    rng = np.random.default_rng(seed)
    u, v = O.rmat_pairs(13, 20000, seed)
    hu = np.repeat(np.arange(hubs, dtype=np.int32), hub_deg)
    hv = np.concatenate([rng.choice(np.arange(hubs, n, dtype=np.int32), hub_deg, replace=False)
                         for _ in range(hubs)])
    a = np.concatenate([u % n, hu]).astype(np.int32)
    b = np.concatenate([v % n, hv]).astype(np.int32)
    return O.apply_batch_csr(n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32),
                             ins_u=a, ins_v=b, del_u=np.zeros(0, np.int32),
                             del_v=np.zeros(0, np.int32))


@pytest.mark.parametrize("n,hub_deg", [(16384, 6000), (131072, 40000)])
@pytest.mark.parametrize("slack", [0, 100])
def test_hub_rows_merge_and_compact(slack, n, hub_deg):
    """Rows of thousands of entries (the block-per-row merge and compaction;
    40,000 entries take three passes of a block): inserts spread over a hub
    row, inserts that all sort after its entries, inserts that all sort before
    them, and deletions from it; with headroom (slack 100) the rows merge in
    place, without it they move.  Store and counts against the oracle after
    every batch."""
    off, adj = _hub_graph(n, 3, hub_deg, 5)
    assert np.diff(off).max() >= hub_deg
    rng = np.random.default_rng(9)
    g = T.cuStinger(off, adj, slack_pct=slack)
    delta = T.DeviceArray(n, np.uint32).zero()
    base = O.custinger_count(off, adj)
    total = int(base.astype(np.uint64).sum())
    cur_off, cur_adj = off, adj
    plans = ["spread", "after", "before", "delete", "mixed"]
    for bi, plan in enumerate(plans):
        hub = bi % 3
        row = cur_adj[cur_off[hub]:cur_off[hub + 1]]
        if plan == "spread":
            v = rng.integers(3, n, 3000).astype(np.int32)
            u = np.full(len(v), hub, np.int32)
            dl = np.zeros(len(v), np.uint8)
        elif plan == "after":
            hi = int(row.max()) + 1
            v = np.arange(hi, min(n, hi + 500), dtype=np.int32)
            u = np.full(len(v), hub, np.int32)
            dl = np.zeros(len(v), np.uint8)
        elif plan == "before":
            lo = int(row.min())
            v = np.arange(3, lo, dtype=np.int32)[:400]
            u = np.full(len(v), hub, np.int32)
            dl = np.zeros(len(v), np.uint8)
        elif plan == "delete":
            v = rng.choice(row, min(len(row), 2500), replace=False).astype(np.int32)
            u = np.full(len(v), hub, np.int32)
            dl = np.ones(len(v), np.uint8)
        else:
            vi = rng.integers(3, n, 1500).astype(np.int32)
            vd = rng.choice(row, 1500, replace=False).astype(np.int32)
            v = np.concatenate([vi, vd])
            u = np.full(len(v), hub, np.int32)
            dl = np.concatenate([np.zeros(1500, np.uint8), np.ones(1500, np.uint8)])
        keep = u != v
        u, v, dl = u[keep], v[keep], dl[keep]
        dg, _ = T.batch_apply(g, u, v, dl, delta)
        total += dg
        want_off, want_adj = _host_apply(n, cur_off, cur_adj, u, v, dl)
        got_off, got_adj = g.to_csr()
        assert np.array_equal(got_off, want_off), f"batch {bi} ({plan}): offsets"
        assert np.array_equal(got_adj, want_adj), f"batch {bi} ({plan}): adjacency"
        cur_off, cur_adj = want_off, want_adj
        now = O.custinger_count(cur_off, cur_adj)
        assert np.array_equal(base + delta.to_host(), now), f"batch {bi} ({plan}): per-vertex"
        assert total == int(now.astype(np.uint64).sum()), f"batch {bi} ({plan}): global"
    g.close()


def test_merge_path_delta_against_oracle():
    """A batch above the one-block size whose hub-hub pairs (both rows 3,000+
    entries, two merge-path segments each) go through the merge-path delta
    and whose other pairs through the per-entry searches: the per-vertex
    delta and the store against the oracle, for the insertions and for the
    deletions of the same pairs."""
    n, hubs = 20000, 16
    rng = np.random.default_rng(41)
    bu = rng.integers(0, n, 60000)
    bv = rng.integers(0, n, 60000)
    hu = np.repeat(np.arange(hubs), 3000)
    hv = np.concatenate([rng.choice(np.arange(hubs, n), 3000, replace=False) for _ in range(hubs)])
    pairs = [(a, b) for a in range(hubs) for b in range(a + 1, hubs)]
    present = [p for i, p in enumerate(pairs) if i % 2 == 0]
    absent = [p for i, p in enumerate(pairs) if i % 2 == 1]
    a = np.concatenate([bu, hu, [p[0] for p in present]]).astype(np.int32)
    b = np.concatenate([bv, hv, [p[1] for p in present]]).astype(np.int32)
    k = a != b
    off, adj = O.apply_batch_csr(n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32),
                                 ins_u=a[k], ins_v=b[k])
    assert np.diff(off)[:hubs].min() >= 3000
    g = T.cuStinger(off, adj)
    delta = T.DeviceArray(n, np.uint32).zero()
    try:
        base = O.custinger_count(off, adj)
        # insert the absent hub pairs and 1,500 background pairs
        xu = rng.integers(hubs, n, 1500).astype(np.int32)
        xv = rng.integers(hubs, n, 1500).astype(np.int32)
        ku = xu != xv
        u = np.concatenate([[p[0] for p in absent], xu[ku]]).astype(np.int32)
        v = np.concatenate([[p[1] for p in absent], xv[ku]]).astype(np.int32)
        dl = np.zeros(len(u), np.uint8)
        assert len(u) > 1024
        dg, _ = T.batch_apply(g, u, v, dl, delta)
        cur_off, cur_adj = _host_apply(n, off, adj, u, v, dl)
        got = g.to_csr()
        assert np.array_equal(got[0], cur_off) and np.array_equal(got[1], cur_adj)
        now = O.custinger_count(cur_off, cur_adj)
        assert np.array_equal(base + delta.to_host(), now)
        assert dg == int(now.astype(np.int64).sum()) - int(base.astype(np.int64).sum())
        # delete the same pairs again: back to the start
        dg2, _ = T.batch_apply(g, u, v, np.ones(len(u), np.uint8), delta)
        got = g.to_csr()
        assert np.array_equal(got[0], off) and np.array_equal(got[1], adj)
        assert not delta.to_host().any() and dg2 == -dg
    finally:
        delta.free()
        g.close()


def test_large_batch_matches_two_halves():
    """A batch of over 2^18 pairs, a third of them at six hub rows of 9,000+
    entries, against the same batch applied to a twin store in two halves:
    the same store and the same per-vertex delta, the running total equal to
    a recount, and the store equal to the oracle's."""
    off, adj = _hub_graph(65536, 6, 9000, 31)
    n = len(off) - 1
    assert np.diff(off).max() >= 4096
    rng = np.random.default_rng(32)
    deg = np.diff(off)
    rows = np.repeat(np.arange(n, dtype=np.int32), deg)
    # insertions: half uniform, half at the hubs; deletions: present edges
    ni = 280000
    iu = np.concatenate([rng.integers(0, n, ni // 2), rng.integers(0, 6, ni // 2)]).astype(np.int32)
    iv = rng.integers(0, n, ni).astype(np.int32)
    e = rng.choice(len(adj), 60000, replace=False)
    keys_i = np.unique((np.minimum(iu, iv).astype(np.int64) << 32) | np.maximum(iu, iv))
    keys_i = keys_i[(keys_i >> 32) != (keys_i & 0xFFFFFFFF)]
    keys_d = np.unique((np.minimum(rows[e], adj[e]).astype(np.int64) << 32) |
                       np.maximum(rows[e], adj[e]))
    keys_i = keys_i[~np.isin(keys_i, keys_d)]
    keys = np.concatenate([keys_i, keys_d])
    u = (keys >> 32).astype(np.int32)
    v = (keys & 0xFFFFFFFF).astype(np.int32)
    dl = np.concatenate([np.zeros(len(keys_i), np.uint8), np.ones(len(keys_d), np.uint8)])
    assert len(u) >= 1 << 18
    g = T.cuStinger(off, adj)
    twin = T.cuStinger(off, adj)
    d1 = T.DeviceArray(n, np.uint32).zero()
    d2 = T.DeviceArray(n, np.uint32).zero()
    try:
        _, total0, _ = T.static_count(g, per_vertex=False)
        dg, _ = T.batch_apply(g, u, v, dl, d1)
        perm = rng.permutation(len(u))
        h = len(u) // 2
        dga, _ = T.batch_apply(twin, u[perm[:h]], v[perm[:h]], dl[perm[:h]], d2)
        dgb, _ = T.batch_apply(twin, u[perm[h:]], v[perm[h:]], dl[perm[h:]], d2)
        assert dg == dga + dgb
        assert np.array_equal(d1.to_host(), d2.to_host())
        assert g.compare(twin) == {"size_mismatch": 0, "entry_mismatch": 0, "first_vertex": -1}
        _, total1, _ = T.static_count(g, per_vertex=False)
        assert total1 == total0 + dg
        want_off, want_adj = _host_apply(n, off, adj, u, v, dl)
        got_off, got_adj = g.to_csr()
        assert np.array_equal(got_off, want_off) and np.array_equal(got_adj, want_adj)
    finally:
        d1.free()
        d2.free()
        g.close()
        twin.close()


@pytest.mark.parametrize("slack", [0, 100])
def test_rows_grow_past_block_threshold(slack):
    """Rows just below the length from which a whole block merges them (1024)
    receive enough new entries to end above it, and a row above it loses
    entries to end below it: every row is merged or compacted exactly once."""
    n = 8192
    rng = np.random.default_rng(12)
    hu = np.repeat(np.arange(4, dtype=np.int32), 900)
    hv = np.concatenate([rng.choice(np.arange(4, n, dtype=np.int32), 900, replace=False)
                         for _ in range(4)])
    bu, bv = rng.integers(4, n, size=(2, 30000)).astype(np.int32)  # background, no hubs
    off, adj = O.apply_batch_csr(n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32),
                                 ins_u=np.concatenate([hu, bu]), ins_v=np.concatenate([hv, bv]),
                                 del_u=np.zeros(0, np.int32), del_v=np.zeros(0, np.int32))
    assert np.diff(off)[:4].tolist() == [900] * 4
    g = T.cuStinger(off, adj, slack_pct=slack)
    delta = T.DeviceArray(n, np.uint32).zero()
    base = O.custinger_count(off, adj)
    total = int(base.astype(np.uint64).sum())
    cur_off, cur_adj = off, adj
    for bi in range(3):
        us, vs, ds = [], [], []
        for hub in range(4):
            row = cur_adj[cur_off[hub]:cur_off[hub + 1]]
            if bi == 2:  # shrink the grown rows back below the threshold
                v = rng.choice(row, 300, replace=False).astype(np.int32)
                d = np.ones(len(v), np.uint8)
            else:
                cand = np.setdiff1d(np.arange(4, n, dtype=np.int32), row)
                v = rng.choice(cand, 200 if bi == 0 else 100, replace=False).astype(np.int32)
                d = np.zeros(len(v), np.uint8)
            us.append(np.full(len(v), hub, np.int32))
            vs.append(v)
            ds.append(d)
        u, v, dl = np.concatenate(us), np.concatenate(vs), np.concatenate(ds)
        dg, _ = T.batch_apply(g, u, v, dl, delta)
        total += dg
        want_off, want_adj = _host_apply(n, cur_off, cur_adj, u, v, dl)
        got_off, got_adj = g.to_csr()
        assert np.array_equal(got_off, want_off), f"batch {bi}: offsets"
        assert np.array_equal(got_adj, want_adj), f"batch {bi}: adjacency"
        cur_off, cur_adj = want_off, want_adj
        now = O.custinger_count(cur_off, cur_adj)
        assert np.array_equal(base + delta.to_host(), now), f"batch {bi}: per-vertex"
        assert total == int(now.astype(np.uint64).sum()), f"batch {bi}: global"
    assert np.diff(cur_off)[:4].tolist() == [900] * 4
    g.close()


@pytest.mark.parametrize("slack", [0, 50])
def test_recorded_batches(slack):
    """tc_schedule.capture: batches padded to their size class; the second
    batch of a class records the sequence, later ones replay it.  After every
    batch the store equals the oracle's and a plainly updated twin store
    (tc_graph_compare), and the per-vertex delta equals the oracle's count
    difference; the pool also outgrows its headroom (slack 0), which sends a
    recorded batch back to the host to grow it."""
    off, adj = O.rmat_graph(13, 8, 17)
    n = len(off) - 1
    rng = np.random.default_rng(23)
    g = T.cuStinger(off, adj, slack_pct=slack)
    twin = T.cuStinger(off, adj, slack_pct=slack)
    delta = T.DeviceArray(n, np.uint32).zero()
    dtwin = T.DeviceArray(n, np.uint32).zero()
    base = O.custinger_count(off, adj)
    sched = T.default_schedule(capture=1)
    cur_off, cur_adj = off, adj
    replayed = []
    try:
        for bi in range(12):
            # 900..1000 pairs: one size class (1024); hub-heavy R-MAT inserts
            bu, bv, dl = _batch(rng, n, cur_off, cur_adj, 800 + int(rng.integers(0, 50)),
                                100 + int(rng.integers(0, 50)), rmat_scale=13 if bi % 2 else None)
            if bi in (4, 7, 8):
                # an invalid batch of the same class (one id out of range, or
                # one self-loop) replays the recorded sequence and changes
                # nothing: the checks below see only the valid batches
                xu, xv = bu.copy(), bv.copy()
                k = len(xu) // 2
                if bi == 4:
                    xv[k] = n + 5
                elif bi == 7:
                    xu[k] = -1
                else:
                    xv[k] = xu[k]
                with pytest.raises(T.TCError):
                    T.batch_apply(g, xu, xv, dl, delta, schedule=sched)
            dg, st = T.batch_apply(g, bu, bv, dl, delta, schedule=sched)
            dg2, st2 = T.batch_apply(twin, bu, bv, dl, dtwin)
            replayed.append(st["replayed"])
            assert dg == dg2 and st["inserted"] == st2["inserted"] and st["deleted"] == st2["deleted"]
            assert st["dup_inserts"] == st2["dup_inserts"]
            assert st["missing_deletes"] == st2["missing_deletes"]
            cur_off, cur_adj = _host_apply(n, cur_off, cur_adj, bu, bv, dl)
            got_off, got_adj = g.to_csr()
            assert np.array_equal(got_off, cur_off) and np.array_equal(got_adj, cur_adj), bi
            assert g.compare(twin) == {"size_mismatch": 0, "entry_mismatch": 0, "first_vertex": -1}
            now = O.custinger_count(cur_off, cur_adj)
            assert np.array_equal(base + delta.to_host(), now), f"batch {bi}: per-vertex"
            assert np.array_equal(delta.to_host(), dtwin.to_host())
        assert replayed[0] == 0 and sum(replayed) >= 6, replayed
    finally:
        for x in (delta, dtwin):
            x.free()
        g.close()
        twin.close()
