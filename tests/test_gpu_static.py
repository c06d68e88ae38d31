"""Static triangle counting on the GPU against the oracle (bit-exact)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import custingeralg_amd as T
import oracle as O
from tests import graphs as GR

pytestmark = pytest.mark.gpu

POLICIES = [T.POLICY_FIXED, T.POLICY_MERGE, T.POLICY_BSEARCH, T.POLICY_HASH, T.POLICY_MERGEPATH]


def count(off, adj, **sched):
    g = T.cuStinger(off, adj)
    try:
        pv, glob, st = T.static_count(g, T.default_schedule(**sched))
    finally:
        g.close()
    return pv, glob, st


@pytest.mark.parametrize("name", sorted(GR.closed_form_cases()))
@pytest.mark.parametrize("policy", POLICIES)
def test_closed_forms(name, policy):
    off, adj, t = GR.closed_form_cases()[name]
    pv, glob, st = count(off, adj, policy=policy)
    assert np.array_equal(pv, (2 * t).astype(np.uint32)), name
    assert glob == 2 * int(t.sum())
    assert st["method_used"] == T.METHOD_ORIENTED


@pytest.mark.parametrize("name", sorted(GR.closed_form_cases()))
def test_closed_forms_task_lists_from_rows(name):
    """Every owner's task list built from its own row (hub degree 1)."""
    off, adj, t = GR.closed_form_cases()[name]
    pv, glob, st = count(off, adj, hub_degree=1)
    assert np.array_equal(pv, (2 * t).astype(np.uint32)), name
    assert glob == 2 * int(t.sum())
    assert st["method_used"] == T.METHOD_ORIENTED


@pytest.mark.parametrize("name", sorted(GR.closed_form_cases()))
def test_closed_forms_fallback_build(name):
    """The oriented layout's fallback paths (per-entry appends, mask by
    degree) give the same counts as the default build."""
    off, adj, t = GR.closed_form_cases()[name]
    pv, glob, st = count(off, adj, build=1)
    assert np.array_equal(pv, (2 * t).astype(np.uint32)), name
    assert glob == 2 * int(t.sum())
    assert st["method_used"] == T.METHOD_ORIENTED


@pytest.mark.parametrize("name", sorted(GR.fixture_cases()))
@pytest.mark.parametrize("method", [T.METHOD_AUTO, T.METHOD_DIRECT])
def test_fixtures_networkx(name, method):
    off, adj, t = GR.fixture_cases()[name]
    pv, glob, _ = count(off, adj, method=method)
    assert np.array_equal(pv, (2 * t).astype(np.uint32))
    assert glob == 2 * int(t.sum())
    assert np.array_equal(pv, O.custinger_count(off, adj))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("loops", [False, True])
def test_asymmetric_and_loops_direct(seed, loops):
    """Inputs the oriented method cannot take fall back to the reference formula."""
    off, adj = GR.random_lists(90, 0.12, seed, loops=loops)
    ref = GR.brute_custinger(off, adj)
    assert np.array_equal(ref, O.custinger_count(off, adj))
    for policy in POLICIES:
        pv, glob, st = count(off, adj, policy=policy)
        assert st["method_used"] == T.METHOD_DIRECT
        assert np.array_equal(pv, ref)
        assert glob == int(ref.astype(np.uint64).sum())


def test_csr_variant_3t():
    off, adj = O.rmat_graph(12, 8, 5)
    pv, glob = T.csr_count_device(off, adj)
    ref = O.csr_count(off, adj)
    assert np.array_equal(pv, ref)
    Tt, _ = O.forward_count(off, adj)
    assert glob == 3 * Tt


def test_er16_plain_config():
    """n=2^16, d=16, seed 1: the committed per-vertex golden vector."""
    z = np.load(os.path.join(GR.GOLDEN, "fixtures.npz"))
    off, adj = O.er_graph(1 << 16, 16, 1)
    for method in (T.METHOD_AUTO, T.METHOD_DIRECT):
        pv, glob, _ = count(off, adj, method=method)
        assert np.array_equal(pv, z["er16_pv"])
        assert glob == 6 * json.load(open(os.path.join(GR.GOLDEN, "fixtures.json")))["er16"]["T"]


@pytest.mark.parametrize("scale", [12, 16])
@pytest.mark.parametrize("method", [T.METHOD_AUTO, T.METHOD_DIRECT])
def test_rmat_against_reference_formula(scale, method):
    # DIRECT streams the mid bin's long neighbour lists in wave-wide chunks
    # whose lanes past a list's end must read -1, not 0: vertex 0 is the
    # R-MAT hub and sits in most lists
    off, adj = O.rmat_graph(scale, 16, 2)
    ref = O.custinger_count(off, adj)
    for policy in POLICIES:
        pv, glob, st = count(off, adj, policy=policy, method=method)
        assert np.array_equal(pv, ref)
        assert glob == int(ref.astype(np.uint64).sum())
        if method == T.METHOD_DIRECT:
            assert st["method_used"] == T.METHOD_DIRECT


def test_rmat18_schedules_identical():
    off, adj = O.rmat_graph(18, 16, 2)
    Tt, tv = O.forward_count(off, adj, per_vertex=True)
    want = (2 * tv % (1 << 32)).astype(np.uint32)
    variants = [dict(), dict(small_max=16, mid_max=128, job_items=1024),
                dict(small_max=64, mid_max=512, job_items=65536), dict(policy=T.POLICY_HASH),
                dict(policy=T.POLICY_BSEARCH), dict(policy=T.POLICY_MERGE),
                dict(policy=T.POLICY_MERGEPATH),
                dict(mpath_min=4), dict(mpath_min=48),  # balanced pairs merge-path
                dict(ratio_x16=64)]  # the fixed heuristic with its merges on
    for v in variants:
        pv, glob, _ = count(off, adj, **v)
        assert np.array_equal(pv, want), v
        assert glob == 6 * Tt


@pytest.mark.parametrize("mpath_min", [1, 16, 200])
@pytest.mark.parametrize("method", [T.METHOD_AUTO, T.METHOD_DIRECT])
def test_mixed_merge_path_tasks(mpath_min, method):
    """The fixed heuristic sends balanced pairs to the wavefront merge path
    while the other tasks of the same jobs probe the LDS table: third-vertex
    hits of merge-path tasks then take their slot from the table."""
    off, adj = O.rmat_graph(15, 24, 9)
    ref = O.custinger_count(off, adj)
    pv, glob, st = count(off, adj, method=method, mpath_min=mpath_min)
    assert np.array_equal(pv, ref)
    assert glob == int(ref.astype(np.uint64).sum())
    assert st["strategy_tasks"]["mergepath"] > 0
    assert sum(st["strategy_tasks"][k] for k in ("hash", "bsearch")) > 0


@pytest.mark.parametrize("n", [63, 64, 65, 127, 128, 129])
def test_merge_path_window_boundaries(n):
    """K_n and K_n minus a perfect matching put list lengths on and around the
    64-entry windows of the merge path."""
    off, adj, t = GR.complete_fast(n)
    for method in (T.METHOD_AUTO, T.METHOD_DIRECT):
        pv, glob, st = count(off, adj, method=method, policy=T.POLICY_MERGEPATH)
        assert np.all(pv == (2 * t).astype(np.uint32)), method
        assert st["strategy_tasks"]["mergepath"] > 0
    keep = np.ones(len(adj), bool)
    rows = np.repeat(np.arange(n), np.diff(off))
    keep[(adj == (rows ^ 1))] = False  # drop the edges {2i, 2i+1}
    off2 = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int64)
    adj2 = adj[keep]
    ref = O.custinger_count(off2, adj2)
    for method in (T.METHOD_AUTO, T.METHOD_DIRECT):
        pv, glob, _ = count(off2, adj2, method=method, policy=T.POLICY_MERGEPATH)
        assert np.array_equal(pv, ref), method


@pytest.mark.parametrize("n", [130, 300, 600])
@pytest.mark.parametrize("method", [T.METHOD_AUTO, T.METHOD_DIRECT])
def test_complete_graphs_chunk_boundaries(n, method):
    """K_n: the oriented task suffixes take every length up to n - 2, so the
    wave-wide chunk streams end on and off 64-entry boundaries in every bin;
    the direct rows all hold vertex 0."""
    off, adj, t = GR.complete_fast(n)
    pv, glob, st = count(off, adj, method=method)
    assert np.all(pv == (2 * t).astype(np.uint32))
    assert glob == 2 * int(t.sum())
    if method == T.METHOD_DIRECT:
        assert st["method_used"] == T.METHOD_DIRECT


def test_tiled_owner_lists():
    # oriented: vertex 0 of K_2100 has |N+| = 2099 > the LDS tile
    off, adj, t = GR.complete_fast(2100)
    pv, glob, _ = count(off, adj)
    assert np.all(pv == (2 * t).astype(np.uint32))
    assert glob == 2 * int(t.sum())
    # direct: the wheel hub has degree 3000
    off, adj, t = GR.wheel(3000)
    pv, glob, _ = count(off, adj, method=T.METHOD_DIRECT)
    assert np.array_equal(pv, (2 * t).astype(np.uint32))


def test_edge_cases():
    for off, adj in [(np.zeros(1, np.int64), np.zeros(0, np.int32)),
                     (np.zeros(6, np.int64), np.zeros(0, np.int32)),
                     O.csr_from_edges(2, [(0, 1)]),
                     O.csr_from_edges(10, [(0, 1), (1, 2), (0, 2)])]:
        pv, glob, _ = count(off, adj)
        assert np.array_equal(pv, O.custinger_count(off, adj))
        assert glob == int(O.custinger_count(off, adj).astype(np.uint64).sum())


def test_device_generators_match_oracle():
    for scale, ef, seed in [(10, 8, 3), (16, 16, 2)]:
        g = T.rmat_graph_device(scale, ef, seed)
        off, adj = g.to_csr()
        g.close()
        roff, radj = O.rmat_graph(scale, ef, seed)
        assert np.array_equal(off, roff) and np.array_equal(adj, radj)
    g = T.er_graph_device(1 << 16, 16, 1)
    off, adj = g.to_csr()
    g.close()
    roff, radj = O.er_graph(1 << 16, 16, 1)
    assert np.array_equal(off, roff) and np.array_equal(adj, radj)


def test_class_surface():
    off, adj, t = GR.wheel(64)
    g = T.cuStinger(off, adj)
    alg = T.StaticTriangleCounting()
    alg.Init(g)
    alg.Run(g)
    alg.SyncHostWithDevice()
    assert np.array_equal(alg.host_out, (2 * t).astype(np.uint32))
    assert alg.getGlobal() == 2 * int(t.sum())
    alg.Reset()
    assert not alg.copyPerVertexToHost().any()
    alg.Release()
    g.freecuStinger()


def test_validation_rejects_malformed():
    off = np.array([0, 2, 3], np.int64)
    adj = np.array([1, 1, 0], np.int32)  # duplicate in row 0
    with pytest.raises(T.TCError):
        T.cuStinger(off, adj)
    adj = np.array([1, 7, 0], np.int32)  # out of range
    with pytest.raises(T.TCError):
        T.cuStinger(off, adj)


def _hub_graph(n, hubs, hub_deg, seed):
    """Sparse R-MAT background plus a few hubs of hub_deg neighbours each (hubs
    adjacent to each other), so some owners pass the hub degree of the
    oriented task construction."""
    rng = np.random.default_rng(seed)
    u, v = O.rmat_pairs(14, 60000, seed)
    hu = np.repeat(np.arange(hubs, dtype=np.int32), hub_deg)
    hv = np.concatenate([rng.choice(np.arange(n, dtype=np.int32), hub_deg, replace=False)
                         for _ in range(hubs)])
    return O.canon_csr(n, np.concatenate([u % n, hu]).astype(np.int32),
                       np.concatenate([v % n, hv]).astype(np.int32))


@pytest.mark.parametrize("hub_degree", [0, 1, 3, 16384])
@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("build", [0, 1])
def test_hub_owners(policy, hub_degree, build):
    """Task lists built by the partitioned transpose (build 0) or by per-entry
    appends (build 1, the fallback), and from the owners' own rows (every owner
    from degree 1, from degree 3, only the four hubs of degree 20000+, none at
    the default) give the same bit-exact counts."""
    off, adj = _hub_graph(40000, 4, 20000, 3)
    assert np.sort(np.diff(off))[-4] >= 16384
    want = O.custinger_count(off, adj)
    pv, glob, st = count(off, adj, policy=policy, hub_degree=hub_degree, build=build)
    assert st["method_used"] == T.METHOD_ORIENTED
    assert np.array_equal(pv, want)
    assert glob == int(want.astype(np.uint64).sum())


@pytest.mark.parametrize("build", [0, 1])
@pytest.mark.parametrize("hub_degree", [0, 1, 16384])
@pytest.mark.parametrize("drop", ["hub_side", "leaf_side"])
def test_declared_symmetric_hub_rows_that_are_not(drop, hub_degree, build):
    """Rows declared symmetric (check skipped) with one entry of a hub pair
    missing: the task construction notices (appends or hub rows) and AUTO
    takes the direct method, whose result is the reference formula on the
    rows as given."""
    off, adj = _hub_graph(40000, 2, 20000, 4)
    hub = int(np.argmax(np.diff(off)))
    row = adj[off[hub]:off[hub + 1]]
    leaf = int(row[len(row) // 2])
    if drop == "hub_side":
        r, x = hub, leaf
    else:
        r, x = leaf, hub
    lst = adj[off[r]:off[r + 1]]
    k = off[r] + int(np.searchsorted(lst, x))
    adj2 = np.delete(adj, k)
    off2 = off.copy()
    off2[r + 1:] -= 1
    want = O.custinger_count(off2, adj2)
    g = T.cuStinger(off2, adj2, assume_symmetric=True)
    try:
        pv, glob, st = T.static_count(g, T.default_schedule(hub_degree=hub_degree, build=build))
        assert st["method_used"] == T.METHOD_DIRECT
        assert np.array_equal(pv, want)
        assert glob == int(want.astype(np.uint64).sum())
        with pytest.raises(T.TCError):
            T.static_count(g, T.default_schedule(method=T.METHOD_ORIENTED, hub_degree=hub_degree,
                                                 build=build))
    finally:
        g.close()


@pytest.mark.parametrize("method", [T.METHOD_AUTO, T.METHOD_DIRECT])
def test_rmat20_golden_hashes(method):
    """Scale-20 config from the device generator: its CSR and the per-vertex
    counts hash to the committed fixture values (bit-exact at 31.4M entries)."""
    import hashlib
    meta = json.load(open(os.path.join(GR.GOLDEN, "fixtures.json")))["rmat20"]
    g = T.rmat_graph_device(20, 16, 2)
    try:
        off, adj = g.to_csr()
        assert hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest() == meta["csr_sha256"]
        pv, glob, st = T.static_count(g, T.default_schedule(method=method))
        assert hashlib.sha256(pv.tobytes()).hexdigest() == meta["pv_sha256"]
        assert glob == 6 * meta["T"]
    finally:
        g.close()
