"""The three full-size workloads on the GPU against the oracle's goldens
(tests/golden/big_fixtures.json, written by make_big_fixtures.py on the CPU):

* static count of R-MAT scale 24, edge factor 32, seed 4 under the fixed
  heuristic and the learned strategy table: T, the canonical CSR's and the
  per-vertex vector's SHA-256, and the per-vertex values that wrapped past
  2^32 (the reference's int32 counts, include/static_triangle_counting/cct.hpp:12,
  wrap the same way);
* the streaming workload (R-MAT 22 base, 20 batches of 10^3..10^6 pairs):
  after every batch the delta equals the change of the oracle's T and a
  recount on the GPU, and the final per-vertex vector (base count plus the
  accumulated deltas) hashes to the oracle's (tests/cct_upd.cpp:390-414 runs
  the same recount after its batches);
* the 4096 batched R-MAT graphs: every graph's count and the concatenated
  per-vertex vector.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import custingeralg_amd as T
from custingeralg_amd import learned as Lm
from custingeralg_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "big_fixtures.json")


def golden(section):
    with open(GOLDEN) as f:
        return json.load(f)[section]


def sha_csr(off, adj):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, np.int64).tobytes())
    h.update(np.ascontiguousarray(adj, np.int32).tobytes())
    return h.hexdigest()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.timeout(600)
def test_rmat24_static_fixed_and_learned():
    gd = golden("rmat24")
    g = T.rmat_graph_device(24, 32, 4)
    try:
        info = g.info()
        assert info["nnz"] == gd["nnz"] and info["max_degree"] == gd["max_degree"]
        off, adj = g.to_csr()
        assert sha_csr(off, adj) == gd["csr_sha256"]
        del off, adj
        out = T.DeviceArray(g.nv, np.uint32)
        for policy in (T.POLICY_FIXED, T.POLICY_LEARNED):
            if policy == T.POLICY_LEARNED:
                Lm.load_selector(Lm.default_model_path())
            _, glob, st = T.static_count(g, T.default_schedule(policy=policy), out=out)
            pv = out.to_host()
            print(f"policy {policy}: global {glob} T {st['triangles']} "
                  f"{st['ms_total']:.1f} ms", flush=True)
            assert glob == gd["global"] and st["triangles"] == gd["T"]
            assert sha(pv) == gd["pv_sha256"]
            assert int((pv >= (1 << 31)).sum()) == gd["pv_negative_int32"]
            assert int(pv[gd["argmax_2t"]]) == gd["pv_at_argmax"]
            # the global is the exact sum; the uint32 values wrapped pv_wrap_sum
            # times 2^32 in all
            assert int(pv.astype(np.uint64).sum()) + (gd["pv_wrap_sum"] << 32) == glob
        out.free()
    finally:
        g.close()


@pytest.mark.timeout(600)
def test_stream_rmat22_twenty_batches():
    gd = golden("stream")
    b = W.STREAM_BASE
    g = T.rmat_graph_device(b["scale"], b["edge_factor"], b["seed"])
    nv = g.nv
    try:
        assert g.info()["nnz"] == gd["base"]["nnz"]
        base_pv, total, _ = T.static_count(g)
        assert total == 6 * gd["base"]["T"]
        delta = T.DeviceArray(nv, np.uint32).zero()
        rng = np.random.default_rng(W.STREAM_RNG_SEED)

        def er(n, m, seed):
            du, dv = T.DeviceArray(m, np.int32), T.DeviceArray(m, np.int32)
            T.check(T.lib().tc_gen_er_pairs(n, m, seed, du.ptr, dv.ptr, None))
            return du.to_host(), dv.to_host()

        def rmat(scale, m, seed):
            du, dv = T.DeviceArray(m, np.int32), T.DeviceArray(m, np.int32)
            T.check(T.lib().tc_gen_rmat_pairs(scale, m, 0.57, 0.19, 0.19, seed, du.ptr, dv.ptr,
                                              None))
            return du.to_host(), dv.to_host()

        prev_T = gd["base"]["T"]
        for i, (B, want) in enumerate(zip(W.STREAM_SIZES, gd["batches"])):
            off, adj = g.to_csr()
            u, v, dl = W.stream_batch(er, rmat, off, adj, nv, B, b["scale"],
                                      W.stream_batch_seed(i), rng)
            del off, adj
            assert sha(np.concatenate([u.view(np.uint8), v.view(np.uint8), dl])) == \
                want["pairs_sha256"], f"batch {i}: the recipe drew other pairs"
            dg, st = T.batch_apply(g, u, v, dl, delta)
            total += dg
            _, recount, _ = T.static_count(g, per_vertex=False)
            print(f"batch {i}: B={len(u)} delta {dg} ins {st['inserted']} del {st['deleted']} "
                  f"update {st['ms_update']:.2f} ms delta {st['ms_delta']:.2f} ms", flush=True)
            assert st["inserted"] == want["inserted"] and st["deleted"] == want["deleted"]
            assert st["dup_inserts"] == (len(u) - want["deletions"]) - want["inserted"]
            assert st["missing_deletes"] == want["deletions"] - want["deleted"]
            assert dg == 6 * (want["T"] - prev_T), f"batch {i}: delta"
            assert recount == total == 6 * want["T"], f"batch {i}: running total"
            assert g.info()["nnz"] == want["nnz"]
            prev_T = want["T"]
        final = (base_pv + delta.to_host()).astype(np.uint32)  # mod 2^32, as accumulated
        assert sha(final) == gd["final"]["pv_sha256"]
        pv, glob, _ = T.static_count(g)
        assert sha(pv) == gd["final"]["pv_sha256"] and glob == 6 * gd["final"]["T"]
        off, adj = g.to_csr()
        assert sha_csr(off, adj) == gd["final"]["csr_sha256"]
        delta.free()
    finally:
        g.close()


@pytest.mark.timeout(600)
def test_batched_4096_graphs():
    gd = golden("batched")
    n, ef, seeds = T.batched_spec()
    assert len(n) == gd["graphs"] == 4096
    dvoff, doff, dadj = T.batched_graph_device(n, ef, seeds)
    assert dadj.n == gd["nnz"]
    pg = T.DeviceArray(len(n), np.uint64)
    pv = T.DeviceArray(int(n.astype(np.int64).sum()), np.uint32)
    bg = T.BatchedGraph(dvoff, doff, dadj)
    try:
        for policy in (T.POLICY_FIXED, T.POLICY_HASH):
            st = bg.count(pg, pv, T.default_schedule(policy=policy))
            got = pg.to_host()
            want = np.array(gd["T"], np.uint64) * np.uint64(6)
            bad = np.nonzero(got != want)[0]
            assert len(bad) == 0, f"policy {policy}: graphs {bad[:10].tolist()} differ"
            assert sha(pv.to_host()) == gd["pv_sha256"]
            assert st["triangles"] == gd["T_total"]
    finally:
        bg.close()
        pg.free()
        pv.free()
        for a in (dvoff, doff, dadj):
            a.free()
