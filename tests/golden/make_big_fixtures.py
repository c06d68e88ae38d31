"""Writes tests/golden/big_fixtures.json: expected results of the three
full-size workloads, computed on the CPU by the oracle alone.

  rmat24   static count of R-MAT scale 24, edge factor 32,
           (a,b,c,d) = (0.57,0.19,0.19,0.05), seed 4: T, the SHA-256 of the
           canonical CSR (int64 offsets then int32 ids) and of the uint32
           per-vertex vector 2*t(v) mod 2^32, and how many of those values
           wrapped past 2^32 or read negative as int32;
  stream   R-MAT scale 22 ef 16 seed 3 base and its 20 update batches
           (custingeralg_amd.workloads): per batch the pairs, deletions,
           undirected pairs inserted / deleted and T after the batch; the
           SHA-256 of the final per-vertex vector;
  batched  the 4096 R-MAT graphs of the batched workload: T of every graph,
           SHA-256 of the concatenated per-vertex vector.

The counts come from the oracle's degree-oriented forward count
(tco_forward_count), pinned against networkx and scipy on the small fixtures
and against the reference-formula restatement at R-MAT 16 and 20.  The graphs
come from the oracle's generators, bit-identical to the device generators.

Run in the build container (tens of minutes; rmat24 needs about 30 GB of
host memory):
    python tests/golden/make_big_fixtures.py [rmat24] [stream] [batched]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
OUT = os.path.join(HERE, "big_fixtures.json")


def sha_csr(off, adj) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, np.int64).tobytes())
    h.update(np.ascontiguousarray(adj, np.int32).tobytes())
    return h.hexdigest()


def pv_of(t) -> np.ndarray:
    """2*t(v) mod 2^32, the reference's per-vertex value."""
    return ((2 * t.astype(np.uint64)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def log(msg):
    print(msg, flush=True)


def stage_rmat24():
    import oracle as O
    t0 = time.time()
    off, adj = O.rmat_graph(24, 32, 4)
    log(f"rmat24: nnz={len(adj)} generated in {time.time() - t0:.0f}s")
    deg = np.diff(off)
    rec = {"scale": 24, "edge_factor": 32, "seed": 4, "nv": len(off) - 1, "nnz": int(len(adj)),
           "max_degree": int(deg.max()), "csr_sha256": sha_csr(off, adj)}
    del deg
    t0 = time.time()
    T, t = O.forward_count(off, adj, per_vertex=True)
    log(f"rmat24: T={T} counted in {time.time() - t0:.0f}s")
    pv = pv_of(t)
    two_t = 2 * t.astype(np.uint64)
    rec.update({"T": int(T), "global": int(6 * T), "pv_sha256": hashlib.sha256(pv.tobytes()).hexdigest(),
                "pv_wrapped": int((two_t >= (1 << 32)).sum()),
                "pv_wrap_sum": int((two_t >> np.uint64(32)).sum()),
                "pv_negative_int32": int((pv >= (1 << 31)).sum()),
                "max_2t": int(two_t.max()), "argmax_2t": int(two_t.argmax()),
                "pv_at_argmax": int(pv[two_t.argmax()])})
    return rec


def stage_stream():
    import oracle as O
    from custingeralg_amd import workloads as W
    b = W.STREAM_BASE
    off, adj = O.rmat_graph(b["scale"], b["edge_factor"], b["seed"])
    nv = len(off) - 1
    T0 = O.forward_count(off, adj)[0]
    log(f"stream: base nnz={len(adj)} T={T0}")
    rec = {"base": dict(b, nnz=int(len(adj)), T=int(T0), csr_sha256=sha_csr(off, adj)),
           "batches": []}
    rng = np.random.default_rng(W.STREAM_RNG_SEED)
    for i, B in enumerate(W.STREAM_SIZES):
        t0 = time.time()
        u, v, dl = W.stream_batch(O.er_pairs, lambda s, m, sd: O.rmat_pairs(s, m, sd), off, adj, nv,
                                  B, b["scale"], W.stream_batch_seed(i), rng)
        off, adj, ins, dele = O.apply_batch_sorted(nv, off, adj, u, v, dl)
        Ti = O.forward_count(off, adj)[0]
        rec["batches"].append({"B": B, "pairs": int(len(u)), "deletions": int(dl.sum()),
                               "inserted": int(ins), "deleted": int(dele), "nnz": int(len(adj)),
                               "T": int(Ti),
                               "pairs_sha256": hashlib.sha256(u.tobytes() + v.tobytes() +
                                                              dl.tobytes()).hexdigest()})
        log(f"stream: batch {i} B={B} pairs={len(u)} ins={ins} del={dele} T={Ti} "
            f"({time.time() - t0:.0f}s)")
    T, t = O.forward_count(off, adj, per_vertex=True)
    rec["final"] = {"T": int(T), "nnz": int(len(adj)), "csr_sha256": sha_csr(off, adj),
                    "pv_sha256": hashlib.sha256(pv_of(t).tobytes()).hexdigest()}
    return rec


def stage_batched():
    import oracle as O
    n, ef, seeds = O.batched_spec()
    Ts = []
    h = hashlib.sha256()
    nnz = 0
    t0 = time.time()
    for i, (ni, ei, si) in enumerate(zip(n, ef, seeds)):
        sc = max(1, int(np.ceil(np.log2(ni))))
        u, v = O.rmat_pairs(sc, int(ei) * int(ni), int(si))
        off, adj = O.canon_csr(int(ni), u, v)
        T, t = O.forward_count(off, adj, per_vertex=True)
        Ts.append(int(T))
        h.update(pv_of(t).tobytes())
        nnz += len(adj)
        if i % 512 == 0:
            log(f"batched: graph {i} n={ni} ef={ei} T={T} ({time.time() - t0:.0f}s)")
    return {"graphs": len(n), "nv": int(n.astype(np.int64).sum()), "nnz": int(nnz),
            "T_total": int(sum(Ts)), "T": Ts, "pv_sha256": h.hexdigest()}


def main(argv):
    stages = argv or ["rmat24", "stream", "batched"]
    fns = {"rmat24": stage_rmat24, "stream": stage_stream, "batched": stage_batched}
    for s in stages:
        t0 = time.time()
        rec = fns[s]()
        rec["cpu_seconds"] = round(time.time() - t0, 1)
        data = {}
        if os.path.exists(OUT):
            with open(OUT) as f:
                data = json.load(f)
        data[s] = rec
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        log(f"{s}: written ({rec['cpu_seconds']} s)")


if __name__ == "__main__":
    main(sys.argv[1:])
