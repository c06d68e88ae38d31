"""Writes tests/golden/fixtures.npz: small graphs with per-vertex triangle
counts computed by independent libraries (networkx.triangles and scipy's
diag(A^3)/2), plus the ER n=2^16 / d=16 / seed 1 per-vertex vector in the
reference convention.

Run in the build container only (networkx is a checker, never shipped):
    python tests/golden/make_fixtures.py
The reference ships no graph files, goldens or asserts for this path, so
every expected value here comes from those libraries or closed forms.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def csr_of(G, n):
    import networkx as nx  # noqa: F401
    lists = [sorted(int(u) for u in G.neighbors(v) if u != v) for v in range(n)]
    off = np.zeros(n + 1, np.int64)
    for i, l in enumerate(lists):
        off[i + 1] = off[i] + len(l)
    adj = np.array([x for l in lists for x in l], np.int32)
    return off, adj


def main():
    import networkx as nx
    import scipy.sparse as sp

    graphs = {
        "karate": nx.karate_club_graph(),
        "les_miserables": nx.les_miserables_graph(),
        "florentine": nx.florentine_families_graph(),
        "davis": nx.davis_southern_women_graph(),
        "petersen": nx.petersen_graph(),
        "dodecahedral": nx.dodecahedral_graph(),
        "barbell": nx.barbell_graph(6, 3),
        "lollipop": nx.lollipop_graph(7, 4),
        "powerlaw_cluster": nx.powerlaw_cluster_graph(300, 4, 0.3, seed=7),
        "gnp_dense": nx.gnp_random_graph(120, 0.3, seed=3),
        "ws": nx.watts_strogatz_graph(400, 8, 0.1, seed=5),
        "ba": nx.barabasi_albert_graph(2000, 6, seed=11),
    }
    out = {}
    summary = {}
    for name, G in graphs.items():
        G = nx.convert_node_labels_to_integers(nx.Graph(G))
        n = G.number_of_nodes()
        off, adj = csr_of(G, n)
        tri = nx.triangles(G)
        t = np.array([tri[v] for v in range(n)], np.int64)
        # independent second checker: diag(A^3)/2
        A = sp.csr_matrix((np.ones(len(adj)), adj, off), shape=(n, n))
        d3 = (A @ A @ A).diagonal() / 2
        assert np.array_equal(np.rint(d3).astype(np.int64), t), name
        out[f"{name}_off"] = off
        out[f"{name}_adj"] = adj
        out[f"{name}_t"] = t
        summary[name] = {"n": n, "nnz": int(len(adj)), "T": int(t.sum() // 3)}

    # ER n=2^16, average degree 16, seed 1 (the plain-method configuration)
    import oracle as O
    off, adj = O.er_graph(1 << 16, 16, 1)
    pv = O.custinger_count(off, adj)
    T, tv = O.forward_count(off, adj, per_vertex=True)
    assert np.array_equal(pv.astype(np.uint64), 2 * tv)
    out["er16_pv"] = pv
    summary["er16"] = {
        "n": 1 << 16, "nnz": int(len(adj)), "T": int(T),
        "csr_sha256": hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest(),
        "pv_sha256": hashlib.sha256(pv.tobytes()).hexdigest(),
    }
    # R-MAT scale 16, ef 16, seed 2: hashes of the canonical CSR and counts
    off, adj = O.rmat_graph(16, 16, 2)
    T, tv = O.forward_count(off, adj, per_vertex=True)
    pv = (2 * tv).astype(np.uint32)
    summary["rmat16"] = {
        "nnz": int(len(adj)), "T": int(T),
        "csr_sha256": hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest(),
        "pv_sha256": hashlib.sha256(pv.tobytes()).hexdigest(),
    }
    # R-MAT scale 20, ef 16, seed 2: the same hashes at the scale-20 config;
    # the per-vertex vector is the forward count's, checked here against the
    # reference-formula restatement (about a minute)
    off, adj = O.rmat_graph(20, 16, 2)
    T, tv = O.forward_count(off, adj, per_vertex=True)
    pv = (2 * tv).astype(np.uint32)
    assert np.array_equal(pv, O.custinger_count(off, adj))
    summary["rmat20"] = {
        "nnz": int(len(adj)), "T": int(T),
        "csr_sha256": hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest(),
        "pv_sha256": hashlib.sha256(pv.tobytes()).hexdigest(),
    }
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **out)
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(summary, f, indent=1, sort_keys=True)
    print(json.dumps(summary, indent=1))


if __name__ == "__main__":
    main()
