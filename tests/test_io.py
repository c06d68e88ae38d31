"""Graph file readers (host code in the library, no GPU): DIMACS/METIS,
SNAP and MatrixMarket inputs against hand-built CSRs, and round trips of a
generated R-MAT graph through each format."""
from __future__ import annotations

import numpy as np
import pytest

import custingeralg_amd as T
import oracle as O


def _csr(n, pairs):
    s = sorted({(a, b) for a, b in pairs if a != b} | {(b, a) for a, b in pairs if a != b})
    off = np.zeros(n + 1, np.int64)
    for a, _ in s:
        off[a + 1] += 1
    return np.cumsum(off), np.array([b for _, b in s], np.int32)


def _check(path, n, pairs, **kw):
    off, adj = T.read_graph(str(path), **kw)
    want_off, want_adj = _csr(n, pairs)
    assert np.array_equal(off, want_off) and np.array_equal(adj, want_adj)


def test_dimacs_plain_and_weighted(tmp_path):
    p = tmp_path / "tri.graph"
    p.write_text("% a triangle with a tail\n4 4\n2 3\n1 3 4\n1 2\n2\n")
    _check(p, 4, [(0, 1), (0, 2), (1, 2), (1, 3)])
    # fmt 1: an edge weight after every neighbour; fmt 11: a vertex weight first
    p = tmp_path / "w.graph"
    p.write_text("3 3 1\n2 5 3 7\n1 5 3 2\n1 7 2 2\n")
    _check(p, 3, [(0, 1), (0, 2), (1, 2)])
    p = tmp_path / "vw.graph"
    p.write_text("3 2 11\n9 2 4\n8 1 4 3 1\n7 2 1\n")
    _check(p, 3, [(0, 1), (1, 2)])


def test_snap_loops_and_duplicates(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# FromNodeId\tToNodeId\n0\t1\n1\t2\n2\t0\n2\t2\n1\t0\n\n4 2\n")
    _check(p, 5, [(0, 1), (1, 2), (0, 2), (2, 4)])
    off, adj = T.read_graph(str(p), as_is=True)  # directed pairs as given, loop kept
    assert off.tolist() == [0, 1, 3, 5, 5, 6]
    assert adj.tolist() == [1, 0, 2, 0, 2, 2]


def test_matrix_market_symmetric_and_general(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% c\n4 4 4\n2 1\n3 1\n3 2\n4 4\n")
    _check(p, 4, [(0, 1), (0, 2), (1, 2)])
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 2 0.5\n3 2 -1e3\n")
    _check(p, 3, [(0, 1), (2, 1)])
    off, adj = T.read_graph(str(p), as_is=True)
    assert off.tolist() == [0, 1, 1, 2] and adj.tolist() == [1, 1]


def test_errors(tmp_path):
    with pytest.raises(T.TCError):
        T.read_graph(str(tmp_path / "missing.graph"))
    p = tmp_path / "bad.graph"
    p.write_text("2 1\n3\n1\n")
    with pytest.raises(T.TCError):
        T.read_graph(str(p))
    p = tmp_path / "x.dat"
    p.write_text("0 1\n")
    with pytest.raises(T.TCError):
        T.read_graph(str(p))  # no format from the name
    off, adj = T.read_graph(str(p), fmt=T.FMT_SNAP)
    assert adj.tolist() == [1, 0]


@pytest.mark.parametrize("fmt", ["graph", "txt", "mtx"])
def test_round_trip_rmat(tmp_path, fmt):
    off, adj = O.rmat_graph(10, 8, 3)
    n = len(off) - 1
    rows = np.repeat(np.arange(n), np.diff(off))
    p = tmp_path / f"kron_g.{fmt}"
    with open(p, "w") as f:
        if fmt == "graph":
            f.write(f"{n} {len(adj) // 2}\n")
            for v in range(n):
                f.write(" ".join(str(x + 1) for x in adj[off[v]:off[v + 1]]) + "\n")
        elif fmt == "txt":
            f.write("# directed listing of both directions\n")
            for a, b in zip(rows, adj):
                f.write(f"{a}\t{b}\n")
        else:
            keep = rows > adj  # lower triangle only
            f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
            f.write(f"{n} {n} {int(keep.sum())}\n")
            for a, b in zip(rows[keep], adj[keep]):
                f.write(f"{a + 1} {b + 1}\n")
    got_off, got_adj = T.read_graph(str(p))
    # trailing isolated vertices are not visible in an edge list
    m = len(got_off) - 1
    assert np.array_equal(got_off, off[:m + 1]) and np.all(off[m:] == off[-1])
    assert np.array_equal(got_adj, adj)
