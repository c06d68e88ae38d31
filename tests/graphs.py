"""Small graphs with closed-form triangle counts, shared by the CPU and GPU tests."""
from __future__ import annotations

import os
from math import comb

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def complete(n):
    """K_n: t(v) = C(n-1, 2)."""
    e = [(i, j) for i in range(n) for j in range(i + 1, n)]
    off, adj = O.csr_from_edges(n, e)
    return off, adj, np.full(n, comb(n - 1, 2), np.int64)


def complete_fast(n):
    """K_n built directly (large n)."""
    off = np.arange(n + 1, dtype=np.int64) * (n - 1)
    rows = [np.concatenate([np.arange(0, v), np.arange(v + 1, n)]) for v in range(n)]
    adj = np.concatenate(rows).astype(np.int32)
    return off, adj, np.full(n, comb(n - 1, 2), np.int64)


def star(k):
    off, adj = O.csr_from_edges(k + 1, [(0, i) for i in range(1, k + 1)])
    return off, adj, np.zeros(k + 1, np.int64)


def cycle(n):
    off, adj = O.csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)])
    t = np.zeros(n, np.int64) if n > 3 else np.ones(n, np.int64)
    return off, adj, t


def wheel(n):
    """Hub 0 joined to a cycle 1..n: hub t = n, rim t = 2 (n >= 4)."""
    e = [(0, i) for i in range(1, n + 1)] + [(i, i % n + 1) for i in range(1, n + 1)]
    off, adj = O.csr_from_edges(n + 1, e)
    t = np.full(n + 1, 2, np.int64)
    t[0] = n
    return off, adj, t


def friendship(k):
    """F_k: k triangles sharing vertex 0."""
    e = []
    for i in range(k):
        a, b = 2 * i + 1, 2 * i + 2
        e += [(0, a), (0, b), (a, b)]
    off, adj = O.csr_from_edges(2 * k + 1, e)
    t = np.ones(2 * k + 1, np.int64)
    t[0] = k
    return off, adj, t


def bipartite(m, n):
    e = [(i, m + j) for i in range(m) for j in range(n)]
    off, adj = O.csr_from_edges(m + n, e)
    return off, adj, np.zeros(m + n, np.int64)


def grid(r, c):
    e = []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                e.append((v, v + 1))
            if i + 1 < r:
                e.append((v, v + c))
    off, adj = O.csr_from_edges(r * c, e)
    return off, adj, np.zeros(r * c, np.int64)


def closed_form_cases():
    return {
        "K3": complete(3), "K4": complete(4), "K10": complete(10), "K70": complete(70),
        "star50": star(50), "C3": cycle(3), "C9": cycle(9), "W8": wheel(8), "W100": wheel(100),
        "F5": friendship(5), "F300": friendship(300), "K7_9": bipartite(7, 9),
        "grid12x9": grid(12, 9),
    }


def fixture_cases():
    z = np.load(os.path.join(GOLDEN, "fixtures.npz"))
    names = sorted({k[:-4] for k in z.files if k.endswith("_off")})
    return {n: (z[f"{n}_off"], z[f"{n}_adj"], z[f"{n}_t"]) for n in names}


def random_lists(n, p, seed, loops=False, symmetric=False):
    """Random rows (sorted, unique); optionally asymmetric and with self-loops."""
    rng = np.random.default_rng(seed)
    A = rng.random((n, n)) < p
    if symmetric:
        A = np.triu(A, 1)
        A = A | A.T
    if not loops:
        np.fill_diagonal(A, False)
    lists = [list(np.nonzero(A[i])[0]) for i in range(n)]
    return O.csr_from_lists(lists)


def brute_custinger(off, adj):
    """The static kernel's formula by Python sets (cct.cu:233-268, tpi=1)."""
    n = len(off) - 1
    N = [set(adj[off[v]:off[v + 1]].tolist()) for v in range(n)]
    deg = np.diff(off)
    out = np.zeros(n, np.uint64)
    for s in range(n):
        tot = 0
        for d in adj[off[s]:off[s + 1]]:
            d = int(d)
            if s == d or deg[d] < 2 or deg[s] < 2:
                continue
            tot += len(N[s] & N[d])
        out[s] = tot
    return (out % (1 << 32)).astype(np.uint32)
