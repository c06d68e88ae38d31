"""CPU oracle for the triangle-counting path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package; it is the checker, never the
thing measured or shipped.  It wraps ``liboracle.so`` (``tc_oracle.c``), a
plain-C restatement of the reference routines named in ``tc_oracle.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.tco_rng.restype = C.c_uint64
        L.tco_rng.argtypes = [C.c_uint64, C.c_uint64]
        L.tco_gen_er_pairs.argtypes = [C.c_int64, C.c_int64, C.c_uint64, i32p, i32p]
        L.tco_rmat_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, u64p]
        L.tco_gen_rmat_pairs.argtypes = [C.c_int, C.c_int64, u64p, C.c_uint64, i32p, i32p]
        L.tco_canon_csr.restype = C.c_int64
        L.tco_canon_csr.argtypes = [C.c_int64, C.c_int64, i32p, i32p, i64p, i32p]
        L.tco_host_single_intersection.restype = C.c_int32
        L.tco_host_single_intersection.argtypes = [C.c_int32, C.c_int64, i32p,
                                                   C.c_int32, C.c_int64, i32p]
        L.tco_host_count_triangles.restype = C.c_int64
        L.tco_host_count_triangles.argtypes = [C.c_int32, i64p, i32p, C.c_void_p, C.c_int]
        L.tco_host_count_rows.restype = C.c_int64
        L.tco_host_count_rows.argtypes = [C.c_int32, i64p, i32p, C.c_int32, C.c_int32, C.c_int]
        L.tco_host_count_entries.restype = C.c_int64
        L.tco_host_count_entries.argtypes = [C.c_int32, i64p, i32p, i64p, C.c_int64, C.c_int]
        L.tco_custinger_count.argtypes = [C.c_int32, i64p, i32p, u32p, C.c_int]
        L.tco_csr_count.argtypes = [C.c_int32, i64p, i32p, u32p, C.c_int]
        L.tco_sum_triangle_array.restype = C.c_int64
        L.tco_sum_triangle_array.argtypes = [i32p, C.c_int32]
        L.tco_forward_count.restype = C.c_uint64
        L.tco_forward_count.argtypes = [C.c_int32, i64p, i32p, C.c_void_p, C.c_int]
        L.tco_batch_prepare.argtypes = [C.c_int32, C.c_int64, i32p, i32p, i64p, i32p,
                                        i32p, i32p, i32p, i64p]
        L.tco_stream_delta.argtypes = [C.c_int32, i64p, i32p, C.c_int64, i32p, i32p, i32p,
                                       i64p, u32p, C.c_int]
        L.tco_lb_partition.restype = C.c_int64
        L.tco_lb_partition.argtypes = [C.c_int64, i64p, i64p, i64p]
        _LIB = L
    return _LIB


def nthreads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


# ----------------------------------------------------------------- generators
def rng(seed: int, ctr: int) -> int:
    return int(lib().tco_rng(seed, ctr))


def er_pairs(n: int, m: int, seed: int):
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    lib().tco_gen_er_pairs(n, m, seed, u, v)
    return u, v


def rmat_thresholds(a: float, b: float, c: float) -> np.ndarray:
    t = np.zeros(3, np.uint64)
    lib().tco_rmat_thresholds(a, b, c, t)
    return t


def rmat_pairs(scale: int, m: int, seed: int, abc=(0.57, 0.19, 0.19)):
    t = rmat_thresholds(*abc)
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    lib().tco_gen_rmat_pairs(scale, m, t, seed, u, v)
    return u, v


def canon_csr(n: int, u: np.ndarray, v: np.ndarray):
    """Symmetrise, drop self-loops, dedup, sort -> (off int64[n+1], adj int32[nnz])."""
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    off = np.zeros(n + 1, np.int64)
    adj = np.empty(max(1, 2 * len(u)), np.int32)
    nnz = lib().tco_canon_csr(n, len(u), u, v, off, adj)
    return off, adj[:nnz].copy()


def er_graph(n: int, avg_deg: int, seed: int):
    u, v = er_pairs(n, n * avg_deg // 2, seed)
    return canon_csr(n, u, v)


def rmat_graph(scale: int, edge_factor: int, seed: int, abc=(0.57, 0.19, 0.19)):
    n = 1 << scale
    u, v = rmat_pairs(scale, edge_factor * n, seed, abc)
    return canon_csr(n, u, v)


def batched_spec(G: int = 4096, seed_root: int = 1000, nmin: int = 1024, nmax: int = 65536,
                 efmin: int = 4, efmax: int = 32):
    """Sizes of the batched workload: graph i has n_i = floor(exp(U(ln nmin, ln nmax)))
    vertices and edge factor ef_i ~ U{efmin..efmax}, both drawn from rng(seed_root+i, .),
    and is generated with seed seed_root+i."""
    import math
    n = np.empty(G, np.int32)
    ef = np.empty(G, np.int32)
    seeds = np.empty(G, np.uint64)
    lo, hi = math.log(nmin), math.log(nmax)
    for i in range(G):
        u01 = (rng(seed_root + i, 0) >> 11) / float(1 << 53)
        n[i] = min(nmax, max(nmin, int(math.floor(math.exp(lo + u01 * (hi - lo))))))
        ef[i] = efmin + rng(seed_root + i, 1) % (efmax - efmin + 1)
        seeds[i] = seed_root + i
    return n, ef, seeds


def batched_graph(n, ef, seeds, abc=(0.57, 0.19, 0.19)):
    """Host build of the concatenated CSR: graph i is R-MAT on 2^ceil(log2 n_i) with
    pairs having an endpoint >= n_i rejected; ids offset by the previous n."""
    offs, adjs, voff = [], [], [0]
    base_e = 0
    for ni, ei, si in zip(n, ef, seeds):
        sc = max(1, int(np.ceil(np.log2(ni))))
        u, v = rmat_pairs(sc, int(ei) * int(ni), int(si), abc)
        o, a = canon_csr(int(ni), u, v)
        offs.append(o[:-1] + base_e)
        adjs.append(a + voff[-1])
        base_e += len(a)
        voff.append(voff[-1] + int(ni))
    off = np.concatenate(offs + [np.array([base_e], np.int64)])
    adj = np.concatenate(adjs).astype(np.int32) if adjs else np.zeros(0, np.int32)
    return np.array(voff, np.int64), off, adj


def csr_from_edges(n: int, edges) -> tuple:
    """Symmetric simple CSR from an undirected edge iterable."""
    e = np.asarray(list(edges), dtype=np.int64).reshape(-1, 2)
    if len(e) == 0:
        return np.zeros(n + 1, np.int64), np.zeros(0, np.int32)
    return canon_csr(n, e[:, 0].astype(np.int32), e[:, 1].astype(np.int32))


def csr_from_lists(lists) -> tuple:
    """CSR exactly as given (rows may be asymmetric or hold self-loops)."""
    n = len(lists)
    off = np.zeros(n + 1, np.int64)
    for i, l in enumerate(lists):
        off[i + 1] = off[i] + len(l)
    adj = np.array([x for l in lists for x in l], dtype=np.int32)
    return off, adj


# -------------------------------------------------------------------- counts
def host_count_triangles(off, adj, per_edge=False, threads=1):
    """hostCountTriangles (tests/cct.cpp:71-89) -> (int64 total, triNE or None)."""
    n = len(off) - 1
    tri = np.zeros(len(adj), np.int32) if per_edge else None
    ptr = tri.ctypes.data if per_edge and len(adj) else None
    total = lib().tco_host_count_triangles(n, off, _nz(adj), ptr, threads)
    return int(total), tri


def host_count_rows(off, adj, v0, v1, threads=1) -> int:
    return int(lib().tco_host_count_rows(len(off) - 1, off, _nz(adj), v0, v1, threads))


def host_count_entries(off, adj, entries, threads=1) -> int:
    """hostCountTriangles' per-entry intersections on a sample of entry indices."""
    entries = np.ascontiguousarray(entries, np.int64)
    return int(lib().tco_host_count_entries(len(off) - 1, off, _nz(adj), _nz(entries),
                                            len(entries), threads))


def custinger_count(off, adj, threads=None) -> np.ndarray:
    """Per-vertex uint32 counts with the static kernel's semantics (cct.cu:216-270)."""
    n = len(off) - 1
    out = np.zeros(max(n, 1), np.uint32)
    lib().tco_custinger_count(n, off, _nz(adj), out, threads or nthreads())
    return out[:n]


def csr_count(off, adj, threads=None) -> np.ndarray:
    """Per-vertex uint32 counts with the CSR kernel's semantics (tests/cct_csr.cu:228-276)."""
    n = len(off) - 1
    out = np.zeros(max(n, 1), np.uint32)
    lib().tco_csr_count(n, off, _nz(adj), out, threads or nthreads())
    return out[:n]


def sum_triangle_array(h) -> int:
    h = np.ascontiguousarray(h).view(np.int32)
    return int(lib().tco_sum_triangle_array(h, len(h)))


def forward_count(off, adj, per_vertex=False, threads=None):
    """Degree-oriented count on a symmetric simple CSR -> (T, t[v] uint64 or None)."""
    n = len(off) - 1
    t = np.zeros(max(n, 1), np.uint64) if per_vertex else None
    T = lib().tco_forward_count(n, off, _nz(adj), t.ctypes.data if per_vertex else None,
                                threads or nthreads())
    return int(T), (t[:n] if per_vertex else None)


# ----------------------------------------------------------------- streaming
def batch_prepare(n, src, dst, g_off, g_adj):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    nb = len(src)
    s_src = np.empty(max(nb, 1), np.int32)
    s_dst = np.empty(max(nb, 1), np.int32)
    dup = np.empty(max(nb, 1), np.int32)
    boff = np.zeros(n + 1, np.int64)
    lib().tco_batch_prepare(n, nb, _nz(src), _nz(dst), g_off, _nz(g_adj), s_src, s_dst, dup,
                            boff)
    return s_src[:nb], s_dst[:nb], dup[:nb], boff


def stream_delta(gp_off, gp_adj, s_src, s_dst, dup, boff, pass_mask=7, out=None):
    n = len(gp_off) - 1
    if out is None:
        out = np.zeros(max(n, 1), np.uint32)
    lib().tco_stream_delta(n, gp_off, _nz(gp_adj), len(s_src), _nz(s_src), _nz(s_dst),
                           _nz(dup), boff, out, pass_mask)
    return out[:n]


def apply_batch_csr(n, off, adj, ins_u=(), ins_v=(), del_u=(), del_v=()):
    """Reference graph update on the host: remove the undirected pairs in
    del, add those in ins (no-ops for present/absent pairs)."""
    s = set()
    for a in range(n):
        for b in adj[off[a]:off[a + 1]]:
            s.add((a, int(b)))
    for a, b in zip(del_u, del_v):
        s.discard((int(a), int(b)))
        s.discard((int(b), int(a)))
    for a, b in zip(ins_u, ins_v):
        if a != b:
            s.add((int(a), int(b)))
            s.add((int(b), int(a)))
    lists = [[] for _ in range(n)]
    for a, b in s:
        lists[a].append(b)
    return csr_from_lists([sorted(l) for l in lists])


def apply_batch_sorted(n, off, adj, u, v, is_delete):
    """apply_batch_csr on sorted int64 keys (row << 32 | column), for graphs
    too large for Python sets: removes the present pairs of the deletions,
    then adds the absent pairs of the insertions (self-loops ignored), both
    directions of each.  Returns (off, adj, inserted, deleted) with the
    numbers of undirected pairs actually inserted and deleted."""
    off = np.asarray(off, np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    keys = (rows << 32) | np.asarray(adj, np.int64)
    del rows
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    d = np.asarray(is_delete, bool) if is_delete is not None else np.zeros(len(u), bool)

    def both(a, b):
        keep = a != b
        a, b = a[keep], b[keep]
        return np.unique(np.concatenate([(a << 32) | b, (b << 32) | a]))

    def present(k):
        if len(keys) == 0 or len(k) == 0:
            return np.zeros(len(k), bool), np.zeros(len(k), np.int64)
        pos = np.searchsorted(keys, k)
        hit = pos < len(keys)
        hit[hit] = keys[pos[hit]] == k[hit]
        return hit, pos

    dk = both(u[d], v[d])
    hit, pos = present(dk)
    if hit.any():
        keep = np.ones(len(keys), bool)
        keep[pos[hit]] = False
        keys = keys[keep]
    deleted = int(hit.sum()) // 2
    ik = both(u[~d], v[~d])
    hit, _ = present(ik)
    ik = ik[~hit]
    if len(ik):
        keys = np.insert(keys, np.searchsorted(keys, ik), ik)
    inserted = len(ik) // 2
    new_off = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(keys >> 32, minlength=n), out=new_off[1:])
    return new_off, (keys & 0xFFFFFFFF).astype(np.int32), inserted, deleted


# -------------------------------------------------------------- load balance
def lb_partition(work):
    work = np.ascontiguousarray(work, np.int64)
    n = len(work)
    prefix = np.zeros(n + 1, np.int64)
    parts = np.zeros(1000, np.int64)
    P = lib().tco_lb_partition(n, work, prefix, parts)
    return prefix, parts[:P].copy()


def _nz(a):
    """ndpointer rejects empty arrays' null data pointers on some numpy builds."""
    a = np.ascontiguousarray(a)
    if a.size == 0:
        return np.zeros(1, a.dtype)
    return a
