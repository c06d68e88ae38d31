"""Synthetic workload recipes shared by the benchmark and the tests.

The streaming workload is an R-MAT scale-22, edge-factor-16 base graph
(seed 3) and 20 update batches of 10^3, 10^4, 10^5 and 10^6 undirected pairs
(5 of each), in that order.  Every batch is 80% insertions, half drawn
uniformly and half from the same R-MAT distribution, and 20% deletions drawn
uniformly from the entries present before the batch.  Pairs are canonical
(min, max) and deduplicated; an insertion equal to a deletion of the same
batch is dropped.

The recipe takes its pair generators as arguments: the device generators of
this package (``tc_gen_er_pairs`` / ``tc_gen_rmat_pairs``) or any
bit-identical restatement of them, so the same batches can be rebuilt on a
host without a GPU.
"""
from __future__ import annotations

import numpy as np

STREAM_BASE = {"scale": 22, "edge_factor": 16, "seed": 3}
STREAM_SIZES = [10 ** 3] * 5 + [10 ** 4] * 5 + [10 ** 5] * 5 + [10 ** 6] * 5
STREAM_RNG_SEED = 3


def stream_batch_seed(i: int) -> int:
    """Generator seed of batch i (its R-MAT half uses seed + 1)."""
    return 5000 + 7 * i


def stream_batch(er_pairs, rmat_pairs, off, adj, nv: int, B: int, scale: int, seed: int, rng):
    """One update batch of about B undirected pairs against the graph (off, adj).

    er_pairs(n, m, seed) and rmat_pairs(scale, m, seed) return host int32
    arrays (u, v); rng is a numpy Generator that picks the deleted entries.
    Returns host arrays (u int32, v int32, is_delete uint8): insertions first,
    then deletions, each sorted by (u, v)."""
    n_ins = int(0.8 * B)
    n_del = B - n_ins
    nu = n_ins // 2
    du, dv = er_pairs(nv, nu, seed)
    ru, rv = rmat_pairs(scale, n_ins - nu, seed + 1)
    iu = np.concatenate([np.asarray(du, np.int32), np.asarray(ru, np.int32)])
    iv = np.concatenate([np.asarray(dv, np.int32), np.asarray(rv, np.int32)])
    e = rng.integers(0, len(adj), n_del)
    xu = (np.searchsorted(off, e, side="right") - 1).astype(np.int32)
    xv = adj[e]

    def canon(a, b):
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        keep = lo != hi
        return np.unique((lo[keep].astype(np.int64) << 32) | hi[keep].astype(np.int64))

    ki, kd = canon(iu, iv), canon(xu, xv)
    ki = ki[~np.isin(ki, kd)]
    keys = np.concatenate([ki, kd])
    u = (keys >> 32).astype(np.int32)
    v = (keys & 0xFFFFFFFF).astype(np.int32)
    dl = np.concatenate([np.zeros(len(ki), np.uint8), np.ones(len(kd), np.uint8)])
    return u, v, dl
