"""Learned intersection-strategy selector and schedule autotuning.

The cost model is a 4 -> 32 -> 32 -> 3 MLP over per-task features
(log2 |owner tile| / 16, log2 |task list| / 16, log2 ratio / 16, bin / 2)
predicting log2(cycles) of the merge, binary-search and hash strategies.
Timings come from the method itself (``collect_timings``), training runs on
the device (``tc_mlp_train_step``: FP32 forward/backward + Adam), and
``load`` rounds the weights to BF16 and fills the strategy table that
``POLICY_LEARNED`` reads.  Counts never depend on the strategy.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import (MLP_PARAMS, POLICY_FIXED, POLICY_LEARNED, DeviceArray, Schedule, check,
               cuStinger, lib, static_count)


def collect_timings(g: cuStinger, max_samples: int = 4096, seed: int = 1):
    """Per-task features (n, 4) and cycles per strategy (n, 3: merge, bsearch, hash)."""
    feat = np.zeros((max_samples, 4), np.float32)
    cost = np.zeros((max_samples, 3), np.float32)
    n = C.c_int64(0)
    check(lib().tc_collect_timings(g.h, max_samples, seed, feat.ctypes.data, cost.ctypes.data,
                                   C.byref(n)))
    return feat[:n.value].copy(), cost[:n.value].copy()


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases, in the device layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape, w in (((32, 4), True), ((32,), False), ((32, 32), True), ((32,), False),
                     ((3, 32), True), ((3,), False)):
        parts.append(rng.normal(0, np.sqrt(2.0 / shape[1]), shape).ravel() if w
                     else np.zeros(shape))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == MLP_PARAMS
    return p


class CostModel:
    """Device-trained cost model."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        self.params = DeviceArray.from_host(p)
        self.m = DeviceArray(MLP_PARAMS, np.float32).zero()
        self.v = DeviceArray(MLP_PARAMS, np.float32).zero()
        self.step = 0

    def host_params(self) -> np.ndarray:
        return self.params.to_host()

    def train_step(self, feat_dev: DeviceArray, target_dev: DeviceArray, n: int, lr: float,
                   adam: bool = True, grads: DeviceArray | None = None) -> float:
        self.step += 1
        loss = C.c_float(0)
        check(lib().tc_mlp_train_step(self.params.ptr, self.m.ptr if adam else None,
                                      self.v.ptr if adam else None, self.step, feat_dev.ptr,
                                      target_dev.ptr, n, lr, grads.ptr if grads else None,
                                      C.byref(loss)))
        return float(loss.value)

    def fit(self, feat: np.ndarray, target: np.ndarray, steps: int = 300, lr: float = 0.01):
        fd = DeviceArray.from_host(np.ascontiguousarray(feat, np.float32))
        td = DeviceArray.from_host(np.ascontiguousarray(target, np.float32))
        losses = [self.train_step(fd, td, len(feat), lr) for _ in range(steps)]
        return losses

    def predict(self, feat: np.ndarray) -> np.ndarray:
        fd = DeviceArray.from_host(np.ascontiguousarray(feat, np.float32))
        out = DeviceArray(len(feat) * 3, np.float32)
        check(lib().tc_mlp_forward(self.params.ptr, fd.ptr, len(feat), out.ptr))
        return out.to_host().reshape(-1, 3)

    def load(self) -> np.ndarray:
        """Make this model the learned policy; returns the strategy table."""
        p = self.host_params()
        check(lib().tc_model_load(p.ctypes.data))
        return model_table()

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump({"params": self.host_params().tolist(), "step": self.step}, f)

    @classmethod
    def from_file(cls, path: str) -> "CostModel":
        with open(path) as f:
            d = json.load(f)
        return cls(np.array(d["params"], np.float32))


def targets_from_cycles(cost: np.ndarray) -> np.ndarray:
    """Training target: log2(cycles) / 16 per strategy."""
    return (np.log2(np.maximum(cost, 1.0)) / 16.0).astype(np.float32)


def model_table() -> np.ndarray:
    dim = C.c_int32(0)
    check(lib().tc_model_table(None, C.byref(dim)))
    t = np.zeros(dim.value * dim.value, np.uint8)
    check(lib().tc_model_table(t.ctypes.data, C.byref(dim)))
    return t.reshape(dim.value, dim.value)


def unload() -> None:
    check(lib().tc_model_load(None))


def train_selector(graphs, samples_per_graph: int = 4096, steps: int = 400, lr: float = 0.01,
                   seed: int = 0):
    """Collect timings on each graph, train a model on the device and load it."""
    feats, costs = [], []
    for i, g in enumerate(graphs):
        f, c = collect_timings(g, samples_per_graph, seed + 17 * i + 1)
        feats.append(f)
        costs.append(c)
    feat = np.concatenate(feats)
    cost = np.concatenate(costs)
    model = CostModel(seed=seed)
    losses = model.fit(feat, targets_from_cycles(cost), steps, lr)
    table = model.load()
    return model, table, losses, (feat, cost)


def autotune(g: cuStinger):
    """Sweep bin thresholds and job grains; the best becomes the default schedule."""
    s = Schedule()
    ms = C.c_double(0)
    check(lib().tc_autotune(g.h, C.byref(s), C.byref(ms)))
    return s, float(ms.value)


def set_default_schedule(s: Schedule | None) -> None:
    """Make s the process default schedule (None restores the built-in one)."""
    check(lib().tc_schedule_set_default(C.byref(s) if s is not None else None))


_SCHED_KEYS = ("small_max", "mid_max", "job_items", "ratio_x16", "persistent", "flush_items")


def save_selector(path: str, model: CostModel, tuned: Schedule | None, meta: dict) -> None:
    """Write the trained cost model and the tuned schedule as JSON."""
    d = {"params": [float(x) for x in model.host_params()],
         "schedule": ({k: int(getattr(tuned, k)) for k in _SCHED_KEYS} if tuned is not None
                      else None),
         "meta": meta}
    with open(path, "w") as f:
        json.dump(d, f, indent=1)


def load_selector(path: str) -> dict:
    """Load a saved selector: its BF16 strategy table becomes the learned
    policy and its schedule the process default."""
    from . import default_schedule
    with open(path) as f:
        d = json.load(f)
    model = CostModel(np.array(d["params"], np.float32))
    table = model.load()
    sched = None
    if d.get("schedule"):
        set_default_schedule(None)
        sched = default_schedule()
        for k, v in d["schedule"].items():
            setattr(sched, k, int(v))
        set_default_schedule(sched)
    return {"model": model, "table": table, "schedule": sched, "meta": d.get("meta", {})}


def train_default(train_scales=(18, 20, 22), edge_factor: int = 16, samples_per_graph: int = 4096,
                  steps: int = 400, seed: int = 0):
    """Train the selector on R-MAT graphs of the given scales (generated on the
    device); returns (model, table, losses, number of timed tasks)."""
    from . import rmat_graph_device
    graphs = [rmat_graph_device(s, edge_factor, 100 + s) for s in train_scales]
    try:
        model, table, losses, (feat, _) = train_selector(graphs, samples_per_graph, steps,
                                                         seed=seed)
    finally:
        for g in graphs:
            g.close()
    return model, table, losses, len(feat)


def compare_policies(g: cuStinger, reps: int = 3):
    """Best-of-reps time of the fixed heuristic and the learned table on one graph."""
    from . import default_schedule
    out = {}
    ref = None
    for name, pol in (("fixed", POLICY_FIXED), ("learned", POLICY_LEARNED)):
        best = None
        for _ in range(reps):
            _, glob, st = static_count(g, default_schedule(policy=pol), per_vertex=False)
            if ref is None:
                ref = glob
            assert glob == ref, "policies disagree on the count"
            best = st if best is None or st["ms_total"] < best["ms_total"] else best
        out[name] = best
    return out


# ------------------------------------------------------------ in-situ training
# The timings above are one wavefront per sampled task and cannot see what a
# strategy costs inside a running count (lockstep lanes, per-job tiles,
# occupancy).  The in-situ path measures whole counts instead: starting from
# a table that encodes the fixed heuristic cell by cell, it switches one group
# of cells to another strategy, counts, and keeps the count's time relative to
# the unmodified table.  The model is then trained on the device on those
# relative costs, and its BF16 table is the learned policy.

# owner-tile buckets (row i = floor(log2 L) + 1) of the small and mid bins and
# task-list buckets (column j = floor(log2 pl) + 1) the sweep varies together
SWEEP_ROWS = ((1, 3), (4, 5), (6, 7), (8, 8), (9, 10))
SWEEP_COLS = ((1, 3), (4, 6), (7, 9), (10, 13), (14, 31))
TABLE_DIM = 32


def cell_features(small_max: int = 64, mid_max: int = 512) -> np.ndarray:
    """Features of every table cell (the device's cell_features): row i and
    column j stand for lengths 2^(i-0.5) and 2^(j-0.5)."""
    x = np.zeros((TABLE_DIM * TABLE_DIM, 4), np.float32)
    for i in range(TABLE_DIM):
        fl = 0.0 if i == 0 else i - 0.5
        L = 2.0 ** fl
        b = 0 if L <= small_max else (1 if L <= mid_max else 2)
        for j in range(TABLE_DIM):
            fp = 0.0 if j == 0 else j - 0.5
            x[i * TABLE_DIM + j] = (fl / 16, fp / 16, (fp - fl) / 16, b / 2)
    return x


def fixed_table() -> np.ndarray:
    """The fixed heuristic cell by cell (at the cell's representative
    lengths): binary search when L * log2(pl) * 4 < pl, else hash."""
    t = np.full((TABLE_DIM, TABLE_DIM), 2, np.uint8)
    for i in range(1, TABLE_DIM):
        L = 2.0 ** (i - 0.5)
        for j in range(1, TABLE_DIM):
            pl = 2.0 ** (j - 0.5)
            if L * j * 4 < pl:
                t[i, j] = 1
    return t


def set_table(table) -> None:
    t = np.ascontiguousarray(table, np.uint8)
    check(lib().tc_strategy_table_set(t.ctypes.data, TABLE_DIM))


def _best_ms(g: cuStinger, reps: int, out) -> tuple:
    from . import default_schedule
    best, glob = None, None
    for _ in range(reps):
        _, gl, st = static_count(g, default_schedule(policy=POLICY_LEARNED), out=out)
        glob = gl if glob is None else glob
        if gl != glob:
            raise RuntimeError("strategy tables disagree on the count")
        best = st["ms_total"] if best is None else min(best, st["ms_total"])
    return best, glob


def insitu_sweep(g: cuStinger, reps: int = 3, strategies=(0, 1, 2)):
    """Relative count time of every (cell group, strategy): the count with
    that group of cells switched to the strategy over the count with the
    fixed table (best of reps each).  Returns {(rows, cols): {s: rel}}."""
    base = fixed_table()
    out = DeviceArray(g.nv, np.uint32)
    try:
        set_table(base)
        t0, glob0 = _best_ms(g, reps, out)
        res = {}
        for rows in SWEEP_ROWS:
            for cols in SWEEP_COLS:
                rel = {}
                for st in strategies:
                    t = base.copy()
                    t[rows[0]:rows[1] + 1, cols[0]:cols[1] + 1] = st
                    if np.array_equal(t, base):
                        rel[st] = 1.0
                        continue
                    set_table(t)
                    ms, glob = _best_ms(g, reps, out)
                    if glob != glob0:
                        raise RuntimeError("a strategy table changed the count")
                    rel[st] = ms / t0
                res[(rows, cols)] = rel
        return res
    finally:
        out.free()
        unload()


def train_insitu(graphs, reps: int = 3, steps: int = 400, lr: float = 0.01, seed: int = 0,
                 min_gain: float = 0.01):
    """In-situ sweeps on each graph, then the model trained on the device on
    the relative costs and loaded.  Every cell's target is the fixed
    heuristic's strategy (the others 0.5 above it), except in a group where
    switching every cell to one strategy saved at least min_gain of the count
    on average: there that strategy is the target, the others 10x their
    measured extra cost above it (at least 0.5, wider than the network's fit
    error).  Returns (model, table, losses, per-graph sweeps)."""
    sweeps = [insitu_sweep(g, reps) for g in graphs]
    base = fixed_table()
    x = cell_features()
    tgt = np.full((TABLE_DIM * TABLE_DIM, 3), 0.5, np.float32)
    tgt[np.arange(TABLE_DIM * TABLE_DIM), base.ravel()] = 0.0
    for key in sweeps[0]:
        rows, cols = key
        rel = {st: float(np.mean([sw[key][st] for sw in sweeps])) for st in sweeps[0][key]}
        best = min(rel, key=rel.get)
        if rel[best] > 1.0 - min_gain:
            continue  # (the unmodified table is 1.0)
        for i in range(rows[0], rows[1] + 1):
            for j in range(cols[0], cols[1] + 1):
                c = i * TABLE_DIM + j
                for st, r in rel.items():
                    tgt[c, st] = 0.0 if st == best else max(10.0 * (r - rel[best]), 0.5)
    model = CostModel(seed=seed)
    losses = model.fit(x, tgt, steps, lr)
    table = model.load()
    return model, table, losses, sweeps


def default_model_path() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_model.json")
