"""The learned selector on the GPU: forward/backward kernels against the CPU
oracle, BF16 table against its restatement, counts unchanged by the policy."""
from __future__ import annotations

import numpy as np
import pytest

import custingeralg_amd as T
from custingeralg_amd import learned as Lm
import oracle as O
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, size=(n, 4)).astype(np.float32)
    t = np.stack([x[:, 0] + x[:, 1], 2 * x[:, 2] - x[:, 3], x[:, 3] * x[:, 0]], 1).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    x, _ = _data(1000, 1)
    p = Lm.init_params(2)
    model = Lm.CostModel(p)
    got = model.predict(x)
    want, _ = M.forward(p.astype(np.float64), x)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_gradient_matches_oracle(n):
    """FP32 device gradient vs the float64 oracle: |g - g*| <= 1e-5 + 1e-3 |g*|."""
    x, t = _data(n, n)
    p = Lm.init_params(4)
    model = Lm.CostModel(p)
    grads = T.DeviceArray(T.MLP_PARAMS, np.float32)
    fd, td = T.DeviceArray.from_host(x), T.DeviceArray.from_host(t)
    loss = model.train_step(fd, td, n, 0.0, adam=False, grads=grads)
    want_loss, want_g = M.loss_and_grad(p.astype(np.float64), x, t)
    got = grads.to_host()
    assert abs(loss - want_loss) <= 1e-5 + 1e-4 * want_loss
    assert np.all(np.abs(got - want_g) <= 1e-5 + 1e-3 * np.abs(want_g))
    # lr = 0 leaves the parameters untouched
    assert np.array_equal(model.host_params(), p)


def test_adam_step_matches_oracle():
    x, t = _data(300, 7)
    p = Lm.init_params(6)
    model = Lm.CostModel(p)
    fd, td = T.DeviceArray.from_host(x), T.DeviceArray.from_host(t)
    model.train_step(fd, td, len(x), 0.01)
    _, g = M.loss_and_grad(p.astype(np.float64), x, t)
    want, _, _ = M.adam_step(p.astype(np.float64), np.zeros(M.NP), np.zeros(M.NP), g, 1, 0.01)
    assert np.allclose(model.host_params(), want, rtol=1e-4, atol=1e-6)


def test_training_reduces_loss():
    x, t = _data(4096, 3)
    model = Lm.CostModel(seed=1)
    losses = model.fit(x, t, steps=200, lr=0.01)
    assert losses[-1] < 0.1 * losses[0]


def test_bf16_table_matches_restatement():
    p = Lm.init_params(9)
    table = Lm.CostModel(p).load()
    want, y = M.table_bf16(p)
    ys = np.sort(y, axis=1)
    clear = (ys[:, 1] - ys[:, 0]) > 1e-3  # skip near-ties (summation order)
    assert np.array_equal(table.ravel()[clear], want[clear])
    Lm.unload()


def test_timings_training_and_counts_unchanged():
    off, adj = O.rmat_graph(14, 16, 2)
    g = T.cuStinger(off, adj)
    feat, cost = Lm.collect_timings(g, 2000, 5)
    assert len(feat) > 500 and np.all(cost > 0)
    assert np.all((feat >= -2) & (feat <= 2))
    model, table, losses, _ = Lm.train_selector([g], samples_per_graph=2000, steps=150)
    assert losses[-1] < losses[0]
    assert set(np.unique(table)) <= {0, 1, 2}
    ref = O.custinger_count(off, adj)
    for pol in (T.POLICY_FIXED, T.POLICY_LEARNED):
        pv, glob, _ = T.static_count(g, T.default_schedule(policy=pol))
        assert np.array_equal(pv, ref)
        assert glob == int(ref.astype(np.uint64).sum())
    res = Lm.compare_policies(g, reps=2)
    assert res["fixed"]["ms_total"] > 0 and res["learned"]["ms_total"] > 0
    Lm.unload()
    g.close()


def test_learned_without_model_falls_back():
    Lm.unload()
    off, adj = O.rmat_graph(10, 8, 2)
    g = T.cuStinger(off, adj)
    pv, glob, st = T.static_count(g, T.default_schedule(policy=T.POLICY_LEARNED))
    assert np.array_equal(pv, O.custinger_count(off, adj))
    g.close()


def test_autotune_keeps_counts():
    off, adj = O.rmat_graph(12, 16, 3)
    g = T.cuStinger(off, adj)
    s, ms = Lm.autotune(g)
    assert ms > 0 and s.flush_items in (256, 1024, 4096)
    pv, glob, _ = T.static_count(g, T.default_schedule())
    assert np.array_equal(pv, O.custinger_count(off, adj))
    g.close()


def test_insitu_training_on_held_out_graph():
    """The selector trained in situ (whole counts with one group of table
    cells switched at a time) on R-MAT 18 and 20, then compared with the fixed
    heuristic on a held-out R-MAT 22 (another seed): equal counts, and the
    learned table's count kernels no more than 5% slower (best of 5 counts
    each; run-to-run spread on one box is about 3%)."""
    train = [T.rmat_graph_device(s, 16, 300 + s) for s in (18, 20)]
    try:
        model, table, losses, sweeps = Lm.train_insitu(train, reps=2, steps=300)
    finally:
        for g in train:
            g.close()
    assert losses[-1] < losses[0]
    assert set(np.unique(table)) <= {0, 1, 2}
    g = T.rmat_graph_device(22, 16, 3)
    try:
        res = Lm.compare_policies(g, reps=5)  # asserts equal counts
        # the table picks the count kernels' strategies: compare their time
        # (the orientation build is the same under both policies)
        fx, lr = (sum(res[k]["ms_bin"][:3]) for k in ("fixed", "learned"))
        print(f"held-out R-MAT 22: count bins fixed {fx:.2f} ms, learned {lr:.2f} ms "
              f"(whole counts {res['fixed']['ms_total']:.2f} / {res['learned']['ms_total']:.2f} ms)",
              flush=True)
        assert lr <= 1.05 * fx
    finally:
        g.close()
        Lm.unload()


def test_strategy_table_set_drives_the_policy():
    """A table given directly: every cell merge path, then every cell binary
    search; the counts stay the oracle's and the strategy counters follow the
    table in the small and mid bins."""
    off, adj = O.rmat_graph(13, 16, 8)
    ref = O.custinger_count(off, adj)
    g = T.cuStinger(off, adj)
    try:
        for s, key in ((3, "mergepath"), (1, "bsearch")):
            Lm.set_table(np.full((32, 32), s, np.uint8))
            pv, glob, st = T.static_count(g, T.default_schedule(policy=T.POLICY_LEARNED))
            assert np.array_equal(pv, ref)
            assert st["strategy_tasks"][key] > 0
        with pytest.raises(T.TCError):
            Lm.set_table(np.full((32, 32), 4, np.uint8))
    finally:
        Lm.unload()
        g.close()
