"""Pins the CPU MLP oracle: gradients against central finite differences,
BF16 rounding against known values, Adam against its closed form."""
from __future__ import annotations

import numpy as np

from oracle import mlp as M


def test_layout_size():
    assert M.NP == 4 * 32 + 32 + 32 * 32 + 32 + 3 * 32 + 3 == 1315


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(3).astype(np.float64)
    x = rng.normal(size=(40, 4))
    t = rng.normal(size=(40, 3))
    _, g = M.loss_and_grad(p, x, t)
    h = 1e-6
    idx = rng.choice(M.NP, 60, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (M.loss_and_grad(pp, x, t)[0] - M.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), i


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.1415927], np.float32)
    r = M.bf16(v)
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1+2^-7: ties to even -> 1.0
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-7


def test_adam_first_step_is_sign_times_lr():
    p = np.zeros(5)
    g = np.array([1.0, -2.0, 0.5, 3.0, -0.1])
    p1, m, v = M.adam_step(p, np.zeros(5), np.zeros(5), g, 1, 0.01)
    assert np.allclose(p1, -0.01 * np.sign(g), atol=1e-7)


def test_training_reduces_loss_on_cpu():
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, size=(256, 4))
    t = np.stack([x[:, 0] + x[:, 1], 2 * x[:, 2], x[:, 3] - x[:, 0]], 1)
    p = M.init_params(5).astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0 = M.loss_and_grad(p, x, t)[0]
    for s in range(1, 201):
        _, g = M.loss_and_grad(p, x, t)
        p, m, v = M.adam_step(p, m, v, g, s, 0.01)
    assert M.loss_and_grad(p, x, t)[0] < 0.1 * l0
