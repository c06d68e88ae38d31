"""CPU oracle for the strategy cost model (TEST INFRASTRUCTURE ONLY).

A 4 -> 32 -> 32 -> 3 ReLU MLP with the parameter layout of the device
kernels: W1[32][4], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]; the loss
is the mean over samples and outputs of (y - t)^2.  Float64 throughout; the
gradient is pinned by central finite differences in tests/test_oracle_mlp.py.
Also restates the BF16 table inference (round-to-nearest-even weights and
activations, float accumulation) and the Adam update.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 4, 32, 3
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
         ("W3", (OUT, HID)), ("b3", (OUT,))]
NP = sum(int(np.prod(s)) for _, s in SIZES)
TABLE_DIM = 32


def unpack(p):
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], np.float64).reshape(shape)
        k += n
    return out


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("W"):
            fan_in = shape[1]
            parts.append(rng.normal(0, np.sqrt(2.0 / fan_in), shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def forward(p, x):
    P = unpack(p)
    x = np.asarray(x, np.float64)
    a1 = x @ P["W1"].T + P["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ P["W3"].T + P["b3"]
    return y, (x, a1, h1, a2, h2)


def loss_and_grad(p, x, t):
    y, (x, a1, h1, a2, h2) = forward(p, x)
    t = np.asarray(t, np.float64)
    n = len(x)
    e = y - t
    loss = float((e ** 2).sum() / (n * OUT))
    dy = 2 * e / (n * OUT)
    P = unpack(p)
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    d2 = (dy @ P["W3"]) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ P["W2"]) * (a1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam_step(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and back."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def cell_features(small_max=64, mid_max=512):
    f = np.zeros((TABLE_DIM * TABLE_DIM, IN), np.float32)
    for i in range(TABLE_DIM):
        for j in range(TABLE_DIM):
            fl = 0.0 if i == 0 else i - 0.5
            fp = 0.0 if j == 0 else j - 0.5
            L = 2.0 ** fl
            b = 0 if L <= small_max else (1 if L <= mid_max else 2)
            f[i * TABLE_DIM + j] = [fl / 16, fp / 16, (fp - fl) / 16, b / 2]
    return f


def table_bf16(p, small_max=64, mid_max=512):
    """Arg-min strategy per (owner, task) log2 bucket with BF16 weights and
    activations and float32 accumulation, as the device table kernel does.
    Returns (table, predictions) so callers can skip near-ties."""
    P = {k: bf16(v.astype(np.float32)) for k, v in unpack(p).items()}
    x = bf16(cell_features(small_max, mid_max))
    h1 = bf16(np.maximum(x @ P["W1"].T.astype(np.float32) + P["b1"], 0))
    h2 = bf16(np.maximum(h1 @ P["W2"].T.astype(np.float32) + P["b2"], 0))
    y = h2 @ P["W3"].T.astype(np.float32) + P["b3"]
    return np.argmin(y, axis=1).astype(np.uint8), y
