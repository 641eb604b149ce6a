"""FP64 numpy restatement of the cost model's network (test infrastructure).

Same packed parameter layout and loss as finufft_amd/csrc/mlp.hip:
h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3,
loss = mean_b 0.5 sum_j m_bj (y_bj - t_bj)^2. Pinned by central finite
differences of loss() in tests/test_oracle_pinning.py.
"""

from __future__ import annotations

import numpy as np

IN, H, OUT = 11, 32, 8
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)),
          ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def forward(p, x):
    w = unpack(p)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (z1, h1, z2, h2)


def loss(p, x, t, m):
    y, _ = forward(p, x)
    return float(np.mean(0.5 * np.sum(m * (y - t) ** 2, axis=1)))


def grad(p, x, t, m):
    """Gradient of loss() with respect to the packed parameters."""
    w = unpack(p)
    n = x.shape[0]
    y, (z1, h1, z2, h2) = forward(p, x)
    dy = m * (y - t) / n
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = dy @ w["W3"]
    dz2 = dh2 * (z2 > 0)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ w["W2"]
    dz1 = dh1 * (z1 > 0)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    return np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
