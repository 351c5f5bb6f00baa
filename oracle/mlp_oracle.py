"""Float64 CPU reference of the planner MLP (test infrastructure only).

Same model as bigint_amd/csrc/planner_kernels.hip: x[3] -> 64 ReLU -> 64 ReLU -> K, parameters
W1[3][64] b1 W2[64][64] b2 W3[64][K] b3 in one vector; loss = sum(mask * (y - t)^2) / sum(mask).
Pinned against central finite differences in tests/test_planner.py.
"""
from __future__ import annotations

import numpy as np

H = 64


def unpack(p: np.ndarray, K: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in [(3, H), (H,), (H, H), (H,), (H, K), (K,)]:
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def forward(p, x, K):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    x = np.asarray(x, np.float64)
    a1 = x @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ W3 + b3
    return y, (x, a1, h1, a2, h2)


def loss_grad(p, x, t, mask, K):
    y, (x, a1, h1, a2, h2) = forward(p, x, K)
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    cnt = mask.sum()
    d = (y - t) * mask
    loss = float((d * (y - t)).sum() / cnt)
    dy = 2 * d / cnt
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dh2 = (dy @ W3.T) * (a2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
