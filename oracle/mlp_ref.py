"""Float64 NumPy restatement of the cost-model MLP (forward, loss, gradient).

TEST INFRASTRUCTURE ONLY: the checker for libnq's FP32 MLP kernels.  Layout
and loss are those of include/nq.h: W1[64x24] b1[64] W2[64x64] b2[64] W3[64]
b3, ReLU, loss = mean((y - t)^2).  The gradient itself is pinned by central
finite differences (tests/test_mlp.py).
"""
import numpy as np

H, I = 64, 24
NPARAM = H * I + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * I].reshape(H, I); o += H * I
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3, h1, h2


def loss(p, x, t):
    y, _, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_grad(p, x, t):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y, h1, h2 = forward(p, x)
    m = x.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    dz3 = 2.0 * e / m
    gW3 = h2.T @ dz3
    gb3 = dz3.sum()
    dz2 = np.outer(dz3, W3) * (h2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return float(np.mean(e ** 2)), g
