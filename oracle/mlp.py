"""Oracle (test infrastructure only) for the rank predictor's network: a numpy restatement of the
forward pass, the mean-squared-error gradient, one Adam step and the BF16 inference of the
device kernels (hcorepp_amd/csrc/mlp.hip, learned.hpp).  The network has no counterpart in the
reference; these functions are pinned by finite differences and closed forms in tests/test_oracle.py.

Parameter layout: W1 [F][H], b1 [H], W2 [H][H], b2 [H], w3 [H], b3 (row-major, FP32).
"""
from __future__ import annotations

import numpy as np

F, H = 24, 64


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ w3 + b3, (X, a1, h1, a2, h2)


def loss_grad(p, X, y):
    """mean((out - y)^2) and its gradient in the packed layout."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    out, (X, a1, h1, a2, h2) = forward(p, X)
    n = X.shape[0]
    e = out - y
    loss = float(np.mean(e * e))
    dout = 2.0 * e / n
    gw3 = h2.T @ dout
    gb3 = dout.sum()
    d2 = np.outer(dout, w3) * (a2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def predict_bf16(p, X):
    """The in-kernel inference: BF16 weights and activations, FP32 accumulation."""
    W1, b1, W2, b2, w3, b3 = [bf16(np.asarray(t, np.float32)) for t in unpack(np.asarray(p, np.float32))]
    Xb = bf16(np.asarray(X, np.float32)).astype(np.float64)
    h1 = bf16(np.maximum(Xb @ W1.astype(np.float64) + b1, 0.0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ W2.astype(np.float64) + b2, 0.0).astype(np.float32)).astype(np.float64)
    return h2 @ w3.astype(np.float64) + float(b3)
