"""CPU oracle of the cost model's MLP (test infrastructure only).

numpy restatement, in float64, of the network the library trains on the GPU:
nin -> H -> H -> 1 with ReLU, parameters flattened as W1 [H x nin], b1 [H],
W2 [H x H], b2 [H], W3 [1 x H], b3 [1]; loss = mean squared error; Adam with
beta1 0.9, beta2 0.999, eps 1e-8 and bias correction.  The network has no
counterpart in the reference; its gradients are pinned by central finite
differences (tests/test_cost_model.py).
"""
from __future__ import annotations

import numpy as np


def unpack(p, nin, h):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + h * nin].reshape(h, nin); o += h * nin
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + h].reshape(1, h); o += h
    b3 = p[o:o + 1]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def param_count(nin, h):
    return h * nin + h + h * h + h + h + 1


def forward(p, X, nin, h):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, h)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1.T + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2.T + b2
    A2 = np.maximum(Z2, 0)
    y = (A2 @ W3.T + b3)[:, 0]
    return y, (X, Z1, A1, Z2, A2)


def loss(p, X, t, nin, h):
    y, _ = forward(p, X, nin, h)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(p, X, t, nin, h):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, h)
    y, (X, Z1, A1, Z2, A2) = forward(p, X, nin, h)
    B = X.shape[0]
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / B
    dW3 = dy[None, :] @ A2
    db3 = np.array([dy.sum()])
    dZ2 = np.outer(dy, W3[0]) * (Z2 > 0)
    dW2 = dZ2.T @ A1
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    dW1 = dZ1.T @ X
    db1 = dZ1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
