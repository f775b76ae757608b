"""CPU restatement (FP64 numpy) of the fusion cost model's perceptron.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the reference the
GPU forward / backward kernels in ``csrc/mlp.hip`` are tested against.  No reference
counterpart exists; the oracle itself is pinned by central finite differences
(tests/test_oracle_mlp.py).

Network: x -> relu(x W1 + b1) -> relu(h1 W2 + b2) -> h2 W3 + b3; loss = mean
squared error.  Parameter vector: W1 [F x H] row-major, b1 [H], W2 [H x H],
b2 [H], W3 [H], b3 [1].
"""
import numpy as np


def unpack(p, F, H):
    p = np.asarray(p, dtype=np.float64)
    i = 0
    W1 = p[i:i + F * H].reshape(F, H); i += F * H
    b1 = p[i:i + H]; i += H
    W2 = p[i:i + H * H].reshape(H, H); i += H * H
    b2 = p[i:i + H]; i += H
    W3 = p[i:i + H]; i += H
    b3 = p[i]; i += 1
    assert i == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, F, H, X):
    W1, b1, W2, b2, W3, b3 = unpack(p, F, H)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ W3 + b3, (h1, h2)


def loss(p, F, H, X, y):
    out, _ = forward(p, F, H, X)
    return float(np.mean((out - y) ** 2))


def loss_grad(p, F, H, X, y):
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, F, H)
    out, (h1, h2) = forward(p, F, H, X)
    B = X.shape[0]
    dout = 2.0 * (out - y) / B
    gW3 = h2.T @ dout
    gb3 = dout.sum()
    d2 = np.outer(dout, W3) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.reshape(-1), gb1, gW2.reshape(-1), gb2, gW3, [gb3]])
    return float(np.mean((out - y) ** 2)), g
