"""Pins the perceptron oracle (oracle/mlp.py) by central finite differences."""
import numpy as np

from oracle import mlp


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    F, H, B = 7, 5, 11
    P = F * H + H + H * H + H + H + 1
    p = rng.normal(size=P) * 0.5
    X = rng.normal(size=(B, F))
    y = rng.normal(size=B)
    _, g = mlp.loss_grad(p, F, H, X, y)
    eps = 1e-6
    num = np.zeros(P)
    for i in range(P):
        d = np.zeros(P)
        d[i] = eps
        num[i] = (mlp.loss(p + d, F, H, X, y) - mlp.loss(p - d, F, H, X, y)) / (2 * eps)
    assert np.max(np.abs(num - g)) < 1e-7


def test_oracle_forward_shapes_and_relu():
    F, H = 3, 4
    P = F * H + H + H * H + H + H + 1
    p = np.zeros(P)
    p[-1] = 2.5  # b3
    out, (h1, h2) = mlp.forward(p, F, H, np.ones((2, F)))
    assert out.shape == (2,) and np.all(out == 2.5) and np.all(h1 == 0) and np.all(h2 == 0)
