"""The cost model's GPU forward / backward / training kernels against the
FP64 numpy oracle (FP32 kernels: relative tolerance 1e-4)."""
import numpy as np
import pytest

from dmsim import native
from oracle import mlp

pytestmark = pytest.mark.gpu

F, H = 24, 64


def data(seed, B=300):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, F)).astype(np.float32)
    w = rng.normal(size=F)
    y = (np.tanh(X @ w / 4) + 0.1 * X[:, 0] ** 2).astype(np.float32)
    return X, y


def test_param_count_and_init():
    P = native.mlp_param_count(F, H)
    assert P == F * H + H + H * H + H + H + 1
    p = native.mlp_init(F, H, 7)
    assert p.shape == (P,) and np.isfinite(p).all() and np.std(p[:F * H]) > 0.1


def test_forward_matches_oracle():
    X, _ = data(1)
    p = native.mlp_init(F, H, 3)
    got = native.mlp_predict(p, F, H, X)
    want, _ = mlp.forward(p.astype(np.float64), F, H, X.astype(np.float64))
    assert np.max(np.abs(got - want)) <= 1e-4 * (1 + np.max(np.abs(want)))


def test_gradient_matches_oracle():
    X, y = data(2, 97)
    p = native.mlp_init(F, H, 5)
    loss, g = native.mlp_loss_grad(p, F, H, X, y)
    lo, go = mlp.loss_grad(p.astype(np.float64), F, H, X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.max(np.abs(g - go)) <= 1e-4 * np.max(np.abs(go))


def test_training_reduces_loss():
    X, y = data(3, 512)
    p0 = native.mlp_init(F, H, 11)
    l0 = mlp.loss(p0.astype(np.float64), F, H, X.astype(np.float64), y.astype(np.float64))
    p1, curve = native.mlp_train(p0, F, H, X, y, 60, 64, 3e-3, 1)
    l1 = mlp.loss(p1.astype(np.float64), F, H, X.astype(np.float64), y.astype(np.float64))
    assert l1 < 0.2 * l0
    assert curve[-1] < curve[0]
