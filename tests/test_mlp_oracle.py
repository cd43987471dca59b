"""Pin the emulator's CPU oracle (oracle/mlp_oracle.py) before the GPU
kernels are compared with it.

  loss_grad   every gradient entry of a small random network against
              central differences of the oracle's own loss
  forward     the output tangent dY along an input direction dX against
              central differences (Y(X + h dX) - Y(X - h dX)) / 2h
Both checks fail on a planted mistake (a wrong SiLU derivative).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def rand_weights(seed):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in M.SIZES:
        fan = shape[1] if len(shape) == 2 else 1
        p[name] = rng.uniform(-1, 1, shape) * (np.sqrt(3.0 / fan) if len(shape) == 2 else 0.1)
    return M.pack(p)


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.uniform(0, np.pi / 2, n)
    X = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), np.cos(psi), np.sin(psi)], 1)
    dX = np.stack([rng.normal(size=n), rng.normal(size=n), -np.sin(psi), np.cos(psi)], 1)
    return X, dX


def tangent_error(w, X, dX, h=1e-5):
    _, dy = M.forward(w, X, dX)
    fd = (M.forward(w, X + h * dX)[0] - M.forward(w, X - h * dX)[0]) / (2 * h)
    return np.max(np.abs(dy - fd)) / np.abs(fd).max()


def test_forward_tangent_matches_central_differences():
    w = rand_weights(11)
    X, dX = inputs(64, 12)
    assert tangent_error(w, X, dX) < 1e-8


def test_loss_gradient_matches_central_differences():
    w = rand_weights(13)
    X, _ = inputs(16, 14)
    Y = np.random.default_rng(15).normal(0, 0.3, (16, 2))
    _, g = M.loss_grad(w, X, Y)
    rng = np.random.default_rng(16)
    # a sample of entries from every layer's weights and biases
    idx, o = [], 0
    for _, shape in M.SIZES:
        n = int(np.prod(shape))
        idx += list(o + rng.choice(n, min(n, 12), replace=False))
        o += n
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAMS)
        e[i] = h
        fd = (M.loss_grad(w + e, X, Y)[0] - M.loss_grad(w - e, X, Y)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-10), i


def test_checks_catch_a_wrong_silu_derivative(monkeypatch):
    w = rand_weights(17)
    X, dX = inputs(32, 18)
    Y = np.random.default_rng(19).normal(0, 0.3, (32, 2))
    real = M.dsilu
    # the derivative of z * sigmoid(z) without its z * s * (1 - s) term
    monkeypatch.setattr(M, "dsilu", lambda z: 1 / (1 + np.exp(-z)))
    assert tangent_error(w, X, dX) > 1e-3
    _, g = M.loss_grad(w, X, Y)
    i = M.NPARAMS - 2 - 2 * 128 - 5  # a b3 entry, behind one SiLU derivative
    e = np.zeros(M.NPARAMS)
    e[i] = 1e-6
    # the loss itself does not use the derivative: its differences stay right
    fd = (M.loss_grad(w + e, X, Y)[0] - M.loss_grad(w - e, X, Y)[0]) / 2e-6
    assert abs(fd - g[i]) > 1e-2 * abs(fd), (fd, g[i])
    monkeypatch.setattr(M, "dsilu", real)
