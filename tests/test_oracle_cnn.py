"""Pins the FP64 numpy CNN oracle: finite-difference gradients and the
periodic-convolution identity (a constant field maps to bias + Σ_o W[o]·c)."""
import numpy as np

from oracle import oracle_cnn as oc


def small_problem(seed=0, B=1, M=4):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(oc.nparams()) * 0.05
    x = rng.standard_normal((B, M, M, M, 3))
    t = rng.standard_normal((B, M, M, M, 6)) * 0.1
    return p, x, t


def test_constant_field_identity():
    p, _, _ = small_problem()
    W, b = oc.unpack(p)[0]
    x = np.ones((1, 4, 4, 4, 3)) * 0.7
    out = oc.conv(x, W, b)
    want = b + 0.7 * W.sum(axis=(0, 2))
    assert np.abs(out - want).max() < 1e-12


def test_shift_equivariance():
    p, x, _ = small_problem(1)
    y, _, _ = oc.forward(p, x)
    ys, _, _ = oc.forward(p, np.roll(x, 1, axis=3))
    assert np.abs(np.roll(y, 1, axis=3) - ys).max() < 1e-12


def test_gradient_finite_differences():
    p, x, t = small_problem(2)
    loss, g = oc.loss_grad(p, x, t)
    rng = np.random.default_rng(3)
    idx = rng.choice(len(p), 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = oc.loss_grad(p + e, x, t)
        lm, _ = oc.loss_grad(p - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
