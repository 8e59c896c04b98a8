"""Pins the FP64 MLP oracle (CPU): analytic gradients against central finite differences,
Adam against its closed form on a one-parameter case, the feature layout against the
library's own feature function."""
import ctypes

import numpy as np

from oracle import mlp


def _data(seed, B=37):
    rng = np.random.default_rng(seed)
    X = np.concatenate([rng.uniform(0, 9, (B, 5)), np.eye(16)[rng.integers(0, 16, B)]], axis=1)
    T = rng.normal(3.0, 1.0, B)
    return X, T


def test_param_count():
    assert mlp.NPARAM == 5633


def test_analytic_gradient_matches_finite_differences():
    X, T = _data(1)
    p = mlp.init_params(3).astype(np.float64)
    _, g = mlp.loss_grad(p, X, T)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(mlp.NPARAM, 120, replace=False), [mlp.NPARAM - 1, 0, 64 * 21 + 5]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_first_step_closed_form():
    # first bias-corrected Adam step moves every parameter by lr * sign(g) (|g| >> eps)
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    p1, m, v = mlp.adam_step(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g), rtol=0, atol=1e-7)


def test_features_match_library(wc):
    L = wc.load_library()
    pb = wc.make_problem(4, 96, 27, 27, 128)
    out = (ctypes.c_float * 21)()
    assert L.conv2d_mlp_features(ctypes.byref(pb), 7, out) == wc.OK
    np.testing.assert_array_equal(np.array(out, dtype=np.float32), mlp.features(4, 96, 27, 27, 128, 7))
    assert L.conv2d_mlp_param_count() == mlp.NPARAM
    assert L.conv2d_mlp_features(ctypes.byref(pb), 16, out) == wc.ERR_INVALID
