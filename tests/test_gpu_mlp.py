"""GPU tests of the learned cost model's HIP kernels against the FP64 oracle (oracle/mlp.py)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import mlp

pytestmark = pytest.mark.gpu


def _data(seed, B):
    rng = np.random.default_rng(seed)
    X = np.concatenate([rng.uniform(0, 9, (B, 5)), np.eye(16)[rng.integers(0, 16, B)]], axis=1).astype(np.float32)
    T = rng.normal(3.0, 1.0, B).astype(np.float32)
    return X, T


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_forward_and_gradient_match_oracle(wc, cuda):
    L = wc.load_library()
    X, T = _data(1, 53)
    p = mlp.init_params(2)
    P = torch.from_numpy(p).to(cuda)
    Xd, Td = torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda)
    y = torch.empty(53, dtype=torch.float32, device=cuda)
    assert L.conv2d_mlp_forward(_ptr(P), _ptr(Xd), 53, _ptr(y), _stream()) == wc.OK
    torch.cuda.synchronize()
    _, _, y_ref = mlp.forward(p, X)
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-4)
    G = torch.empty(mlp.NPARAM, dtype=torch.float32, device=cuda)
    loss = torch.empty(1, dtype=torch.float32, device=cuda)
    assert L.conv2d_mlp_loss_grad(_ptr(P), _ptr(Xd), _ptr(Td), 53, _ptr(G), _ptr(loss), _stream()) == wc.OK
    l_ref, g_ref = mlp.loss_grad(p, X, T)
    g = G.cpu().numpy().astype(np.float64)
    assert abs(loss.item() - l_ref) <= 1e-4 * max(1.0, l_ref)
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * scale, np.max(np.abs(g - g_ref)) / scale
    # every block of the layout is exercised (non-zero gradient somewhere in each)
    o = 0
    for name, shape in mlp.SIZES:
        n = int(np.prod(shape))
        assert np.any(g_ref[o:o + n] != 0), name
        o += n


def test_adam_step_matches_oracle(wc, cuda):
    L = wc.load_library()
    rng = np.random.default_rng(3)
    p = rng.normal(size=mlp.NPARAM).astype(np.float32)
    g = rng.normal(size=mlp.NPARAM).astype(np.float32)
    m = rng.normal(size=mlp.NPARAM).astype(np.float32) * 0.1
    v = np.abs(rng.normal(size=mlp.NPARAM)).astype(np.float32) * 0.1
    P, G, M, V = (torch.from_numpy(a.copy()).to(cuda) for a in (p, g, m, v))
    assert L.conv2d_mlp_adam_step(_ptr(P), _ptr(G), _ptr(M), _ptr(V), 5, ctypes.c_float(1e-3), _stream()) == wc.OK
    torch.cuda.synchronize()
    p1, m1, v1 = mlp.adam_step(p.astype(np.float64), g.astype(np.float64), m.astype(np.float64), v.astype(np.float64), 5, 1e-3)
    np.testing.assert_allclose(P.cpu().numpy(), p1, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(M.cpu().numpy(), m1, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(V.cpu().numpy(), v1, rtol=1e-5, atol=1e-7)


def test_training_fits_a_smooth_target_and_select_is_argmin(wc, cuda):
    L = wc.load_library()
    B = 2048
    rng = np.random.default_rng(5)
    X = np.concatenate([rng.uniform(0, 9, (B, 5)), np.eye(16)[rng.integers(0, 16, B)]], axis=1).astype(np.float32)
    T = (0.3 * X[:, 0] + 0.2 * X[:, 1] + 0.25 * X[:, 4] + 0.5 * np.argmax(X[:, 5:], axis=1) / 16).astype(np.float32)
    p = mlp.init_params(7)
    P = torch.from_numpy(p).to(cuda)
    Mm = torch.zeros_like(P)
    Vv = torch.zeros_like(P)
    Xd, Td = torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda)
    l0 = mlp.loss(p, X, T)
    lf = ctypes.c_float()
    assert L.conv2d_mlp_train(_ptr(P), _ptr(Mm), _ptr(Vv), 0, _ptr(Xd), _ptr(Td), B, 400, ctypes.c_float(3e-3),
                              ctypes.byref(lf), _stream()) == wc.OK
    assert lf.value < 0.02 * l0, (l0, lf.value)
    # the reported loss is the oracle's loss of the trained parameters
    assert abs(mlp.loss(P.cpu().numpy(), X, T) - lf.value) <= 1e-3 * max(1.0, lf.value) + 1e-5
    pb = wc.make_problem(4, 96, 27, 27, 128)
    best = ctypes.c_int()
    pred = (ctypes.c_float * L.conv2d_num_tile_configs())()
    assert L.conv2d_mlp_select(_ptr(P), ctypes.byref(pb), ctypes.byref(best), pred, _stream()) == wc.OK
    preds = np.array(pred)
    assert best.value == int(np.argmin(preds))
    feats = np.stack([mlp.features(4, 96, 27, 27, 128, c) for c in range(len(preds))])
    _, _, y_ref = mlp.forward(P.cpu().numpy(), feats)
    np.testing.assert_allclose(np.log(preds), y_ref, rtol=1e-4, atol=1e-4)
