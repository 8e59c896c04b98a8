"""GPU tests of the output epilogue (out = act(conv + bias[k]) in FP32 before the FP16
rounding), on the Winograd path (staged and fused forms) and the general-geometry path,
through the C-ABI.

Bar: |y - act(y* + b)| <= 2^-11 (|y* + b| + e) + e + 2^-25 + 2^-23 |y* + b|, with y* and e the
FP64 model and FP32-accumulation bound of the plain convolution (as in the parity tests) and
the last term the bias addition's own FP32 rounding; ReLU is 1-Lipschitz, so it keeps the
bound. The four-slot protocol with the epilogue set in its kernel argument equals the stage
call bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import direct as dg
from oracle import winograd as wg

pytestmark = pytest.mark.gpu

CASES = [
    # kind, n, c, h, w, k, r, stride, pad
    ("winograd", 2, 40, 14, 22, 70, 3, 1, 1),
    ("winograd", 1, 64, 16, 16, 64, 3, 1, 1),
    ("direct", 1, 40, 11, 13, 33, 3, 2, 1),
    ("direct", 2, 5, 9, 8, 7, 5, 1, 2),
]


def _ids(c):
    return "x".join(map(str, c))


def _model(kind, x, w, n, c, h, wd, k, r, st, p, fused=False):
    if kind == "winograd":
        geo = wg.Geometry(n, c, h, wd, k, p, p)
        y, scale = (wg.fused_gpu_model if fused else wg.winograd_gpu_model)(x, w, geo)
        e = (c + 16) * 2.0 ** -24 * scale
    else:
        geo = dg.DirectGeometry(n, c, h, wd, k, r, r, st, st, p, p)
        y = dg.gemm_output_f64(w, x, geo)
        e = (32 * geo.nsteps + 16) * 2.0 ** -24 * dg.gemm_output_f64(w, x, geo, absolute=True)
    return y, e


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("act,with_bias", [(0, True), (1, True), (1, False)])
def test_epilogue_matches_model(case, act, with_bias, runner, wc, cuda):
    kind, n, c, h, wd, k, r, st, p = case
    print(f"case {case} act={act} bias={with_bias}", flush=True)
    x = wg.uniform_pm1(300 + c, (n, c, h, wd))
    w = wg.uniform_pm1(301 + k, (k, c, r, r))
    b = wg.uniform_pm1(302 + k, (k,))
    X = torch.from_numpy(x).to(cuda)
    W = torch.from_numpy(w).to(cuda)
    B = torch.from_numpy(b).to(cuda) if with_bias else None
    conv = runner.WinogradConv(X, W, pad=(p, p), stride=(st, st))
    assert conv.method() == (wc.METHOD_WINOGRAD if kind == "winograd" else wc.METHOD_DIRECT)
    conv.prepare_weights()
    fused = conv.form() == wc.FORM_FUSED
    # cfg -1: the form getkernelInfo uses; an explicit configuration: the staged form
    runs = [(-1, fused)] + ([(conv.selected_config(), False)] if fused else [])
    for cfg, is_fused in runs:
        print(f"  cfg {cfg} fused={is_fused}", flush=True)
        conv.out.fill_(float("nan"))
        conv.forward_epilogue(B, act, cfg=cfg)
        torch.cuda.synchronize()
        y = conv.out.cpu().numpy().astype(np.float64)
        y_model, e = _model(kind, x, w, n, c, h, wd, k, r, st, p, fused=is_fused)
        z = y_model + (b.astype(np.float64)[None, :, None, None] if with_bias else 0.0)
        want = np.maximum(z, 0.0) if act == wc.ACT_RELU else z
        tol = 2.0 ** -11 * (np.abs(z) + e) + e + 2.0 ** -25 + 2.0 ** -23 * np.abs(z)
        assert np.all(np.abs(y - want) <= tol), f"max excess {np.max(np.abs(y - want) - tol)}"
        if act == wc.ACT_RELU:
            assert np.all(y >= 0.0)
    conv.forward_epilogue(B, act)
    torch.cuda.synchronize()

    # the reference protocol with the epilogue in its kernel argument: same bits
    same = conv.out.clone()
    conv.set_slot_epilogue(B, act)
    conv.out.fill_(float("nan"))
    conv.launch_slots()
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), same.view(torch.int16))
