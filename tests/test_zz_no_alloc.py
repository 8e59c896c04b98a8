"""GPU test of the allocation-free forward pass conv2d_forward_static: its transformed input
lives in the library's static device array, not in the caller's transform_pin. Its output
must equal conv2d_forward's bit for bit (same kernels, same order), on two calls in a row with
the second one smaller (the array keeps the first call's contents), on the staged Winograd
form and the general-geometry path; shapes whose transformed input exceeds the array are
refused before any launch."""
import ctypes
import importlib

import pytest
import torch

pytestmark = pytest.mark.gpu

wc = importlib.import_module("2024pra_conv_amd")
runner = importlib.import_module("2024pra_conv_amd.runner")


def _static_fn(L):
    return L.conv2d_forward_static, L.conv2d_static_scratch_bytes()


def _conv(cuda, n, c, h, w, k, r=3, s=3, stride=1, pad=1, seed=0):
    gen = torch.Generator(device=cuda)
    gen.manual_seed(seed)
    x = (torch.rand((n, c, h, w), generator=gen, device=cuda) * 2 - 1).half()
    wt = (torch.rand((k, c, r, s), generator=gen, device=cuda) * 2 - 1).half()
    return runner.WinogradConv(x, wt, pad=(pad, pad), stride=(stride, stride))


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("pair", [
    # staged Winograd form (configuration forced), then a smaller ragged layer
    (dict(n=4, c=64, h=28, w=28, k=128), dict(n=1, c=40, h=13, w=11, k=20), 9),
    # general-geometry path (5x5 stride 2, then 3x3 stride 2 over a smaller input)
    (dict(n=2, c=16, h=20, w=20, k=24, r=5, s=5, stride=2, pad=2), dict(n=1, c=8, h=9, w=9, k=12, stride=2), 0),
    # 1x1 filters, whose GEMM reads the input itself (8-byte row pieces, then the per-channel
    # gather of a strided layer)
    (dict(n=2, c=24, h=8, w=12, k=40, r=1, s=1, pad=0), dict(n=1, c=13, h=9, w=9, k=8, r=1, s=1, stride=2, pad=0), 1),
], ids=["staged", "direct", "pointwise"])
def test_static_forward_equals_forward(pair, cuda):
    L = wc.load_library()
    fwd_static, cap = _static_fn(L)
    L.conv2d_tuned_clear()
    big, small, cfg = pair
    for i, shape in enumerate((big, small)):
        print(f"case static {shape} cfg {cfg}", flush=True)
        conv = _conv(cuda, seed=10 + i, **shape)
        assert conv.v_bytes <= cap
        conv.prepare_weights()
        conv.out.fill_(float("nan"))
        conv.forward(cfg)
        torch.cuda.synchronize()
        want = conv.out.clone()
        conv.out.fill_(float("nan"))
        conv.V.fill_(0x55)  # the static form must not read transform_pin
        assert fwd_static(ctypes.byref(conv.pb), cfg, _stream()) == wc.OK
        torch.cuda.synchronize()
        assert torch.equal(conv.out.view(torch.int16), want.view(torch.int16)), f"call {i}"


def test_static_forward_fused_form_and_limits(cuda):
    L = wc.load_library()
    fwd_static, cap = _static_fn(L)
    L.conv2d_tuned_clear()
    # the library's default form for this layer is the fused one (no scratch): same output as forward(-1)
    conv = _conv(cuda, 2, 32, 16, 16, 64, seed=3)
    assert conv.form() == wc.FORM_FUSED
    conv.prepare_weights()
    conv.forward(-1)
    torch.cuda.synchronize()
    want = conv.out.clone()
    conv.out.fill_(float("nan"))
    assert fwd_static(ctypes.byref(conv.pb), -1, _stream()) == wc.OK
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), want.view(torch.int16))
    # a staged layer whose V exceeds the array is refused, and nothing is written
    pb = wc.make_problem(32, 256, 56, 56, 256, ptrs={"in_": conv.x.data_ptr(), "weight": conv.w.data_ptr(),
                                                      "out": conv.out.data_ptr(),
                                                      "transform_filter": conv.U.data_ptr()})
    v_bytes, _, _ = wc.workspace_size(pb)
    assert v_bytes > cap
    before = conv.out.clone()
    assert fwd_static(ctypes.byref(pb), 9, _stream()) == wc.ERR_WORKSPACE
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), before.view(torch.int16))
    assert fwd_static(None, -1, _stream()) == wc.ERR_NULL
