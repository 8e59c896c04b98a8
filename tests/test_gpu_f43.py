"""GPU parity tests of the Winograd F(4x4,3x3) form (CONV2D_FORM_F43), through the C-ABI, against
the oracle (oracle/winograd43.py).

Bars
  S1 U, S2 V : bit-exact against the oracle (U: the exact FP64 transform rounded once to FP16;
               V: the kernel's FP32 operation order, rounded once), packed layouts and padding
               included, on poisoned workspaces.
  S3 M       : every GEMM configuration bit-identical; the stored FP16(M / 4) within the FP32
               accumulation bound of the oracle's FP64 GEMM on the same FP16 U and V:
               |m - M*/4| <= 2^-11 (|M*/4| + e) + e + 2^-25, e = (C + 16) 2^-24 (|U|.|V|) / 4.
  S4 Y       : within the FP32 bound of the exact output transform of the GPU's own M:
               |y - y*| <= 2^-11 (|y*| + e) + e + 2^-25, e = 32 2^-24 (|A'| 4|M| |A'|^T).
  end to end : normwise max relative error <= 1e-2 against the FP64 direct convolution; the four
               slots, conv2d_forward and conv2d_prepare_weights of a shape tuned to the form equal
               the stage calls bit for bit; NHWC input bit-identical to NCHW; the autotuner's
               F(4x4) timing and the U it leaves behind.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import cdirect
from oracle import winograd as wg
from oracle import winograd43 as w43

pytestmark = pytest.mark.gpu

SHAPES = [
    # n, c, h, w, k, p, q
    (1, 32, 16, 16, 64, 1, 1),     # whole tiles
    (2, 40, 14, 22, 70, 1, 1),     # ragged C, K and tiles
    (1, 3, 13, 10, 5, 2, 2),       # pad 2, C = 3, K = 5
    (1, 16, 9, 12, 8, 0, 0),       # pad 0
    (3, 64, 8, 64, 96, 1, 1),      # wide
    (1, 96, 30, 30, 130, 1, 0),    # K > 128, q = 0
    (2, 256, 12, 12, 64, 1, 1),    # CB = 8 (the 8-stage configuration's full ring)
    (1, 8, 4, 300, 16, 1, 1),      # one tile row of staging in 119 KiB of LDS
    (1, 33, 11, 7, 17, 1, 1),      # odd everything
]


def _ids(s):
    return "x".join(map(str, s))


def _log(msg):
    print(msg, flush=True)


def _dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _bits(t):
    return t.contiguous().view(torch.int16).cpu().numpy()


def _conv(runner, shape, cuda, seed=21):
    n, c, h, w, k, p, q = shape
    x = wg.uniform_pm1(seed, (n, c, h, w))
    wt = wg.uniform_pm1(seed + 1, (k, c, 3, 3))
    conv = runner.WinogradConv(_dev(x, cuda), _dev(wt, cuda), pad=(p, q))
    assert conv.M is not None and conv.m43_bytes > 0
    return x, wt, conv


def _poison(conv):
    conv.V.fill_(0xFF)
    conv.U.fill_(0xFF)
    conv.M.fill_(0xFF)
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_f43_stagewise(shape, wc, runner, cuda):
    n, c, h, w, k, p, q = shape
    _log(f"case f43 stagewise {_ids(shape)}")
    x, wt, conv = _conv(runner, shape, cuda)
    geo = w43.Geometry43(n, c, h, w, k, p, q)
    _poison(conv)
    # S1, S2: bit-exact
    conv.f43_prepare_weights()
    conv.f43_input_transform()
    torch.cuda.synchronize()
    u16 = w43.filter_transform43(wt)
    v16 = w43.input_transform43(x, geo)
    assert np.array_equal(conv.u43_halves().cpu().numpy().view(np.uint16), w43.pack_U43(u16, geo).view(np.uint16))
    assert np.array_equal(conv.v43_halves().cpu().numpy().view(np.uint16), w43.pack_V43(v16, geo).view(np.uint16))
    # S3: every configuration bit-identical; M within the FP32 accumulation bound
    ncfg = wc.load_library().conv2d_f43_num_tile_configs()
    first = None
    for cfg in range(ncfg):
        conv.f43_gemm(cfg)
        torch.cuda.synchronize()
        m = w43.unpack_M43(conv.m43_halves().cpu().numpy(), geo)
        if first is None:
            first = m.copy()
        else:
            assert np.array_equal(m.view(np.uint16), first.view(np.uint16)), cfg
    mstar = w43.gemm43_f64(u16, v16) * w43.M_SCALE
    e = (c + 16) * 2.0 ** -24 * w43.gemm43_f64(u16, v16, absolute=True) * w43.M_SCALE
    mg = first.astype(np.float64)
    assert np.isfinite(mg).all()
    err = np.abs(mg - mstar)
    bound = 2.0 ** -11 * (np.abs(mstar) + e) + e + 2.0 ** -25
    assert (err <= bound).all(), float((err / bound).max())
    # S4: within the FP32 bound of the exact transform of this M
    conv.f43_output_transform()
    torch.cuda.synchronize()
    y = conv.out.double().cpu().numpy()
    munscaled = mg / w43.M_SCALE
    ystar = w43.output_transform43_f64(munscaled, geo)
    e4 = 32 * 2.0 ** -24 * w43.output_transform43_f64(munscaled, geo, absolute=True)
    bound4 = 2.0 ** -11 * (np.abs(ystar) + e4) + e4 + 2.0 ** -25
    assert (np.abs(y - ystar) <= bound4).all(), float((np.abs(y - ystar) / bound4).max())
    # end to end against the exact convolution
    ref = cdirect.conv2d_f64(x, wt, pad=(p, q))
    max_rel = np.abs(y - ref).max() / np.abs(ref).max()
    _log(f"  max_rel {max_rel:.3e}")
    assert max_rel <= 1e-2
    # S2 + S3 + S4 in one call equals the stages
    before = conv.out.clone()
    conv.out.zero_()
    conv.f43_forward(0)
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), before.view(torch.int16))


def test_f43_exact_zeros(wc, runner, cuda):
    """Inputs half zero (as after a ReLU, -0 included) and all-zero filters: U's exact zeros are
    stored as +0 and V's zeros carry the signs of the FP32 operation order, bit-exact with the
    oracle."""
    n, c, h, w, k = 2, 24, 10, 14, 20
    x = wg.uniform_pm1(51, (n, c, h, w))
    x = np.where(x < 0, np.float16(0.0), x).astype(np.float16)
    x[:, 3] = np.float16(-0.0)
    wt = wg.uniform_pm1(52, (k, c, 3, 3))
    wt[::3] = 0
    conv = runner.WinogradConv(_dev(x, cuda), _dev(wt, cuda))
    geo = w43.Geometry43(n, c, h, w, k)
    _poison(conv)
    conv.f43_prepare_weights()
    conv.f43_input_transform()
    torch.cuda.synchronize()
    assert np.array_equal(conv.u43_halves().cpu().numpy().view(np.uint16),
                          w43.pack_U43(w43.filter_transform43(wt), geo).view(np.uint16))
    assert np.array_equal(conv.v43_halves().cpu().numpy().view(np.uint16),
                          w43.pack_V43(w43.input_transform43(x, geo), geo).view(np.uint16))


@pytest.mark.parametrize("shape", [(2, 40, 14, 22, 70, 1, 1), (1, 64, 28, 28, 96, 1, 1)], ids=_ids)
def test_f43_selected_through_the_tuned_table(shape, wc, runner, cuda):
    """A shape recorded as CONV2D_FORM_F43: conv2d_prepare_weights writes the F(4x4) U, and the four
    slots and conv2d_forward(-1) give the stage calls' output bit for bit; NHWC input gives the
    same bits too."""
    L = wc.load_library()
    _log(f"case f43 tuned {_ids(shape)}")
    x, wt, conv = _conv(runner, shape, cuda, seed=31)
    n, c, h, w, k, p, q = shape
    cfg = 2
    L.conv2d_tuned_clear()
    try:
        conv.f43_prepare_weights()
        conv.f43_forward(cfg)
        torch.cuda.synchronize()
        want = conv.out.clone()
        u_want = conv.u43_halves().clone()
        conv.set_form(wc.FORM_F43, cfg)
        assert conv.form() == wc.FORM_F43
        _poison(conv)
        conv.prepare_weights()
        torch.cuda.synchronize()
        assert torch.equal(conv.u43_halves().view(torch.int16), u_want.view(torch.int16))
        conv.out.zero_()
        conv.launch_slots()
        torch.cuda.synchronize()
        assert conv.param().form == wc.FORM_F43
        assert torch.equal(conv.out.view(torch.int16), want.view(torch.int16))
        conv.out.zero_()
        conv.forward(-1)
        torch.cuda.synchronize()
        assert torch.equal(conv.out.view(torch.int16), want.view(torch.int16))
        # NHWC input and output through the same form: bit-identical after the permutation
        xh = _dev(x, cuda).permute(0, 2, 3, 1).contiguous()
        nhwc = runner.WinogradConv(xh, _dev(wt, cuda), pad=(p, q), io=wc.IO_IN_NHWC | wc.IO_OUT_NHWC)
        assert nhwc.form() == wc.FORM_F43
        nhwc.prepare_weights()
        nhwc.forward(-1)
        torch.cuda.synchronize()
        assert torch.equal(nhwc.out_nchw().view(torch.int16), want.view(torch.int16))
        # bfloat16 output: the rounding of the same FP32 value (checked against the FP16 output)
        bf = runner.WinogradConv(_dev(x, cuda), _dev(wt, cuda), pad=(p, q), io=wc.IO_OUT_BF16)
        bf.prepare_weights()
        bf.forward(-1)
        torch.cuda.synchronize()
        d = (bf.out.double() - want.double()).abs()
        assert (d <= 2.0 ** -8 * want.double().abs() + 2.0 ** -20).all()
        # the epilogue (bias + ReLU) through the form: act(y + b) rounded once
        bias = _dev(wg.uniform_pm1(33, (k,)), cuda)
        conv.forward_epilogue(bias, wc.ACT_RELU)
        torch.cuda.synchronize()
        yb = (want.double() + bias.double().view(1, k, 1, 1)).clamp_min(0)
        assert ((conv.out.double() - yb).abs() <= 2.0 ** -10 * (yb.abs() + want.double().abs()) + 2.0 ** -20).all()
        assert (conv.out.double() >= 0).all()
    finally:
        L.conv2d_tuned_clear()


BIG = {
    "cfg2_N32_C256_56_K256": (32, 256, 56, 56, 256),
    "cfg2r_N32_C96_55_K96": (32, 96, 55, 55, 96),
    "cfg3_N64_C512_28_K512": (64, 512, 28, 28, 512),
}


@pytest.mark.parametrize("name", list(BIG))
def test_f43_full_size(name, wc, runner, cuda):
    n, c, h, w, k = BIG[name]
    _log(f"case f43 full {name}")
    x = wg.normal(1000 + c, (n, c, h, w))
    wt = wg.normal(3000 + c, (k, c, 3, 3), sigma=float(np.sqrt(2.0 / (9 * c))))
    conv = runner.WinogradConv(_dev(x, cuda), _dev(wt, cuda))
    conv.f43_prepare_weights()
    conv.f43_forward()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    y = conv.out.double()
    d = (y - ref).abs()
    max_rel = (d.max() / ref.abs().max()).item()
    rms_rel = (d.norm() / ref.norm()).item()
    _log(f"  {name}: max_rel={max_rel:.3e} rms_rel={rms_rel:.3e}")
    assert bool(torch.isfinite(y).all()) and max_rel <= 1e-2 and rms_rel <= 1e-2
    # U and the first image's V bit-exact (size-independent slices)
    geo = w43.Geometry43(n, c, h, w, k)
    geo1 = w43.Geometry43(1, c, h, w, k)
    u = w43.unpack_U43(conv.u43_halves().cpu().numpy(), geo)
    assert np.array_equal(u.view(np.uint16), w43.filter_transform43(wt).view(np.uint16))
    v = w43.unpack_V43(conv.v43_halves().cpu().numpy(), geo)
    assert np.array_equal(v[:, :, :geo.T].view(np.uint16), w43.input_transform43(x[:1], geo1).view(np.uint16))
    first = conv.out.clone()
    conv.f43_forward()
    torch.cuda.synchronize()
    assert torch.equal(first.view(torch.int16), conv.out.view(torch.int16))


def test_autotuner_times_the_f43_form(wc, runner, cuda):
    """With mid_result set the autotuner times the F(4x4) form too and records the fastest form;
    transform_filter then holds that form's U, so the library's own forward is correct."""
    L = wc.load_library()
    shape = (16, 256, 28, 28, 256, 1, 1)
    x, wt, conv = _conv(runner, shape, cuda, seed=41)
    L.conv2d_tuned_clear()
    try:
        conv.prepare_weights()
        conv.input_transform()
        best, times, t_s2, t_fused, t_f43 = conv.autotune(reps=3, f43=True)
        form = conv.form()
        _log(f"  best {best} staged {t_s2 + times[best]:.1f} fused {t_fused:.1f} f43 {t_f43:.1f} -> form {form}")
        assert t_f43 > 0
        choices = {wc.FORM_STAGED: t_s2 + times[best], wc.FORM_FUSED: t_fused, wc.FORM_F43: t_f43}
        assert form == min((t, f) for f, t in choices.items() if t > 0)[1]
        conv.launch_slots()
        ref = conv.direct_f64()
        torch.cuda.synchronize()
        max_rel = ((conv.out.double() - ref).abs().max() / ref.abs().max()).item()
        assert max_rel <= 1e-2
    finally:
        L.conv2d_tuned_clear()
