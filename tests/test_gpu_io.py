"""GPU tests of the input/output layout and type options (conv2d_forward_io, getkernelInfo_io,
conv2d_direct_f64_io), through the C-ABI.

Bars
  NHWC input / NHWC output : bit-identical to the NCHW run on the same form and configuration
                             (the arithmetic does not depend on where a value is loaded from or
                             stored to), on the fused form, the staged form, the four slots and
                             the general-geometry path; the staged and direct paths' transformed
                             inputs are bit-identical too, and so is the FP64 validation kernel.
  bfloat16 output          : the epilogue rounds the same FP32 value y that the FP16 output
                             rounds, once, to nearest even. From the FP16 result h of the same
                             form: every bfloat16 rounding boundary is an FP16 number in the
                             normal range, so for |h| >= 2^-14, bf16(y) = bf16(h) unless h is
                             itself a boundary (then either neighbour); for FP16 subnormals
                             |bf16(y) - h| <= 2^-25 + 2^-9 |y|.
  rounding, exactly        : a pair-sum convolution (out[k] = x[2k] + x[2k+1], every
                             intermediate exact) on grids that put about a quarter of the
                             outputs on exact FP16 ties and many on bfloat16 ties: FP16 and
                             bfloat16 outputs equal the correctly rounded sum (nearest even;
                             bfloat16 by the oracle's rounding) on the fused, staged and direct
                             paths, in every layout.
  end to end               : normwise max relative error <= 1e-2 against the FP64 direct
                             convolution, bfloat16 output included.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import bf16
from oracle import winograd as wg

pytestmark = pytest.mark.gpu

IN_NHWC, OUT_NHWC, OUT_BF16 = 1, 2, 4
IOS = [1, 2, 3, 4, 5, 6, 7]

WINO = [
    # n, c, h, w, k, p, q              fused block shape, NHWC output stores
    (2, 32, 16, 16, 64, 1, 1),      # 8x8, 8-byte stores (K % 4 == 0)
    (1, 64, 28, 28, 96, 1, 1),      # 4x16
    (1, 8, 4, 130, 32, 1, 1),       # 2x32: 4 NHWC staging units per thread
    (1, 24, 70, 6, 16, 1, 0),       # 16x4, q = 0
    (2, 40, 14, 22, 70, 2, 2),      # pad 2, K = 70: 2-byte NHWC stores, ragged k-block
    (1, 16, 9, 11, 12, 1, 1),       # odd width and height
    (1, 36, 17, 19, 40, 1, 1),      # C % 8 != 0: NHWC input has no fused form (staged)
]
DIRECT = [
    # n, c, h, w, k, r, s, u, v, p, q
    (2, 12, 15, 17, 20, 5, 5, 2, 2, 2, 2),
    (2, 24, 9, 9, 40, 1, 1, 1, 1, 0, 0),    # 1x1: the GEMM gathers the input in every layout
    (2, 16, 13, 11, 24, 3, 3, 2, 2, 1, 1),  # C % 8 == 0: one 16-byte load per packed value group
]


def _ids(s):
    return "x".join(map(str, s))


def _log(msg):
    print(msg, flush=True)


def _bits(t):
    return t.contiguous().view(torch.int16).cpu().numpy()


def _lib(wc):
    return wc.load_library()


def _conv(runner, x_nchw, wt, io, pad, stride=(1, 1)):
    x = x_nchw.permute(0, 2, 3, 1).contiguous() if io & IN_NHWC else x_nchw
    return runner.WinogradConv(x, wt, pad=pad, stride=stride, io=io)


def _run(wc, conv, route, cfg):
    """route: 'fused' (tuned form), 'staged' (explicit configuration), 'slots' (the four
    slots, default form) or 'direct' (general geometry)."""
    L = _lib(wc)
    conv.prepare_weights()
    if route == "fused":
        L.conv2d_tuned_clear()
        assert L.conv2d_tuned_set_form(ctypes.byref(conv.pb), cfg, wc.FORM_FUSED) == wc.OK
        conv.forward(-1)
        L.conv2d_tuned_clear()
    elif route == "staged":
        conv.forward(cfg)
    elif route == "slots":
        conv.launch_slots()
    else:
        conv.forward(-1)
    torch.cuda.synchronize()
    return conv.out_nchw()


def _check_bf16(b, h):
    """b: bfloat16 output, h: FP16 output of the same form, both N,K,Oh,Ow."""
    bv = b.float().cpu().numpy()
    hv = h.float().cpu().numpy()
    assert np.isfinite(hv).all()
    hb = bf16.round_bf16_bits(hv)
    bb = bv.view(np.uint32) >> 16
    normal = np.abs(hv) >= 2.0 ** -14
    tie = (hv.view(np.uint32) & 0xFFFF) == 0x8000
    lo = hv.view(np.uint32) >> 16
    ok = np.where(normal & ~tie, bb == hb, True)
    ok &= np.where(normal & tie, (bb == lo) | (bb == lo + 1), True)
    ok &= np.where(~normal, np.abs(bv.astype(np.float64) - hv) <= 2.0 ** -25 + 2.0 ** -9 * np.abs(hv) * 1.01, True)
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, (bad[:8], bv.ravel()[bad[:8]], hv.ravel()[bad[:8]])
    return int(tie[normal].sum())


def _normwise(y, ref):
    return float((y.double() - ref).abs().max() / ref.abs().max().clamp_min(1e-30))


@pytest.mark.parametrize("shape", WINO, ids=_ids)
def test_winograd_io_matches_nchw(shape, wc, runner, cuda):
    n, c, h, w, k, p, q = shape
    x = torch.from_numpy(wg.uniform_pm1(71, (n, c, h, w))).to(cuda)
    wt = torch.from_numpy(wg.uniform_pm1(72, (k, c, 3, 3))).to(cuda)
    base = _conv(runner, x, wt, 0, (p, q))
    cfg = base.selected_config()
    routes = ["fused", "staged", "slots"]
    ref = {}
    for r in routes:
        ref[r] = _run(wc, base, r, cfg).clone()
        if r == "staged":
            v_ref = _bits(base.v_halves())  # the staged run's transformed input
    f64 = base.direct_f64()
    torch.cuda.synchronize()
    for io in IOS:
        conv = _conv(runner, x, wt, io, (p, q))
        fused_ok = not (io & IN_NHWC) or c % 8 == 0
        for r in routes:
            _log(f"case {_ids(shape)} io={io} route={r} cfg={cfg}")
            y = _run(wc, conv, r, cfg)
            # without a fused form (NHWC input, C % 8 != 0) the tuned fused entry and the
            # four slots run staged: compare with the staged NCHW run
            want = ref[r] if fused_ok or r == "staged" else ref["staged"]
            if io & OUT_BF16:
                ties = _check_bf16(y, want)
                e = _normwise(y, f64)
                _log(f"  bf16 ok, {ties} FP16 results on a bf16 boundary, max_rel {e:.3e}")
                assert e <= 1e-2
            else:
                assert np.array_equal(_bits(y), _bits(want))
            if r == "staged" and io & IN_NHWC:
                assert np.array_equal(_bits(conv.v_halves()), v_ref)
        if io & IN_NHWC:
            f64_io = conv.direct_f64()
            torch.cuda.synchronize()
            assert torch.equal(f64_io, f64)


@pytest.mark.parametrize("shape", DIRECT, ids=_ids)
def test_direct_path_io_matches_nchw(shape, wc, runner, cuda):
    n, c, h, w, k, r, s, u, v, p, q = shape
    x = torch.from_numpy(wg.uniform_pm1(73, (n, c, h, w))).to(cuda)
    wt = torch.from_numpy(wg.uniform_pm1(74, (k, c, r, s))).to(cuda)
    base = _conv(runner, x, wt, 0, (p, q), (u, v))
    assert base.method() == wc.METHOD_DIRECT
    want = _run(wc, base, "direct", -1).clone()
    v_ref = _bits(base.v_halves())  # the packed (im2col) input of that run
    slots = _run(wc, base, "slots", -1).clone()
    assert np.array_equal(_bits(slots), _bits(want))
    f64 = base.direct_f64()
    torch.cuda.synchronize()
    for io in IOS:
        _log(f"case {_ids(shape)} io={io} route=direct")
        conv = _conv(runner, x, wt, io, (p, q), (u, v))
        for route in ("direct", "slots"):
            y = _run(wc, conv, route, -1)
            if io & OUT_BF16:
                _check_bf16(y, want)
                assert _normwise(y, f64) <= 1e-2
            else:
                assert np.array_equal(_bits(y), _bits(want))
            if route == "direct" and io & IN_NHWC:
                assert np.array_equal(_bits(conv.v_halves()), v_ref)
        if io & IN_NHWC:
            f64_io = conv.direct_f64()
            torch.cuda.synchronize()
            assert torch.equal(f64_io, f64)


def _pair_sum(k, r, dev):
    """K filters over C = 2K channels: filter k takes channels 2k and 2k + 1 at the centre tap."""
    wt = torch.zeros((k, 2 * k, r, r), dtype=torch.float16)
    for i in range(k):
        wt[i, 2 * i, r // 2, r // 2] = 1.0
        wt[i, 2 * i + 1, r // 2, r // 2] = 1.0
    return wt.to(dev)


@pytest.mark.parametrize("route", ["fused", "staged", "direct"])
def test_pair_sum_convolution_rounds_ties_to_even(route, wc, runner, cuda):
    """out[k] = x[2k] + x[2k+1] exactly before the output rounding: x[2k] = i 2^-8 (|x| < 2),
    x[2k+1] = j 2^-12 (|x| < 2^-4). On the Winograd path U is 0 or +-1/4, V stays on those
    grids below 8 and 2^-2 (exact in FP16), and every FP32 sum is exact. The sum is a multiple
    of 2^-12 below 2.07, so FP16 output (ulp 2^-10 on [1, 2)) and bfloat16 output both meet
    exact ties; the output must be the correctly rounded sum (nearest even) in every layout."""
    n, k, h, w = 2, 8, 20, 22
    c = 2 * k
    r = 1 if route == "direct" else 3
    g = np.random.default_rng(5)
    x = np.empty((n, c, h, w), dtype=np.float64)
    x[:, 0::2] = g.integers(-511, 512, size=(n, k, h, w)) * 2.0 ** -8
    x[:, 1::2] = g.integers(-255, 256, size=(n, k, h, w)) * 2.0 ** -12
    xv = x.astype(np.float16)
    assert np.array_equal(xv.astype(np.float64), x)
    y = x[:, 0::2] + x[:, 1::2]
    want_f16 = y.astype(np.float16).view(np.int16)  # correctly rounded, ties to even
    want_bf16 = bf16.round_bf16_bits(y.astype(np.float32)).astype(np.int16)
    assert np.array_equal(y.astype(np.float32).astype(np.float64), y)
    ulp16 = np.spacing(np.abs(y).astype(np.float16)).astype(np.float64)
    assert ((y / ulp16) % 1 == 0.5).sum() > 1000  # exact FP16 ties (1764 of 7040)
    xd = torch.from_numpy(xv).to(cuda)
    wt = _pair_sum(k, r, cuda)
    for io in [0] + IOS:
        _log(f"case pair-sum route={route} io={io}")
        conv = _conv(runner, xd, wt, io, (r // 2, r // 2))
        assert conv.method() == (wc.METHOD_DIRECT if route == "direct" else wc.METHOD_WINOGRAD)
        cfg = conv.selected_config() if route != "direct" else -1
        _run(wc, conv, route, cfg)
        got = conv.out
        if io & OUT_NHWC:
            got = got.permute(0, 3, 1, 2)
        got = _bits(got)
        want = want_bf16 if io & OUT_BF16 else want_f16
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (io, bad[:8], y.ravel()[bad[:8]])


@pytest.mark.parametrize("route", ["fused", "staged", "direct"])
def test_epilogue_with_io_words(route, wc, runner, cuda):
    """Bias + ReLU set in a getkernelInfo_io blob (conv2d_set_epilogue), four slots: NHWC
    results bit-identical to the NCHW epilogue result, bfloat16 the rounding of the same
    value (the form forced through the tuned table)."""
    L = _lib(wc)
    if route == "direct":
        n, c, h, w, k, r, st, p = 2, 16, 15, 17, 24, 5, 2, 2
    else:
        n, c, h, w, k, r, st, p = 2, 32, 14, 18, 40, 3, 1, 1
    x = torch.from_numpy(wg.uniform_pm1(75, (n, c, h, w))).to(cuda)
    wt = torch.from_numpy(wg.uniform_pm1(76, (k, c, r, r))).to(cuda)
    bias = torch.from_numpy(wg.uniform_pm1(77, (k,))).to(cuda)
    outs = {}
    for io in [0] + IOS:
        _log(f"case epilogue route={route} io={io}")
        conv = _conv(runner, x, wt, io, (p, p), (st, st))
        L.conv2d_tuned_clear()
        if route != "direct":
            form = wc.FORM_FUSED if route == "fused" else wc.FORM_STAGED
            assert L.conv2d_tuned_set_form(ctypes.byref(conv.pb), conv.selected_config(), form) == wc.OK
        conv.prepare_weights()
        conv.kernel_info()
        if route != "direct":
            assert conv.param().form == (wc.FORM_FUSED if route == "fused" else wc.FORM_STAGED)
        conv.set_slot_epilogue(bias, wc.ACT_RELU)
        conv.launch_slots()
        torch.cuda.synchronize()
        L.conv2d_tuned_clear()
        outs[io] = conv.out_nchw().clone()
        y = outs[io]
        if io == 0:
            assert (y.float() >= 0).all() and (y == 0).any() and (y > 0).any()
        elif io & OUT_BF16:
            _check_bf16(y, outs[0])
        else:
            assert np.array_equal(_bits(y), _bits(outs[0]))
