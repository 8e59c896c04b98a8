"""GPU parity tests of the fused form (input transform inside the GEMM, conv2d_fused_gemm),
through the C-ABI, against the oracle.

Bars
  output     : within the FP32-accumulation bound of the oracle's FP64 GEMM on the same FP16
               U and FP16-arithmetic V (oracle/winograd.py fused_gpu_model):
               |y - y*| <= 2^-11 (|y*| + e) + e + 2^-25, e = (C + 16) 2^-24 sum_pos |coef| |U|.|V|.
  end to end : normwise max relative error <= 1e-2 against the FP64 direct convolution.
"""
import numpy as np
import pytest
import torch

from oracle import cdirect
from oracle import winograd as wg

pytestmark = pytest.mark.gpu

SHAPES = [
    # n, c, h, w, k, p, q                   tile grid -> block shape
    (1, 32, 16, 16, 64, 1, 1),     # 8x8 tiles -> one 8x8 block
    (2, 40, 14, 22, 70, 1, 1),     # 7x11 tiles, C and K ragged, two k-blocks
    (1, 16, 9, 12, 8, 0, 0),       # pad 0, odd Oh (ragged bottom row)
    (1, 3, 13, 10, 5, 2, 2),       # pad 2, C = 3, K = 5
    (3, 64, 8, 64, 96, 1, 1),      # wide: 4x32 tiles -> 4x16 blocks
    (1, 33, 70, 6, 16, 1, 0),      # tall: 35x2 tiles -> 16x4 blocks, q = 0
    (2, 96, 28, 28, 256, 1, 1),    # four k-blocks, ragged blocks of a 14x14 tile grid
    (1, 8, 4, 130, 32, 1, 1),      # 2x65 tiles -> 2x32 blocks (the widest patch: 816 units)
    (1, 512, 12, 12, 64, 1, 1),    # 16 channel chunks (the double-buffered patch alternates 16x)
    # odd widths: pixel-pair loads at 2-byte alignment, the pair at the right edge loading the
    # dword that ends with its pixel (widths 3 and 5: that dword starts one pixel earlier in the
    # same row; odd H*W puts every other channel plane on an odd offset)
    (1, 16, 9, 11, 8, 1, 1),
    (2, 40, 5, 3, 24, 1, 1),
    (1, 16, 7, 5, 8, 0, 0),
    (1, 16, 3, 3, 8, 2, 2),
    (2, 96, 27, 27, 96, 1, 1),
    (1, 40, 13, 13, 20, 0, 0),
    (1, 8, 4, 131, 32, 2, 2),      # odd width with 2x32 blocks and pad 2
    # the remaining kernel instances (block shape x column shift q & 1 x staging width)
    (3, 64, 8, 64, 96, 0, 0),      # 4x16, shift 0, 4-byte
    (2, 32, 8, 63, 64, 1, 1),      # 4x16, shift 1, 2-byte
    (1, 33, 70, 6, 16, 1, 1),      # 16x4, shift 1, 4-byte
    (1, 24, 70, 7, 16, 0, 0),      # 16x4, shift 0, 2-byte
    (1, 24, 70, 7, 16, 1, 1),      # 16x4, shift 1, 2-byte
    (1, 8, 4, 130, 32, 0, 0),      # 2x32, shift 0, 4-byte
    (1, 8, 4, 131, 32, 0, 0),      # 2x32, shift 0, 2-byte
    (1, 8, 4, 131, 32, 1, 1),      # 2x32, shift 1, 2-byte
    # 8-byte staging (width a multiple of 4; shift 0, 3, 2 for q = 0, 1, 2) and the 4-byte
    # cases that remain with it (width 2 mod 4)
    (1, 8, 4, 132, 32, 0, 0),      # 2x32, shift 0, 8-byte
    (1, 8, 2, 132, 32, 2, 2),      # 2x32, shift 2, 8-byte
    (1, 8, 4, 132, 32, 1, 1),      # 2x32, shift 3, 8-byte
    (3, 64, 8, 66, 96, 0, 0),      # 4x16, shift 0, 4-byte
    (3, 64, 8, 62, 96, 1, 1),      # 4x16, shift 1, 4-byte
    (3, 64, 8, 60, 96, 1, 2),      # 4x16, shift 2, 8-byte
    (1, 32, 16, 16, 64, 1, 2),     # 8x8, shift 2, 8-byte
    (1, 33, 70, 8, 16, 1, 0),      # 16x4, shift 0, 8-byte
    (1, 33, 70, 4, 16, 1, 2),      # 16x4, shift 2, 8-byte
]


def _load_width(w, prm):
    """Staging load width label: 2 for odd widths (pixel-pair loads at 2-byte alignment), else 8 when the patch takes one quad
    unit per thread (fnu = 1) and 4 for pixel pairs (always two units per thread)."""
    return 2 if w % 2 else (8 if prm.fnu == 1 else 4)


def _ids(s):
    return "x".join(map(str, s))


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_fused_matches_oracle(shape, runner, cuda):
    n, c, h, w, k, p, q = shape
    print(f"case {shape}", flush=True)
    geo = wg.Geometry(n, c, h, w, k, p, q)
    x = wg.uniform_pm1(900 + c, (n, c, h, w))
    wt = wg.uniform_pm1(907 + c, (k, c, 3, 3))
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    prm = conv.param()
    print(f"  block {prm.fbh}x{prm.fbw} units {prm.fnu}", flush=True)
    conv.prepare_weights()
    conv.out.fill_(float("nan"))
    conv.fused_gemm()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy().astype(np.float64)
    y_model, scale = wg.fused_gpu_model(x, wt, geo)
    e = (c + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"
    ref = cdirect.conv2d_f64(x, wt, pad=(p, q))
    m = wg.error_metrics(y, ref)
    print(f"  max_rel={m['max_rel']:.3e} rms_rel={m['rms_rel']:.3e}", flush=True)
    assert m["finite"] and m["max_rel"] <= 1e-2


@pytest.mark.parametrize("shift,fnu", [(1, 2), (2, 2)], ids=["base2mod4", "base4mod8"])
def test_fused_unaligned_input_base_takes_narrower_loads(shift, fnu, runner, wc, cuda):
    """A width-24 input (8-byte quads possible) whose base address is 2 bytes past a 4-byte
    boundary takes the 2-byte staging path, and one 4 bytes past an 8-byte boundary the
    4-byte pair path (fnu = 2, not the quads' 1); both meet the oracle bound."""
    n, c, h, w, k, p, q = 2, 40, 14, 24, 70, 1, 1
    print(f"case unaligned base +{2 * shift} bytes", flush=True)
    geo = wg.Geometry(n, c, h, w, k, p, q)
    x = wg.uniform_pm1(950, (n, c, h, w))
    wt = wg.uniform_pm1(951, (k, c, 3, 3))
    flat = torch.empty(n * c * h * w + shift, dtype=torch.float16, device=cuda)
    X = flat[shift:].view(n, c, h, w)
    X.copy_(_to_dev(x, cuda))
    assert X.data_ptr() % 8 == 2 * shift
    conv = runner.WinogradConv(X, _to_dev(wt, cuda), pad=(p, q))
    assert conv.form() == wc.FORM_FUSED
    assert conv.param().fnu == fnu
    conv.prepare_weights()
    conv.out.fill_(float("nan"))
    conv.launch_slots()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy().astype(np.float64)
    y_model, scale = wg.fused_gpu_model(x, wt, geo)
    e = (c + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"


def test_wide_image_runs_fused_only(runner, wc, cuda):
    """An image too wide for the input transform's LDS staging (w = 700: one tile row needs
    181 KiB) stays on the Winograd path in the fused form: the four-slot protocol and
    forward(-1) run it and meet the oracle bound."""
    n, c, h, w, k, p, q = 1, 24, 6, 700, 40, 1, 1
    print("case wide 700", flush=True)
    geo = wg.Geometry(n, c, h, w, k, p, q)
    x = wg.uniform_pm1(960, (n, c, h, w))
    wt = wg.uniform_pm1(961, (k, c, 3, 3))
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    assert conv.method() == wc.METHOD_WINOGRAD and conv.form() == wc.FORM_FUSED and conv.param().band == 0
    y_model, scale = wg.fused_gpu_model(x, wt, geo)
    e = (c + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    for run in ("slots", "forward"):
        conv.out.fill_(float("nan"))
        if run == "slots":
            conv.launch_slots()
        else:
            conv.prepare_weights()
            conv.forward()
        torch.cuda.synchronize()
        y = conv.out.cpu().numpy().astype(np.float64)
        assert np.all(np.abs(y - y_model) <= tol), f"{run}: max excess {np.max(np.abs(y - y_model) - tol)}"


def test_fused_block_shapes_cover_every_variant(runner, cuda):
    """The shapes above run every kernel instance: 4 tile-block shapes x staging width x
    column shift: 2- and 4-byte loads with shift q & 1 (2 pixel-pair units per thread,
    648-816 units of 8 channels x 2 pixels), 8-byte loads with shift (4 - q % 4) % 4 (one
    quad unit per thread, 360-432 units of 8 channels x 4 pixels; the 16x4 block with shift 3
    would need 544 and takes 4-byte loads)."""
    seen = set()
    for (n, c, h, w, k, p, q) in SHAPES:
        x = torch.zeros((n, c, h, w), dtype=torch.float16, device=cuda)
        wt = torch.zeros((k, c, 3, 3), dtype=torch.float16, device=cuda)
        prm = runner.WinogradConv(x, wt, pad=(p, q)).param()
        lw = _load_width(w, prm)
        assert prm.fnu == (1 if lw == 8 else 2)
        assert prm.fsh == ((4 - q % 4) % 4 if lw == 8 else q & 1)
        seen.add((prm.fbh, prm.fbw, prm.fsh, lw))
    blocks = [(8, 8), (4, 16), (16, 4), (2, 32)]
    want = {(bh, bw, sh, lw) for (bh, bw) in blocks for sh in (0, 1) for lw in (2, 4)}
    want |= {(bh, bw, sh, 8) for (bh, bw) in blocks for sh in (0, 2, 3) if (bh, bw, sh) != (16, 4, 3)}
    assert seen == want


def test_fused_full_size_against_fp64_gpu(runner, cuda):
    """N=32 C=256 56x56 K=256: the fused form against the GPU FP64 kernel, and close to the
    staged form on the same data."""
    n, c, h, k = 32, 256, 56, 256
    print("case full size", flush=True)
    g = torch.Generator(device=cuda)
    g.manual_seed(77)
    x = torch.randn((n, c, h, h), generator=g, device=cuda).half()
    w = (torch.randn((k, c, 3, 3), generator=g, device=cuda) * (2.0 / (9 * c)) ** 0.5).half()
    conv = runner.WinogradConv(x, w)
    conv.prepare_weights()
    conv.fused_gemm()
    fused = conv.out.clone()
    conv.forward()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    for name, y in (("fused", fused), ("staged", conv.out)):
        d = (y.double() - ref).abs()
        max_rel = (d.max() / ref.abs().max()).item()
        rms_rel = (d.norm() / ref.norm()).item()
        print(f"  {name}: max_rel={max_rel:.3e} rms_rel={rms_rel:.3e}", flush=True)
        assert bool(torch.isfinite(y).all()) and max_rel <= 1e-2
    # deterministic: a second run gives the same bits
    conv.fused_gemm()
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), fused.view(torch.int16))


# Layers whose last 64-channel block (1..32 channels) runs as wide items, 32 channels x two
# 64-tile blocks side by side: pair staging (2-byte for odd widths, 4-byte for even ones) and
# quad staging (8-byte, widths a multiple of 4), block shapes 8x8, 4x16 and 16x4, column shifts
# 0 to 3, K below 64 and above it, enough blocks that the wide items finish sooner on 256
# compute units.
WIDE_SHAPES = [
    # n, c, h, w, k, p, q
    (40, 16, 33, 33, 32, 1, 1),    # 8x8, 2-byte, shift 1, wide items only
    (48, 40, 15, 61, 160, 1, 1),   # 8x8, 2-byte, two full blocks + wide
    (32, 96, 55, 55, 96, 1, 1),    # 4x16, 2-byte
    (64, 16, 31, 33, 24, 1, 1),    # 16x4, 2-byte, 24 channels in the last block
    (40, 40, 30, 62, 96, 1, 1),    # 8x8, 4-byte, shift 1
    (96, 24, 34, 34, 32, 0, 0),    # 4-byte, shift 0
    (32, 96, 56, 56, 96, 1, 1),    # 4x16, 8-byte, shift 3
    (40, 40, 32, 60, 96, 1, 2),    # 8x8, 8-byte, shift 2
    (64, 16, 64, 24, 32, 1, 0),    # 16x4, 8-byte, shift 0
]


@pytest.mark.parametrize("shape", WIDE_SHAPES)
def test_wide_tail_items_match_the_oracle(shape, runner, cuda):
    """The single-layer launch takes wide tail items (its grid is smaller than one item per
    64-channel block and 64-tile block): within the FP32 bound of the oracle's fused model, like
    the half-block items of the smaller layers, and bit-identical to the grouped kernel's wide
    items (a one-layer batch)."""
    n, c, h, w, k, p, q = shape
    print(f"case wide {shape}", flush=True)
    x = wg.uniform_pm1(1300 + c, (n, c, h, w))
    wt = wg.uniform_pm1(1307 + c, (k, c, 3, 3))
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    prm = conv.param()
    ki, _ = conv.kernel_info()
    half = n * prm.fnbh * prm.fnbw * -(-k // 64)
    wide = n * prm.fnbh * (prm.fnbw * (k // 64) + (prm.fnbw + 1) // 2)
    assert torch.cuda.get_device_properties(0).multi_processor_count == 256
    assert prm.form == 1 and prm.fnu in (1, 2) and ki.slot(2)["grid"][0] == wide < half, (prm.fbh, prm.fbw)
    conv.prepare_weights()
    conv.out.fill_(float("nan"))
    conv.fused_gemm()
    torch.cuda.synchronize()
    y_wide = conv.out.clone()
    b = runner.Batch([conv])
    b.prepare_weights()
    conv.out.fill_(float("nan"))
    b.forward()
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), y_wide.view(torch.int16))
    b.close()
    y = y_wide.cpu().numpy().astype(np.float64)
    y_model, scale = wg.fused_gpu_model(x, wt, wg.Geometry(n, c, h, w, k, p, q))
    e = (c + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"
    y64 = conv.direct_f64()
    d = (y_wide.double() - y64).abs()
    assert torch.isfinite(y_wide).all() and (d.max() / y64.abs().max()).item() <= 1e-2


def test_wide_tail_items_epilogue_and_nhwc_output(runner, cuda):
    """Wide items through the epilogue paths: NHWC FP16 output equal to the NCHW output
    transposed, bit for bit; bias + ReLU within the end-to-end bound of the FP64 reference."""
    n, c, h, w, k, p, q = 64, 16, 31, 33, 24, 1, 1
    print("case wide epilogue", flush=True)
    x = _to_dev(wg.uniform_pm1(1400, (n, c, h, w)), cuda)
    wt = _to_dev(wg.uniform_pm1(1401, (k, c, 3, 3)), cuda)
    conv = runner.WinogradConv(x, wt, pad=(p, q))
    conv.prepare_weights()
    conv.forward()
    torch.cuda.synchronize()
    y = conv.out.clone()
    nhwc = runner.WinogradConv(x, wt, pad=(p, q), io=2)
    prm = nhwc.param()
    assert prm.form == 1 and prm.fnu == 2
    nhwc.prepare_weights()
    nhwc.out.fill_(float("nan"))
    nhwc.forward()
    torch.cuda.synchronize()
    assert torch.equal(nhwc.out.view(torch.int16), y.permute(0, 2, 3, 1).contiguous().view(torch.int16))
    bias = (torch.arange(k, device=cuda, dtype=torch.float32) / k - 0.5).half()
    conv.out.fill_(float("nan"))
    conv.forward_epilogue(bias, 1)
    torch.cuda.synchronize()
    y64 = torch.relu(conv.direct_f64() + bias.double().view(1, k, 1, 1))
    d = (conv.out.double() - y64).abs()
    assert torch.isfinite(conv.out).all() and (d.max() / y64.abs().max()).item() <= 1e-2
