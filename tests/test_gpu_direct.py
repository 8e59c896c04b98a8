"""GPU parity tests of the general-geometry (implicit-GEMM direct) path, through the C-ABI,
against the oracle (oracle/direct.py packing, oracle/direct_conv.c convolution).

Bars
  packed W and X    : bit-exact against the oracle's packing (workspaces poisoned first).
  GEMM output       : within the FP32-accumulation bound of the oracle's FP64 GEMM on the same
                      packed operands:  |y - y*| <= 2^-11 (|y*| + e) + e + 2^-25,
                      e = (32 * nsteps + 16) 2^-24 * (|W| . |X|).
  end to end        : normwise max relative error <= 1e-2 against the FP64 direct conv.
  configurations    : bit-identical output across the direct tile configurations; the
                      four-slot protocol equals the stage API.
A 1x1 filter has no packing pass (the GEMM gathers its input operand from the input itself):
its tests check that nothing is written to the input workspace and hold the GEMM output to the
same bound, which is computed from the same operands.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import cdirect
from oracle import direct as dg
from oracle import winograd as wg

pytestmark = pytest.mark.gpu

GEOS = [
    dg.DirectGeometry(2, 5, 9, 8, 7, 5, 5, 1, 1, 2, 2),      # 5x5 pad 2
    dg.DirectGeometry(1, 40, 11, 13, 33, 3, 3, 2, 2, 1, 1),   # 3x3 stride 2, ragged C and K
    dg.DirectGeometry(2, 3, 23, 23, 8, 7, 7, 2, 2, 3, 3),     # stem-like 7x7 stride 2
    dg.DirectGeometry(1, 16, 6, 7, 16, 1, 1, 1, 1, 0, 0),     # 1x1
    dg.DirectGeometry(1, 8, 7, 5, 4, 3, 1, 1, 1, 1, 0),       # 3x1
    dg.DirectGeometry(1, 8, 8, 8, 8, 3, 3, 1, 1, 3, 3),       # pad 3
    dg.DirectGeometry(1, 70, 6, 6, 5, 2, 2, 2, 2, 0, 0),      # 2x2 stride 2, three chunks
    dg.DirectGeometry(3, 64, 14, 14, 130, 3, 3, 2, 2, 1, 1),  # K > 128, columns straddle images
]


def _ids(g):
    return f"n{g.n}c{g.c}h{g.h}w{g.w}k{g.k}r{g.r}s{g.s}u{g.u}p{g.p}"


def _dev(a, cuda):
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


@pytest.mark.parametrize("geo", GEOS, ids=_ids)
def test_direct_packing_gemm_and_accuracy(geo, runner, wc, cuda):
    print(f"case {_ids(geo)}", flush=True)
    x = wg.uniform_pm1(700 + geo.c, (geo.n, geo.c, geo.h, geo.w))
    w = wg.uniform_pm1(701 + geo.r, (geo.k, geo.c, geo.r, geo.s))
    conv = runner.WinogradConv(_dev(x, cuda), _dev(w, cuda), pad=(geo.p, geo.q), stride=(geo.u, geo.v))
    assert conv.method() == wc.METHOD_DIRECT
    conv.V.fill_(0x55)
    conv.U.fill_(0x55)
    conv.prepare_weights()
    conv.input_transform()
    torch.cuda.synchronize()
    assert np.array_equal(conv.u_halves().cpu().numpy().view(np.uint16), dg.pack_weights(w, geo).view(np.uint16))
    if geo.r == 1 and geo.s == 1:
        assert conv.v_bytes == 0 and bool((conv.V == 0x55).all())  # no packed input
    else:
        assert np.array_equal(conv.v_halves().cpu().numpy().view(np.uint16), dg.pack_input(x, geo).view(np.uint16))

    conv.out.fill_(float("nan"))
    conv.gemm()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy().astype(np.float64)
    y_model = dg.gemm_output_f64(w, x, geo)
    e = (32 * geo.nsteps + 16) * 2.0 ** -24 * dg.gemm_output_f64(w, x, geo, absolute=True)
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"
    ref = cdirect.conv2d_f64(x, w, stride=(geo.u, geo.v), pad=(geo.p, geo.q))
    m = wg.error_metrics(y, ref)
    print(f"  max_rel={m['max_rel']:.3e} rms_rel={m['rms_rel']:.3e}", flush=True)
    assert m["finite"] and m["max_rel"] <= 1e-2


@pytest.mark.parametrize("geo", [GEOS[1], GEOS[7]], ids=_ids)
def test_direct_configs_identical_and_slots_equal_stages(geo, runner, wc, cuda):
    x = wg.uniform_pm1(800 + geo.c, (geo.n, geo.c, geo.h, geo.w))
    w = wg.uniform_pm1(801 + geo.c, (geo.k, geo.c, geo.r, geo.s))
    conv = runner.WinogradConv(_dev(x, cuda), _dev(w, cuda), pad=(geo.p, geo.q), stride=(geo.u, geo.v))
    conv.prepare_weights()
    conv.input_transform()
    outs = []
    for cfg in range(wc.load_library().conv2d_direct_num_tile_configs()):
        conv.out.fill_(float("nan"))
        conv.gemm(cfg)
        torch.cuda.synchronize()
        outs.append(conv.out.clone())
    for i, o in enumerate(outs[1:], 1):
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16)), f"config {i}"
    # the reference protocol (getParamsize / getkernelInfo / four slots) on the same problem
    conv.out.fill_(float("nan"))
    conv.V.fill_(0)
    conv.U.fill_(0)
    conv.launch_slots()
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), outs[conv.selected_config()].view(torch.int16))
    assert conv.param().nsteps == geo.nsteps


@pytest.mark.parametrize("shape", [
    # n, c, h, k, r, stride, pad
    (8, 3, 224, 64, 7, 2, 3),     # ResNet-style stem
    (8, 128, 56, 256, 3, 2, 1),   # stride-2 3x3
    (4, 64, 56, 64, 5, 1, 2),     # 5x5
    (8, 256, 28, 512, 1, 1, 0),   # 1x1
], ids=lambda s: "x".join(map(str, s)))
def test_direct_full_size_against_fp64_gpu(shape, runner, cuda):
    n, c, h, k, r, st, p = shape
    print(f"case {shape}", flush=True)
    g = torch.Generator(device=cuda)
    g.manual_seed(31 + r)
    x = torch.randn((n, c, h, h), generator=g, device=cuda).half()
    w = (torch.randn((k, c, r, r), generator=g, device=cuda) * (2.0 / (r * r * c)) ** 0.5).half()
    conv = runner.WinogradConv(x, w, pad=(p, p), stride=(st, st))
    conv.prepare_weights()
    conv.forward()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    d = (conv.out.double() - ref).abs()
    max_rel = (d.max() / ref.abs().max()).item()
    rms_rel = (d.norm() / ref.norm()).item()
    print(f"  max_rel={max_rel:.3e} rms_rel={rms_rel:.3e}", flush=True)
    assert bool(torch.isfinite(conv.out).all()) and max_rel <= 1e-2


POINT_GEOS = [
    # n, c, h, w, k, r, s, u, v, p, q
    dg.DirectGeometry(2, 16, 6, 7, 16, 1, 1, 1, 1, 0, 0),     # columns straddle images
    dg.DirectGeometry(1, 13, 9, 9, 40, 1, 1, 2, 2, 0, 0),     # stride 2, C not a multiple of 8
    dg.DirectGeometry(1, 8, 5, 5, 8, 1, 1, 1, 1, 1, 1),       # pad 1: a border of zero inputs
    dg.DirectGeometry(3, 72, 7, 7, 130, 1, 1, 1, 1, 0, 0),    # three channel chunks, K > 128
    # H*W % 4 == 0, stride 1, no padding: N,C,H,W input read in 8-byte pieces along its rows
    dg.DirectGeometry(2, 16, 6, 8, 24, 1, 1, 1, 1, 0, 0),
    dg.DirectGeometry(1, 13, 4, 9, 40, 1, 1, 1, 1, 0, 0),     # channel quads cut by C = 13
    dg.DirectGeometry(3, 72, 14, 14, 130, 1, 1, 1, 1, 0, 0),  # 32-column blocks straddle images
]
QUAD_GEOS = [g for g in POINT_GEOS if g.u == 1 and g.p == 0 and (g.h * g.w) % 4 == 0]


@pytest.mark.parametrize("geo", POINT_GEOS, ids=_ids)
@pytest.mark.parametrize("nhwc", [False, True], ids=["nchw", "nhwc"])
def test_pointwise_gemm_reads_the_input(geo, nhwc, runner, wc, cuda):
    """1x1 filters: every tile configuration gives the same bits, within the FP32 bound of the
    FP64 GEMM on the same operands and 1e-2 of the FP64 convolution, for N,C,H,W and
    channel-contiguous input (16-byte loads when C % 8 == 0, element loads otherwise)."""
    print(f"case {_ids(geo)} nhwc={nhwc}", flush=True)
    x = wg.uniform_pm1(900 + geo.c, (geo.n, geo.c, geo.h, geo.w))
    w = wg.uniform_pm1(901 + geo.k, (geo.k, geo.c, 1, 1))
    xd = _dev(x.transpose(0, 2, 3, 1) if nhwc else x, cuda)
    conv = runner.WinogradConv(xd, _dev(w, cuda), pad=(geo.p, geo.q), stride=(geo.u, geo.v), io=1 if nhwc else 0)
    assert conv.method() == wc.METHOD_DIRECT and conv.v_bytes == 0
    conv.prepare_weights()
    outs = []
    for cfg in range(wc.load_library().conv2d_direct_num_tile_configs()):
        conv.out.fill_(float("nan"))
        conv.forward(cfg)  # the layout-aware entry point (the stage API reads N,C,H,W input)
        torch.cuda.synchronize()
        outs.append(conv.out.clone())
    for i, o in enumerate(outs[1:], 1):
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16)), f"config {i}"
    y = outs[0].cpu().numpy().astype(np.float64)
    y_model = dg.gemm_output_f64(w, x, geo)
    e = (32 * geo.nsteps + 16) * 2.0 ** -24 * dg.gemm_output_f64(w, x, geo, absolute=True)
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"
    ref = cdirect.conv2d_f64(x, w, stride=(geo.u, geo.v), pad=(geo.p, geo.q))
    m = wg.error_metrics(y, ref)
    print(f"  max_rel={m['max_rel']:.3e}", flush=True)
    assert m["finite"] and m["max_rel"] <= 1e-2
    # the reference protocol on the same problem: four slots, the second one empty
    conv.out.fill_(float("nan"))
    conv.launch_slots()
    torch.cuda.synchronize()
    assert torch.equal(conv.out.view(torch.int16), outs[conv.selected_config()].view(torch.int16))


@pytest.mark.parametrize("geo", QUAD_GEOS, ids=_ids)
def test_pointwise_row_pieces_match_gathered_input(geo, runner, wc, cuda):
    """The 8-byte row pieces (N,C,H,W, H*W % 4 == 0, 8-byte aligned input) and the per-channel
    gather (the same input 2 bytes past an 8-byte boundary) give the same bits."""
    print(f"case {_ids(geo)}", flush=True)
    x = _dev(wg.uniform_pm1(950 + geo.c, (geo.n, geo.c, geo.h, geo.w)), cuda)
    w = _dev(wg.uniform_pm1(951 + geo.k, (geo.k, geo.c, 1, 1)), cuda)
    shifted = torch.empty(x.numel() + 4, dtype=torch.float16, device=cuda)[1:1 + x.numel()].view(x.shape)
    shifted.copy_(x)
    assert shifted.data_ptr() % 8 == 2 and x.data_ptr() % 8 == 0
    outs = []
    for xi in (x, shifted):
        conv = runner.WinogradConv(xi, w, pad=(0, 0))
        conv.prepare_weights()
        for cfg in range(wc.load_library().conv2d_direct_num_tile_configs()):
            conv.out.fill_(float("nan"))
            conv.forward(cfg)
            torch.cuda.synchronize()
            outs.append(conv.out.clone())
    for i, o in enumerate(outs[1:], 1):
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16)), f"run {i}"
