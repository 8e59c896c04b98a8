"""GPU parity tests of the FP8 (e4m3) compensated form, through the C-ABI, against the
oracle's bit-exact FP8 emulation (oracle/fp8.py).

Bars
  U and V records     : bit-exact against the oracle's quantisation and packing (codes,
                        scale bytes and padding; the buffers are poisoned first).
  GEMM (+ fused S4)   : within the FP32-accumulation bound of the oracle's FP64 three-pass
                        product on the same records:  |y - y*| <= 2^-11 (|y*| + e) + e + 2^-25,
                        e = (3 * 64 * CQ + 16) 2^-24 * sum_pos |coef| sum_passes |U_a|.|V_b|.
  end to end          : normwise max relative error <= 1e-2 against the FP64 direct conv.
  tile configurations : bit-identical output across configurations.
"""
import numpy as np
import pytest
import torch

from oracle import cdirect
from oracle import fp8
from oracle import winograd as wg

pytestmark = pytest.mark.gpu

SMALL = [
    # n, c, h, w, k, p, q
    (1, 64, 56, 56, 64, 1, 1),    # one 64-channel record per position
    (2, 3, 7, 7, 5, 1, 1),        # C=3: an odd channel-block count (zero second half)
    (1, 33, 9, 6, 17, 1, 1),      # two channel blocks, the second nearly empty
    (3, 16, 5, 5, 16, 1, 1),      # ragged, tiles straddle images
    (2, 96, 15, 15, 96, 1, 1),    # three channel blocks (odd), ragged
    (1, 8, 5, 7, 8, 2, 2),        # pad 2
    (1, 130, 12, 12, 140, 1, 1),  # three 64-channel records, K > 128
    (1, 256, 14, 14, 256, 1, 1),  # deeper reduction
]


def _ids(s):
    return "x".join(map(str, s))


def _log(msg):
    print(msg, flush=True)


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _oracle_records(x, wt, geo):
    qu = fp8.quantize_U(wg.filter_transform_fp32(wt))
    qv = fp8.quantize_V(fp8.input_transform_f32(x, geo))
    return qu, qv


@pytest.mark.parametrize("shape", SMALL, ids=_ids)
def test_fp8_records_and_gemm(shape, runner, cuda):
    n, c, h, w, k, p, q = shape
    _log(f"case {shape}")
    geo = wg.Geometry(n, c, h, w, k, p, q)
    x = wg.uniform_pm1(300 + c, (n, c, h, w))
    wt = wg.uniform_pm1(307 + c, (k, c, 3, 3))
    conv = runner.WinogradConvFP8(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    conv.V.fill_(0x55)
    conv.U.fill_(0x55)
    conv.prepare_weights()
    conv.input_transform()
    torch.cuda.synchronize()
    qu, qv = _oracle_records(x, wt, geo)
    u_ref = fp8.pack_operand(qu, geo.KB32)
    v_ref = fp8.pack_operand(qv, geo.G32)
    u_gpu = conv.u8_bytes().cpu().numpy()
    v_gpu = conv.v8_bytes().cpu().numpy()
    assert u_gpu.size == u_ref.size and v_gpu.size == v_ref.size
    bad_u = np.nonzero(u_gpu != u_ref)[0]
    bad_v = np.nonzero(v_gpu != v_ref)[0]
    assert bad_u.size == 0, f"U records differ at {bad_u.size} bytes, first offset {bad_u[:4]}"
    assert bad_v.size == 0, f"V records differ at {bad_v.size} bytes, first offset {bad_v[:4]}"

    conv.out.fill_(float("nan"))
    conv.gemm()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy().astype(np.float64)
    y_model = fp8.gemm_output_fp8(qu, qv, geo)
    scale = fp8.gemm_output_fp8(qu, qv, geo, absolute=True)
    cq = (c + 63) // 64
    e = (3 * 64 * cq + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"

    ref = cdirect.conv2d_f64(x, wt, pad=(p, q))
    m = wg.error_metrics(y, ref)
    _log(f"  fp8 max_rel={m['max_rel']:.3e} rms_rel={m['rms_rel']:.3e}")
    assert m["finite"] and m["max_rel"] <= 1e-2


@pytest.mark.parametrize("shape", [SMALL[0], SMALL[3], SMALL[6]], ids=_ids)
def test_fp8_tile_configs_bit_identical(shape, runner, wc, cuda):
    n, c, h, w, k, p, q = shape
    _log(f"case {shape}")
    x = wg.uniform_pm1(400 + c, (n, c, h, w))
    wt = wg.uniform_pm1(407 + c, (k, c, 3, 3))
    conv = runner.WinogradConvFP8(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    conv.prepare_weights()
    conv.input_transform()
    outs = []
    for cfg in wc.fp8_tile_configs():
        conv.out.fill_(float("nan"))
        conv.gemm(cfg["cfg"])
        torch.cuda.synchronize()
        outs.append(conv.out.clone())
    for i, o in enumerate(outs[1:], 1):
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16)), f"config {i}"


def test_fp8_forward_equals_stages(runner, cuda):
    shape = SMALL[4]
    n, c, h, w, k, p, q = shape
    x = wg.uniform_pm1(501, (n, c, h, w))
    wt = wg.uniform_pm1(502, (k, c, 3, 3))
    conv = runner.WinogradConvFP8(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    conv.prepare_weights()
    conv.input_transform()
    conv.gemm()
    torch.cuda.synchronize()
    a = conv.out.clone()
    conv.out.fill_(float("nan"))
    conv.forward()
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int16), conv.out.view(torch.int16))


def test_fp8_large_outliers_vs_fp64_and_fp16(runner, cuda):
    """N=64 C=512 28x28 K=512 with 1% of inputs scaled by 50: the FP8 and FP16 paths on the
    same data against the FP64 direct convolution, plus a bit-exact slice of the V records."""
    n, c, h, w, k = 64, 512, 28, 28, 512
    _log(f"case N{n} C{c} {h}x{w} K{k} outliers")
    x = wg.with_outliers(wg.normal(61, (n, c, h, w)), 62)
    wt = wg.normal(63, (k, c, 3, 3), sigma=float(np.sqrt(2.0 / (9 * c))))
    X, Wt = _to_dev(x, cuda), _to_dev(wt, cuda)
    conv = runner.WinogradConvFP8(X, Wt)
    conv.prepare_weights()
    conv.forward()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    y8 = conv.out.double()
    d8 = (y8 - ref).abs()
    m8 = ((d8.max() / ref.abs().max()).item(), (d8.norm() / ref.norm()).item())
    assert bool(torch.isfinite(y8).all())

    c16 = runner.WinogradConv(X, Wt)
    c16.prepare_weights()
    c16.forward()
    torch.cuda.synchronize()
    y16 = c16.out.double()
    d16 = (y16 - ref).abs()
    m16 = ((d16.max() / ref.abs().max()).item(), (d16.norm() / ref.norm()).item())
    _log(f"  fp8 max_rel={m8[0]:.3e} rms_rel={m8[1]:.3e} | fp16 max_rel={m16[0]:.3e} rms_rel={m16[1]:.3e}")
    assert m8[0] <= 1e-2 and m16[0] <= 1e-2

    # V records of image 0: tiles 0..191 (six whole 32-column blocks), all positions / blocks
    geo1 = wg.Geometry(1, c, h, w, k)
    qv = fp8.quantize_V(fp8.input_transform_f32(x[:1], geo1))
    c1, s1 = fp8.split_regions(fp8.pack_operand(qv, geo1.G32), c, geo1.G32)
    G32 = (n * geo1.T + 31) // 32
    cg, sg = fp8.split_regions(conv.v8_bytes().cpu().numpy(), c, G32)
    assert np.array_equal(cg[:, :, :6], c1[:, :, :6]) and np.array_equal(sg[:, :, :6], s1[:, :, :6])
