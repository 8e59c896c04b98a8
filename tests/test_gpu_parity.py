"""GPU parity tests: the gfx950 kernels, called through the C-ABI, against the oracle.

Bars
  S1 U, S2 V            : bit-exact against the oracle's FP32-order model, packed layout
                          included (padding zeros too).
  S3 (+ fused S4) output: within the FP32-accumulation bound of the oracle's FP64 GEMM on the
                          same FP16 U and V:  |y - y*| <= 2^-11 (|y*| + e) + e + 2^-25,
                          e = (C + 16) 2^-24 * sum_pos |coef| |U|.|V|  (per element).
  end to end            : normwise max relative error <= 1e-2 against the FP64 direct conv
                          (the tolerance the task states); RMS and max are printed.
  tile configurations   : every configuration gives bit-identical output (same per-element
                          summation order), and the four-slot protocol equals the stage API.
  FP64 validation kernel: bit-identical to the CPU FP64 oracle (same order, exact products).
"""
import numpy as np
import pytest
import torch

from oracle import cdirect
from oracle import winograd as wg

pytestmark = pytest.mark.gpu

SMALL = [
    # n, c, h, w, k, p, q
    (1, 64, 56, 56, 64, 1, 1),   # the small single-layer configuration
    (2, 3, 7, 7, 5, 1, 1),       # C=3 (padded reduction), K=5
    (1, 33, 9, 6, 17, 1, 1),     # odd channel counts, non-square
    (1, 8, 6, 6, 8, 0, 0),       # pad 0
    (1, 8, 5, 7, 8, 2, 2),       # pad 2, ragged
    (3, 16, 5, 5, 16, 1, 1),     # ragged, tiles straddle images
    (2, 96, 15, 15, 96, 1, 1),   # ragged 96-channel
    (1, 40, 9, 9, 20, 1, 1),
    (1, 130, 12, 12, 140, 1, 1), # K > 128 (two k-blocks), C not multiple of 32
    (1, 16, 6, 300, 16, 1, 1),   # wide: one tile row of input staging needs 79 KiB of LDS
    (1, 40, 5, 627, 24, 2, 2),   # near the widest Winograd staging (158 KiB of LDS), pad 2
]


def _ids(s):
    return "x".join(map(str, s))


def _inputs(shape, seed):
    n, c, h, w, k, p, q = shape
    x = wg.uniform_pm1(seed, (n, c, h, w))
    wt = wg.uniform_pm1(seed + 7, (k, c, 3, 3))
    return x, wt


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape", SMALL, ids=_ids)
def test_small_shapes_stagewise(shape, runner, cuda):
    n, c, h, w, k, p, q = shape
    _log(f"case {shape}")
    geo = wg.Geometry(n, c, h, w, k, p, q)
    x, wt = _inputs(shape, 100 + c)
    X, Wt = _to_dev(x, cuda), _to_dev(wt, cuda)
    conv = runner.WinogradConv(X, Wt, pad=(p, q))
    conv.V.fill_(0x55)  # poison: every byte of V and U must be written
    conv.U.fill_(0x55)
    conv.prepare_weights()
    conv.input_transform()
    torch.cuda.synchronize()
    u_ref = wg.pack_U(wg.filter_transform_fp32(wt), geo)
    v_ref = wg.pack_V(wg.input_transform_fp32(x, geo), geo)
    u_gpu = conv.u_halves().cpu().numpy()
    v_gpu = conv.v_halves().cpu().numpy()
    assert np.array_equal(u_gpu.view(np.uint16), u_ref.view(np.uint16)), "U differs"
    assert np.array_equal(v_gpu.view(np.uint16), v_ref.view(np.uint16)), "V differs"

    conv.out.fill_(float("nan"))
    conv.gemm()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy().astype(np.float64)
    y_model, scale = wg.winograd_gpu_model(x, wt, geo)
    e = (c + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"

    ref = cdirect.conv2d_f64(x, wt, pad=(p, q))
    m = wg.error_metrics(y, ref)
    _log(f"  max_rel={m['max_rel']:.3e} rms_rel={m['rms_rel']:.3e}")
    assert m["finite"] and m["max_rel"] <= 1e-2


@pytest.mark.parametrize("shape", [SMALL[0], SMALL[5], SMALL[8]], ids=_ids)
def test_all_tile_configs_bit_identical(shape, runner, wc, cuda):
    n, c, h, w, k, p, q = shape
    _log(f"case {shape}")
    x, wt = _inputs(shape, 200 + c)
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda), pad=(p, q))
    conv.prepare_weights()
    conv.input_transform()
    outs = []
    for cfg in wc.tile_configs():
        conv.out.fill_(float("nan"))
        conv.gemm(cfg["cfg"])
        torch.cuda.synchronize()
        outs.append(conv.out.clone())
    for i, o in enumerate(outs[1:], 1):
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16)), f"config {i}"


def test_reference_protocol_slots_equal_stage_api(runner, cuda):
    shape = (2, 96, 15, 15, 96, 1, 1)
    _log(f"case {shape}")
    x, wt = _inputs(shape, 300)
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda))
    conv.launch_slots()
    torch.cuda.synchronize()
    a = conv.out.clone()
    conv.out.fill_(float("nan"))
    conv.prepare_weights()
    conv.forward()
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int16), conv.out.view(torch.int16))
    # repeated launches are idempotent
    conv.launch_slots()
    conv.launch_slots()
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int16), conv.out.view(torch.int16))


def test_positive_data_passes_reference_criterion(runner, cuda):
    # the reference harness's own data range and acceptance test (main.cpp:40-48, :140-151)
    n, c, h, w, k = 1, 64, 56, 56, 64
    idx_x = np.arange(n * c * h * w, dtype=np.uint64)
    idx_w = np.arange(k * c * 9, dtype=np.uint64)
    x = ((wg.splitmix64(1, idx_x) % np.uint64(255)).astype(np.float64) / 255.0).astype(np.float16).reshape(n, c, h, w)
    wt = ((wg.splitmix64(2, idx_w) % np.uint64(255)).astype(np.float64) / 255.0).astype(np.float16).reshape(k, c, 3, 3)
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda))
    conv.launch_slots()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy()
    ref_checker = cdirect.conv2dcpu_faithful(x, wt).astype(np.float64)
    assert wg.reference_criterion_failures(y, ref_checker) == 0


@pytest.mark.parametrize("geom", [
    ((2, 5, 9, 8), (4, 5, 3, 3), (1, 1), (1, 1)),
    ((1, 4, 9, 9), (3, 4, 5, 5), (2, 2), (2, 2)),
    ((1, 6, 7, 10), (2, 6, 1, 1), (1, 1), (0, 0)),
    ((2, 3, 11, 7), (5, 3, 3, 3), (2, 1), (0, 1)),
], ids=["3x3", "5x5s2p2", "1x1", "s2x1"])
def test_direct_f64_kernel_bit_identical_to_cpu(geom, runner, cuda):
    xs, ws, stride, pad = geom
    _log(f"case {geom}")
    x = wg.normal(400, xs)
    wt = wg.normal(401, ws)
    y = runner.direct_f64(_to_dev(x, cuda), _to_dev(wt, cuda), stride=stride, pad=pad).cpu().numpy()
    assert np.array_equal(y, cdirect.conv2d_f64(x, wt, stride=stride, pad=pad))


# ---- full-size configurations: GPU FP64 oracle + sampled CPU points + properties ----------
def _sample_check(x, wt, y64, rng, count=64, pad=(1, 1)):
    """Pins the GPU FP64 kernel at full size: recompute sampled outputs on the CPU."""
    n, c, h, w = x.shape
    k = wt.shape[0]
    oh, ow = y64.shape[2], y64.shape[3]
    for _ in range(count):
        b, kk, i, j = rng.integers(n), rng.integers(k), rng.integers(oh), rng.integers(ow)
        r0, c0 = i - pad[0], j - pad[1]
        rs, re = max(r0, 0), min(r0 + 3, h)
        cs, ce = max(c0, 0), min(c0 + 3, w)
        patch = x[b, :, rs:re, cs:ce].astype(np.float64)
        filt = wt[kk, :, rs - r0:re - r0, cs - c0:ce - c0].astype(np.float64)
        exact = float(np.sum(patch * filt))
        assert abs(exact - y64[b, kk, i, j]) <= 1e-9 * max(1.0, np.sum(np.abs(patch * filt)))


BIG = {
    "cfg2_N32_C256_56_K256": dict(shape=(32, 256, 56, 56, 256), dist="normal", outliers=False),
    "cfg2r_N32_C96_55_K96": dict(shape=(32, 96, 55, 55, 96), dist="normal", outliers=False),
    "cfg3_N64_C512_28_K512": dict(shape=(64, 512, 28, 28, 512), dist="normal", outliers=True),
}


@pytest.mark.parametrize("name", list(BIG))
def test_full_size_configs(name, runner, wc, cuda):
    spec = BIG[name]
    n, c, h, w, k = spec["shape"]
    _log(f"case {name}")
    x = wg.normal(1000 + c, (n, c, h, w))
    if spec["outliers"]:
        x = wg.with_outliers(x, 2000 + c)
    wt = wg.normal(3000 + c, (k, c, 3, 3), sigma=float(np.sqrt(2.0 / (9 * c))))
    X, Wt = _to_dev(x, cuda), _to_dev(wt, cuda)
    conv = runner.WinogradConv(X, Wt)
    conv.launch_slots()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    y = conv.out.double()
    d = (y - ref).abs()
    max_rel = (d.max() / ref.abs().max()).item()
    rms_rel = (d.norm() / ref.norm()).item()
    _log(f"  {name}: max_rel={max_rel:.3e} rms_rel={rms_rel:.3e} finite={bool(torch.isfinite(y).all())}")
    assert bool(torch.isfinite(y).all())
    assert max_rel <= 1e-2 and rms_rel <= 1e-2
    # pin the GPU FP64 kernel at this size on sampled points
    _sample_check(x, wt, ref.cpu().numpy(), np.random.default_rng(5))
    # U/V of the first image bit-exact against the oracle (size-independent slice); the fused
    # form does not write V, so the input transform runs on its own for the check
    if conv.form() == wc.FORM_FUSED:
        conv.input_transform()
        torch.cuda.synchronize()
    geo1 = wg.Geometry(1, c, h, w, k)
    geo = wg.Geometry(n, c, h, w, k)
    v_all = wg.unpack_V(conv.v_halves().cpu().numpy(), geo)
    v0 = wg.input_transform_fp32(x[:1], geo1)
    assert np.array_equal(v_all[:, :, :geo.T].view(np.uint16), v0.view(np.uint16))
    u_all = wg.unpack_U(conv.u_halves().cpu().numpy(), geo)
    assert np.array_equal(u_all.view(np.uint16), wg.filter_transform_fp32(wt).view(np.uint16))
    # determinism: a second run is bit-identical
    first = conv.out.clone()
    conv.launch_slots()
    torch.cuda.synchronize()
    assert torch.equal(first.view(torch.int16), conv.out.view(torch.int16))


@pytest.mark.parametrize("name", list(BIG))
def test_full_size_all_ones_closed_form(name, runner, cuda):
    # x = 1, w = 2^-10: every transform, product and sum is an exact dyadic, so the output
    # must equal C * taps * 2^-10 bit for bit (taps = 4 corner, 6 edge, 9 interior, pad 1)
    n, c, h, w, k = BIG[name]["shape"]
    _log(f"case {name} all-ones")
    X = torch.ones((n, c, h, w), dtype=torch.float16, device=cuda)
    Wt = torch.full((k, c, 3, 3), 2.0 ** -10, dtype=torch.float16, device=cuda)
    conv = runner.WinogradConv(X, Wt)
    conv.launch_slots()
    torch.cuda.synchronize()
    rows = torch.full((h,), 3.0, dtype=torch.float64)
    rows[0] = rows[-1] = 2.0
    cols = torch.full((w,), 3.0, dtype=torch.float64)
    cols[0] = cols[-1] = 2.0
    expect = (rows[:, None] * cols[None, :]) * c * 2.0 ** -10
    out = conv.out.double().cpu()
    assert torch.equal(out, expect.expand(n, k, h, w))


def test_autotuner_gates_times_and_records(runner, wc, cuda):
    shape = (2, 96, 15, 15, 96, 1, 1)
    _log(f"case {shape} autotune")
    x, wt = _inputs(shape, 500)
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda))
    conv.L.conv2d_tuned_clear()
    conv.prepare_weights()
    conv.input_transform()
    best, times = conv.autotune(reps=3)
    assert all(t > 0 for t in times), times  # every configuration passed the parity gate
    assert times[best] == min(times)
    assert conv.selected_config() == best and conv.L.conv2d_tuned_count() == 1
    # the tuned configuration is the one getkernelInfo now puts in slot 3
    conv._ki = None
    assert conv.param().cfg == best
    conv.L.conv2d_tuned_clear()


def test_autotuner_keeps_the_faster_form(runner, wc, cuda):
    """On the N=32 C=256 56x56 K=256 layer the autotuner times the staged configurations, the
    input transform and the fused kernel; the form it records (and getkernelInfo then fills
    into the slots) is the fused one exactly when the fused kernel beat input transform + GEMM."""
    n, c, h, k = 32, 256, 56, 256
    _log("case autotune form 32x256x56x256")
    g = torch.Generator(device=cuda)
    g.manual_seed(78)
    x = torch.randn((n, c, h, h), generator=g, device=cuda).half()
    w = (torch.randn((k, c, 3, 3), generator=g, device=cuda) * (2.0 / (9 * c)) ** 0.5).half()
    conv = runner.WinogradConv(x, w)
    conv.L.conv2d_tuned_clear()
    conv.prepare_weights()
    conv.input_transform()
    best, times, t_s2, t_fused = conv.autotune(reps=3, forms=True)
    _log(f"  best cfg {best} gemm {times[best]:.1f} us, input transform {t_s2:.1f} us, fused {t_fused:.1f} us")
    assert all(t > 0 for t in times) and t_s2 > 0 and t_fused > 0
    fused_faster = t_fused < t_s2 + times[best]
    assert conv.L.conv2d_tuned_count() == 1
    assert conv.form() == (wc.FORM_FUSED if fused_faster else wc.FORM_STAGED)
    conv._ki = None
    assert conv.param().form == conv.form()
    # the slots run the recorded form and meet the accuracy bound
    conv.out.fill_(float("nan"))
    conv.launch_slots()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    d = (conv.out.double() - ref).abs()
    assert (d.max() / ref.abs().max()).item() <= 1e-2
    conv.L.conv2d_tuned_clear()


# ---- multi-layer (batched) form --------------------------------------------------------------
BATCH_SMALL = [(1, 3, 7, 5), (4, 32, 13, 64), (2, 64, 14, 32), (1, 96, 27, 128), (2, 128, 7, 3),
               (1, 40, 9, 20), (3, 130, 12, 140), (1, 16, 28, 16)]


def _batch_convs(runner, cuda, shapes, seed=600):
    convs = []
    for i, (n, c, h, k) in enumerate(shapes):
        x, wt = _inputs((n, c, h, h, k, 1, 1), seed + i)
        convs.append(runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda)))
    return convs


def test_batched_equals_separate_and_graph_replays(runner, cuda):
    _log("case batched small")
    convs = _batch_convs(runner, cuda, BATCH_SMALL)
    b = runner.Batch(convs)
    b.prepare_weights()
    b.forward_separate()
    torch.cuda.synchronize()
    ref = [c.out.clone() for c in convs]
    for c in convs:
        c.out.fill_(float("nan"))
        c.V.fill_(0x55)
    b.forward()
    torch.cuda.synchronize()
    for i, c in enumerate(convs):
        assert torch.equal(c.out.view(torch.int16), ref[i].view(torch.int16)), f"layer {i}"
    # per-layer accuracy against the FP64 direct conv
    for i, c in enumerate(convs):
        y64 = c.direct_f64()
        d = (c.out.double() - y64).abs()
        assert (d.max() / y64.abs().max()).item() <= 1e-2, f"layer {i}"
    # hipGraph: capture once, replay several times with eager work in between
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.capture_graph(s)
    for rep in range(3):
        for c in convs:
            c.out.fill_(float("nan"))
        torch.cuda.synchronize()
        b.launch_graph(s)
        s.synchronize()
        for i, c in enumerate(convs):
            assert torch.equal(c.out.view(torch.int16), ref[i].view(torch.int16)), f"replay {rep} layer {i}"
        _ = torch.ones(1000, device=cuda).sum().item()  # eager work between replays
    b.close()


def test_batched_autotune_keeps_results(runner, wc, cuda):
    """After autotuning (tile configurations, and fused or staged per layer) the batched form
    still equals separate launches in the same forms, and layers whose form did not change
    keep their bits (every staged configuration gives the same output)."""
    _log("case batched autotune")
    convs = _batch_convs(runner, cuda, BATCH_SMALL[:4], seed=700)
    b = runner.Batch(convs)
    b.prepare_weights()
    b.forward_separate()
    torch.cuda.synchronize()
    ref = [c.out.clone() for c in convs]
    forms0 = b.layer_forms()
    b.autotune(reps=2)
    cfgs, forms = b.layer_configs(), b.layer_forms()
    _log(f"  configs {cfgs} forms {forms0} -> {forms}")
    assert all(0 <= c < len(wc.tile_configs()) for c in cfgs)
    assert all(f in (wc.FORM_STAGED, wc.FORM_FUSED) for f in forms)
    b.forward()
    torch.cuda.synchronize()
    got = [c.out.clone() for c in convs]
    b.forward_separate()
    torch.cuda.synchronize()
    for i, c in enumerate(convs):
        assert torch.equal(got[i].view(torch.int16), c.out.view(torch.int16)), f"layer {i}"
        if forms[i] == forms0[i]:
            assert torch.equal(got[i].view(torch.int16), ref[i].view(torch.int16)), f"layer {i} changed bits"
        y64 = c.direct_f64()
        assert ((got[i].double() - y64).abs().max() / y64.abs().max()).item() <= 1e-2, f"layer {i}"
    b.close()


def test_batched_mixes_fused_and_staged_layers(runner, wc, cuda):
    """A batch with both forms (small, odd-width and K <= 384 layers fused; an 8x256x56x56
    K=512 layer, N*T*K*C = 8.2e8 with C < K, staged) equals separate launches, in eager and
    graph launches."""
    _log("case batched mixed forms")
    shapes = [(1, 64, 14, 64), (8, 256, 56, 512), (1, 96, 27, 128), (3, 130, 13, 140), (1, 16, 28, 16)]
    convs = _batch_convs(runner, cuda, shapes, seed=800)
    b = runner.Batch(convs)
    forms = b.layer_forms()
    _log(f"  forms {forms}")
    assert wc.FORM_FUSED in forms and wc.FORM_STAGED in forms
    b.prepare_weights()
    b.forward_separate()
    torch.cuda.synchronize()
    ref = [c.out.clone() for c in convs]
    for c in convs:
        c.out.fill_(float("nan"))
    b.forward()
    torch.cuda.synchronize()
    for i, c in enumerate(convs):
        assert torch.equal(c.out.view(torch.int16), ref[i].view(torch.int16)), f"layer {i}"
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.capture_graph(s)
    for rep in range(2):
        for c in convs:
            c.out.fill_(float("nan"))
        torch.cuda.synchronize()
        b.launch_graph(s)
        s.synchronize()
        for i, c in enumerate(convs):
            assert torch.equal(c.out.view(torch.int16), ref[i].view(torch.int16)), f"replay {rep} layer {i}"
    b.close()
