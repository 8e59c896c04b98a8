"""Pins the oracle (CPU only): every oracle function is checked against something other
than its own formula — closed forms, an independent library (torch, via committed
fixtures), the reference's own checker, brute force on small tiles, or invariants."""
import os

import numpy as np
import pytest

from oracle import cdirect
from oracle import winograd as wg

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- FP16 rounding ----------------------------------------------------------------------
def test_round_half_matches_numpy_rne():
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 5, 20000),
        # exact ties between FP16 neighbours must go to even
        (np.arange(1, 2000) + 0.5) * 2.0 ** -10,
        np.array([65504.0, 65519.99, 65520.0, 1e6, -1e6, 2.0 ** -25, 3 * 2.0 ** -26, 0.0]),
    ])
    for x in xs:
        assert cdirect.round_half(x) == float(np.float16(x)) or (np.isinf(np.float16(x)) and np.isinf(cdirect.round_half(x)))


def test_splitmix64_known_answer():
    # SplitMix64 with state 0: the first output is 0xE220A8397B1DCDAF (Steele et al. 2014)
    assert int(wg.splitmix64(0, np.array([0], dtype=np.uint64))[0]) == 0xE220A8397B1DCDAF
    a = wg.uniform_pm1(5, (3, 4))
    assert a.dtype == np.float16 and np.array_equal(a, wg.uniform_pm1(5, (3, 4)))
    assert np.all(np.abs(a.astype(np.float64)) <= 1.0)
    z = wg.normal(3, (200000,)).astype(np.float64)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_outliers_fraction():
    x = np.ones(200000, dtype=np.float16)
    y = wg.with_outliers(x, 9)
    frac = np.count_nonzero(y == 50.0) / x.size
    assert 0.009 < frac < 0.011


# ---- direct convolution -------------------------------------------------------------------
def test_direct_all_ones_closed_form():
    # all-ones input and filter, pad 1: output = C * (valid taps) = 4C corners, 6C edges, 9C inside
    c = 7
    x = np.ones((1, c, 4, 5), dtype=np.float16)
    w = np.ones((2, c, 3, 3), dtype=np.float16)
    y = cdirect.conv2d_f64(x, w)
    taps = np.array([[4, 6, 6, 6, 4], [6, 9, 9, 9, 6], [6, 9, 9, 9, 6], [4, 6, 6, 6, 4]], dtype=np.float64)
    assert np.array_equal(y[0, 0], c * taps) and np.array_equal(y[0, 1], c * taps)


def test_direct_matches_torch_fixtures():
    d = np.load(os.path.join(GOLDEN, "torch_conv_f64.npz"))
    ncase = len([k for k in d.files if k.endswith("_shape")])
    assert ncase >= 7
    for i in range(ncase):
        n, c, h, w, k, p, q = d[f"case{i}_shape"]
        x = d[f"case{i}_x"].view(np.float16)
        wt = d[f"case{i}_w"].view(np.float16)
        y = cdirect.conv2d_f64(x, wt, pad=(int(p), int(q)))
        np.testing.assert_allclose(y, d[f"case{i}_y"], rtol=1e-12, atol=1e-12)


def test_direct_general_geometry_vs_brute_force():
    # stride 2, 5x5 filter, pad 2: independent numpy brute force on a tiny case
    x = wg.uniform_pm1(1, (1, 2, 7, 6))
    w = wg.uniform_pm1(2, (3, 2, 5, 5))
    y = cdirect.conv2d_f64(x, w, stride=(2, 2), pad=(2, 2))
    xp = np.pad(x.astype(np.float64), ((0, 0), (0, 0), (2, 2), (2, 2)))
    ref = np.zeros(y.shape)
    for k in range(3):
        for i in range(y.shape[2]):
            for j in range(y.shape[3]):
                ref[0, k, i, j] = np.sum(xp[0, :, 2 * i:2 * i + 5, 2 * j:2 * j + 5] * w[k].astype(np.float64))
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_direct_linearity_exact():
    # small integers: all sums exact in FP64, so conv(x1 + x2) == conv(x1) + conv(x2) bit for bit
    rng = np.random.default_rng(4)
    x1 = rng.integers(-8, 8, (2, 5, 6, 7)).astype(np.float16)
    x2 = rng.integers(-8, 8, (2, 5, 6, 7)).astype(np.float16)
    w = rng.integers(-4, 4, (3, 5, 3, 3)).astype(np.float16)
    assert np.array_equal(cdirect.conv2d_f64(x1 + x2, w), cdirect.conv2d_f64(x1, w) + cdirect.conv2d_f64(x2, w))


def test_threaded_direct_bit_identical():
    x = wg.normal(7, (2, 9, 11, 10))
    w = wg.normal(8, (13, 9, 3, 3), sigma=0.2)
    assert np.array_equal(cdirect.conv2d_f64(x, w), cdirect.conv2d_f64_threads(x, w, threads=4))


# ---- faithful restatement of the reference checker ----------------------------------------
def test_faithful_restatement_matches_reference_checker_fixtures():
    d = np.load(os.path.join(GOLDEN, "ref_checker.npz"))
    ncase = len([k for k in d.files if k.endswith("_shape")])
    assert ncase >= 4
    for i in range(ncase):
        n, c, h, w, k, r, s, u, v, p, q = (int(t) for t in d[f"case{i}_shape"])
        x = d[f"case{i}_x"].view(np.float16)
        wt = d[f"case{i}_w"].view(np.float16)
        y = cdirect.conv2dcpu_faithful(x, wt, stride=(u, v), pad=(p, q))
        assert np.array_equal(y.view(np.uint16), d[f"case{i}_y"]), f"case {i}"


@pytest.mark.skipif(cdirect.ref_lib() is None, reason="reference checker not built (no reference source here)")
def test_faithful_restatement_matches_reference_checker_live():
    for seed, shape in ((31, (1, 7, 6, 9, 5)), (32, (2, 3, 5, 5, 4))):
        n, c, h, w, k = shape
        x = wg.normal(seed, (n, c, h, w))
        wt = wg.normal(seed + 1, (k, c, 3, 3))
        assert np.array_equal(cdirect.conv2dcpu_faithful(x, wt).view(np.uint16),
                              cdirect.ref_conv2dcpu(x, wt).view(np.uint16))


def test_faithful_rounds_products():
    # (1 + 2^-10)^2 = 1 + 2^-9 + 2^-20: the FP16-rounded product is 1 + 2^-9 (verfiy.h:32)
    a = np.full((1, 1, 1, 1), 1 + 2 ** -10, dtype=np.float16)
    w = np.zeros((1, 1, 3, 3), dtype=np.float16)
    w[0, 0, 1, 1] = a[0, 0, 0, 0]
    y = cdirect.conv2dcpu_faithful(a, w)
    assert float(y[0, 0, 0, 0]) == 1 + 2 ** -9
    assert cdirect.conv2d_f64(a, w)[0, 0, 0, 0] == (1 + 2 ** -10) ** 2


# ---- Winograd model -------------------------------------------------------------------------
def _correlate_tile(d, g):
    y = np.zeros((2, 2))
    for i in range(2):
        for j in range(2):
            for r in range(3):
                for s in range(3):
                    y[i, j] += d[i + r, j + s] * g[r, s]
    return y


def test_winograd_identity_brute_force_exact():
    # A^T [(G g G^T) . (B^T d B)] A == 2x2 valid correlation, exactly, on dyadic inputs
    rng = np.random.default_rng(1)
    for _ in range(200):
        d = rng.integers(-64, 64, (4, 4)) / 8.0
        g = rng.integers(-64, 64, (3, 3)) / 16.0
        u = wg.G @ g @ wg.G.T
        v = wg.BT @ d @ wg.BT.T
        y = wg.AT @ (u * v) @ wg.AT.T
        assert np.array_equal(y, _correlate_tile(d, g))


def test_transform_functions_match_matrix_definition():
    rng = np.random.default_rng(2)
    w = rng.integers(-50, 50, (3, 4, 3, 3)).astype(np.float16) / np.float16(8)
    u = wg.filter_transform_f64(w)
    for k in range(3):
        for c in range(4):
            ref = wg.G @ w[k, c].astype(np.float64) @ wg.G.T
            assert np.array_equal(u[:, k, c].reshape(4, 4), ref)
    geo = wg.Geometry(1, 2, 5, 6, 1)
    x = rng.integers(-50, 50, (1, 2, 5, 6)).astype(np.float16)
    v = wg.input_transform_f64(x, geo)
    xp = np.zeros((1, 2, 2 * geo.tiles_h + 2, 2 * geo.tiles_w + 2))
    xp[:, :, 1:6, 1:7] = x
    for c in range(2):
        for th in range(geo.tiles_h):
            for tw in range(geo.tiles_w):
                d = xp[0, c, 2 * th:2 * th + 4, 2 * tw:2 * tw + 4]
                ref = wg.BT @ d @ wg.BT.T
                assert np.array_equal(v[:, c, th * geo.tiles_w + tw].reshape(4, 4), ref)


@pytest.mark.parametrize("shape", [
    (1, 64, 8, 8, 64, 1, 1), (2, 3, 7, 7, 5, 1, 1), (1, 33, 9, 6, 17, 1, 1),
    (1, 8, 6, 6, 8, 0, 0), (1, 8, 5, 7, 8, 2, 2), (3, 16, 5, 5, 16, 1, 1)])
def test_winograd_f64_equals_direct(shape):
    n, c, h, w, k, p, q = shape
    geo = wg.Geometry(n, c, h, w, k, p, q)
    x = wg.uniform_pm1(40 + n, (n, c, h, w))
    wt = wg.uniform_pm1(50 + c, (k, c, 3, 3))
    y = wg.winograd_f64(x, wt, geo)
    ref = cdirect.conv2d_f64(x, wt, pad=(p, q))
    assert y.shape == ref.shape
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_fp32_transforms_within_one_ulp_of_exact():
    w = wg.normal(3, (64, 48, 3, 3), sigma=0.3)
    u32 = wg.filter_transform_fp32(w).astype(np.float64)
    u64 = wg.filter_transform_f64(w).astype(np.float16).astype(np.float64)
    ulp = np.spacing(np.abs(u64).astype(np.float16)).astype(np.float64)
    assert np.all(np.abs(u32 - u64) <= ulp)
    assert np.mean(u32 == u64) > 0.999
    geo = wg.Geometry(2, 20, 11, 9, 4)
    x = wg.normal(4, (2, 20, 11, 9))
    v32 = wg.input_transform_fp32(x, geo).astype(np.float64)
    v64 = wg.input_transform_f64(x, geo).astype(np.float16).astype(np.float64)
    ulp = np.spacing(np.abs(v64).astype(np.float16)).astype(np.float64)
    assert np.all(np.abs(v32 - v64) <= ulp)
    assert np.mean(v32 == v64) > 0.999


def test_fp16_input_transform_exact_on_small_integers():
    """With integer inputs every intermediate of the fused form's FP16 transform is a small
    integer, so it must equal the exact B^T d B: a wrong row, column or sign fails."""
    geo = wg.Geometry(2, 9, 11, 12, 4, 1, 1)
    x = np.random.default_rng(8).integers(-6, 7, (2, 9, 11, 12)).astype(np.float16)
    assert np.array_equal(wg.input_transform_fp16(x, geo).astype(np.float64), wg.input_transform_f64(x, geo))
    for p in (0, 2):
        geo = wg.Geometry(1, 5, 9, 8, 3, p, p)
        x = np.random.default_rng(9 + p).integers(-6, 7, (1, 5, 9, 8)).astype(np.float16)
        assert np.array_equal(wg.input_transform_fp16(x, geo).astype(np.float64), wg.input_transform_f64(x, geo))


def test_fp16_input_transform_within_two_roundings():
    """Two FP16 roundings (s, then V): |V - V*| <= 2^-10 (1 + 2^-12) (|B^T| |d| |B|) + 2 * 2^-24,
    and the bound is not loose: some values sit more than one ulp away from the exact V
    (so the transform really rounds twice, unlike the FP32-order form)."""
    geo = wg.Geometry(2, 20, 11, 10, 4)
    x = wg.normal(14, (2, 20, 11, 10))
    v16 = wg.input_transform_fp16(x, geo).astype(np.float64)
    v64 = wg.input_transform_f64(x, geo)
    d = np.abs(wg.input_windows(x, geo).astype(np.float64))
    scale = wg._to_g(np.einsum("ai,nctsij,bj->abncts", np.abs(wg.BT), d, np.abs(wg.BT)).reshape(16, *d.shape[:4]), geo)
    assert np.all(np.abs(v16 - v64) <= 2.0 ** -10 * (1 + 2.0 ** -12) * scale + 2.0 ** -23)
    one_ulp = np.spacing(np.abs(v64).astype(np.float16)).astype(np.float64)
    assert np.any(np.abs(v16 - v64) > one_ulp)
    assert np.mean(v16 == v64.astype(np.float16).astype(np.float64)) > 0.5


def test_numpy_float16_add_is_correctly_rounded():
    """The oracle's FP16 arithmetic: ties go to even, and a sum needing more than 24 bits of
    range still rounds once (1 + 2^-11 -> 1, 1 + 3*2^-11 -> 1 + 2^-9, 2048 + 1 -> 2048,
    2048 + 3 -> 2052, 65504 + 2^-24 -> 65504); x - x is +0."""
    f = np.float16
    assert f(1) + f(2.0 ** -11) == f(1)
    assert f(1) + f(3 * 2.0 ** -11) == f(1 + 2.0 ** -9)
    assert f(2048) + f(1) == f(2048)
    assert f(2048) + f(3) == f(2052)
    assert f(65504) + f(2.0 ** -24) == f(65504)
    assert f(2.0 ** -24) - f(2.0 ** -24) == f(0) and not np.signbit(f(2.0 ** -24) - f(2.0 ** -24))


def test_pack_roundtrip_and_known_offsets():
    geo = wg.Geometry(2, 40, 7, 9, 20)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((16, geo.k, geo.c)).astype(np.float16)
    v = rng.standard_normal((16, geo.c, geo.NT)).astype(np.float16)
    pu, pv = wg.pack_U(u, geo), wg.pack_V(v, geo)
    assert pu.size == geo.u_halves and pv.size == geo.v_halves
    assert np.array_equal(wg.unpack_U(pu, geo), u) and np.array_equal(wg.unpack_V(pv, geo), v)
    # hand-computed offsets: layout [pos][c16][blk][lane][e], lane = (c%16//8)*32 + r%32
    C16 = 2 * geo.CB
    pos, k, c = 3, 17, 38
    off = ((pos * C16 + c // 16) * geo.KB32 + k // 32) * 512 + (((c % 16) // 8) * 32 + k % 32) * 8 + c % 8
    assert pu[off] == u[pos, k, c]
    pos, g, c = 14, 37, 9
    off = ((pos * C16 + c // 16) * geo.G32 + g // 32) * 512 + (((c % 16) // 8) * 32 + g % 32) * 8 + c % 8
    assert pv[off] == v[pos, c, g]
    # padding (channels >= C, rows >= K, columns >= NT) is zero
    assert np.count_nonzero(pu) == np.count_nonzero(u)
    assert np.count_nonzero(pv) == np.count_nonzero(v)


def test_fp16_method_error_is_small_cfg1():
    # the FP16 data path (U, V rounded to FP16) stays far inside 1e-2 on the small config
    geo = wg.Geometry(1, 64, 56, 56, 64)
    x = wg.uniform_pm1(1, (1, 64, 56, 56))
    w = wg.uniform_pm1(2, (64, 64, 3, 3))
    y, _ = wg.winograd_gpu_model(x, w, geo)
    ref = cdirect.conv2d_f64_threads(x, w)
    m = wg.error_metrics(y.astype(np.float16), ref)
    assert m["finite"] and m["max_rel"] < 2e-3 and m["rms_rel"] < 1e-3


def test_error_metrics_definitions():
    ref = np.array([1.0, -2.0, 4.0])
    y = np.array([1.5, -2.0, 4.0])
    m = wg.error_metrics(y, ref)
    assert m["max_rel"] == 0.5 / 4.0
    assert abs(m["rms_rel"] - 0.5 / np.sqrt(21.0)) < 1e-15
    # the reference's elementwise criterion (main.cpp:144) flags > 1% with the GPU value as denominator
    assert wg.reference_criterion_failures(np.array([1.0, 1.02], dtype=np.float16), np.array([1.0, 1.0])) == 1


# ---- general-geometry (implicit-GEMM direct) path ------------------------------------------
from oracle import direct as dg  # noqa: E402

DIRECT_GEOS = [
    dg.DirectGeometry(2, 5, 9, 8, 7, 5, 5, 1, 1, 2, 2),     # 5x5 pad 2
    dg.DirectGeometry(1, 40, 11, 13, 33, 3, 3, 2, 2, 1, 1),  # stride 2, C and K ragged
    dg.DirectGeometry(2, 3, 15, 15, 8, 7, 7, 2, 2, 3, 3),    # stem-like 7x7 stride 2
    dg.DirectGeometry(1, 16, 6, 7, 16, 1, 1, 1, 1, 0, 0),    # 1x1
    dg.DirectGeometry(1, 8, 7, 5, 4, 3, 1, 1, 1, 1, 0),      # 3x1
    dg.DirectGeometry(1, 8, 8, 8, 8, 3, 3, 1, 1, 3, 3),      # pad 3 (> 2)
    dg.DirectGeometry(1, 70, 6, 6, 5, 2, 2, 2, 2, 0, 0),     # 2x2 stride 2, three channel chunks
]


@pytest.mark.parametrize("geo", DIRECT_GEOS, ids=lambda g: f"{g.r}x{g.s}s{g.u}p{g.p}c{g.c}")
def test_direct_packing_reproduces_the_direct_conv(geo):
    """The implicit GEMM over the packed operands (filter taps x im2col) is the convolution:
    exact on integer-valued data against the C direct-convolution oracle."""
    rng = np.random.default_rng(geo.c + 7 * geo.r)
    x = rng.integers(-3, 4, (geo.n, geo.c, geo.h, geo.w)).astype(np.float16)
    w = rng.integers(-2, 3, (geo.k, geo.c, geo.r, geo.s)).astype(np.float16)
    y = dg.gemm_output_f64(w, x, geo)
    ref = cdirect.conv2d_f64(x, w, stride=(geo.u, geo.v), pad=(geo.p, geo.q))
    assert y.shape == ref.shape and np.array_equal(y, ref)


def test_direct_padding_is_positive_zero():
    """Taps outside the image are +0 (the kernel's zero), also next to negative inputs."""
    geo = DIRECT_GEOS[0]
    x = -np.ones((geo.n, geo.c, geo.h, geo.w), dtype=np.float16)
    px = dg.pack_input(x, geo).view(np.uint16)
    assert not np.any(px == 0x8000)


def test_direct_packed_offsets():
    geo = DIRECT_GEOS[1]
    rng = np.random.default_rng(1)
    x = rng.standard_normal((geo.n, geo.c, geo.h, geo.w)).astype(np.float16)
    w = rng.standard_normal((geo.k, geo.c, geo.r, geo.s)).astype(np.float16)
    pw, px = dg.pack_weights(w, geo), dg.pack_input(x, geo)
    assert pw.size == geo.nsteps * 2 * geo.KB32 * 512 and px.size == geo.nsteps * 2 * geo.P32 * 512
    # C = 40: tap t covers j = 40t .. 40t+39, so taps share 16-value blocks (j = 79 | 80)
    for (k, c, ri, si) in [(0, 0, 0, 0), (32, 39, 2, 1), (17, 33, 1, 2), (5, 39, 0, 1), (6, 0, 0, 2)]:
        j = (ri * geo.s + si) * geo.c + c
        L, lane = j // 16, (j % 16 // 8) * 32 + k % 32
        assert pw[(L * geo.KB32 + k // 32) * 512 + lane * 8 + j % 8] == w[k, c, ri, si]
    for (n, c, oh, ow, ri, si) in [(0, 5, 2, 3, 1, 1), (0, 35, 5, 6, 2, 2), (0, 0, 0, 0, 0, 0), (0, 39, 1, 1, 0, 1)]:
        j = (ri * geo.s + si) * geo.c + c
        g = (n * geo.oh + oh) * geo.ow + ow
        L, lane = j // 16, (j % 16 // 8) * 32 + g % 32
        ih, iw = oh * geo.u - geo.p + ri, ow * geo.v - geo.q + si
        want = x[n, c, ih, iw] if (0 <= ih < geo.h and 0 <= iw < geo.w) else 0
        assert px[(L * geo.P32 + g // 32) * 512 + lane * 8 + j % 8] == want
    # the reduction tail j >= r*s*C is zero in both operands
    jn = geo.r * geo.s * geo.c
    tail_w = pw.reshape(geo.nsteps * 2, geo.KB32, 2, 32, 8).transpose(1, 3, 0, 2, 4).reshape(geo.KB32 * 32, -1)
    assert jn < tail_w.shape[1] and not np.any(tail_w[:, jn:])


def test_direct_few_channel_layers_pack_densely():
    """A 7x7 stem over 3 channels reduces over 147 values: 5 steps of 32, padded to 8, not one
    32-channel step per tap."""
    geo = dg.DirectGeometry(1, 3, 16, 16, 64, 7, 7, 2, 2, 3, 3)
    assert geo.nsteps == 8
    assert dg.DirectGeometry(1, 64, 8, 8, 8, 3, 3).nsteps == 24  # 576 values = 18 steps -> 24
    assert dg.DirectGeometry(1, 256, 8, 8, 8, 1, 1).nsteps == 8
