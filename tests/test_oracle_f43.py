"""CPU pins of the F(4x4,3x3) oracle (oracle/winograd43.py)."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import cdirect
from oracle import winograd as wg
from oracle import winograd43 as w43


def _frac(mat):
    return [[Fr(x) for x in row] for row in mat]


def test_cook_toom_reproduces_lavin_gray_f4x4():
    """Lavin and Gray (CVPR 2016) publish F(4x4,3x3) for the points 0, +-1, +-2."""
    at, g, bt = w43.cook_toom((0, 1, -1, 2, -2), 4, 3)
    assert bt == _frac([[4, 0, -5, 0, 1, 0], [0, -4, -4, 1, 1, 0], [0, 4, -4, -1, 1, 0],
                        [0, -2, -1, 2, 1, 0], [0, 2, -1, -2, 1, 0], [0, 4, 0, -5, 0, 1]])
    assert g == [[Fr(1, 4), 0, 0], [Fr(-1, 6), Fr(-1, 6), Fr(-1, 6)], [Fr(-1, 6), Fr(1, 6), Fr(-1, 6)],
                 [Fr(1, 24), Fr(1, 12), Fr(1, 6)], [Fr(1, 24), Fr(-1, 12), Fr(1, 6)], [0, 0, 1]]
    assert at == _frac([[1, 1, 1, 1, 1, 0], [0, 1, -1, 2, -2, 0], [0, 1, 1, 4, 4, 0], [0, 1, -1, 8, -8, 1]])


def _mm(a, b):
    return [[sum(a[i][k] * b[k][j] for k in range(len(b))) for j in range(len(b[0]))] for i in range(len(a))]


def _t(a):
    return [list(r) for r in zip(*a)]


@pytest.mark.parametrize("mats", ["lavin", "chosen", "scaled"])
def test_winograd_identity_exact(mats):
    """A^T [(G g G^T) * (B^T d B)] A equals the 4x4 correlation of a 6x6 tile with a 3x3
    filter exactly (rationals), and the 1-D form too."""
    if mats == "lavin":
        at, g, bt = w43.cook_toom((0, 1, -1, 2, -2), 4, 3)
    elif mats == "chosen":
        at, g, bt = w43.cook_toom(w43.POINTS, 4, 3)
    else:
        at, g, bt = w43.AT43_EXACT, w43.G43_EXACT, w43.BT43_EXACT
    rng = np.random.default_rng(1)
    for _ in range(4):
        f = [[Fr(int(v)) for v in row] for row in rng.integers(-9, 10, (3, 3))]
        d = [[Fr(int(v), 8) for v in row] for row in rng.integers(-40, 41, (6, 6))]
        u = _mm(_mm(g, f), _t(g))
        v = _mm(_mm(bt, d), _t(bt))
        m = [[u[i][j] * v[i][j] for j in range(6)] for i in range(6)]
        y = _mm(_mm(at, m), _t(at))
        want = [[sum(f[r][s] * d[a + r][b + s] for r in range(3) for s in range(3)) for b in range(4)] for a in range(4)]
        assert y == want
        # 1-D: y_i = sum_k f_k d_{i+k}
        f1, d1 = f[0], d[0]
        u1 = [sum(g[i][k] * f1[k] for k in range(3)) for i in range(6)]
        v1 = [sum(bt[i][k] * d1[k] for k in range(6)) for i in range(6)]
        y1 = [sum(at[a][i] * u1[i] * v1[i] for i in range(6)) for a in range(4)]
        assert y1 == [sum(f1[k] * d1[a + k] for k in range(3)) for a in range(4)]


def test_scaled_matrices_are_dyadic():
    """G' and B^T have power-of-two denominators (the exactness of the FP64 transforms)."""
    for mat in (w43.G43_EXACT, w43.BT43_EXACT):
        for row in mat:
            for x in row:
                d = Fr(x).denominator
                assert d & (d - 1) == 0, x


def _wide_fp16(rng, shape):
    """FP16 values with exponents across the whole range (subnormals to 2^15)."""
    mant = rng.uniform(1, 2, shape) * rng.choice([-1, 1], shape)
    e = rng.integers(-24, 15, shape)
    return (mant * np.exp2(e)).astype(np.float16)


def test_fp64_transforms_are_exact():
    """The FP64 transforms equal the rational ones on inputs spanning the FP16 range."""
    rng = np.random.default_rng(2)
    w = _wide_fp16(rng, (2, 3, 3, 3))
    x = _wide_fp16(rng, (1, 2, 6, 10))
    geo = w43.Geometry43(1, 2, 6, 10, 2, 1, 1)
    u = w43.filter_transform43_f64(w)
    for k in range(2):
        for c in range(3):
            f = [[Fr(float(w[k, c, r, s])) for s in range(3)] for r in range(3)]
            ue = _mm(_mm(w43.G43_EXACT, f), _t(w43.G43_EXACT))
            got = u[:, k, c].reshape(6, 6)
            assert all(Fr(float(got[i, j])) == ue[i][j] for i in range(6) for j in range(6))
    v = w43.input_transform43_f64(x, geo)
    win = w43.input_windows43(x, geo)
    for c in range(2):
        for tw in range(geo.tiles_w):
            d = [[Fr(float(win[0, c, 0, tw, i, j])) for j in range(6)] for i in range(6)]
            ve = _mm(_mm(w43.BT43_EXACT, d), _t(w43.BT43_EXACT))
            got = v[:, c, tw].reshape(6, 6)
            assert all(Fr(float(got[i, j])) == ve[i][j] for i in range(6) for j in range(6))


SHAPES = [(1, 3, 9, 9, 5, 1, 1), (2, 8, 7, 13, 6, 0, 0), (1, 5, 10, 6, 7, 2, 2), (2, 4, 16, 16, 3, 1, 2),
          (1, 6, 5, 5, 4, 2, 1)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_exact_path_equals_direct_convolution(shape):
    n, c, h, w, k, p, q = shape
    geo = w43.Geometry43(n, c, h, w, k, p, q)
    x = wg.uniform_pm1(11, (n, c, h, w))
    wt = wg.uniform_pm1(12, (k, c, 3, 3))
    ref = cdirect.conv2d_f64(x, wt, pad=(p, q))
    y = w43.winograd43_exact(x, wt, geo)
    assert y.shape == ref.shape
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dist", ["uniform", "normal"])
def test_fp16_model_error(dist):
    """With FP16 U, V and M the normwise error stays well inside 1e-2 (measured 3.4e-3 and
    3.9e-3 on C = 256 layers)."""
    n, c, h, w, k = 1, 128, 12, 12, 8
    geo = w43.Geometry43(n, c, h, w, k)
    if dist == "uniform":
        x = wg.uniform_pm1(13, (n, c, h, w))
        wt = wg.uniform_pm1(14, (k, c, 3, 3))
    else:
        x = wg.normal(13, (n, c, h, w))
        wt = wg.normal(14, (k, c, 3, 3), sigma=float(np.sqrt(2 / (9 * c))))
    ref = cdirect.conv2d_f64(x, wt)
    e = np.abs(w43.winograd43_model(x, wt, geo) - ref).max() / np.abs(ref).max()
    assert e <= 6e-3, e


def test_packing_round_trips():
    geo = w43.Geometry43(2, 40, 9, 11, 70, 1, 1)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((36, geo.k, geo.c)).astype(np.float16)
    v = rng.standard_normal((36, geo.c, geo.NT)).astype(np.float16)
    pu, pv = w43.pack_U43(u, geo), w43.pack_V43(v, geo)
    assert pu.size == geo.u_halves and pv.size == geo.v_halves
    assert np.array_equal(w43.unpack_U43(pu, geo), u) and np.array_equal(w43.unpack_V43(pv, geo), v)
    m = np.zeros((36, geo.KB32 * 8, geo.G32 * 32, 4), dtype=np.float16)
    mk = rng.standard_normal((36, geo.k, geo.NT)).astype(np.float16)
    for kk in range(geo.k):
        m[:, kk // 4, :geo.NT, kk % 4] = mk[:, kk, :]
    assert np.array_equal(w43.unpack_M43(m.reshape(-1), geo), mk)


def test_float64_to_float16_conversion_rounds_to_nearest_even():
    """The oracle rounds exact FP64 transforms with numpy's float64 -> float16 conversion: one
    correct rounding (ties to even, subnormals, overflow), checked against a brute-force
    nearest search over the FP16 neighbours."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal(20000) * np.exp2(rng.uniform(-26, 16, 20000))
    h = x.astype(np.float16)
    lo = np.nextafter(h, np.float16(-np.inf)).astype(np.float64)
    hi = np.nextafter(h, np.float16(np.inf)).astype(np.float64)
    hv = h.astype(np.float64)
    fin = np.isfinite(hv)
    assert (np.abs(x - hv)[fin] <= np.minimum(np.abs(x - lo), np.abs(x - hi))[fin]).all()
    # exact ties between neighbours go to the even significand
    base = rng.integers(1024, 2047, 2000).astype(np.float64)
    ties = (2 * base + 1) * 2.0 ** -12  # odd multiples of 2^-12 on [0.5, 1): half the FP16 ulp
    t16 = ties.astype(np.float16)
    assert (t16.view(np.uint16) % 2 == 0).all()
    assert np.array_equal(np.abs(t16.astype(np.float64) - ties), np.full(ties.size, 2.0 ** -12))
    assert np.isinf(np.float64(65520.0).astype(np.float16)) and np.float64(65519.99).astype(np.float16) == 65504


def test_zero_transforms_are_positive_zero():
    """A filter or input that is zero (or cancels exactly) transforms to +0 in every position."""
    w = np.zeros((2, 3, 3, 3), dtype=np.float16)
    w[1] = -0.0
    u = w43.filter_transform43(w)
    assert (u.view(np.uint16) == 0).all()
    geo = w43.Geometry43(1, 2, 8, 8, 1, 1, 1)
    x = np.zeros((1, 2, 8, 8), dtype=np.float16)
    assert (w43.input_transform43(x, geo).view(np.uint16) == 0).all()


def test_fp32_input_transform_within_one_ulp_of_exact():
    """The FP32-order V differs from the correctly rounded exact V by at most one FP16 ulp, and
    on ordinary data in few elements."""
    rng = np.random.default_rng(6)
    geo = w43.Geometry43(2, 16, 14, 18, 1, 1, 1)
    # exponents up to 2^9: |V| <= 42.25 * 2^10 stays inside the FP16 range
    wide = _wide_fp16(rng, (2, 16, 14, 18))
    wide = np.where(np.abs(wide) >= 2 ** 10, wide / np.float16(64), wide).astype(np.float16)
    for x in (wg.uniform_pm1(61, (2, 16, 14, 18)), wide):
        v32 = w43.input_transform43(x, geo).astype(np.float64)
        ve = w43.input_transform43_f64(x, geo)
        v16 = ve.astype(np.float16).astype(np.float64)
        ulp = np.spacing(np.abs(v16).astype(np.float16)).astype(np.float64)
        assert (np.abs(v32 - v16) <= ulp * 1.0000001).all()
    # ordinary data: almost always the correctly rounded value
    x = wg.uniform_pm1(62, (2, 16, 14, 18))
    v32 = w43.input_transform43(x, geo)
    v16 = w43.input_transform43_f64(x, geo).astype(np.float16)
    assert (v32 != v16).mean() < 1e-3
