"""Pins the FP8 emulation (CPU): e4m3fn rounding against brute force over the code table,
the block-scale rule against its defining inequality, hi/lo reconstruction bounds, packing."""
import numpy as np

from oracle import fp8
from oracle import winograd as wg

TAB = fp8.e4m3_table()
FINITE = np.array([c for c in range(256) if np.isfinite(TAB[c])])


def test_table_known_values():
    assert TAB[0x7E] == 448.0 and TAB[0x08] == 2.0 ** -6 and TAB[0x01] == 2.0 ** -9
    assert TAB[0x38] == 1.0 and TAB[0xB8] == -1.0 and np.isnan(TAB[0x7F]) and np.isnan(TAB[0xFF])
    assert len(FINITE) == 254


def test_encode_is_nearest_ties_to_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-448, 448, 5000), rng.uniform(-0.05, 0.05, 5000),
                        (TAB[FINITE][:-1] + np.diff(TAB[FINITE])) / 2])  # midpoints (sorted codes below)
    x = x[np.abs(x) <= 448]
    codes = fp8.e4m3_encode(x)
    vals = TAB[codes.astype(int)]
    pos = np.sort(np.unique(np.abs(TAB[FINITE])))
    for xi, vi, ci in zip(x, vals, codes):
        d = np.abs(pos - abs(xi))
        best = d.min()
        assert abs(abs(vi) - abs(xi)) == best
        ties = pos[d == best]
        if len(ties) == 2:  # exact midpoint: the even code wins
            assert (int(ci) & 1) == 0
        if vi != 0:
            assert np.sign(vi) == np.sign(xi)


def test_encode_roundtrip_all_codes():
    for c in FINITE:
        if c == 0x80:
            continue
        assert fp8.e4m3_encode(np.array([TAB[c]]))[0] == c


def test_block_exponent_is_minimal():
    rng = np.random.default_rng(1)
    m = np.concatenate([10.0 ** rng.uniform(-12, 6, 4000), [448.0, 449.0, 224.0, 0.875 * 2 ** 9, 0.0]])
    E = fp8.block_exponent(m)
    nz = m > 0
    assert np.all(m[nz] <= 448.0 * 2.0 ** E[nz])
    assert np.all(m[nz] > 448.0 * 2.0 ** (E[nz] - 1))
    assert E[-1] == -127 and E[-2] == 0 and E[-5] == 0 and E[-4] == 1


def test_hi_lo_reconstruction_bound():
    rng = np.random.default_rng(2)
    v = (rng.standard_normal((500, 32)) * 10.0 ** rng.uniform(-3, 3, (500, 1))).astype(np.float32)
    hi, lo, eh, el = fp8.quantize_blocks(v)
    rec = fp8.dequant(hi, eh) + fp8.dequant(lo, el)
    r = v.astype(np.float64) - fp8.dequant(hi, eh)
    bound = 2.0 ** -4 * np.max(np.abs(r), axis=1, keepdims=True) + 2.0 ** (el[:, None] - 10.0)
    assert np.all(np.abs(v - rec) <= bound)
    # relative to the block maximum the compensated pair is ~2^-8 accurate
    rel = np.max(np.abs(v - rec), axis=1) / np.max(np.abs(v), axis=1)
    assert np.max(rel) < 2.0 ** -7


def test_packing_offsets():
    rng = np.random.default_rng(3)
    P, R, NB = 16, 20, 8  # two 128-channel blocks
    codes = rng.integers(0, 120, (P, R, NB, 32)).astype(np.uint8)
    flat = fp8.pack_codes(codes, rows_blocks=2)
    pos, r, c = 5, 17, 200
    cq, cb, cbyte = c // 128, (c % 128) // 32, c % 32
    off = ((pos * 2 + cq) * 2 + r // 16) * 2048 + (cb * 16 + r % 16) * 32 + cbyte
    assert flat[off] == codes[pos, r, c // 32, cbyte]
    E = rng.integers(-20, 20, (P, R, NB))
    s = fp8.pack_scales(E, rows_blocks=2)
    off = ((pos * 2 + cq) * 2 + r // 16) * 64 + cb * 16 + r % 16
    assert s[off] == E[pos, r, c // 32] + 127


def test_three_pass_error_within_tolerance_small():
    geo = wg.Geometry(2, 256, 10, 10, 64)
    x = wg.with_outliers(wg.normal(5, (2, 256, 10, 10)), 6)
    w = wg.normal(7, (64, 256, 3, 3), sigma=float(np.sqrt(2 / (9 * 256))))
    qu = fp8.quantize_U(wg.filter_transform_fp32(w))
    qv = fp8.quantize_V(fp8.input_transform_f32(x, geo))
    y = fp8.gemm_output_fp8(qu, qv, geo)
    ref = wg.winograd_f64(x, w, geo)
    m = wg.error_metrics(y, ref)
    assert m["max_rel"] < 1e-2 and m["rms_rel"] < 5e-3
