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
        assert np.signbit(vi) == np.signbit(xi)  # -0 for negatives that round to zero


def test_encode_roundtrip_all_codes():
    for c in FINITE:
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
    v[7] = 0.0  # an all-zero block
    v[8, :31] = 0.0  # a block with one nonzero value
    hi, lo, eh, el = fp8.quantize_blocks(v)
    assert np.all(el == np.maximum(eh - 4, -127)) and eh[7] == -127 and el[7] == -127
    # lo codes never saturate: |lo| <= 256 (the hi residual is at most 16 * 2^E_hi)
    assert np.max(np.abs(TAB[lo.astype(int)])) <= 256.0
    rec = fp8.dequant(hi, eh) + fp8.dequant(lo, el)
    r = v.astype(np.float64) - fp8.dequant(hi, eh)
    assert np.all(np.abs(r) <= 16.0 * 2.0 ** eh[:, None])
    # lo rounding: half an e4m3 spacing (relative 2^-4) or half the subnormal quantum
    bound = 2.0 ** -4 * np.abs(r) + 2.0 ** (el[:, None] - 10.0)
    assert np.all(np.abs(v - rec) <= bound)
    nz = np.max(np.abs(v), axis=1) > 0
    rel = np.max(np.abs(v - rec), axis=1)[nz] / np.max(np.abs(v), axis=1)[nz]
    assert np.max(rel) < 2.0 ** -7
    assert np.all(rec[7] == 0)


def test_record_layout_offsets():
    rng = np.random.default_rng(3)
    P, R, C = 16, 40, 150  # three 64-channel blocks (the last ragged), two 32-row blocks
    vals = rng.standard_normal((P, R, C)).astype(np.float32)
    q = fp8.quantize_U(vals.astype(np.float16))
    flat = fp8.pack_operand(q, rows_blocks=2)
    CQ = 3
    nrec = P * CQ * 2
    assert flat.size == fp8.operand_bytes(C, 2) == nrec * (4096 + 128) == nrec * fp8.REC_BYTES
    scale0 = nrec * 4096  # the scale region follows every record's codes
    for (pos, r, c) in [(5, 17, 130), (0, 0, 0), (15, 39, 149), (3, 33, 63), (9, 2, 64)]:
        cq, kb, b = c // 64, (c % 64) // 32, c % 32
        lane = (b // 16) * 32 + r % 32
        ri = (pos * CQ + cq) * 2 + r // 32
        off = ri * 4096 + kb * 1024 + lane * 16 + b % 16
        assert flat[off] == q["hi"][pos, r, c // 32, b]
        assert flat[off + 2048] == q["lo"][pos, r, c // 32, b]
        slane = kb * 32 + r % 32
        assert flat[scale0 + ri * 128 + 2 * slane] == q["eh"][pos, r, c // 32] + 127
        assert flat[scale0 + ri * 128 + 2 * slane + 1] == q["el"][pos, r, c // 32] + 127
    # padding rows (40..63) and channels (160..191: a whole padding 32-block) are zero codes
    # with zero scale bytes
    ri = (2 * CQ + 2) * 2 + 1  # pos 2, cq 2, rb 1
    for lane in range(64):
        assert not flat[ri * 4096 + 1024 + lane * 16:ri * 4096 + 1024 + lane * 16 + 16].any()
    assert flat[scale0 + ri * 128 + 2 * (32 + 45 % 32)] == 0
    codes, scales = fp8.split_regions(flat, C, 2)
    assert codes.shape == (16, 3, 2, 4096) and scales.shape == (16, 3, 2, 128)
    assert scales[5, 2, 0, 2 * 17] == q["eh"][5, 17, 130 // 32] + 127  # c = 130: block kb = 0 of cq = 2


def test_two_passes_miss_and_three_meet_the_bound():
    """The design claim behind the third pass: with the hi/lo split of V only, the filter's
    own e4m3 rounding leaves a few percent of error; compensating U too meets 1e-2."""
    geo = wg.Geometry(1, 128, 8, 8, 32)
    x = wg.normal(11, (1, 128, 8, 8))
    w = wg.normal(12, (32, 128, 3, 3), sigma=float(np.sqrt(2 / (9 * 128))))
    qu = fp8.quantize_U(wg.filter_transform_fp32(w))
    qv = fp8.quantize_V(fp8.input_transform_f32(x, geo))
    ref = wg.winograd_f64(x, w, geo)
    uh = fp8.dequant_operand(qu, geo.c, "hi")
    vh = fp8.dequant_operand(qv, geo.c, "hi").transpose(0, 2, 1)
    vl = fp8.dequant_operand(qv, geo.c, "lo").transpose(0, 2, 1)
    two = wg.gemm_output_f64(uh, vh, geo) + wg.gemm_output_f64(uh, vl, geo)
    three = fp8.gemm_output_fp8(qu, qv, geo)
    assert wg.error_metrics(two, ref)["rms_rel"] > 1e-2
    assert wg.error_metrics(three, ref)["max_rel"] < 1e-2


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
