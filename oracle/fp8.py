"""TEST INFRASTRUCTURE ONLY — bit-exact CPU emulation of the FP8 (OCP e4m3fn) compensated
GEMM stage: E8M0 block scales over 32 channels, hi/lo split, packed layouts, and the
three-pass product U_hi.V_hi + U_hi.V_lo + U_lo.V_hi.

No reference counterpart (the FP8 form extends the method). Pinned by
tests/test_oracle_fp8.py: e4m3 rounding against an exhaustive decode table of all 256 codes
(nearest representable value, ties to even code), scale choice against its defining
inequality, hi + lo reconstruction error bounds, and packing round trips.

Formats
  e4m3fn: 1 sign, 4 exponent (bias 7), 3 mantissa bits; max 448; subnormal quantum 2^-9;
          code 0x7f / 0xff is NaN (never produced here: values are pre-scaled to <= 448).
  E8M0:   scale 2^(e - 127), e in [0, 254].
Block scale of 32 values with max |v| = m: the smallest E with m <= 448 * 2^E
(E = -127 when m = 0), clamped to [-127, 127].
"""
from __future__ import annotations

import numpy as np

from . import winograd as wg


def e4m3_table() -> np.ndarray:
    """Value of every e4m3fn code (NaN codes -> nan)."""
    vals = np.zeros(256)
    for code in range(256):
        s = -1.0 if code & 0x80 else 1.0
        e = (code >> 3) & 0xF
        m = code & 7
        if e == 15 and m == 7:
            vals[code] = np.nan
        elif e == 0:
            vals[code] = s * m * 2.0 ** -9
        else:
            vals[code] = s * (1 + m / 8.0) * 2.0 ** (e - 7)
    return vals


_TAB = e4m3_table()


def e4m3_encode(x: np.ndarray) -> np.ndarray:
    """Round float values (|x| <= 448) to the nearest e4m3fn code, ties to even mantissa.
    The sign bit is kept when the magnitude rounds to zero (-0 is code 0x80), as gfx950's
    v_cvt_pk_fp8_f32 does (probed)."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    assert np.all(a <= 448.0), "values must be pre-scaled into the e4m3 range"
    # exponent of the binade (normal range) or the subnormal quantum
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -12)))
    e = np.clip(e, -6, 8)
    q = 2.0 ** (e - 3)
    k = np.round(a / q)  # numpy rounds half to even -> ties to even mantissa
    r = k * q
    # rounding up into the next binade is fine (exactly representable)
    # encode
    code = np.zeros(x.shape, dtype=np.int64)
    sub = r < 2.0 ** -6
    code[sub] = (r[sub] / 2.0 ** -9).astype(np.int64)
    ne = ~sub & (r > 0)
    er = np.floor(np.log2(r[ne])).astype(np.int64)
    mant = np.round((r[ne] / 2.0 ** er - 1.0) * 8).astype(np.int64)
    code[ne] = ((er + 7) << 3) | mant
    code = code | np.where(np.signbit(x), 0x80, 0)
    return code.astype(np.uint8)


def e4m3_decode(code: np.ndarray) -> np.ndarray:
    return _TAB[np.asarray(code, dtype=np.int64)]


def block_exponent(m: np.ndarray) -> np.ndarray:
    """Smallest E with m <= 448 * 2^E (E = -127 for m = 0), clamped to [-127, 127]."""
    m = np.asarray(m, dtype=np.float64)
    f, x = np.frexp(m)  # m = f * 2^x, f in [0.5, 1)
    E = x - 9 + (f > 0.875)
    E = np.where(m == 0, -127, E)
    return np.clip(E, -127, 127).astype(np.int64)


LO_OFFSET = 4  # E_lo = E_hi - 4: a hi residual is at most 16 * 2^E_hi, so lo codes stay <= 256


def quantize_blocks(v32: np.ndarray):
    """v32: (..., 32) float32 blocks. Returns (hi codes, lo codes, E_hi, E_lo) with
    v ~= dec(hi) 2^E_hi + dec(lo) 2^E_lo, E_lo = max(E_hi - 4, -127).

    Every intermediate is exact in FP32 (power-of-two scalings; the hi residual is the
    difference between v and v rounded to 4 significant bits), so the FP32 kernels and this
    FP64 restatement produce identical codes."""
    v = np.asarray(v32, dtype=np.float32).astype(np.float64)
    E_hi = block_exponent(np.max(np.abs(v), axis=-1))
    hi = e4m3_encode(v * 2.0 ** -E_hi[..., None])
    r = v - e4m3_decode(hi) * 2.0 ** E_hi[..., None]
    E_lo = np.maximum(E_hi - LO_OFFSET, -127)
    lo = e4m3_encode(r * 2.0 ** -E_lo[..., None])
    return hi, lo, E_hi, E_lo


def dequant(codes, E):
    return e4m3_decode(codes) * 2.0 ** np.asarray(E)[..., None]


# ---- operands ---------------------------------------------------------------------------
def _blocks(a: np.ndarray, C: int):
    """(P, R, C) -> (P, R, CQ*2, 32) with channels padded to a multiple of 64."""
    P, R, _ = a.shape
    CQ = (C + 63) // 64
    full = np.zeros((P, R, CQ * 64), dtype=np.float32)
    full[:, :, :C] = a
    return full.reshape(P, R, CQ * 2, 32)


def quantize_U(u16: np.ndarray):
    """u16: (16, K, C) FP16 transformed filter -> dict of hi/lo codes (16, K, CQ*2, 32) and
    exponents (16, K, CQ*2)."""
    hi, lo, eh, el = quantize_blocks(_blocks(u16.astype(np.float32), u16.shape[2]))
    return {"hi": hi, "lo": lo, "eh": eh, "el": el}


def quantize_V(v32: np.ndarray):
    """v32: (16, C, NT) FP32 transformed input -> blocks over (16, NT, CQ*2, 32)."""
    vt = np.ascontiguousarray(v32.astype(np.float32).transpose(0, 2, 1))
    hi, lo, eh, el = quantize_blocks(_blocks(vt, v32.shape[1]))
    return {"hi": hi, "lo": lo, "eh": eh, "el": el}


def dequant_operand(q: dict, C: int, part: str) -> np.ndarray:
    """(16, R, C) float64 values of the hi or lo part."""
    codes, E = (q["hi"], q["eh"]) if part == "hi" else (q["lo"], q["el"])
    d = dequant(codes, E)
    P, R = d.shape[:2]
    return d.reshape(P, R, -1)[:, :, :C]


def gemm_output_fp8(qu: dict, qv: dict, geo: wg.Geometry, absolute: bool = False) -> np.ndarray:
    """Three-pass product U_hi.V_hi + U_hi.V_lo + U_lo.V_hi, then A^T M A, in FP64."""
    C = geo.c
    uh, ul = dequant_operand(qu, C, "hi"), dequant_operand(qu, C, "lo")
    vh = dequant_operand(qv, C, "hi").transpose(0, 2, 1)
    vl = dequant_operand(qv, C, "lo").transpose(0, 2, 1)
    return (wg.gemm_output_f64(uh, vh, geo, absolute) + wg.gemm_output_f64(uh, vl, geo, absolute)
            + wg.gemm_output_f64(ul, vh, geo, absolute))


def input_transform_f32(x16: np.ndarray, geo: wg.Geometry) -> np.ndarray:
    """The input transform in the kernels' FP32 order, kept in FP32: (16, C, NT)."""
    return wg.input_transform_fp32(x16, geo, keep_fp32=True)


# ---- packed layout ---------------------------------------------------------------------------
# The v_mfma_scale_f32_32x32x64_f8f6f4 operand (32 rows x 64 channels; probed on gfx950):
# lane l holds row l % 32; its bytes 0-15 are channels 16 * (l // 32) + 0..15 of the first
# 32-channel block, its bytes 16-31 the same channels of the second block; the E8M0 scale
# of (row r, 32-channel block kb) is supplied by lane r + 32 * kb (op_sel picks the byte of
# that lane's scale register). One record per (pos, cq, rb) (cq = 64-channel block, rb =
# 32-row block): 4096 code bytes
#   [0, 1024)     codes of block kb = 0: lane * 16 + (c % 16), lane = (c % 32 // 16) * 32 + r % 32
#   [1024, 2048)  codes of block kb = 1, same arrangement            (hi part)
#   [2048, 4096)  lo codes, same arrangement
# and 128 scale bytes: lane * 2 -> E_hi + 127, lane * 2 + 1 -> E_lo + 127, lane = kb * 32 + r % 32.
# The codes of all records come first, ordered [pos][cq][rb] (each 1 KiB unit one contiguous run
# across the wave, so a lane's operand is conflict-free 16-byte LDS reads and a record moves as
# four whole 1 KiB LDS-DMA units), then the scales of all records in the same order.
CODE_BYTES = 4096
SCALE_BYTES = 128
REC_BYTES = CODE_BYTES + SCALE_BYTES  # bytes per record over both regions


def pack_operand(q: dict, rows_blocks: int) -> np.ndarray:
    """Workspace image of a quantised operand (dict from quantize_U / quantize_V): the code
    region, then the scale region."""
    P, R, NB, _ = q["hi"].shape
    CQ = NB // 2
    RB = rows_blocks
    codes = np.zeros((P, CQ, RB, CODE_BYTES), dtype=np.uint8)

    def halves(c8):  # (P, R, NB, 32) -> (P, CQ, RB, kb 2, lane 64, 16 bytes)
        full = np.zeros((P, RB * 32, NB, 32), dtype=np.uint8)
        full[:, :R] = c8
        # (P, rb, r32, cq, kb, hl, b16) -> (P, cq, rb, kb, hl, r32, b16)
        x = full.reshape(P, RB, 32, CQ, 2, 2, 16).transpose(0, 3, 1, 4, 5, 2, 6)
        return x.reshape(P, CQ, RB, 2048)

    def lane_exps(E):
        full = np.zeros((P, RB * 32, NB), dtype=np.int64)
        full[:, :R] = E + 127
        x = full.reshape(P, RB, 32, CQ, 2).transpose(0, 3, 1, 4, 2)  # P cq rb kb r32
        return x.reshape(P, CQ, RB, 64).astype(np.uint8)

    codes[..., 0:2048] = halves(q["hi"])
    codes[..., 2048:4096] = halves(q["lo"])
    sc = np.stack([lane_exps(q["eh"]), lane_exps(q["el"])], axis=-1)  # (P, CQ, RB, 64, 2)
    return np.concatenate([codes.reshape(-1), sc.reshape(-1)])


def split_regions(flat: np.ndarray, c: int, rows_blocks: int):
    """(codes (P, CQ, RB, 4096), scales (P, CQ, RB, 128)) views of an operand image."""
    CQ = (c + 63) // 64
    n = 16 * CQ * rows_blocks
    return (flat[:n * CODE_BYTES].reshape(16, CQ, rows_blocks, CODE_BYTES),
            flat[n * CODE_BYTES:n * REC_BYTES].reshape(16, CQ, rows_blocks, SCALE_BYTES))


def operand_bytes(c: int, rows_blocks: int) -> int:
    return 16 * ((c + 63) // 64) * rows_blocks * REC_BYTES
