"""TEST INFRASTRUCTURE ONLY — bfloat16 rounding, the checker for the library's bfloat16 output
(CONV2D_IO_OUT_BF16). Nothing in the library calls it.

A bfloat16 value is the upper 16 bits of an IEEE binary32; rounding a binary32 to it is round
to nearest, ties to even on bit 16 (a NaN stays a NaN, quieted). The reference has no bfloat16
output (its stage 4 stores FP16, conv2d.cpp:37-47); the outputs are pinned by this function,
itself pinned in tests/test_oracle_bf16.py against known values and a brute-force nearest
search over the two neighbouring bfloat16 values.
"""
from __future__ import annotations

import numpy as np


def round_bf16_bits(x) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of float32(x), rounded to nearest even."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    b = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16) & 0xFFFF
    r = np.where(nan, ((b >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_to_float(bits) -> np.ndarray:
    """float32 values of bfloat16 bit patterns."""
    u = np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16
    return u.view(np.float32)


def round_bf16(x) -> np.ndarray:
    """float32(x) rounded to the nearest bfloat16 (ties to even), as float32."""
    return bf16_to_float(round_bf16_bits(x))
