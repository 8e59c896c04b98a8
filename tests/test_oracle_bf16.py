"""CPU pins of the oracle's bfloat16 rounding (oracle/bf16.py)."""
import numpy as np

from oracle import bf16


def test_known_values():
    cases = [
        (1.0, 0x3F80),
        (-2.5, 0xC020),
        (1.0 + 2.0 ** -8, 0x3F80),            # tie between 1 and 1 + 2^-7: even mantissa (1)
        (1.0 + 3 * 2.0 ** -8, 0x3F82),        # tie between 1 + 2^-7 and 1 + 2^-6: even (1 + 2^-6)
        (1.0 + 2.0 ** -8 + 2.0 ** -20, 0x3F81),  # above the tie: up
        (65504.0, 0x4780),                    # FP16 max rounds to 65536
        (0.0, 0x0000),
        (-0.0, 0x8000),
        (2.0 ** -130, 0x0008),                # binary32 subnormal 0x00080000, exactly representable
        (3.3895313892515355e38, 0x7F7F),      # the largest bfloat16
        (float("inf"), 0x7F80),
    ]
    got = bf16.round_bf16_bits(np.array([x for x, _ in cases], dtype=np.float32))
    assert [int(g) for g in got] == [b for _, b in cases]
    assert np.isnan(bf16.round_bf16(np.float32("nan")))


def test_nearest_even_by_brute_force():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-40, 40, 20000))).astype(np.float32)
    # add exact ties: bfloat16 values plus half an ulp
    base = bf16.round_bf16(x[:2000])
    x[:2000] = (base.view(np.uint32) + 0x8000).view(np.float32)
    got = bf16.round_bf16(x).astype(np.float64)
    lo = (x.view(np.uint32) & 0xFFFF0000).view(np.float32).astype(np.float64)  # truncation toward zero
    hi = ((x.view(np.uint32) & 0xFFFF0000) + 0x10000).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dlo, dhi = np.abs(xd - lo), np.abs(hi - xd)
    lo_even = ((lo.astype(np.float32).view(np.uint32) >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo_even, lo, hi)))
    assert np.array_equal(got, want)
