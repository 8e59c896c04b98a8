"""Writes the small golden fixtures under tests/golden/ (run in the development container).

  torch_conv_f64.npz  : seeded FP16 inputs and torch.nn.functional.conv2d in float64 on the
                        CPU (an independent library), for small shapes covering even,
                        ragged (odd H/W), C=3, K not a multiple of 16, pads 0/1/2 and a
                        non-square image.
  ref_checker.npz     : the reference's own CPU checker conv2dcpu (include/verfiy.h:5-43,
                        compiled from the reference source by oracle/Makefile into
                        oracle/_ref) on seeded inputs, positive (the reference's
                        distribution, main.cpp:40-48) and signed; FP16 outputs.
Inputs are stored as raw FP16 bits; outputs as float64 / FP16 bits.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import cdirect, winograd as wg  # noqa: E402

TORCH_CASES = [
    # n, c, h, w, k, p, q, seed
    (1, 16, 6, 6, 16, 1, 1, 11),
    (1, 16, 5, 5, 16, 1, 1, 12),
    (2, 3, 7, 7, 5, 1, 1, 13),
    (1, 40, 9, 9, 20, 1, 1, 14),
    (1, 8, 6, 6, 8, 0, 0, 15),
    (1, 8, 5, 7, 8, 2, 2, 16),
    (2, 33, 13, 10, 17, 1, 1, 17),
]

REF_CASES = [
    # n, c, h, w, k, r, s, u, v, p, q, seed, positive
    (1, 16, 8, 8, 16, 3, 3, 1, 1, 1, 1, 21, True),
    (1, 16, 8, 8, 16, 3, 3, 1, 1, 1, 1, 22, False),
    (2, 5, 7, 9, 3, 3, 3, 1, 1, 1, 1, 23, False),
    (1, 4, 9, 9, 4, 5, 5, 2, 2, 2, 2, 24, False),
]


def torch_fixtures() -> dict:
    import torch
    out = {}
    for i, (n, c, h, w, k, p, q, seed) in enumerate(TORCH_CASES):
        x = wg.uniform_pm1(seed, (n, c, h, w))
        wt = wg.uniform_pm1(seed + 1000, (k, c, 3, 3))
        y = torch.nn.functional.conv2d(torch.from_numpy(x.astype(np.float64)),
                                       torch.from_numpy(wt.astype(np.float64)), padding=(p, q)).numpy()
        out[f"case{i}_shape"] = np.array([n, c, h, w, k, p, q], dtype=np.int64)
        out[f"case{i}_x"] = x.view(np.uint16)
        out[f"case{i}_w"] = wt.view(np.uint16)
        out[f"case{i}_y"] = y
    return out


def ref_fixtures() -> dict:
    if cdirect.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built (needs the reference source): make -C oracle ref")
    out = {}
    for i, (n, c, h, w, k, r, s, u, v, p, q, seed, positive) in enumerate(REF_CASES):
        if positive:
            x = ((wg.splitmix64(seed, np.arange(n * c * h * w, dtype=np.uint64)) % np.uint64(255)).astype(np.float64) / 255.0)
            wt = ((wg.splitmix64(seed + 1, np.arange(k * c * r * s, dtype=np.uint64)) % np.uint64(255)).astype(np.float64) / 255.0)
            x = x.astype(np.float16).reshape(n, c, h, w)
            wt = wt.astype(np.float16).reshape(k, c, r, s)
        else:
            x = wg.uniform_pm1(seed, (n, c, h, w))
            wt = wg.uniform_pm1(seed + 1, (k, c, r, s))
        y = cdirect.ref_conv2dcpu(x, wt, stride=(u, v), pad=(p, q))
        out[f"case{i}_shape"] = np.array([n, c, h, w, k, r, s, u, v, p, q], dtype=np.int64)
        out[f"case{i}_x"] = x.view(np.uint16)
        out[f"case{i}_w"] = wt.view(np.uint16)
        out[f"case{i}_y"] = y.view(np.uint16)
    return out


def main() -> None:
    cdirect.build()
    np.savez_compressed(os.path.join(HERE, "torch_conv_f64.npz"), **torch_fixtures())
    np.savez_compressed(os.path.join(HERE, "ref_checker.npz"), **ref_fixtures())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
