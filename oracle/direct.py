"""TEST INFRASTRUCTURE ONLY — CPU restatement of the general-geometry (implicit-GEMM direct)
path's packed operands.

No reference counterpart: the reference's plugin ignores r, s, u, v, p, q (conv2d.cpp:254-271);
its CPU checker (include/verfiy.h:5-43) is the convolution this path must reproduce, restated
in oracle/direct_conv.c. Here:

  j:      flattened reduction index j = tap * C + c (tap = ri * s + si), zero-padded to
          nsteps * 32 with nsteps = ceil(ceil(r*s*C / 32) / 8) * 8; 16-value block L = j // 16.
  W:      fragment (L, kb): lane = (j % 16 // 8) * 32 + k % 32, element j % 8, value
          w[k][c][ri][si] (zero for k >= K and j >= r*s*C).
  X:      fragment (L, pb): lane = (j % 16 // 8) * 32 + g % 32, element j % 8, value
          x[n][c][oh*u - p + ri][ow*v - q + si] (zero outside the image), g = n*Oh*Ow + oh*Ow + ow.

Pinned by tests/test_oracle.py: the GEMM over the unpacked operands equals oracle/direct_conv.c's
direct convolution exactly on integer-valued data, for odd filter sizes, strides and pads.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class DirectGeometry:
    n: int
    c: int
    h: int
    w: int
    k: int
    r: int
    s: int
    u: int = 1
    v: int = 1
    p: int = 0
    q: int = 0

    @property
    def oh(self) -> int:
        return (self.h - self.r + 2 * self.p) // self.u + 1

    @property
    def ow(self) -> int:
        return (self.w - self.s + 2 * self.q) // self.v + 1

    @property
    def CB(self) -> int:
        return (self.c + 31) // 32

    @property
    def KB32(self) -> int:
        return (self.k + 31) // 32

    @property
    def P(self) -> int:
        return self.n * self.oh * self.ow

    @property
    def P32(self) -> int:
        return (self.P + 31) // 32

    @property
    def nsteps(self) -> int:
        return -(-(-(-(self.r * self.s * self.c) // 32)) // 8) * 8


def weight_matrix(w16: np.ndarray, geo: DirectGeometry) -> np.ndarray:
    """(K, nsteps*32) with column j = (ri*S + si)*C + c (zero padding beyond r*s*C)."""
    K, C, R, S = w16.shape
    out = np.zeros((K, geo.nsteps * 32), dtype=np.float16)
    out[:, : R * S * C] = w16.transpose(0, 2, 3, 1).reshape(K, R * S * C)
    return out


def im2col(x16: np.ndarray, geo: DirectGeometry) -> np.ndarray:
    """(nsteps*32, P): row j = (ri*S + si)*C + c, column g = n*Oh*Ow + oh*Ow + ow."""
    N, C, H, W = x16.shape
    out = np.zeros((geo.nsteps * 32, geo.P), dtype=np.float16)
    oh = np.arange(geo.oh)[:, None]
    ow = np.arange(geo.ow)[None, :]
    for ri in range(geo.r):
        for si in range(geo.s):
            tap = ri * geo.s + si
            ih = oh * geo.u - geo.p + ri
            iw = ow * geo.v - geo.q + si
            ok = (ih >= 0) & (ih < H) & (iw >= 0) & (iw < W)
            ihc, iwc = np.clip(ih, 0, H - 1), np.clip(iw, 0, W - 1)
            vals = np.where(ok, x16[:, :, ihc, iwc], np.float16(0))  # (N, C, oh, ow); +0 outside
            out[tap * C: tap * C + C] = vals.transpose(1, 0, 2, 3).reshape(C, -1)
    return out


def _pack(mat: np.ndarray, rows_blocks: int, nsteps: int) -> np.ndarray:
    """mat (R, nsteps*32) -> [L = nsteps*2][rb][lane][e], lane = (j%16//8)*32 + r%32."""
    R = mat.shape[0]
    full = np.zeros((rows_blocks * 32, nsteps * 32), dtype=np.float16)
    full[:R] = mat
    # (rb, r32, L, half, e) -> (L, rb, half, r32, e)
    x = full.reshape(rows_blocks, 32, nsteps * 2, 2, 8).transpose(2, 0, 3, 1, 4)
    return np.ascontiguousarray(x).reshape(-1)


def pack_weights(w16: np.ndarray, geo: DirectGeometry) -> np.ndarray:
    return _pack(weight_matrix(w16, geo), geo.KB32, geo.nsteps)


def pack_input(x16: np.ndarray, geo: DirectGeometry) -> np.ndarray:
    return _pack(np.ascontiguousarray(im2col(x16, geo).T), geo.P32, geo.nsteps)


def gemm_output_f64(w16: np.ndarray, x16: np.ndarray, geo: DirectGeometry, absolute: bool = False) -> np.ndarray:
    """W . X in FP64, reshaped to (N, K, Oh, Ow); absolute=True gives |W| . |X| (error scale)."""
    a = weight_matrix(w16, geo).astype(np.float64)
    b = im2col(x16, geo).astype(np.float64)
    if absolute:
        a, b = np.abs(a), np.abs(b)
    y = a @ b  # (K, P)
    return np.ascontiguousarray(y.reshape(geo.k, geo.n, geo.oh, geo.ow).transpose(1, 0, 2, 3))
