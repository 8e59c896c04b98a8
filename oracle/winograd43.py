"""TEST INFRASTRUCTURE ONLY — numpy model of the Winograd F(4x4,3x3) form.

No reference counterpart: the reference implements F(2x2,3x3) only (B^T preprocessing.cuh:50-54,
G :77-81, A^T conv2d.cpp:18-22). F(4x4,3x3) comes from the Cook-Toom (Lagrange interpolation)
construction of Winograd's minimal filtering algorithms (S. Winograd, "Arithmetic Complexity of
Computations", 1980; A. Lavin and S. Gray, "Fast Algorithms for Convolutional Neural Networks",
CVPR 2016), restated in `cook_toom` with exact rationals, on the interpolation points
{0, 1, -1, 2, -1/2} and infinity. Lavin and Gray's points {0, 1, -1, 2, -2} give the larger
transform growth; a reciprocal pair (2, -1/2) balances it (B. Barabasz, A. Anderson, K. Soodhalter,
D. Gregg, "Error analysis and improving the accuracy of Winograd convolution for deep neural
networks", ACM TOMS 2020). Measured with FP16 U, V and M on C = 256 layers, the normwise error
against the exact convolution is 3.4e-3 (uniform data) / 3.9e-3 (normal data) here, against
8.7e-3 / 8.9e-3 with Lavin and Gray's points.

Stages (36 positions pos = 6 i + j of the 6x6 transform domain):
  S1  U = G' g G'^T, G' = diag(gamma) G: gamma = (1, 3/4, 3/4, 15/4, 15/16, 1) makes every row
      of G' dyadic. Every entry is then a sum of 9 FP16 values times dyadic constants; FP16
      values are multiples of 2^-24 below 2^16, so the sum is exact in FP64 (fewer than 53
      significant bits) in any order. It is rounded once to FP16; a zero is stored as +0 (the
      sign of an exact zero sum depends on the operation order, so both sides add +0.0).
  S2  V = B^T d B over the 6x6 input window at (4 th - p, 4 tw - q) of output tile (th, tw)
      (zero outside the image), in FP32 in a fixed operation order (restated exactly by
      `input_transform43`: columns first, each product and sum rounded once), rounded once to
      FP16; within one FP16 ulp of the exact transform (`input_transform43_f64`, exact in FP64:
      B^T's entries are dyadic, 36 terms below 2^22 on a grid of 2^-26).
  S3  per position M = U . V over the channels, FP32 accumulation, stored as FP16(M / 4) (the
      power-of-two scale keeps the stored values within the output's range: max|M| stayed below
      4.3 max|y| on the measured layers).
  S4  Y = A'^T (4 M) A', A' = A diag(1/gamma) (the compensation of gamma), FP32.

Pinned by tests/test_oracle_f43.py: `cook_toom` reproduces Lavin and Gray's published F(4x4,3x3)
matrices for their points; the 1-D and 2-D Winograd identities hold exactly (rationals) for the
chosen points and for the gamma-scaled matrices; the FP64 transforms equal the exact rational
transforms on inputs with a wide exponent spread (the exactness argument above); the FP32-order
input transform within one FP16 ulp of the exact one; the exact
F(4x4) path equals the direct convolution to FP64 rounding on ragged shapes with pads 0/1/2;
packing round trips.
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction as Fr

import numpy as np

from oracle.winograd import FRAG, _pack, _unpack

POINTS = (Fr(0), Fr(1), Fr(-1), Fr(2), Fr(-1, 2))
GAMMA = (Fr(1), Fr(3, 4), Fr(3, 4), Fr(15, 4), Fr(15, 16), Fr(1))
M_SCALE = 0.25  # M is stored as FP16(M * M_SCALE)
NPOS = 36


def _poly_mul(a, b):
    out = [Fr(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def cook_toom(points, m: int, r: int):
    """Exact A^T (m x n), G (n x r), B^T (n x n), n = m + r - 1, for F(m, r) on the n - 1 finite
    points and infinity. Finite point p_i: G row = (1, p_i, .., p_i^(r-1)) / prod_{j != i}(p_i - p_j),
    A^T column = (1, p_i, .., p_i^(m-1)), B^T row = coefficients (ascending) of
    prod_{j != i}(x - p_j); infinity: G row (0, .., 0, 1), A^T column (0, .., 0, 1), B^T row =
    coefficients of prod_j (x - p_j)."""
    n = m + r - 1
    pts = [Fr(p) for p in points]
    assert len(pts) == n - 1
    at = [[Fr(0)] * n for _ in range(m)]
    g = [[Fr(0)] * r for _ in range(n)]
    bt = [[Fr(0)] * n for _ in range(n)]
    for i, p in enumerate(pts):
        den = Fr(1)
        for j, q in enumerate(pts):
            if j != i:
                den *= p - q
        for k in range(m):
            at[k][i] = p ** k
        for k in range(r):
            g[i][k] = p ** k / den
        poly = [Fr(1)]
        for j, q in enumerate(pts):
            if j != i:
                poly = _poly_mul(poly, [-q, Fr(1)])
        bt[i][: len(poly)] = poly
    at[m - 1][n - 1] = Fr(1)
    g[n - 1][r - 1] = Fr(1)
    poly = [Fr(1)]
    for q in pts:
        poly = _poly_mul(poly, [-q, Fr(1)])
    bt[n - 1][: len(poly)] = poly
    return at, g, bt


def scaled_matrices(points=POINTS, gamma=GAMMA):
    """(A'^T, G', B^T) as exact rationals: G' = diag(gamma) G, A'^T = A^T diag(1/gamma)."""
    at, g, bt = cook_toom(points, 4, 3)
    gs = [[x * gamma[i] for x in g[i]] for i in range(6)]
    ats = [[at[a][i] / gamma[i] for i in range(6)] for a in range(4)]
    return ats, gs, bt


def _f64(mat) -> np.ndarray:
    return np.array([[float(x) for x in row] for row in mat], dtype=np.float64)


AT43_EXACT, G43_EXACT, BT43_EXACT = scaled_matrices()
AT43, G43, BT43 = _f64(AT43_EXACT), _f64(G43_EXACT), _f64(BT43_EXACT)


@dataclass(frozen=True)
class Geometry43:
    n: int
    c: int
    h: int
    w: int
    k: int
    p: int = 1
    q: int = 1

    @property
    def oh(self) -> int:
        return self.h - 3 + 2 * self.p + 1

    @property
    def ow(self) -> int:
        return self.w - 3 + 2 * self.q + 1

    @property
    def tiles_h(self) -> int:
        return (self.oh + 3) // 4

    @property
    def tiles_w(self) -> int:
        return (self.ow + 3) // 4

    @property
    def T(self) -> int:
        return self.tiles_h * self.tiles_w

    @property
    def NT(self) -> int:
        return self.n * self.T

    @property
    def G32(self) -> int:
        return (self.NT + 31) // 32

    @property
    def CB(self) -> int:
        return (self.c + 31) // 32

    @property
    def KB32(self) -> int:
        return (self.k + 31) // 32

    @property
    def v_halves(self) -> int:
        return NPOS * 2 * self.CB * self.G32 * FRAG

    @property
    def u_halves(self) -> int:
        return NPOS * 2 * self.CB * self.KB32 * FRAG

    @property
    def m_halves(self) -> int:
        return NPOS * self.KB32 * 32 * self.G32 * 32

    def gemm_flops(self) -> float:
        return 2.0 * NPOS * self.n * self.k * self.c * self.T


# ---- S1 ------------------------------------------------------------------------------------
def filter_transform43_f64(w16: np.ndarray) -> np.ndarray:
    """U = G' g G'^T in FP64 (exact): (36, K, C) float64, pos = 6 i + j."""
    u = np.einsum("ir,kcrs,js->ijkc", G43, w16.astype(np.float64), G43)
    return u.reshape(NPOS, *w16.shape[:2])


def filter_transform43(w16: np.ndarray) -> np.ndarray:
    """S1 as the kernel writes it: the exact U rounded once to FP16, a zero as +0 (+ 0.0)."""
    return (filter_transform43_f64(w16) + 0.0).astype(np.float16)


# ---- S2 ------------------------------------------------------------------------------------
def input_windows43(x16: np.ndarray, geo: Geometry43) -> np.ndarray:
    """6x6 windows d[n, c, th, tw, i, j] = x[n, c, 4th - p + i, 4tw - q + j] (0 outside)."""
    n, c, h, w = x16.shape
    rows = 4 * geo.tiles_h + 2
    cols = 4 * geo.tiles_w + 2
    xp = np.zeros((n, c, rows, cols), dtype=x16.dtype)
    hh = min(h, rows - geo.p)
    ww = min(w, cols - geo.q)
    xp[:, :, geo.p:geo.p + hh, geo.q:geo.q + ww] = x16[:, :, :hh, :ww]
    win = np.lib.stride_tricks.sliding_window_view(xp, (6, 6), axis=(2, 3))
    return win[:, :, ::4, ::4][:, :, :geo.tiles_h, :geo.tiles_w]


def _to_g(arr: np.ndarray, geo: Geometry43) -> np.ndarray:
    """(36, n, c, th, tw) -> (36, c, NT), column g = image * T + th * tiles_w + tw."""
    n, c = arr.shape[1], arr.shape[2]
    return arr.transpose(0, 2, 1, 3, 4).reshape(NPOS, c, n * geo.T)


def input_transform43_f64(x16: np.ndarray, geo: Geometry43) -> np.ndarray:
    """V = B^T d B in FP64 (exact): (36, C, NT) float64."""
    d = input_windows43(x16, geo).astype(np.float64)
    v = np.einsum("ai,nctsij,bj->abncts", BT43, d, BT43)
    return _to_g(v.reshape(NPOS, *v.shape[2:]), geo)


def _bt6_fp32(d):
    """One 6-point B^T transform in the kernel's FP32 operation order (wino43.hip bt6f: no
    contraction, each product and sum rounded once): d is a list of 6 float32 arrays."""
    f = np.float32
    a = d[1] - d[3]
    e = d[4] - d[2]
    return [f(1.5) * a + (f(-2.0) * d[2] + (d[0] + d[4])),
            f(-2.5) * d[2] + (f(-0.5) * d[3] + (d[4] - d[1])),
            f(0.5) * d[2] + (f(-2.5) * d[3] + (d[4] + d[1])),
            f(-0.5) * a + e,
            f(2.0) * a + e,
            f(-1.5) * e + (f(-2.0) * d[3] + (d[1] + d[5]))]


def input_transform43(x16: np.ndarray, geo: Geometry43) -> np.ndarray:
    """S2 as the kernel writes it: B^T d B in the kernel's FP32 operation order (columns first:
    t[:, j] = bt6(d[:, j]), then V[i, :] = bt6(t[i, :])), rounded once to FP16. (36, C, NT)."""
    d = input_windows43(x16, geo).astype(np.float32)
    t = [[None] * 6 for _ in range(6)]
    for j in range(6):
        col = _bt6_fp32([d[..., r, j] for r in range(6)])
        for i in range(6):
            t[i][j] = col[i]
    v = np.empty((NPOS,) + d.shape[:4], dtype=np.float16)
    for i in range(6):
        row = _bt6_fp32(t[i])
        for j in range(6):
            v[6 * i + j] = row[j].astype(np.float16)
    return _to_g(v, geo)


# ---- packed layouts (the same fragment layouts as F(2x2), 36 positions) --------------------
def pack_U43(u: np.ndarray, geo: Geometry43) -> np.ndarray:
    return _pack(u, geo.KB32)


def pack_V43(v: np.ndarray, geo: Geometry43) -> np.ndarray:
    return _pack(np.ascontiguousarray(v.transpose(0, 2, 1)), geo.G32)


def unpack_U43(flat: np.ndarray, geo: Geometry43) -> np.ndarray:
    return _unpack(flat, NPOS, geo.k, geo.c, geo.KB32)


def unpack_V43(flat: np.ndarray, geo: Geometry43) -> np.ndarray:
    return _unpack(flat, NPOS, geo.NT, geo.c, geo.G32).transpose(0, 2, 1)


def unpack_M43(flat: np.ndarray, geo: Geometry43) -> np.ndarray:
    """The stored M (mid_result, [pos][k / 4][column][k % 4], channels padded to KB32*32 and
    columns to G32*32) -> (36, K, NT) as stored (FP16 of M * M_SCALE)."""
    x = flat[: geo.m_halves].reshape(NPOS, geo.KB32 * 8, geo.G32 * 32, 4).transpose(0, 1, 3, 2)
    return np.ascontiguousarray(x).reshape(NPOS, geo.KB32 * 32, geo.G32 * 32)[:, :geo.k, :geo.NT]


# ---- S3, S4 --------------------------------------------------------------------------------
def gemm43_f64(u: np.ndarray, v: np.ndarray, absolute: bool = False) -> np.ndarray:
    """M = U . V per position in FP64 from (36, K, C) U and (36, C, NT) V: (36, K, NT).
    absolute=True: |U| . |V| (the scale of the FP32 accumulation bound)."""
    u = u.astype(np.float64)
    v = v.astype(np.float64)
    if absolute:
        u, v = np.abs(u), np.abs(v)
    return np.einsum("pkc,pcg->pkg", u, v)


def output_transform43_f64(m: np.ndarray, geo: Geometry43, absolute: bool = False) -> np.ndarray:
    """Y = A'^T M A' in FP64 from the unscaled M (36, K, NT) -> (N, K, Oh, Ow).
    absolute=True uses |A'| and |M| (the scale of the FP32 output-transform bound)."""
    at = np.abs(AT43) if absolute else AT43
    mm = np.abs(m) if absolute else m
    mm = mm.reshape(6, 6, geo.k, geo.n, geo.tiles_h, geo.tiles_w)
    y = np.einsum("ai,ijkntr,bj->nktarb", at, mm, at)
    y = y.reshape(geo.n, geo.k, 4 * geo.tiles_h, 4 * geo.tiles_w)
    return y[:, :, :geo.oh, :geo.ow]


def winograd43_exact(x16: np.ndarray, w16: np.ndarray, geo: Geometry43) -> np.ndarray:
    """The F(4x4) path in FP64 without any FP16 rounding: equals the direct convolution."""
    u = filter_transform43_f64(w16)
    v = input_transform43_f64(x16, geo)
    return output_transform43_f64(gemm43_f64(u, v), geo)


def winograd43_model(x16: np.ndarray, w16: np.ndarray, geo: Geometry43) -> np.ndarray:
    """The GPU data path with its three FP16 roundings (U, V, stored M) and exact arithmetic
    otherwise: (N, K, Oh, Ow) float64."""
    u = filter_transform43(w16)
    v = input_transform43(x16, geo)
    m = (gemm43_f64(u, v) * M_SCALE).astype(np.float16).astype(np.float64) / M_SCALE
    return output_transform43_f64(m, geo)
