"""TEST INFRASTRUCTURE ONLY — numpy model of the Winograd F(2x2,3x3) path.

Follows the reference's four stages:
  S1 filter transform U = G g G^T      G   from preprocessing.cuh:77-81 (formulas :93-108)
  S2 input transform  V = B^T d B      B^T from preprocessing.cuh:50-54 (formulas :56-71)
  S3 per-position GEMM M = U^T V       matrixmul.cuh:139-556 (16 independent GEMMs)
  S4 output transform Y = A^T M A      A^T from conv2d.cpp:18-22 (formulas :37-40)
with the geometry of readme.txt:17-24 and the window origin of verfiy.h:19-20
(row 2*th - p, column 2*tw - q for output tile (th, tw)).

Two precisions:
  *_f64   exact transforms in FP64 (the mathematical definition);
  *_fp32  the GPU kernels' operation order in FP32 with one rounding to FP16 on store, so
          the packed U and V the kernels write can be compared bit for bit.

Pinned by tests/test_oracle.py: Winograd identity against brute-force correlation on random
dyadic tiles (exact), fp32 transforms within 1 FP16 ulp of the exact FP64 transforms, the
end-to-end FP64 model against the C direct conv (itself pinned by closed forms, torch and
the reference's own checker), packing round trips.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Transform matrices (Lavin & Gray F(2,3)); the same as the reference's comments.
G = np.array([[1, 0, 0], [0.5, 0.5, 0.5], [0.5, -0.5, 0.5], [0, 0, 1]], dtype=np.float64)
BT = np.array([[1, 0, -1, 0], [0, 1, 1, 0], [0, -1, 1, 0], [0, 1, 0, -1]], dtype=np.float64)
AT = np.array([[1, 1, 1, 0], [0, 1, -1, -1]], dtype=np.float64)

FRAG = 512  # halves per packed MFMA fragment (64 lanes x 8)


@dataclass(frozen=True)
class Geometry:
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
        return (self.oh + 1) // 2

    @property
    def tiles_w(self) -> int:
        return (self.ow + 1) // 2

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
        return 16 * 2 * self.CB * self.G32 * FRAG

    @property
    def u_halves(self) -> int:
        return 16 * 2 * self.CB * self.KB32 * FRAG

    def direct_flops(self) -> float:
        return 2.0 * self.n * self.k * self.c * 9 * self.oh * self.ow

    def gemm_flops(self) -> float:
        return 32.0 * self.n * self.k * self.c * self.T


# ---------------------------------------------------------------------------------------
# Seeded generators (counter-based, so any element can be regenerated)
# ---------------------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """SplitMix64 of (seed + (i+1)*golden), vectorised; uint64 wrap-around arithmetic."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def unit(seed: int, count: int, offset: int = 0) -> np.ndarray:
    idx = np.arange(offset, offset + count, dtype=np.uint64)
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_pm1(seed: int, shape) -> np.ndarray:
    """uniform[-1, 1] rounded (RNE) to FP16."""
    cnt = int(np.prod(shape))
    return (2.0 * unit(seed, cnt) - 1.0).astype(np.float16).reshape(shape)


def normal(seed: int, shape, sigma: float = 1.0) -> np.ndarray:
    """N(0, sigma^2) by Box-Muller on consecutive uniform pairs, rounded to FP16."""
    cnt = int(np.prod(shape))
    u = unit(seed, 2 * cnt).reshape(cnt, 2)
    z = np.sqrt(-2.0 * np.log1p(-u[:, 0])) * np.cos(2.0 * np.pi * u[:, 1])
    return (sigma * z).astype(np.float16).reshape(shape)


def with_outliers(x16: np.ndarray, seed: int, frac: float = 0.01, scale: float = 50.0) -> np.ndarray:
    """Element i is multiplied by `scale` iff hash(seed, i) < frac * 2^64."""
    h = splitmix64(seed, np.arange(x16.size, dtype=np.uint64))
    mask = h < np.uint64(int(frac * 2.0 ** 64))
    y = x16.astype(np.float64).ravel()
    y[mask] *= scale
    return y.astype(np.float16).reshape(x16.shape)


# ---------------------------------------------------------------------------------------
# S1 filter transform
# ---------------------------------------------------------------------------------------
def filter_transform_f64(w16: np.ndarray) -> np.ndarray:
    """Exact U = G g G^T. w16: (K, C, 3, 3) -> (16, K, C) float64, position = 4a + b."""
    g = w16.astype(np.float64)
    u = np.einsum("ar,kcrs,bs->abkc", G, g, G)
    return u.reshape(16, *w16.shape[:2])


def filter_transform_fp32(w16: np.ndarray) -> np.ndarray:
    """The kernel's FP32 operation order (wino_transforms.hip filter_tile_fp32), FP16 out."""
    f = np.float32
    g = w16.astype(np.float32)
    half = f(0.5)
    t = [[None] * 3 for _ in range(4)]
    for j in range(3):
        g0, g1, g2 = g[:, :, 0, j], g[:, :, 1, j], g[:, :, 2, j]
        t[0][j] = g0
        t[1][j] = ((g0 + g1) + g2) * half
        t[2][j] = ((g0 - g1) + g2) * half
        t[3][j] = g2
    out = np.empty((16,) + w16.shape[:2], dtype=np.float16)
    for i in range(4):
        a0, a1, a2 = t[i]
        out[i * 4 + 0] = a0.astype(np.float16)
        out[i * 4 + 1] = (((a0 + a1) + a2) * half).astype(np.float16)
        out[i * 4 + 2] = (((a0 - a1) + a2) * half).astype(np.float16)
        out[i * 4 + 3] = a2.astype(np.float16)
    return out


# ---------------------------------------------------------------------------------------
# S2 input transform
# ---------------------------------------------------------------------------------------
def input_windows(x16: np.ndarray, geo: Geometry) -> np.ndarray:
    """4x4 windows d[n, c, th, tw, i, j] = x[n, c, 2th - p + i, 2tw - q + j] (0 outside)."""
    n, c, h, w = x16.shape
    rows = 2 * geo.tiles_h + 2
    cols = 2 * geo.tiles_w + 2
    xp = np.zeros((n, c, rows, cols), dtype=x16.dtype)
    hh = min(h, rows - geo.p)
    ww = min(w, cols - geo.q)
    xp[:, :, geo.p:geo.p + hh, geo.q:geo.q + ww] = x16[:, :, :hh, :ww]
    win = np.lib.stride_tricks.sliding_window_view(xp, (4, 4), axis=(2, 3))
    return win[:, :, ::2, ::2][:, :, :geo.tiles_h, :geo.tiles_w]


def _to_g(arr: np.ndarray, geo: Geometry) -> np.ndarray:
    """(16, n, c, th, tw) -> (16, c, NT) with g = image*T + th*tiles_w + tw."""
    n, c = arr.shape[1], arr.shape[2]
    return arr.transpose(0, 2, 1, 3, 4).reshape(16, c, n * geo.T)


def input_transform_f64(x16: np.ndarray, geo: Geometry) -> np.ndarray:
    """Exact V = B^T d B: (16, C, NT) float64."""
    d = input_windows(x16, geo).astype(np.float64)
    v = np.einsum("ai,nctsij,bj->abncts", BT, d, BT)
    return _to_g(v.reshape(16, *v.shape[2:]), geo)


def input_transform_fp32(x16: np.ndarray, geo: Geometry, keep_fp32: bool = False) -> np.ndarray:
    """The kernel's FP32 operation order (wino_transforms.hip input_tile_fp32), FP16 out
    (FP32 out with keep_fp32, as the FP8 input transform consumes it)."""
    d = input_windows(x16, geo).astype(np.float32)
    s = [[None] * 4 for _ in range(4)]
    for j in range(4):
        d0, d1, d2, d3 = (d[..., i, j] for i in range(4))
        s[0][j] = d0 - d2
        s[1][j] = d1 + d2
        s[2][j] = d2 - d1
        s[3][j] = d1 - d3
    v = np.empty((16,) + d.shape[:4], dtype=np.float32 if keep_fp32 else np.float16)
    for i in range(4):
        v[i * 4 + 0] = (s[i][0] - s[i][2]).astype(v.dtype)
        v[i * 4 + 1] = (s[i][1] + s[i][2]).astype(v.dtype)
        v[i * 4 + 2] = (s[i][2] - s[i][1]).astype(v.dtype)
        v[i * 4 + 3] = (s[i][1] - s[i][3]).astype(v.dtype)
    return _to_g(v, geo)


def input_transform_fp16(x16: np.ndarray, geo: Geometry) -> np.ndarray:
    """The fused kernel's input transform (wino_fused.hip), FP16 arithmetic with one rounding
    per operation, as the reference's own transform is FP16 arithmetic
    (preprocessing.cuh:57-72, in another operation order): rows first,
    s_j = d0-d2 | d1+d2 | d2-d1 | d1-d3 of window column j, then V_b = s0-s2 | s1+s2 | s2-s1 |
    s1-s3. numpy evaluates each float16 operation in float32 and rounds once to float16,
    which equals the correctly rounded float16 result for + and - (float32 carries more than
    2*11+2 significand bits). Returns (16, C, NT) float16."""
    d = input_windows(x16, geo).astype(np.float16)
    v = np.empty((16,) + d.shape[:4], dtype=np.float16)
    for a in range(4):
        rp, rm, add = [(0, 2, False), (1, 2, True), (2, 1, False), (1, 3, False)][a]
        s = [d[..., rp, j] + d[..., rm, j] if add else d[..., rp, j] - d[..., rm, j] for j in range(4)]
        v[4 * a + 0] = s[0] - s[2]
        v[4 * a + 1] = s[1] + s[2]
        v[4 * a + 2] = s[2] - s[1]
        v[4 * a + 3] = s[1] - s[3]
    return _to_g(v, geo)


def fused_gpu_model(x16: np.ndarray, w16: np.ndarray, geo: Geometry):
    """The fused form's data path: FP16 U as the filter transform writes it, FP16-arithmetic
    V, then the GEMM and output transform in FP64. Returns (y_model, bound_scale) as
    winograd_gpu_model does."""
    u = filter_transform_fp32(w16)
    v = input_transform_fp16(x16, geo)
    return gemm_output_f64(u, v, geo), gemm_output_f64(u, v, geo, absolute=True)


# ---------------------------------------------------------------------------------------
# Packed layouts (the v_mfma_f32_32x32x16_f16 operand layouts the kernels use)
#   U[pos][c16][kb][lane][e], kb = k//32, lane = (c%16//8)*32 + k%32, e = c%8
#   V[pos][c16][gb][lane][e], gb = g//32, lane = (c%16//8)*32 + g%32, e = c%8
# ---------------------------------------------------------------------------------------
def _pack(a: np.ndarray, rows_blocks: int) -> np.ndarray:
    """a: (16, R, C) with R = row dim (k or g). Returns the flat packed array: fragments of
    32 rows x 16 channels, [pos][c16][rb][lane][e], lane = (c % 16 // 8) * 32 + r % 32;
    channels padded to a multiple of 32."""
    P, R, C = a.shape
    C16 = 2 * ((C + 31) // 32)
    full = np.zeros((P, rows_blocks * 32, C16 * 16), dtype=a.dtype)
    full[:, :R, :C] = a
    # (P, rb, r32, c16, cg, e) -> (P, c16, rb, cg, r32, e)
    x = full.reshape(P, rows_blocks, 32, C16, 2, 8).transpose(0, 3, 1, 4, 2, 5)
    return np.ascontiguousarray(x).reshape(-1)


def _unpack(flat: np.ndarray, P: int, R: int, C: int, rows_blocks: int) -> np.ndarray:
    C16 = 2 * ((C + 31) // 32)
    x = flat.reshape(P, C16, rows_blocks, 2, 32, 8).transpose(0, 2, 4, 1, 3, 5)
    return np.ascontiguousarray(x).reshape(P, rows_blocks * 32, C16 * 16)[:, :R, :C]


def pack_U(u: np.ndarray, geo: Geometry) -> np.ndarray:
    """u: (16, K, C) -> flat packed U (length geo.u_halves)."""
    return _pack(u, geo.KB32)


def pack_V(v: np.ndarray, geo: Geometry) -> np.ndarray:
    """v: (16, C, NT) -> flat packed V (length geo.v_halves)."""
    return _pack(np.ascontiguousarray(v.transpose(0, 2, 1)), geo.G32)


def unpack_U(flat: np.ndarray, geo: Geometry) -> np.ndarray:
    return _unpack(flat, 16, geo.k, geo.c, geo.KB32)


def unpack_V(flat: np.ndarray, geo: Geometry) -> np.ndarray:
    return _unpack(flat, 16, geo.NT, geo.c, geo.G32).transpose(0, 2, 1)


# ---------------------------------------------------------------------------------------
# S3 + S4: per-position GEMM and output transform
# ---------------------------------------------------------------------------------------
def gemm_output_f64(u: np.ndarray, v: np.ndarray, geo: Geometry, absolute: bool = False) -> np.ndarray:
    """Y = A^T (U . V per position) A in FP64, from transformed (16,K,C) U and (16,C,NT) V.

    absolute=True computes the error-bound scale sum_pos |coef| * (|U| . |V|) instead.
    Returns (N, K, Oh, Ow) float64.
    """
    u = u.astype(np.float64)
    v = v.astype(np.float64)
    if absolute:
        u, v = np.abs(u), np.abs(v)
    m = np.einsum("pkc,pcg->pkg", u, v, optimize=True).reshape(4, 4, geo.k, geo.NT)
    at = np.abs(AT) if absolute else AT
    y = np.einsum("ia,abkg,jb->ijkg", at, m, at, optimize=True)  # (2, 2, K, NT)
    y = y.reshape(2, 2, geo.k, geo.n, geo.tiles_h, geo.tiles_w)
    full = y.transpose(3, 2, 4, 0, 5, 1).reshape(geo.n, geo.k, 2 * geo.tiles_h, 2 * geo.tiles_w)
    return np.ascontiguousarray(full[:, :, :geo.oh, :geo.ow])


def winograd_f64(x16: np.ndarray, w16: np.ndarray, geo: Geometry) -> np.ndarray:
    """The whole method in exact FP64 arithmetic (equals direct conv up to FP64 rounding)."""
    return gemm_output_f64(filter_transform_f64(w16), input_transform_f64(x16, geo), geo)


def winograd_gpu_model(x16: np.ndarray, w16: np.ndarray, geo: Geometry):
    """The GPU method's data path: FP16 U and V exactly as the kernels write them, then the
    GEMM and output transform in FP64. Returns (y_model_f64, bound_scale_f64) where
    bound_scale = sum_pos |coef| |U|.|V| bounds the FP32 accumulation error."""
    u = filter_transform_fp32(w16)
    v = input_transform_fp32(x16, geo)
    return gemm_output_f64(u, v, geo), gemm_output_f64(u, v, geo, absolute=True)


# ---------------------------------------------------------------------------------------
# Error metrics (stated for every result)
#   max_rel = max|y - y*| / max|y*|   (normwise; elementwise ratios are ill-posed for
#                                      signed data whose outputs cross zero)
#   rms_rel = ||y - y*||_2 / ||y*||_2
# ---------------------------------------------------------------------------------------
def error_metrics(y: np.ndarray, ref: np.ndarray) -> dict:
    y = y.astype(np.float64)
    ref = ref.astype(np.float64)
    d = np.abs(y - ref)
    mx = float(np.max(np.abs(ref))) if ref.size else 0.0
    nr = float(np.sqrt(np.sum(ref * ref)))
    return {
        "max_rel": float(np.max(d)) / mx if mx > 0 else float(np.max(d)),
        "rms_rel": float(np.sqrt(np.sum(d * d))) / nr if nr > 0 else float(np.sqrt(np.sum(d * d))),
        "finite": bool(np.all(np.isfinite(y))),
    }


def reference_criterion_failures(y16: np.ndarray, ref: np.ndarray) -> int:
    """The reference harness's acceptance test (main.cpp:140-151): |gpu - cpu| / gpu > 0.01,
    or NaN/Inf; denominator is the GPU value. Meaningful only for positive data."""
    y = y16.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        bad = (np.abs(y - ref) / y > 0.01) | ~np.isfinite(y)
    return int(np.count_nonzero(bad))
