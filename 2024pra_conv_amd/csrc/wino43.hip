// Winograd F(4x4,3x3): filter transform (S1), input transform (S2) and output transform (S4)
// of the four-slot F(4x4) form; S3 is the per-position GEMM of wino_gemm.hip (kGemmPos mode),
// which stores M in FP16 to mid_result.
//
// No reference counterpart (the reference implements F(2x2,3x3): BTdB* preprocessing.cuh:273-424,
// GgGT* :75-270, store* conv2d.cpp:17-250); the matrices come from the Cook-Toom construction on
// the interpolation points {0, 1, -1, 2, -1/2} and infinity, with G's rows rescaled to dyadic
// coefficients (gamma = 1, 3/4, 3/4, 15/4, 15/16, 1) and A^T's columns by 1/gamma:
//   B^T = [[1, 3/2, -2, -3/2, 1, 0], [0, -1, -5/2, -1/2, 1, 0], [0, 1, 1/2, -5/2, 1, 0],
//          [0, -1/2, -1, 1/2, 1, 0], [0, 2, -1, -2, 1, 0], [0, 1, 3/2, -2, -3/2, 1]]
//   G'  = [[1, 0, 0], [-1/4, -1/4, -1/4], [1/4, -1/4, 1/4], [1/4, 1/2, 1], [-1, 1/2, -1/4], [0, 0, 1]]
//   A'^T = [[1, 4/3, 4/3, 4/15, 16/15, 0], [0, 4/3, -4/3, 8/15, -8/15, 0],
//           [0, 4/3, 4/3, 16/15, 4/15, 0], [0, 4/3, -4/3, 32/15, -2/15, 1]]
// S1 computes in FP64, where every sum of FP16 values times these dyadic constants is exact
// (fewer than 53 significant bits for any FP16 input), so U is the exact transform rounded once
// to FP16, whatever the operation order. S2 computes in FP32 in a fixed operation order (no
// contraction; the oracle restates it) and rounds once to FP16. S3 stores FP16(M / 4); S4 undoes
// the scale inside its FP32 coefficients (2 A'^T applied on both sides).
#include "wino_tile_ops.h"

namespace {

// One 6-point B^T transform in FP32, every product and sum rounded once in this order.
__device__ inline void bt6f(const float d[6], float o[6]) {
#pragma clang fp contract(off)
    const float a = d[1] - d[3], e = d[4] - d[2];
    o[0] = 1.5f * a + (-2.0f * d[2] + (d[0] + d[4]));
    o[1] = -2.5f * d[2] + (-0.5f * d[3] + (d[4] - d[1]));
    o[2] = 0.5f * d[2] + (-2.5f * d[3] + (d[4] + d[1]));
    o[3] = -0.5f * a + e;
    o[4] = 2.0f * a + e;
    o[5] = -1.5f * e + (-2.0f * d[3] + (d[1] + d[5]));
}

// One 3-point G' transform (exact in FP64).
__device__ inline void g6(const double g[3], double o[6]) {
    const double s = g[0] + g[2];
    o[0] = g[0];
    o[1] = -0.25 * (s + g[1]);
    o[2] = 0.25 * (s - g[1]);
    o[3] = __builtin_fma(0.25, g[0], __builtin_fma(0.5, g[1], g[2]));
    o[4] = __builtin_fma(-0.25, g[2], __builtin_fma(0.5, g[1], -g[0]));
    o[5] = g[2];
}

// One 6 -> 4 output transform with 2 A'^T (FP32): y_a = sum_i 2 A'^T[a][i] m_i.
__device__ inline void at4(const float m[6], float y[4]) {
    constexpr float c83 = 8.0f / 3.0f, c815 = 8.0f / 15.0f, c1615 = 16.0f / 15.0f, c3215 = 32.0f / 15.0f,
                    c6415 = 64.0f / 15.0f, c415 = 4.0f / 15.0f;
    const float s12 = m[1] + m[2], d12 = m[1] - m[2];
    y[0] = __builtin_fmaf(c3215, m[4], __builtin_fmaf(c815, m[3], __builtin_fmaf(c83, s12, 2.0f * m[0])));
    y[1] = __builtin_fmaf(c1615, m[3] - m[4], c83 * d12);
    y[2] = __builtin_fmaf(c815, m[4], __builtin_fmaf(c3215, m[3], c83 * s12));
    y[3] = __builtin_fmaf(-c415, m[4], __builtin_fmaf(c6415, m[3], __builtin_fmaf(c83, d12, 2.0f * m[5])));
}

}  // namespace

// S1: U = G' g G'^T, one thread per element (k, c) of the packed U[pos][c16][kb][lane][8]
// (lane = (c % 16 / 8) * 32 + k % 32, element c % 8): consecutive threads write consecutive
// halves of a fragment, so each of the 36 stores of a wave is 128 contiguous bytes. Zeros for
// k >= K or c >= C.
__global__ void __launch_bounds__(256) wino43_filter_transform(mykernelParamType prm) {
    const int KB = prm.KB32, C16 = 2 * prm.CB;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= C16 * KB * kFragHalves) return;
    const int frag = idx / kFragHalves, r = idx - frag * kFragHalves;
    const int c16 = frag / KB, kb = frag - c16 * KB;
    const int lane = r >> 3, e = r & 7;
    const int k = kb * 32 + (lane & 31);
    const int c = c16 * 16 + 8 * (lane >> 5) + e;
    const bool ok = k < prm.k && c < prm.c;
    const half_t* wp = (const half_t*)prm.pweight + ((size_t)(ok ? k : 0) * prm.c + (ok ? c : 0)) * 9;
    double t[6][3];  // (G' g)[i][s]
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        double g[3], o[6];
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) g[rr] = ok ? (double)wp[rr * 3 + s] : 0.0;
        g6(g, o);
#pragma unroll
        for (int i = 0; i < 6; ++i) t[i][s] = o[i];
    }
    half_t* dst = (half_t*)prm.transform_filter + (size_t)idx;
    const size_t pos_stride = (size_t)C16 * KB * kFragHalves;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double o[6];
        g6(t[i], o);
#pragma unroll
        for (int j = 0; j < 6; ++j) dst[(size_t)(6 * i + j) * pos_stride] = f64_to_f16(o[j]);
    }
}

// S2: V = B^T d B of the 6x6 window of every 4x4 tile in tile-row band bi (image n, 32-channel
// block cb), staged in LDS like the F(2x2) input transform. One thread per (tile of the band,
// 4-channel quad of the block): it writes its 8 bytes of the lane slot of each of the 36
// position fragments V[pos][c16][g32][lane][8] (lane = (c % 16 / 8) * 32 + g % 32). The last
// band of the last image also writes the padding columns g in [NT, G32*32) as zeros.
__global__ void __launch_bounds__(256) wino43_input_transform(mykernelParamType prm) {
    extern __shared__ __attribute__((aligned(16))) half_t lds[];
    const int n = blockIdx.z, cb = blockIdx.y, bi = blockIdx.x;
    const int TH = prm.band;
    const int th0 = bi * TH;
    const int tiles_w = prm.tiles_w;
    const int th1 = min(th0 + TH, prm.tiles_h);
    const int LR = input43_lds_rows(TH);
    const int LWp = input43_lds_pitch(tiles_w, prm.w, prm.q);
    stage_band(prm, n, cb, 4 * th0 - prm.p, LR, LWp, lds);
    const int T = prm.T, NT = prm.NT, G32 = prm.G32, CB16 = 2 * prm.CB;
    const int gs = n * T + th0 * tiles_w;
    const int ge = n * T + th1 * tiles_w;
    // the last band of the last image also covers the padding columns of the last 32-column block
    const int gend = ge == NT ? G32 * 32 : ge;
    const size_t pos_stride = (size_t)CB16 * G32 * kFragHalves;
    typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
    const int items = (gend - gs) * 8;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int quad = it & 7;  // channels 4*quad .. +3 of the staged 32
        const int g = gs + (it >> 3);
        const int c16 = 2 * cb + (quad >> 2);
        const int lane = (g & 31) + 32 * ((quad >> 1) & 1);
        half_t* dst = (half_t*)prm.transform_pin + ((size_t)c16 * G32 + (g >> 5)) * kFragHalves + lane * 8 +
                      4 * (quad & 1);
        uint32_t outv[36][2];  // [pos][channel pair]
#pragma unroll
        for (int pos = 0; pos < 36; ++pos) outv[pos][0] = outv[pos][1] = 0u;
        if (g < ge) {
            const int t = g - n * T;
            const int th = t / tiles_w;
            const int tw = t - th * tiles_w;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int ch = 4 * quad + e;
                const half_t* base = lds + (ch * LR + 4 * (th - th0)) * LWp + 4 * tw + 8;
                float tt[6][6];  // tt[i][j] = (B^T d)[i][j]
#pragma unroll
                for (int j2 = 0; j2 < 3; ++j2) {
                    float c0[6], c1[6];
#pragma unroll
                    for (int r = 0; r < 6; ++r) {
                        const half2_t pr = *(const half2_t*)(base + r * LWp + 2 * j2);
                        c0[r] = (float)pr[0];
                        c1[r] = (float)pr[1];
                    }
                    float o0[6], o1[6];
                    bt6f(c0, o0);
                    bt6f(c1, o1);
#pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        tt[i][2 * j2] = o0[i];
                        tt[i][2 * j2 + 1] = o1[i];
                    }
                }
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    float o[6];
                    bt6f(tt[i], o);
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        const uint32_t hb = __builtin_bit_cast(uint16_t, (half_t)o[j]);
                        outv[6 * i + j][e >> 1] |= hb << (16 * (e & 1));
                    }
                }
            }
        }
#pragma unroll
        for (int pos = 0; pos < 36; ++pos) {
            const u32x2_t pk = {outv[pos][0], outv[pos][1]};
            *(u32x2_t*)(dst + (size_t)pos * pos_stride) = pk;
        }
    }
}

// S4: Y = (2 A'^T) M (2 A'^T)^T from the stored FP16 M (mid_result, [pos][k/4][column][k%4],
// FP16(M / 4)), in FP32, then the epilogue and one rounding to the output type. Thread = one
// 4x4 tile (GEMM column t, consecutive threads on consecutive tiles) x 4 output channels.
__global__ void __launch_bounds__(256) wino43_output_transform(mykernelParamType prm) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int kq = blockIdx.y;
    if (t >= prm.NT) return;
    const int K = prm.k;
    const size_t KQp = (size_t)prm.KB32 * 8, NTp = (size_t)prm.G32 * 32;
    typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
    const u32x2_t* __restrict__ M = (const u32x2_t*)prm.mid_result + (size_t)kq * NTp + t;
    float m[4][36];  // [channel][pos]
#pragma unroll
    for (int pos = 0; pos < 36; ++pos) {
        const u32x2_t v = M[(size_t)pos * KQp * NTp];
        // the elements go through scalars: hipcc (ROCm 7.2) compiles __builtin_bit_cast of an
        // ext_vector element expression such as v[1] to a read of element 0
        const uint32_t w0 = v[0], w1 = v[1];
        const half2_t lo = __builtin_bit_cast(half2_t, w0), hi = __builtin_bit_cast(half2_t, w1);
        m[0][pos] = (float)lo[0];
        m[1][pos] = (float)lo[1];
        m[2][pos] = (float)hi[0];
        m[3][pos] = (float)hi[1];
    }
    const int img = t / prm.T;
    const int tt = t - img * prm.T;
    const int th = tt / prm.tiles_w;
    const int tw = tt - th * prm.tiles_w;
    const int oh0 = 4 * th, ow0 = 4 * tw;
    const int Oh = prm.Oh, Ow = prm.Ow;
    float y[4][16];  // [channel][4 a + b]
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        float z[4][6];  // z[a][j] = sum_i 2A'^T[a][i] M[i][j]
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            float col[6], o[4];
#pragma unroll
            for (int i = 0; i < 6; ++i) col[i] = m[kk][6 * i + j];
            at4(col, o);
#pragma unroll
            for (int a = 0; a < 4; ++a) z[a][j] = o[a];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            float o[4];
            at4(z[a], o);
#pragma unroll
            for (int b = 0; b < 4; ++b) y[kk][4 * a + b] = o[b];
        }
        const int k = 4 * kq + kk;
        if ((prm.bias || prm.act) && k < K) {
            const float bk = epilogue_bias(prm, k);
#pragma unroll
            for (int i = 0; i < 16; ++i) y[kk][i] = epilogue(prm, y[kk][i], bk);
        }
    }
    const bool bf = (prm.io & CONV2D_IO_OUT_BF16) != 0;
    const bool full = oh0 + 3 < Oh && ow0 + 3 < Ow;
    if (prm.io & CONV2D_IO_OUT_NHWC) {
        // per pixel, the 4 channels in one 8-byte store when they all exist and align
        const bool vec = 4 * kq + 3 < K && (K & 3) == 0 && ((uintptr_t)prm.pout & 7) == 0;
        global_u16_t* o = (global_u16_t*)prm.pout;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int oh = oh0 + a, ow = ow0 + b;
                if (oh >= Oh || ow >= Ow) continue;
                const size_t px = ((size_t)img * Oh + oh) * Ow + ow;
                if (vec) {
                    const u32x2_t pk = {pack_out2(bf, y[0][4 * a + b], y[1][4 * a + b]),
                                        pack_out2(bf, y[2][4 * a + b], y[3][4 * a + b])};
                    *(__attribute__((address_space(1))) u32x2_t*)(o + px * K + 4 * kq) = pk;
                } else {
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
                        if (4 * kq + kk < K) o[px * K + 4 * kq + kk] = out_bits(prm, y[kk][4 * a + b]);
                }
            }
        return;
    }
    // N,K,Oh,Ow: a tile row of 4 pixels in one 8-byte store when the tile is inside and rows
    // start 8-byte aligned
    const bool vec = full && (Ow & 3) == 0 && ((uintptr_t)prm.pout & 7) == 0;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const int k = 4 * kq + kk;
        if (k >= K) continue;
        global_u16_t* o = (global_u16_t*)prm.pout + (((size_t)img * K + k) * Oh + oh0) * Ow + ow0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (vec) {
                const u32x2_t pk = {pack_out2(bf, y[kk][4 * a], y[kk][4 * a + 1]),
                                    pack_out2(bf, y[kk][4 * a + 2], y[kk][4 * a + 3])};
                *(__attribute__((address_space(1))) u32x2_t*)(o + (size_t)a * Ow) = pk;
            } else if (oh0 + a < Oh) {
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (ow0 + b < Ow) o[(size_t)a * Ow + b] = out_bits(prm, y[kk][4 * a + b]);
            }
        }
    }
}
