// General-geometry path (any filter size, stride and padding): packing kernels of the
// implicit-GEMM direct convolution.
//
// The reference plugin handles only 3x3 stride-1 pad-1 and ignores r, s, u, v, p, q
// (src/conv2d.cpp:254-271: they only enter Oh/Ow), returning wrong results for anything
// else. Here such problems run as out[k][g] = sum_j W[k][j] * X[j][g] over the flattened
// reduction index j = tap * C + c (tap = ri * s + si), with g the output pixel of the batch,
// on the same MFMA GEMM pipeline as the Winograd path: these kernels write W and X in the
// packed fragment layouts the GEMM copies verbatim (a GEMM step covers 32 consecutive j;
// see mykernelParamType.h), and the GEMM accumulates all steps into one FP32 accumulator
// whose elements are the outputs (no output transform). Both are pure copies of FP16 values;
// j >= r*s*C (the padding up to nsteps * 32) is zero. Flattening keeps few-channel layers
// (C = 3 stems) from padding every tap to 32 channels.
#include "wino_common.h"

// One lane per (16-value block L = 2*st + kk, 32-row block kb, fragment lane): row
// k = kb*32 + lane%32, reduction indices j0 .. j0+7; zero for k >= K and j >= r*s*C.
// Consecutive lanes write one contiguous 1 KiB fragment.
__global__ void __launch_bounds__(256) direct_pack_weights(mykernelParamType prm) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int KB = prm.KB32;
    const long long frags = (long long)prm.nsteps * 2 * KB;
    if (gt >= frags * 64) return;
    const int lane = (int)(gt & 63);
    const long long frag = gt >> 6;
    const int kb = (int)(frag % KB);
    const int L = (int)(frag / KB);
    const int R = prm.r, S = prm.s, C = prm.c, K = prm.k;
    const int jn = R * S * C;
    const int k = kb * 32 + (lane & 31);
    const int j0 = L * 16 + (lane >> 5) * 8;
    // (tap, c) of j0 by division once, then stepped: c wraps into the next tap
    const int tap0 = j0 / C;
    int c = j0 - tap0 * C, ri = tap0 / S, si = tap0 - ri * S;
    half8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        v[e] = (k < K && j0 + e < jn) ? ((const half_t*)prm.pweight)[(((size_t)k * C + c) * R + ri) * S + si]
                                      : (half_t)0.0f;
        if (++c == C) {
            c = 0;
            if (++si == S) si = 0, ++ri;
        }
    }
    *(half8*)((half_t*)prm.transform_filter + ((size_t)L * KB + kb) * kFragHalves + lane * 8) = v;
}

// One lane per (16-value block L, 32-pixel block pb, fragment lane): output pixel
// g = pb*32 + lane%32, reduction indices j0 .. j0+7 of the im2col input; zero outside the
// image, for j >= r*s*C and g >= NT. Consecutive lanes take consecutive output pixels
// (adjacent input columns for stride 1) and write one contiguous 1 KiB fragment. The input is
// N,C,H,W, or N,H,W,C with CONV2D_IO_IN_NHWC (the same image base, channel-contiguous pixels).
__global__ void __launch_bounds__(256) direct_pack_input(mykernelParamType prm) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int P32 = prm.G32;
    const long long frags = (long long)prm.nsteps * 2 * P32;
    if (gt >= frags * 64) return;
    const int lane = (int)(gt & 63);
    const long long frag = gt >> 6;
    const int pb = (int)(frag % P32);
    const int L = (int)(frag / P32);
    const int S = prm.s, C = prm.c, H = prm.h, W = prm.w;
    const int jn = prm.r * S * C;
    const int g = pb * 32 + (lane & 31);
    const int img = g / prm.T;
    const int pix = g - img * prm.T;
    const int oh = pix / prm.Ow, ow = pix - (pix / prm.Ow) * prm.Ow;
    const int ih0 = oh * prm.u - prm.p, iw0 = ow * prm.v - prm.q;
    const int j0 = L * 16 + (lane >> 5) * 8;
    const half_t* x = (const half_t*)prm.pin + (g < prm.NT ? (size_t)img * C * H * W : 0);
    const int tap0 = j0 / C;
    int c = j0 - tap0 * C, ri = tap0 / S, si = tap0 - ri * S;
    half8 v;
    if ((C & 7) == 0) {
        // C % 8 == 0 (wave-uniform): j0 is a multiple of 8, so the lane's 8 values are channels
        // c .. c+7 of one tap: one bounds check, and for channel-contiguous input one 16-byte load
        // when the pixel's channels are 16-byte aligned
        const int ih = ih0 + ri, iw = iw0 + si;
        const bool ok = g < prm.NT && j0 < jn && ih >= 0 && ih < H && iw >= 0 && iw < W;
        v = half8{};
        if (ok) {
            if (prm.io & CONV2D_IO_IN_NHWC) {
                const half_t* px = x + ((size_t)ih * W + iw) * C + c;
                if ((((uintptr_t)px) & 15) == 0) {
                    v = *(const half8*)px;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = px[e];
                }
            } else {
                const half_t* px = x + ((size_t)c * H + ih) * W + iw;
                const size_t plane = (size_t)H * W;
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = px[e * plane];
            }
        }
        *(half8*)((half_t*)prm.transform_pin + ((size_t)L * P32 + pb) * kFragHalves + lane * 8) = v;
        return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int ih = ih0 + ri, iw = iw0 + si;
        const bool ok = g < prm.NT && j0 + e < jn && ih >= 0 && ih < H && iw >= 0 && iw < W;
        const size_t xi = (prm.io & CONV2D_IO_IN_NHWC) ? ((size_t)ih * W + iw) * C + c : ((size_t)c * H + ih) * W + iw;
        v[e] = ok ? x[xi] : (half_t)0.0f;
        if (++c == C) {
            c = 0;
            if (++si == S) si = 0, ++ri;
        }
    }
    *(half8*)((half_t*)prm.transform_pin + ((size_t)L * P32 + pb) * kFragHalves + lane * 8) = v;
}

// The general gather (C % 8 != 0) with the per-pixel work hoisted: one wave per 32-pixel block,
// each lane keeping its output pixel and its (tap, channel) position as it walks the 16-value
// blocks L (no divisions per fragment slot). Same values, same order as direct_pack_input.
__global__ void __launch_bounds__(256) direct_pack_input_walk(mykernelParamType prm) {
    const int P32 = prm.G32;
    const int pb = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (pb >= P32) return;  // whole waves
    const int lane = threadIdx.x & 63;
    const int S = prm.s, C = prm.c, H = prm.h, W = prm.w, NT = prm.NT;
    const int jn = prm.r * S * C;
    const int g = pb * 32 + (lane & 31);
    const int img = g < NT ? g / prm.T : 0;
    const int pix = g - img * prm.T;
    const int oh = pix / prm.Ow, ow = pix - (pix / prm.Ow) * prm.Ow;
    const int ih0 = oh * prm.u - prm.p, iw0 = ow * prm.v - prm.q;
    const half_t* x = (const half_t*)prm.pin + (size_t)img * C * H * W;
    // (tap, channel) of j = 8 * (lane / 32), stepped 8 values per block and 8 skipped (the other
    // half-wave's)
    int j = (lane >> 5) * 8;
    int c = j % C, ri = (j / C) / S, si = (j / C) % S;
    half_t* dst = (half_t*)prm.transform_pin + (size_t)pb * kFragHalves + lane * 8;
    const int nL = 2 * prm.nsteps;
    for (int L = 0; L < nL; ++L) {
        half8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ih = ih0 + ri, iw = iw0 + si;
            const bool ok = g < NT && j + e < jn && ih >= 0 && ih < H && iw >= 0 && iw < W;
            v[e] = ok ? x[((size_t)c * H + ih) * W + iw] : (half_t)0.0f;
            if (++c == C) {
                c = 0;
                if (++si == S) si = 0, ++ri;
            }
        }
        *(half8*)(dst + (size_t)L * P32 * kFragHalves) = v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (++c == C) {
                c = 0;
                if (++si == S) si = 0, ++ri;
            }
        }
        j += 16;
    }
}
