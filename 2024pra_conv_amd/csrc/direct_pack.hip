// General-geometry path (any filter size, stride and padding): packing kernels of the
// implicit-GEMM direct convolution.
//
// The reference plugin handles only 3x3 stride-1 pad-1 and ignores r, s, u, v, p, q
// (src/conv2d.cpp:254-271: they only enter Oh/Ow), returning wrong results for anything
// else. Here such problems run as out[k][g] = sum_j W[k][j] * X[j][g] over j = (filter tap,
// channel), with g the output pixel of the batch, on the same MFMA GEMM pipeline as the
// Winograd path: these kernels write W and X in the packed fragment layouts the GEMM copies
// verbatim (the Winograd position becomes the GEMM step st = tap * CB + channel chunk; see
// mykernelParamType.h), and the GEMM accumulates all steps into one FP32 accumulator whose
// elements are the outputs (no output transform). Both are pure copies of FP16 values.
#include "wino_common.h"

// One lane per (16-channel block L = 2*st + kk, 32-row block kb, fragment lane): row
// k = kb*32 + lane%32, channels c0 .. c0+7 of tap (ri, si); zero for k >= K, c >= C and the
// padding steps. Consecutive lanes write one contiguous 1 KiB fragment.
__global__ void __launch_bounds__(256) direct_pack_weights(mykernelParamType prm) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int KB = prm.KB32, CB = prm.CB;
    const long long frags = (long long)prm.nsteps * 2 * KB;
    if (gt >= frags * 64) return;
    const int lane = (int)(gt & 63);
    const long long frag = gt >> 6;
    const int kb = (int)(frag % KB);
    const int L = (int)(frag / KB);
    const int st = L >> 1, kk = L & 1;
    const int rs = st / CB, cb = st - rs * CB;
    const int R = prm.r, S = prm.s, C = prm.c, K = prm.k;
    const int ri = rs / S, si = rs - ri * S;
    const int k = kb * 32 + (lane & 31);
    const int c0 = cb * 32 + kk * 16 + (lane >> 5) * 8;
    const bool ok = rs < R * S && k < K;
    half8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int c = c0 + e;
        v[e] = (ok && c < C) ? ((const half_t*)prm.pweight)[(((size_t)k * C + c) * R + ri) * S + si] : (half_t)0.0f;
    }
    *(half8*)((half_t*)prm.transform_filter + ((size_t)L * KB + kb) * kFragHalves + lane * 8) = v;
}

// One lane per (16-channel block L, 32-pixel block pb, fragment lane): output pixel
// g = pb*32 + lane%32, channels c0 .. c0+7 of tap (ri, si) of the im2col input; zero outside
// the image, for c >= C, g >= NT and the padding steps. Consecutive lanes take consecutive
// output pixels (adjacent input columns for stride 1) and write one contiguous 1 KiB
// fragment.
__global__ void __launch_bounds__(256) direct_pack_input(mykernelParamType prm) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int P32 = prm.G32, CB = prm.CB;
    const long long frags = (long long)prm.nsteps * 2 * P32;
    if (gt >= frags * 64) return;
    const int lane = (int)(gt & 63);
    const long long frag = gt >> 6;
    const int pb = (int)(frag % P32);
    const int L = (int)(frag / P32);
    const int st = L >> 1, kk = L & 1;
    const int rs = st / CB, cb = st - rs * CB;
    const int S = prm.s, C = prm.c, H = prm.h, W = prm.w;
    const int ri = rs / S, si = rs - ri * S;
    const int g = pb * 32 + (lane & 31);
    const int img = g / prm.T;
    const int pix = g - img * prm.T;
    const int oh = pix / prm.Ow, ow = pix - (pix / prm.Ow) * prm.Ow;
    const int ih = oh * prm.u - prm.p + ri;
    const int iw = ow * prm.v - prm.q + si;
    const bool ok = rs < prm.r * S && g < prm.NT && ih >= 0 && ih < H && iw >= 0 && iw < W;
    const int c0 = cb * 32 + kk * 16 + (lane >> 5) * 8;
    const size_t base = ok ? ((size_t)img * C * H + ih) * W + iw : 0;  // index of (img, c = 0, ih, iw)
    const half_t* x = (const half_t*)prm.pin + base;
    half8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int c = c0 + e;
        v[e] = (ok && c < C) ? x[(size_t)c * H * W] : (half_t)0.0f;
    }
    *(half8*)((half_t*)prm.transform_pin + ((size_t)L * P32 + pb) * kFragHalves + lane * 8) = v;
}
