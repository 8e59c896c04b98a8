// Validation-only FP64 direct convolution (any r, s, u, v, p, q).
//
// Same cross-correlation as the reference's CPU checker conv2dcpu (include/verfiy.h:5-43):
// window origin (oh*u - p, ow*v - q), zero padding by bounds test (verfiy.h:30), NCHW
// input (or N,H,W,C with CONV2D_IO_IN_NHWC) and KCRS weight indexing (verfiy.h:32), loop
// order c -> r -> s; the FP64 result is N,K,Oh,Ow. It differs in one
// respect, on purpose: each FP16 x FP16 product is formed exactly in FP64 (the reference
// rounds it to FP16 first), and the result stays FP64. Since every product is exact, the
// sum in this order is bit-identical to a CPU FP64 loop in the same order.
//
// This is the only FP64 arithmetic on the GPU; no production stage uses it.
#include "wino_common.h"

__global__ void __launch_bounds__(256) conv_direct_f64_kernel(mykernelParamType prm, double* out) {
    const int N = prm.n, C = prm.c, H = prm.h, W = prm.w, K = prm.k;
    const int R = prm.r, S = prm.s, su = prm.u, sv = prm.v, pp = prm.p, qq = prm.q;
    const int Oh = prm.Oh, Ow = prm.Ow;
    const size_t total = (size_t)N * K * Oh * Ow;
    const half_t* __restrict__ in = (const half_t*)prm.pin;
    const half_t* __restrict__ wt = (const half_t*)prm.pweight;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int ow = (int)(idx % Ow);
        const int oh = (int)((idx / Ow) % Oh);
        const int k = (int)((idx / ((size_t)Ow * Oh)) % K);
        const int n = (int)(idx / ((size_t)Ow * Oh * K));
        const int h0 = oh * su - pp, w0 = ow * sv - qq;
        double sum = 0.0;
        const bool nhwc = (prm.io & CONV2D_IO_IN_NHWC) != 0;
        for (int c = 0; c < C; ++c) {
            // element (ih, iw) of channel c at ip[ih * rs + iw * cs]
            const half_t* ip = nhwc ? in + (size_t)n * C * H * W + c : in + ((size_t)n * C + c) * H * W;
            const size_t rs = nhwc ? (size_t)W * C : (size_t)W, cs = nhwc ? (size_t)C : 1;
            const half_t* wp = wt + ((size_t)k * C + c) * R * S;
            for (int r = 0; r < R; ++r) {
                const int ih = h0 + r;
                if (ih < 0 || ih >= H) continue;
                for (int s = 0; s < S; ++s) {
                    const int iw = w0 + s;
                    if (iw < 0 || iw >= W) continue;
                    sum += (double)ip[ih * rs + iw * cs] * (double)wp[r * S + s];
                }
            }
        }
        out[idx] = sum;
    }
}
