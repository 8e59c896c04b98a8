/*
 * oracle/direct_conv.c — TEST INFRASTRUCTURE ONLY. CPU restatement of the reference's
 * checker conv2dcpu (include/verfiy.h:5-43), written from its behaviour. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this; the product never does.
 *
 * Two functions:
 *   oracle_conv2d_f64       : exact FP64 direct convolution. Every FP16 x FP16 product is
 *                             exact in FP64; summation order c -> r -> s per output
 *                             (verfiy.h:24-36); output kept in FP64.
 *   oracle_conv2dcpu_faithful: the reference's own arithmetic: each product rounded to FP16
 *                             before widening (verfiy.h:32), FP64 sum, FP16 output
 *                             (verfiy.h:38); same loop order n -> k -> oh -> ow -> c -> r -> s
 *                             (verfiy.h:10-36). Used for the CPU baseline and to check the
 *                             reference's acceptance test.
 * FP16 values are passed as raw uint16 bits; FP16 rounding is implemented here (RNE, with
 * subnormals and overflow to infinity), since the host compiler may lack _Float16.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

double oracle_half_to_double(uint16_t h) {
    const int sign = (h >> 15) & 1;
    const int exp = (h >> 10) & 0x1f;
    const int man = h & 0x3ff;
    double v;
    if (exp == 0) v = ldexp((double)man, -24);
    else if (exp == 31) v = man ? NAN : INFINITY;
    else v = ldexp((double)(man | 0x400), exp - 25);
    return sign ? -v : v;
}

/* Round a double to the nearest FP16 value (ties to even), returned as a double. */
double oracle_round_half(double x) {
    if (isnan(x) || isinf(x) || x == 0.0) return x;
    const double ax = fabs(x);
    double q;
    if (ax < 0x1p-14) {
        q = 0x1p-24; /* subnormal quantum */
    } else {
        int e;
        frexp(ax, &e); /* ax = f * 2^e, f in [0.5, 1) -> leading bit 2^(e-1) */
        q = ldexp(1.0, e - 1 - 10);
    }
    double r = rint(ax / q) * q; /* ax/q exact (power-of-two scaling); rint = RNE */
    if (r > 65504.0) r = INFINITY;
    return x < 0 ? -r : r;
}

uint16_t oracle_double_to_half(double x) {
    const double r = oracle_round_half(x);
    uint16_t sign = (r < 0 || (r == 0 && signbit(r))) ? 0x8000 : 0;
    const double ar = fabs(r);
    if (isnan(r)) return 0x7e00;
    if (isinf(ar)) return sign | 0x7c00;
    if (ar == 0.0) return sign;
    if (ar < 0x1p-14) return sign | (uint16_t)(ar / 0x1p-24);
    int e;
    const double f = frexp(ar, &e); /* ar = f*2^e */
    const int be = e - 1 + 15;
    const uint16_t man = (uint16_t)((f * 2.0 - 1.0) * 1024.0);
    return sign | (uint16_t)(be << 10) | man;
}

static inline int out_dim(int x, int k, int pad, int stride) { return (x - k + 2 * pad) / stride + 1; }

void oracle_conv2d_f64(const uint16_t* in, const uint16_t* wt, double* out, int n, int c, int h,
                       int w, int k, int r, int s, int u, int v, int p, int q) {
    const int oh = out_dim(h, r, p, u), ow = out_dim(w, s, q, v);
    for (int in_ = 0; in_ < n; ++in_)
        for (int kk = 0; kk < k; ++kk)
            for (int i = 0; i < oh; ++i)
                for (int j = 0; j < ow; ++j) {
                    double sum = 0.0;
                    const int ph = i * u - p, pw = j * v - q;
                    for (int cc = 0; cc < c; ++cc)
                        for (int kh = 0; kh < r; ++kh)
                            for (int kw = 0; kw < s; ++kw) {
                                const int y = ph + kh, x = pw + kw;
                                if (x >= 0 && y >= 0 && x < w && y < h)
                                    sum += oracle_half_to_double(in[(size_t)in_ * c * h * w + (size_t)cc * h * w + (size_t)y * w + x]) *
                                           oracle_half_to_double(wt[(size_t)kk * r * s * c + (size_t)cc * r * s + kh * s + kw]);
                            }
                    out[(size_t)in_ * k * oh * ow + (size_t)kk * oh * ow + (size_t)i * ow + j] = sum;
                }
}

/* Faster variant of oracle_conv2d_f64 for larger checks: same per-output summation order,
 * products exact, with inputs decoded once into caller-provided FP64 scratch
 * (din: n*c*h*w doubles, dw: k*c*r*s doubles). Results are bit-identical to
 * oracle_conv2d_f64. Work over (n, k) pairs can be split with k0/k1. */
void oracle_conv2d_f64_decoded(const double* din, const double* dw, double* out, int n, int c,
                               int h, int w, int k, int r, int s, int u, int v, int p, int q,
                               int k0, int k1) {
    const int oh = out_dim(h, r, p, u), ow = out_dim(w, s, q, v);
    for (int in_ = 0; in_ < n; ++in_)
        for (int kk = k0; kk < k1; ++kk)
            for (int i = 0; i < oh; ++i)
                for (int j = 0; j < ow; ++j) {
                    double sum = 0.0;
                    const int ph = i * u - p, pw = j * v - q;
                    for (int cc = 0; cc < c; ++cc) {
                        const double* ip = din + ((size_t)in_ * c + cc) * h * w;
                        const double* wp = dw + ((size_t)kk * c + cc) * r * s;
                        for (int kh = 0; kh < r; ++kh) {
                            const int y = ph + kh;
                            if (y < 0 || y >= h) continue;
                            for (int kw = 0; kw < s; ++kw) {
                                const int x = pw + kw;
                                if (x < 0 || x >= w) continue;
                                sum += ip[(size_t)y * w + x] * wp[kh * s + kw];
                            }
                        }
                    }
                    out[(size_t)in_ * k * oh * ow + (size_t)kk * oh * ow + (size_t)i * ow + j] = sum;
                }
}

void oracle_decode_halves(const uint16_t* src, double* dst, size_t count) {
    for (size_t i = 0; i < count; ++i) dst[i] = oracle_half_to_double(src[i]);
}

void oracle_conv2dcpu_faithful(const uint16_t* in, const uint16_t* wt, uint16_t* out, int n,
                               int c, int h, int w, int k, int r, int s, int u, int v, int p,
                               int q) {
    const int oh = out_dim(h, r, p, u), ow = out_dim(w, s, q, v);
    for (int in_ = 0; in_ < n; ++in_)
        for (int kk = 0; kk < k; ++kk)
            for (int i = 0; i < oh; ++i)
                for (int j = 0; j < ow; ++j) {
                    double sum = 0.0;
                    const int ph = i * u - p, pw = j * v - q;
                    for (int cc = 0; cc < c; ++cc)
                        for (int kh = 0; kh < r; ++kh)
                            for (int kw = 0; kw < s; ++kw) {
                                const int y = ph + kh, x = pw + kw;
                                if (x >= 0 && y >= 0 && x < w && y < h) {
                                    const double a = oracle_half_to_double(in[(size_t)in_ * c * h * w + (size_t)cc * h * w + (size_t)y * w + x]);
                                    const double b = oracle_half_to_double(wt[(size_t)kk * r * s * c + (size_t)cc * r * s + kh * s + kw]);
                                    sum += oracle_round_half(a * b); /* product rounded to FP16 */
                                }
                            }
                    out[(size_t)in_ * k * oh * ow + (size_t)kk * oh * ow + (size_t)i * ow + j] = oracle_double_to_half(sum);
                }
}
