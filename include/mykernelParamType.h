/*
 * mykernelParamType.h — by-value kernel argument of every kernel this plugin puts in a
 * launch slot.
 *
 * The first 19 fields keep the reference's struct (src/mykernelParamType.cuh:5-26): six
 * FP16 pointers then thirteen `short` dimensions. The derived Winograd geometry is
 * appended after them; getParamsize reports the full size, so callers that malloc the blob
 * from that size (reference main.cpp:84-86) keep working.
 */
#ifndef MYKERNELPARAMTYPE_MI355X_H_
#define MYKERNELPARAMTYPE_MI355X_H_

#include "conv2d.h"

typedef struct mykernelParamType {
    conv2d_half_t* __restrict__ pin;              /* input, NCHW */
    conv2d_half_t* __restrict__ pweight;          /* weight, KCRS */
    conv2d_half_t* __restrict__ pout;             /* output, N K Oh Ow */
    conv2d_half_t* __restrict__ transform_pin;    /* V, packed (see below) */
    conv2d_half_t* __restrict__ transform_filter; /* U, packed (see below) */
    conv2d_half_t* __restrict__ mid_result;       /* unused */
    short n, c, h, w, k, r, s, u, v, p, q, Oh, Ow;

    /* ---- appended: derived geometry, filled by getkernelInfo ---- */
    int tiles_h; /* ceil(Oh / 2) */
    int tiles_w; /* ceil(Ow / 2) */
    int T;       /* tiles_h * tiles_w, 2x2 output tiles per image */
    int NT;      /* n * T, GEMM columns (tile index g = image * T + tile) */
    int G32;     /* ceil(NT / 32), 32-column blocks of V */
    int CB;      /* ceil(c / 32); channels are padded to 32 * CB (2 * CB 16-channel blocks) */
    int KB32;    /* ceil(k / 32), 32-row blocks of U */
    int band;    /* tile rows per input-transform workgroup */
    int cfg;     /* GEMM tile configuration index */
    int flags;   /* input-row load width: bit 0 16 B (w % 8 == 0, 16-byte aligned base),
                    bit 1 8 B (w % 4 == 0), bit 2 4 B (w % 2 == 0); none: 2 B */
    int nsteps;  /* GEMM steps of 32 reduction values: 16 * CB (Winograd positions x
                    channel chunks) or, on the general-geometry direct path,
                    ceil(r * s * c / 32) rounded up to a multiple of 8 (the padding carries
                    zeros) */
    /* fused form (input transform inside the GEMM), Winograd geometry:
       fbh x fbw = 64 tiles per workgroup, fnbh x fnbw such blocks per image, input patch
       of fpr rows x fpc columns starting fsh columns left of 2*tw0 - q, fnu staging units
       per thread of the 512-thread workgroup: 1 with 8-byte loads (8 channels x 4 pixels;
       w % 4 == 0, 8-byte aligned input, fsh = (4 - q % 4) % 4), else 2 (8 channels x 2
       pixels, 4-byte loads, 2-byte aligned for odd widths, fsh = q & 1) */
    int fbh, fbw, fnbh, fnbw, fpr, fpc, fsh, fnu;
    /* output epilogue, applied in FP32 before the one FP16 rounding of every output:
       out = act(y + bias[k]); bias: K FP16 values or NULL, act: 0 none, 1 ReLU */
    int act;
    int form;  /* Winograd path: 0 staged (slots: S1, S2, S3, empty), 1 fused (S1, fused, empty x2),
                  2 F(4x4,3x3) (S1, S2, the per-position GEMM, output transform) */
    const void* bias;
    /* tensor layouts and output type (CONV2D_IO_* bits of conv2d.h): bit 0 input N,H,W,C (else
       N,C,H,W), bit 1 output N,Oh,Ow,K (else N,K,Oh,Ow), bit 2 bfloat16 output (else FP16) */
    int io;
    /* F(4x4,3x3) form: Winograd positions each GEMM workgroup runs in sequence (divides 36) */
    int f43_npos;
} mykernelParamType;

/*
 * Packed layouts (element = one FP16 value). Both are the operand layouts of
 * v_mfma_f32_32x32x16_f16 (32 rows x 16 channels per 1 KiB fragment), so a GEMM workgroup
 * copies fragments verbatim into LDS and each lane then reads its 16 contiguous bytes.
 *
 *   U (transform_filter): U[pos][c16][kb][lane][e], c16 = c / 16, kb = k / 32,
 *                         lane = (c % 16 / 8) * 32 + k % 32, e = c % 8;
 *                         pos = 4*a + b is the Winograd position (row a, col b).
 *                         Rows k >= K and channels c >= C are zero.
 *   V (transform_pin):    V[pos][c16][gb][lane][e], gb = g / 32,
 *                         lane = (c % 16 / 8) * 32 + g % 32, e = c % 8,
 *                         g = image * T + tile_row * tiles_w + tile_col.
 *                         Channels c >= C and columns g >= NT are zero.
 * Sizes (2 * CB 16-channel blocks): U = 16 * 2CB * KB32 * 512 halves,
 * V = 16 * 2CB * G32 * 512 halves.
 *
 * General-geometry (direct) path, any r, s, u, v, p, q: the same fragment layouts with
 * (pos, c) replaced by the flattened reduction index j = (r_i * s + s_i) * c_total + c
 * (filter tap outer, channel inner; 16-value block j / 16, lane (j % 16 / 8) * 32 + row,
 * element j % 8), zero for j >= r * s * c_total up to nsteps * 32:
 *   U (transform_filter): w[k][c][r_i][s_i], rows k, same lanes as above;
 *   V (transform_pin):    the im2col input x[n][c][oh*u - p + r_i][ow*v - q + s_i] (zero
 *                         outside the image), column g = n * Oh * Ow + oh * Ow + ow;
 * with T = Oh * Ow, NT = n * T, G32 = ceil(NT / 32). Sizes: U = nsteps * 2 * KB32 * 512
 * halves, V = nsteps * 2 * G32 * 512 halves. For c a multiple of 32 this is step
 * st = tap * CB + cb, as in the Winograd layout with the position replaced by the tap.
 */

#endif /* MYKERNELPARAMTYPE_MI355X_H_ */
