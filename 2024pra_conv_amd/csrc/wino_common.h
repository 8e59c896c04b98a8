// Shared device-side definitions for the gfx950 Winograd F(2x2,3x3) kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mykernelParamType.h"

typedef _Float16 half_t;
typedef half_t half8 __attribute__((ext_vector_type(8)));
typedef half_t half2_t __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// 16 B vector of raw bits, used for global <-> LDS copies.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to nearest even and packed (a in the low half): one gfx950
// v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32 instead of two conversions and a pack.
__device__ inline uint32_t pack_f16x2(float a, float b) {
    const float2v v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, half2_t));
}
__device__ inline uint32_t pack_bf16x2(float a, float b) {
    const float2v v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ inline uint32_t pack_out2(bool bf16, float a, float b) {
    return bf16 ? pack_bf16x2(a, b) : pack_f16x2(a, b);
}
// FP64 -> FP16 rounded once to nearest even: FP32 rounded to odd (an inexact result gets an
// odd last bit, moving away from the FP64 value's side it was rounded past), then the FP32 ->
// FP16 conversion; FP32 keeps 13 more bits than FP16, so the two steps equal one correct
// rounding. A zero comes out as +0 whatever its sign (d + 0.0), so exact transforms that sum
// to zero do not depend on the operation order.
__device__ inline half_t f64_to_f16(double d) {
    d = d + 0.0;
    float r = (float)d;
    const double back = (double)r;
    if (back != d) {
        uint32_t b = __builtin_bit_cast(uint32_t, r);
        if ((b & 1u) == 0u) {
            const bool away = (d > back) == !(b >> 31);  // |d| > |r|
            b = away ? b + 1u : b - 1u;
            r = __builtin_bit_cast(float, b);
        }
    }
    return (half_t)r;
}
// A compile-time flag passed to generic lambdas (one instantiation per output type).
template <bool B>
struct BoolTag {
    static constexpr bool value = B;
};

// Elements of one packed MFMA fragment (32 rows x 16 channels of FP16 = 64 lanes x 8 halves).
constexpr int kFragHalves = 512;

// Offset (in halves) of fragment (pos, cb, blk) in a packed U or V buffer whose
// block dimension has nblk entries: layout [pos][cb][blk][lane][8].
__host__ __device__ inline size_t frag_offset(int pos, int cb, int blk, int CB, int nblk) {
    return ((size_t)(pos * CB + cb) * (size_t)nblk + (size_t)blk) * kFragHalves;
}

// ---- kernels (host stubs are taken by getkernelInfo to fill the launch slots) ----
__global__ void wino_filter_transform(mykernelParamType prm);
__global__ void wino_input_transform(mykernelParamType prm);
__global__ void wino_noop(mykernelParamType prm);
// F(4x4,3x3) form: filter, input and output transforms (S3: wino_pos_*kernel)
__global__ void wino43_filter_transform(mykernelParamType prm);
__global__ void wino43_input_transform(mykernelParamType prm);
__global__ void wino43_output_transform(mykernelParamType prm);
__global__ void conv_direct_f64_kernel(mykernelParamType prm, double* out);
// general-geometry path: packing kernels of the implicit-GEMM direct convolution
__global__ void direct_pack_weights(mykernelParamType prm);
__global__ void direct_pack_input(mykernelParamType prm);
// The same packing for C % 8 != 0 with one wave per 32-pixel block walking every 16-value block.
__global__ void direct_pack_input_walk(mykernelParamType prm);

// Per-item bodies shared by the per-layer kernels and the multi-layer persistent kernels.
__device__ void input_transform_block(const mykernelParamType& prm, int n, int cb, int bi, half_t* lds);

// ---- multi-layer (batched) form ---------------------------------------------------------
// One layer of a batch, in device memory. Work items of a stage are numbered globally: layer
// i owns [first, first + items). S2 item -> (band, channel block, image); S3 item ->
// (k-block, column block) of the layer's batched GEMM configuration.
struct BatchLayer {
    mykernelParamType prm;
    int s2_first, s2_items, nbands;
    int s3_first, s3_items, nk;  // GEMM items are numbered within the queue of the layer's tile
                                 // configuration; nk: k-blocks (column blocks = items / nk)
    int cfg;                     // GEMM tile configuration
    int pad_;
};

__global__ void wino_batch_reset(unsigned* counters, int n);
__global__ void wino_batch_input_transform(const BatchLayer* __restrict__ layers, int nlayers,
                                           int total, unsigned* counter);

// Work-queue helper: thread 0 takes the next item index, every thread gets it (as a
// wave-uniform value, so everything derived from it stays in scalar registers).
__device__ inline int batch_next_item(unsigned* counter, int* slot) {
    if (threadIdx.x == 0) *slot = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(*slot);
    __syncthreads();
    return item;
}

// Layer owning global item `item` (items are contiguous per layer, in layer order).
__device__ inline int batch_find_layer(const BatchLayer* layers, int nlayers, int item, bool s3) {
    int L = 0;
    for (int i = 1; i < nlayers; ++i)
        if ((s3 ? layers[i].s3_first : layers[i].s2_first) <= item) L = i;
    return L;
}

// Input-transform LDS staging: a workgroup stages `band` tile rows of one image and one
// 32-channel block. LDS row pitch in halves: column index = input column + q + 8, so the
// 4x4 window of tile column tw starts at the even column 2*tw + 8 (4-byte aligned reads)
// and 16-byte global chunks (input columns 8j..8j+7) map to LDS columns 8j+q+8.
__host__ __device__ inline int input_lds_pitch(int tiles_w, int W, int q) {
    int a = 2 * tiles_w + 10;
    int b = W + q + 8;
    int m = a > b ? a : b;
    return (m + 1) & ~1;
}
__host__ __device__ inline int input_lds_rows(int band) { return 2 * band + 2; }
// F(4x4,3x3) input staging: tile tw's 6-column window starts at LDS column 4*tw + 8, a band of
// `band` tile rows needs 4*band + 2 input rows.
__host__ __device__ inline int input43_lds_pitch(int tiles_w, int W, int q) {
    int a = 4 * tiles_w + 10;
    int b = W + q + 8;
    int m = a > b ? a : b;
    return (m + 1) & ~1;
}
__host__ __device__ inline int input43_lds_rows(int band) { return 4 * band + 2; }
__host__ __device__ inline size_t input43_lds_bytes(int band, int tiles_w, int W, int q) {
    return (size_t)32 * input43_lds_rows(band) * input43_lds_pitch(tiles_w, W, q) * sizeof(uint16_t);
}
__host__ __device__ inline size_t input_lds_bytes(int band, int tiles_w, int W, int q) {
    return (size_t)32 * input_lds_rows(band) * input_lds_pitch(tiles_w, W, q) * sizeof(uint16_t);
}
