// FP8 (OCP e4m3fn) compensated form of the transform and GEMM stages: layouts, kernels and
// tile configurations shared by the device and host translation units.
#pragma once

#include "wino_common.h"
#include "wino_gemm.h"

// One operand record = 32 rows x 64 channels of one Winograd position, for the
// v_mfma_scale_f32_32x32x64_f8f6f4 operand (lane l holds row l % 32; its bytes 0-15 are
// channels 16*(l/32) + 0..15 of the first 32-channel block, bytes 16-31 the same channels
// of the second; the E8M0 scale of (row r, block kb) comes from lane r + 32*kb). Its codes
// are four whole 1 KiB units, so a GEMM step moves them into LDS with one wave-wide 16-byte
// LDS-DMA load each:
//   [0, 1024)     hi codes of block 0: lane * 16 + c % 16, lane = (c % 32 / 16) * 32 + r % 32
//   [1024, 2048)  hi codes of block 1
//   [2048, 4096)  lo codes, same arrangement
// and its 128 scale bytes sit in a separate region after all records' codes:
//   2 * lane -> E_hi + 127, 2 * lane + 1 -> E_lo + 127, lane = kb * 32 + r % 32
// Records are ordered [pos][64-channel block][32-row block] in both regions.
constexpr int kFp8CodeBytes = 4096;
constexpr int kFp8ScaleBytes = 128;

// The lo part's exponent is the hi exponent minus this (a hi residual is at most
// 16 * 2^E_hi, so scaled lo values stay within 256 < 448).
constexpr int kFp8LoOffset = 4;

__host__ __device__ inline int fp8_cq(int c) { return (c + 63) / 64; }

// Records of an operand with rows_blocks32 32-row blocks.
__host__ __device__ inline size_t fp8_records(int c, long long rows_blocks32) {
    return (size_t)16 * fp8_cq(c) * (size_t)rows_blocks32;
}
__host__ __device__ inline size_t fp8_operand_bytes(int c, long long rows_blocks32) {
    return fp8_records(c, rows_blocks32) * (kFp8CodeBytes + kFp8ScaleBytes);
}
// Byte offset of the scale region (after the codes of every record).
__host__ __device__ inline size_t fp8_scale_offset(int c, long long rows_blocks32) {
    return fp8_records(c, rows_blocks32) * kFp8CodeBytes;
}

// The FP8 filter records follow the FP16 transformed filter (which S1 writes first) in the
// transform_filter buffer; this is their byte offset (a multiple of 256).
__host__ __device__ inline size_t fp8_u8_offset(int CB, int KB32) {
    return (size_t)16 * 2 * CB * (size_t)KB32 * kFragHalves * 2;
}

__global__ void wino_filter_fp8(mykernelParamType prm);
__global__ void wino_input_transform_fp8(mykernelParamType prm);

extern const GemmConfig kFp8Configs[];
extern const int kNumFp8Configs;
