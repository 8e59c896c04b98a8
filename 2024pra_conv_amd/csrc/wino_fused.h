// Fused form: input transform, GEMM and output transform in one kernel (wino_fused.hip).
#pragma once

// Threads per workgroup of the fused kernels: 8 waves, two per SIMD.
constexpr int kFusedThreads = 512;
int fused_threads();

// Kernel for 2-byte (d16: odd width or a base that is not 4-byte aligned) or 4-byte staging
// loads, and its multi-layer (grouped) form with arguments (const BatchLayer* layers,
// int nlayers, int total items).
const void* fused_kernel_ptr(bool d16);
const void* fused_grouped_kernel_ptr(bool d16);
