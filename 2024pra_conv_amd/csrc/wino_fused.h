// Fused form: input transform, GEMM and output transform in one kernel (wino_fused.hip).
#pragma once

// Kernel for fnu staging units per thread (3 or 4).
const void* fused_kernel_ptr(int nu);
