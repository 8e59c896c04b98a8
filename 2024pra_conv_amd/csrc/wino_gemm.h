// Tile configurations of the GEMM stage (S3).
#pragma once

struct GemmConfig {
    int bk;              // output channels per workgroup
    int bt;              // tiles (GEMM columns) per workgroup
    int threads;         // threads per workgroup
    int depth;           // global-load steps in flight (register ring); 0 = 2-D grid without XCD remap
    const void* kernel;  // host stub of wino_gemm_kernel<...>
    const void* grouped_kernel;  // host stub of wino_gemm_grouped_kernel<...> (multi-layer form)
};

extern const GemmConfig kGemmConfigs[];
extern const int kNumGemmConfigs;

// Per-position GEMM of the F(4x4,3x3) form (36 positions, M to mid_result): same fields, no
// grouped kernels.
extern const GemmConfig kPosConfigs[];
extern const int kNumPosConfigs;

// Implicit-GEMM direct convolution (general-geometry path): same fields, no grouped kernels.
extern const GemmConfig kDirectConfigs[];
extern const int kNumDirectConfigs;
// 1x1 filters on the general-geometry path: the GEMM reads the input itself; configuration i
// has the tile shape of kDirectConfigs[i].
extern const GemmConfig kPointConfigs[];
// The same for stride 1, no padding, N,C,H,W input with H*W % 4 == 0 and an 8-byte aligned base:
// 8-byte loads along the input rows, transposed in registers.
extern const GemmConfig kPointQuadConfigs[];
