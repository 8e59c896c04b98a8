// Tile configurations of the GEMM stage (S3).
#pragma once

struct GemmConfig {
    int bk;              // output channels per workgroup
    int bt;              // tiles (GEMM columns) per workgroup
    int threads;         // threads per workgroup
    int depth;           // 0: rolled kernel, 2-D grid; D > 0: ring kernel with D steps in flight, 1-D XCD-remapped grid
    const void* kernel;  // host stub of wino_gemm_kernel<...>
};

extern const GemmConfig kGemmConfigs[];
extern const int kNumGemmConfigs;
