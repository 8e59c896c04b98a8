// Tile configurations of the GEMM stage (S3).
#pragma once

struct GemmConfig {
    int bk;              // output channels per workgroup
    int bt;              // tiles (GEMM columns) per workgroup
    int threads;         // threads per workgroup
    const void* kernel;  // host stub of wino_gemm_kernel<...>
};

extern const GemmConfig kGemmConfigs[];
extern const int kNumGemmConfigs;
