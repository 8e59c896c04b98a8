// Fused form: input transform, GEMM and output transform in one kernel (wino_fused.hip).
#pragma once

// Threads per workgroup of the fused kernels: 8 waves, two per SIMD.
constexpr int kFusedThreads = 512;
int fused_threads();

// The 64-tile block shapes (rows x columns of tiles) a fused item covers, by index.
constexpr int kFusedShapes[4][2] = {{8, 8}, {4, 16}, {16, 4}, {2, 32}};
inline int fused_shape_index(int fbh, int fbw) {
    for (int i = 0; i < 4; ++i)
        if (kFusedShapes[i][0] == fbh && kFusedShapes[i][1] == fbw) return i;
    return -1;
}
// Patch columns of a block fbw tiles wide (room for the one-column shift fsh).
inline int fused_patch_cols(int fbw) { return 2 * fbw + 4; }

// Kernel for 2-byte (d16: odd width or a base that is not 4-byte aligned) or 4-byte staging
// loads, block shape index `shape` and column shift fsh (q & 1), and its multi-layer
// (grouped) form with arguments (const BatchLayer* layers, int nlayers, int total items).
const void* fused_kernel_ptr(bool d16, int shape, int fsh);
const void* fused_grouped_kernel_ptr(bool d16, int shape, int fsh);
