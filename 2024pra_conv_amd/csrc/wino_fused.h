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
// Staging load width of the fused kernels, per channel of a staging unit: kFusedLoad2, one
// 4-byte pixel pair at a 2-byte aligned offset (odd width, or an input base that is not 4-byte
// aligned; the pair at the right edge of an odd width loads the dword ending with its one
// pixel); kFusedLoad4, one 4-byte aligned pixel pair; kFusedLoad8, one 8-byte pixel quad (width
// a multiple of 4, base 8-byte aligned, patch within one unit per thread).
// kFusedLoadNHWC: channel-contiguous (N,H,W,C) input, one 16-byte load of 8 channels per pixel
// (C % 8 == 0, 16-byte aligned base).
enum FusedLoad { kFusedLoad2 = 0, kFusedLoad4 = 1, kFusedLoad8 = 2, kFusedLoadNHWC = 3 };
constexpr int kFusedLoadWidths = 4;
// Pixels per staging unit (8 channels each).
constexpr int fused_unit_pixels(int lw) { return lw == kFusedLoad8 ? 4 : (lw == kFusedLoadNHWC ? 1 : 2); }
// Column shift: the patch starts fsh columns left of 2*tw0 - q (tw0 the block's first tile
// column, a multiple of 4), so that its first column is even (q & 1) or, with 8-byte loads,
// a multiple of 4 ((4 - q % 4) % 4); single-pixel NHWC units need no shift.
constexpr int fused_col_shift(int lw, int q) {
    return lw == kFusedLoad8 ? (4 - q % 4) % 4 : (lw == kFusedLoadNHWC ? 0 : (q & 1));
}
// Patch columns of a block fbw tiles wide: the windows reach column 2*fbw + fsh + 1; a whole
// number of staging units per row (and an even count, for the column-parity split).
constexpr int fused_patch_cols(int lw, int fbw, int fsh) {
    return lw == kFusedLoad8 ? (2 * fbw + fsh + 2 + 3) & ~3 : (lw == kFusedLoadNHWC ? 2 * fbw + 2 : 2 * fbw + 4);
}
// NHWC staging units per thread (4 channel groups x patch rows x patch columns: at most 1584
// for the 2 x 32 block).
constexpr int kFusedNhwcUnits = 4;
// Patch rows of a block fbh tiles high.
constexpr int fused_patch_rows(int fbh) { return 2 * fbh + 2; }

// Kernel for staging load width lw, block shape index `shape` and column shift fsh (tail: for a
// last 64-channel block of 1..32 channels, run as wide items, fused_wide_tail, or with the
// half-block body), and its multi-layer
// (grouped) form with arguments (const BatchLayer* layers, int nlayers), one workgroup per
// item, which always has the tail body; null for a combination without an instance (8-byte
// loads with a patch that needs more than one unit per thread).
const void* fused_kernel_ptr(int lw, int shape, int fsh, bool tail);
const void* fused_grouped_kernel_ptr(int lw, int shape, int fsh);
// Whether a layer's last 64-channel block has its second 32 channels beyond K (K % 64 is 1..32,
// i.e. an odd number of 32-row blocks).
inline bool fused_tail(int kb32) { return (kb32 & 1) != 0; }
// Whether the single-layer tail instance for staging load width lw and block shape index
// `shape` can run the last 64-channel block as wide items (32 channels x a block of fbh x 2*fbw
// tiles) instead of half-block items: 2- and 4-byte pair staging (three units per thread for a
// 128-tile patch) and 8-byte quad staging (two), except for the 2 x 32 block (a 2 x 64 patch
// needs four pair units).
constexpr bool fused_wide_tail(int lw, int shape) {
    return (lw == kFusedLoad2 || lw == kFusedLoad4 || lw == kFusedLoad8) && shape != 3;
}
// Wide tail items per image: fnbh block rows of ceil(fnbw / 2) wide blocks.
inline long long fused_wide_items(int fnbh, int fnbw) { return (long long)fnbh * ((fnbw + 1) / 2); }
