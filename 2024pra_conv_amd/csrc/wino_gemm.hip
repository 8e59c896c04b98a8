// S3: the 16-position batched GEMM on CDNA4 FP16 MFMA with the output transform fused
// into the epilogue.
//
// Reference counterpart: matrixMul1/3/4/5/6 (matrixmul.cuh:139-556), which run the DCU
// v_mmac_f32_16x16x16_f16 and store every per-position product M to HBM in FP16, then the
// separate store* kernels (conv2d.cpp:17-250) read M back and apply
//   A^T = [[1,1,1,0],[0,1,-1,-1]]     (conv2d.cpp:18-22)
// Here, for each Winograd position (a,b) a workgroup reduces over all channels into FP32
// MFMA accumulators, then folds them at once into the four output accumulators of the
// 2x2 tile: Y[i][j] += A^T[i][a] * A^T[j][b] * M(a,b). M never leaves registers, and the
// FP16 output is rounded once from FP32.
//
// Workgroup tile: BK output channels x BT tiles (GEMM columns g = image*T + tile). Waves
// form a WK x WT grid; a wave owns (BK/WK) x (BT/WT) as 16x16 sub-tiles of
// v_mfma_f32_16x16x32_f16. One pipeline step = one 32-channel chunk of one position: BK/16
// fragments of U and BT/16 of V, contiguous 1 KiB blocks in the packed layouts, copied into
// one of two LDS buffers while the other feeds the MFMAs.
//
// Global loads run D steps ahead in a register ring. The copies are branch-free: a piece
// whose fragment lies beyond the matrix reads an in-bounds dummy fragment instead (its
// bytes only reach accumulator rows/columns that are never stored), and the last steps
// re-read the final step. Without branches the compiler's counter-based waits
// (s_waitcnt vmcnt(N)) only wait for the oldest step, so D steps of HBM latency overlap
// with MFMA work.
#include "wino_gemm.h"
#include "wino_tile_ops.h"

namespace {

template <int BK, int BT, int WK, int WT>
struct GemmShape {
    static constexpr int kThreads = WK * WT * 64;
    static constexpr int MI = BK / WK / 16;  // 16-row sub-tiles per wave
    static constexpr int NI = BT / WT / 16;  // 16-column sub-tiles per wave
    static constexpr int UF = BK / 16;       // U fragments per step
    static constexpr int VF = BT / 16;       // V fragments per step
    static constexpr int kPieces = (UF + VF) * 64;  // 16 B pieces per step
    static constexpr int kLoads = (kPieces + kThreads - 1) / kThreads;
    // stage size, padded so every thread copies kLoads pieces (the padding is never read)
    static constexpr int kStageHalves = kLoads * kThreads * 8;
    static_assert(MI >= 1 && NI >= 1, "wave tile must be at least 16x16");
    static_assert(BK % (WK * 16) == 0 && BT % (WT * 16) == 0, "tile must divide evenly");
};

}  // namespace

// One GEMM work item (k-block kblk x column block gblk). smem holds 2 stages. Ends after
// the epilogue stores; callers reusing smem for another item must __syncthreads() first.
template <int BK, int BT, int WK, int WT, int D>
__device__ inline void gemm_body(const mykernelParamType& prm, int kblk, int gblk, half_t* smem) {
    using S = GemmShape<BK, BT, WK, WT>;
    constexpr int MI = S::MI, NI = S::NI, UF = S::UF, VF = S::VF, NL = S::kLoads;
    static_assert(16 % D == 0, "ring depth must divide the step count (a multiple of 16)");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int KB = prm.KB, G16 = prm.G16, CB = prm.CB;
    const int gb0 = gblk * VF;  // first 16-column block of V
    const int kb0 = kblk * UF;  // first 16-row block of U
    const int nsteps = 16 * CB;
    const half_t* __restrict__ U = (const half_t*)prm.transform_filter;
    const half_t* __restrict__ V = (const half_t*)prm.transform_pin;

    // copy slot i of this thread: piece pc = tid + i * threads of every step
    const half_t* src[NL];
    size_t stride[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int pc = tid + i * S::kThreads;
        if (pc < UF * 64) {
            const bool ok = kb0 + pc / 64 < KB;
            src[i] = U + (ok ? (size_t)kb0 * kFragHalves + (size_t)pc * 8 : (size_t)(pc % 64) * 8);
            stride[i] = (size_t)KB * kFragHalves;
        } else {
            const int pv = pc < S::kPieces ? pc - UF * 64 : 0;  // padding slots copy a dummy piece
            const bool ok = pc < S::kPieces && gb0 + pv / 64 < G16;
            src[i] = V + (ok ? (size_t)gb0 * kFragHalves + (size_t)pv * 8 : (size_t)(pv % 64) * 8);
            stride[i] = (size_t)G16 * kFragHalves;
        }
    }

    half8 ring[D][NL];
    auto gload = [&](half8 (&slot)[NL], int step) {
        const size_t s = (size_t)min(step, nsteps - 1);
#pragma unroll
        for (int i = 0; i < NL; ++i) slot[i] = *(const half8*)(src[i] + s * stride[i]);
    };
    auto lstore = [&](const half8 (&slot)[NL], int buf) {
#pragma unroll
        for (int i = 0; i < NL; ++i) *(half8*)(smem + buf * S::kStageHalves + (tid + i * S::kThreads) * 8) = slot[i];
    };

    floatx4 acc[MI][NI];
    floatx4 Y[4][MI][NI];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) Y[q][mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};

    // prologue: step 0 into LDS buffer 0, steps 1..D into the ring (slot = step % D)
    gload(ring[0], 0);
    lstore(ring[0], 0);
#pragma unroll
    for (int d = 1; d <= D; ++d) gload(ring[d % D], d);
    __syncthreads();

    int pos = 0, cb = 0;
    for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = s0 + d;
            const half_t* sb = smem + (s & 1) * S::kStageHalves;
            half8 a[MI], b[NI];
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) a[mi] = *(const half8*)(sb + ((wk * MI + mi) * 64 + lane) * 8);
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
                b[ni] = *(const half8*)(sb + UF * kFragHalves + ((wt * NI + ni) * 64 + lane) * 8);
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
            // step s+1 (ring slot (d+1)%D) -> the other LDS buffer, then refill that slot
            lstore(ring[(d + 1) % D], (s + 1) & 1);
            gload(ring[(d + 1) % D], s + 1 + D);
            __syncthreads();
            if (++cb == CB) {
                fold_position<MI, NI>(pos, acc, Y);
#pragma unroll
                for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};
                cb = 0;
                ++pos;
            }
        }
    }

    store_tiles<MI, NI>(prm, Y, gb0 + wt * NI, kb0 + wk * MI, lane);
}

template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) wino_gemm_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB + S::UF - 1) / S::UF;
    const int ng = (prm.G16 + S::VF - 1) / S::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_body<BK, BT, WK, WT, D>(prm, kblk, gblk, smem);
}

// ---------------------------------------------------------------------------------------
// Multi-layer persistent GEMM: workgroups pull (layer, k-block, column block) items from
// one queue; each layer runs the tile configuration chosen for it (256-thread shapes, so
// one kernel covers all; the LDS buffer is sized for the largest).
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) wino_batch_gemm(const BatchLayer* __restrict__ layers, int nlayers,
                                                       int total, unsigned* counter) {
    constexpr int kSmem = 2 * GemmShape<64, 128, 1, 4>::kStageHalves;
    static_assert(2 * GemmShape<128, 64, 2, 2>::kStageHalves <= kSmem, "LDS sizing");
    static_assert(2 * GemmShape<32, 128, 1, 4>::kStageHalves <= kSmem, "LDS sizing");
    __shared__ __attribute__((aligned(16))) half_t smem[kSmem];
    __shared__ int slot;
    for (;;) {
        const int item = batch_next_item(counter, &slot);
        if (item >= total) break;
        const int L = batch_find_layer(layers, nlayers, item, true);
        const BatchLayer& bl = layers[L];
        const int local = item - bl.s3_first;
        const int kblk = local % bl.nk;
        const int gblk = local / bl.nk;
        switch (bl.bcfg) {
            case 1: gemm_body<128, 64, 2, 2, 2>(bl.prm, kblk, gblk, smem); break;
            case 2: gemm_body<32, 128, 1, 4, 2>(bl.prm, kblk, gblk, smem); break;
            default: gemm_body<64, 128, 1, 4, 2>(bl.prm, kblk, gblk, smem); break;
        }
        __syncthreads();
    }
}

// ---- configuration table --------------------------------------------------------------

#define WINO_CFG(BK, BT, WK, WT, D) \
    { BK, BT, WK * WT * 64, D, (const void*)&wino_gemm_kernel<BK, BT, WK, WT, D> }

const GemmConfig kGemmConfigs[] = {
    WINO_CFG(128, 128, 2, 4, 2),  // 0: wave 64x32, 8 waves
    WINO_CFG(64, 128, 1, 4, 2),   // 1: wave 64x32, 4 waves (batched set)
    WINO_CFG(128, 64, 2, 2, 2),   // 2: wave 64x32, 4 waves (batched set)
    WINO_CFG(64, 64, 1, 2, 2),    // 3: wave 64x32, 2 waves
    WINO_CFG(32, 128, 1, 4, 2),   // 4: wave 32x32, 4 waves (batched set)
    WINO_CFG(256, 64, 4, 2, 2),   // 5: wave 64x32, 8 waves
    WINO_CFG(16, 256, 1, 8, 2),   // 6: wave 16x32, 8 waves (tiny K)
    WINO_CFG(64, 256, 1, 8, 2),   // 7: wave 64x32, 8 waves
    WINO_CFG(128, 128, 2, 4, 4),  // 8
    WINO_CFG(64, 128, 1, 4, 4),   // 9
    WINO_CFG(128, 64, 2, 2, 4),   // 10
    WINO_CFG(256, 64, 4, 2, 4),   // 11
    WINO_CFG(128, 128, 2, 2, 2),  // 12: wave 64x64, 4 waves (one wave per SIMD)
    WINO_CFG(128, 128, 2, 2, 4),  // 13
    WINO_CFG(32, 128, 1, 4, 4),   // 14
    WINO_CFG(64, 64, 1, 2, 4),    // 15
};
const int kNumGemmConfigs = (int)(sizeof(kGemmConfigs) / sizeof(kGemmConfigs[0]));
