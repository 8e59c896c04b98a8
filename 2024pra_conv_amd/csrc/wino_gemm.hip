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
// form a WK x WT grid; a wave owns (BK/WK) x (BT/WT) and issues
// v_mfma_f32_16x16x32_f16 on 16x16 sub-tiles. Each pipeline step moves one 32-channel
// chunk of U (BK/16 fragments) and V (BT/16 fragments) -- contiguous 1 KiB fragments in
// the packed layouts -- into one of two LDS buffers while the other is consumed.
#include "wino_common.h"
#include "wino_gemm.h"

namespace {

constexpr int kAT[2][4] = {{1, 1, 1, 0}, {0, 1, -1, -1}};

template <int BK, int BT, int WK, int WT>
struct GemmShape {
    static constexpr int kThreads = WK * WT * 64;
    static constexpr int MI = BK / WK / 16;  // 16-row sub-tiles per wave
    static constexpr int NI = BT / WT / 16;  // 16-column sub-tiles per wave
    static constexpr int UF = BK / 16;       // U fragments per step
    static constexpr int VF = BT / 16;       // V fragments per step
    static constexpr int kPieces = (UF + VF) * 64;  // 16 B pieces per step
    static constexpr int kLoads = (kPieces + kThreads - 1) / kThreads;
    static constexpr int kStageHalves = (UF + VF) * kFragHalves;
    static_assert(MI >= 1 && NI >= 1, "wave tile must be at least 16x16");
    static_assert(BK % (WK * 16) == 0 && BT % (WT * 16) == 0, "tile must divide evenly");
};

}  // namespace

// Epilogue shared by the GEMM kernels: the 2x2 output tile of each (k, g) held by this
// lane, FP32 -> FP16, ragged bottom/right edges guarded. gbw / kbw: the wave's first
// 16-column / 16-row block. MFMA 16x16 accumulator layout: element j of lane l is row
// 4*(l/16) + j, column l % 16.
template <int MI, int NI>
__device__ inline void store_tiles(const mykernelParamType& prm, const floatx4 (&Y)[4][MI][NI], int gbw,
                                   int kbw, int lane) {
    const int T = prm.T, tiles_w = prm.tiles_w, K = prm.k, Oh = prm.Oh, Ow = prm.Ow;
    const int NT = prm.NT;
    half_t* __restrict__ out = (half_t*)prm.pout;
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
        const int g = (gbw + ni) * 16 + (lane & 15);
        if (g >= NT) continue;
        const int img = g / T;
        const int t = g - img * T;
        const int th = t / tiles_w;
        const int tw = t - th * tiles_w;
        const int oh = 2 * th, ow = 2 * tw;
        const bool r1 = oh + 1 < Oh, c1 = ow + 1 < Ow;
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = (kbw + mi) * 16 + 4 * (lane >> 4) + j;
                if (k >= K) continue;
                half_t* o = out + (((size_t)img * K + k) * Oh + oh) * Ow + ow;
                o[0] = (half_t)Y[0][mi][ni][j];
                if (c1) o[1] = (half_t)Y[1][mi][ni][j];
                if (r1) {
                    o[Ow] = (half_t)Y[2][mi][ni][j];
                    if (c1) o[Ow + 1] = (half_t)Y[3][mi][ni][j];
                }
            }
    }
}

// Folds the accumulators of position pos into the output accumulators (wave-uniform
// runtime coefficients in {-1, 0, 1}; zero ones skipped).
template <int MI, int NI>
__device__ inline void fold_position(int pos, const floatx4 (&acc)[MI][NI], floatx4 (&Y)[4][MI][NI]) {
    const int pa = pos >> 2, pb = pos & 3;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int coef = kAT[i][pa] * kAT[j][pb];
            if (coef == 0) continue;
            const float cf = (float)coef;
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) Y[i * 2 + j][mi][ni] += cf * acc[mi][ni];
        }
}

// One GEMM work item (k-block kblk x column block gblk), rolled position loop. smem must
// hold 2 * GemmShape::kStageHalves halves. Ends after the epilogue stores; callers reusing
// smem for another item must __syncthreads() first.
template <int BK, int BT, int WK, int WT>
__device__ inline void gemm_body_rolled(const mykernelParamType& prm, int kblk, int gblk, half_t* smem) {
    using S = GemmShape<BK, BT, WK, WT>;
    constexpr int MI = S::MI, NI = S::NI, UF = S::UF, VF = S::VF;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int gb0 = gblk * VF;  // first 16-column block of V
    const int kb0 = kblk * UF;  // first 16-row block of U
    const int KB = prm.KB, G16 = prm.G16, CB = prm.CB;
    const int nsteps = 16 * CB;
    const half_t* __restrict__ U = (const half_t*)prm.transform_filter;
    const half_t* __restrict__ V = (const half_t*)prm.transform_pin;

    // Per-thread copy slots: piece i of a step is U piece i (i < UF*64) or V piece
    // i - UF*64. A piece whose fragment lies beyond the matrix is not loaded; the stale LDS
    // bytes it leaves only feed accumulator rows/columns that are never stored.
    const half_t* src[S::kLoads];
    size_t sstride[S::kLoads];
    bool sval[S::kLoads];
#pragma unroll
    for (int i = 0; i < S::kLoads; ++i) {
        const int pc = tid + i * S::kThreads;
        if (pc < UF * 64) {
            src[i] = U + (size_t)kb0 * kFragHalves + (size_t)pc * 8;
            sstride[i] = (size_t)KB * kFragHalves;
            sval[i] = (kb0 + pc / 64) < KB;
        } else if (pc < S::kPieces) {
            const int pv = pc - UF * 64;
            src[i] = V + (size_t)gb0 * kFragHalves + (size_t)pv * 8;
            sstride[i] = (size_t)G16 * kFragHalves;
            sval[i] = (gb0 + pv / 64) < G16;
        } else {
            src[i] = U;
            sstride[i] = 0;
            sval[i] = false;
        }
    }

    half8 pre[S::kLoads];
    auto gload = [&](int step) {
#pragma unroll
        for (int i = 0; i < S::kLoads; ++i)
            if (sval[i]) pre[i] = *(const half8*)(src[i] + (size_t)step * sstride[i]);
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < S::kLoads; ++i) {
            const int pc = tid + i * S::kThreads;
            if (sval[i]) *(half8*)(smem + buf * S::kStageHalves + pc * 8) = pre[i];
        }
    };

    floatx4 acc[MI][NI];
    floatx4 Y[4][MI][NI];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) Y[q][mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};

    gload(0);
    lstore(0);
    __syncthreads();

    // The position loop stays rolled: unrolling all 16 positions makes the register
    // allocator spill. The fold coefficients are then wave-uniform runtime values in
    // {-1, 0, 1}; zero ones are skipped with a uniform branch.
#pragma unroll 1
    for (int pos = 0; pos < 16; ++pos) {
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};

        for (int cb = 0; cb < CB; ++cb) {
            const int step = pos * CB + cb;
            const bool has_next = step + 1 < nsteps;
            if (has_next) gload(step + 1);
            const half_t* sb = smem + (step & 1) * S::kStageHalves;
            half8 a[MI], b[NI];
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
                a[mi] = *(const half8*)(sb + ((wk * MI + mi) * 64 + lane) * 8);
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
                b[ni] = *(const half8*)(sb + UF * kFragHalves + ((wt * NI + ni) * 64 + lane) * 8);
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
            if (has_next) lstore((step + 1) & 1);
            __syncthreads();
        }

        fold_position<MI, NI>(pos, acc, Y);
    }

    store_tiles<MI, NI>(prm, Y, gb0 + wt * NI, kb0 + wk * MI, lane);
}

template <int BK, int BT, int WK, int WT>
__global__ void __launch_bounds__(WK * WT * 64) wino_gemm_kernel(mykernelParamType prm) {
    __shared__ __attribute__((aligned(16))) half_t smem[2 * GemmShape<BK, BT, WK, WT>::kStageHalves];
    gemm_body_rolled<BK, BT, WK, WT>(prm, blockIdx.y, blockIdx.x, smem);
}

// ---------------------------------------------------------------------------------------
// Ring variant: the same arithmetic (bit-identical output) with
//   * D pipeline steps of global loads in flight in a register ring (one step = one
//     32-channel chunk of all the workgroup's U and V fragments), so HBM latency is
//     covered by D steps of MFMA work instead of one;
//   * a flat step loop over (position, channel chunk), unrolled by D so ring slots are
//     compile-time registers; the fold runs when a position's last chunk is done;
//   * an XCD-aware workgroup order: the hardware deals workgroup ids round-robin over the
//     8 XCDs, so ids are remapped to give every XCD a contiguous run of the work, with the
//     k-blocks that share one V column block adjacent (they hit the same L2).
// ---------------------------------------------------------------------------------------
__device__ inline void xcd_remap(int L, int total, int nk, int* kblk, int* gblk) {
    const int xcd = L & 7, j = L >> 3;
    const int base = total >> 3, rem = total & 7;
    const int m = xcd * base + min(xcd, rem) + j;
    *kblk = m % nk;
    *gblk = m / nk;
}

template <int BK, int BT, int WK, int WT, int D>
__device__ inline void gemm_body_ring(const mykernelParamType& prm, int kblk, int gblk, half_t* smem) {
    using S = GemmShape<BK, BT, WK, WT>;
    constexpr int MI = S::MI, NI = S::NI, UF = S::UF, VF = S::VF;
    static_assert(16 % D == 0, "ring depth must divide the 16 positions' step count");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int KB = prm.KB, G16 = prm.G16, CB = prm.CB;
    const int gb0 = gblk * VF;
    const int kb0 = kblk * UF;
    const int nsteps = 16 * CB;
    const half_t* __restrict__ U = (const half_t*)prm.transform_filter;
    const half_t* __restrict__ V = (const half_t*)prm.transform_pin;

    const half_t* src[S::kLoads];
    size_t sstride[S::kLoads];
    bool sval[S::kLoads];
#pragma unroll
    for (int i = 0; i < S::kLoads; ++i) {
        const int pc = tid + i * S::kThreads;
        if (pc < UF * 64) {
            src[i] = U + (size_t)kb0 * kFragHalves + (size_t)pc * 8;
            sstride[i] = (size_t)KB * kFragHalves;
            sval[i] = (kb0 + pc / 64) < KB;
        } else if (pc < S::kPieces) {
            const int pv = pc - UF * 64;
            src[i] = V + (size_t)gb0 * kFragHalves + (size_t)pv * 8;
            sstride[i] = (size_t)G16 * kFragHalves;
            sval[i] = (gb0 + pv / 64) < G16;
        } else {
            src[i] = U;
            sstride[i] = 0;
            sval[i] = false;
        }
    }

    half8 ring[D][S::kLoads];
    auto gload = [&](half8 (&slot)[S::kLoads], int step) {
#pragma unroll
        for (int i = 0; i < S::kLoads; ++i)
            if (sval[i]) slot[i] = *(const half8*)(src[i] + (size_t)step * sstride[i]);
    };
    auto lstore = [&](const half8 (&slot)[S::kLoads], int buf) {
#pragma unroll
        for (int i = 0; i < S::kLoads; ++i) {
            const int pc = tid + i * S::kThreads;
            if (sval[i]) *(half8*)(smem + buf * S::kStageHalves + pc * 8) = slot[i];
        }
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

    // prologue: step 0 into LDS, steps 1..D in the ring (slot = step % D)
    gload(ring[0], 0);
    lstore(ring[0], 0);
#pragma unroll
    for (int d = 1; d <= D; ++d)
        if (d < nsteps) gload(ring[d % D], d);
    __syncthreads();

    int pos = 0, cb = 0;
    for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = s0 + d;
            const half_t* sb = smem + (s & 1) * S::kStageHalves;
            half8 a[MI], b[NI];
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
                a[mi] = *(const half8*)(sb + ((wk * MI + mi) * 64 + lane) * 8);
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
                b[ni] = *(const half8*)(sb + UF * kFragHalves + ((wt * NI + ni) * 64 + lane) * 8);
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
            // step s+1 (ring slot (d+1)%D) -> the other LDS buffer; refill the slot
            if (s + 1 < nsteps) lstore(ring[(d + 1) % D], (s + 1) & 1);
            if (s + 1 + D < nsteps) gload(ring[(d + 1) % D], s + 1 + D);
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
__global__ void __launch_bounds__(WK * WT * 64) wino_gemm_ring_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB + S::UF - 1) / S::UF;
    const int ng = (prm.G16 + S::VF - 1) / S::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_body_ring<BK, BT, WK, WT, D>(prm, kblk, gblk, smem);
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
            case 1: gemm_body_rolled<128, 64, 2, 2>(bl.prm, kblk, gblk, smem); break;
            case 2: gemm_body_rolled<32, 128, 1, 4>(bl.prm, kblk, gblk, smem); break;
            default: gemm_body_rolled<64, 128, 1, 4>(bl.prm, kblk, gblk, smem); break;
        }
        __syncthreads();
    }
}

// ---- configuration table --------------------------------------------------------------

#define WINO_CFG(BK, BT, WK, WT) \
    { BK, BT, WK * WT * 64, 0, (const void*)&wino_gemm_kernel<BK, BT, WK, WT> }
#define WINO_RING(BK, BT, WK, WT, D) \
    { BK, BT, WK * WT * 64, D, (const void*)&wino_gemm_ring_kernel<BK, BT, WK, WT, D> }

const GemmConfig kGemmConfigs[] = {
    WINO_CFG(128, 128, 2, 4),  // 0: wave 64x32, 8 waves
    WINO_CFG(64, 128, 1, 4),   // 1: wave 64x32, 4 waves
    WINO_CFG(128, 64, 2, 2),   // 2: wave 64x32, 4 waves
    WINO_CFG(64, 64, 1, 2),    // 3: wave 64x32, 2 waves
    WINO_CFG(32, 128, 1, 4),   // 4: wave 32x32, 4 waves
    WINO_CFG(256, 64, 4, 2),   // 5: wave 64x32, 8 waves
    WINO_CFG(16, 256, 1, 8),   // 6: wave 16x32, 8 waves (tiny K)
    WINO_CFG(64, 256, 1, 8),   // 7: wave 64x32, 8 waves
    WINO_RING(128, 128, 2, 4, 2),  // 8
    WINO_RING(128, 128, 2, 4, 4),  // 9
    WINO_RING(64, 128, 1, 4, 4),   // 10
    WINO_RING(128, 64, 2, 2, 4),   // 11
    WINO_RING(256, 64, 4, 2, 4),   // 12
    WINO_RING(64, 64, 1, 2, 4),    // 13
    WINO_RING(32, 128, 1, 4, 4),   // 14
    WINO_RING(64, 256, 1, 8, 2),   // 15
};
const int kNumGemmConfigs = (int)(sizeof(kGemmConfigs) / sizeof(kGemmConfigs[0]));
