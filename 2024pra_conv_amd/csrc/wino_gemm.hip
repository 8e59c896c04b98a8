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
// form a WK x WT grid; a wave owns (BK/WK) x (BT/WT) as 32x32 sub-tiles of
// v_mfma_f32_32x32x16_f16 (measured on MI355X: ~2.1 PFLOP/s sustained in a register loop,
// against ~1.2 for 16x16x32 with the same accumulator count). One pipeline step = one
// 32-channel chunk of one position = two 16-channel MFMA k-steps: 2 x BK/32 fragments of U
// and 2 x BT/32 of V, contiguous 1 KiB blocks in the packed layouts, copied into one of two
// LDS buffers while the other feeds the MFMAs.
//
// Global loads run D steps ahead in a register ring. The copies are branch-free: a piece
// whose fragment lies beyond the matrix reads an in-bounds dummy fragment instead (its
// bytes only reach accumulator rows/columns that are never stored), and the last steps
// re-read the final step. Without branches the compiler's counter-based waits
// (s_waitcnt vmcnt(N)) only wait for the oldest step, so D steps of HBM latency overlap
// with MFMA work.
//
// A second body (gemm_dma_body, below) stages the fragments with LDS DMA into a deeper LDS
// ring instead of the register ring; both bodies compute every output element in the same
// order, so all configurations give bit-identical results.
#include "wino_gemm.h"
#include "wino_tile_ops.h"

namespace {

// KS 16-channel substeps per pipeline step (one v_mfma_f32_32x32x16_f16 k-step each).
constexpr int KS = 2;

typedef half_t half4v __attribute__((ext_vector_type(4)));

// What a GEMM body computes:
//   kGemmWino   the F(2x2) GEMM: 16 positions, fold per position, 2x2-tile epilogue;
//   kGemmDirect the implicit-GEMM direct convolution of the general-geometry path;
//   kGemmPos    positions fpos .. fpos + f43_npos - 1 of the F(4x4) form (36 positions): their
//               packed U and V in sequence, each M stored in FP16 to mid_result (store_m16);
//   kGemmPoint  the direct convolution of a 1x1 filter: as kGemmDirect, but the B operand is
//               read from the input itself (register-staged body only), with no packing pass.
enum { kGemmWino = 0, kGemmDirect = 1, kGemmPos = 2, kGemmPoint = 3 };

template <int BK, int BT, int WK, int WT>
struct GemmShape {
    static constexpr int kThreads = WK * WT * 64;
    static constexpr int MI = BK / WK / 32;  // 32-row sub-tiles per wave
    static constexpr int NI = BT / WT / 32;  // 32-column sub-tiles per wave
    static constexpr int UF = BK / 32;       // U fragments per substep
    static constexpr int VF = BT / 32;       // V fragments per substep
    static constexpr int kPieces = KS * (UF + VF) * 64;  // 16 B pieces per step
    static constexpr int kLoads = (kPieces + kThreads - 1) / kThreads;
    // stage size, padded so every thread copies kLoads pieces (the padding is never read)
    static constexpr int kStageHalves = kLoads * kThreads * 8;
    static_assert(MI >= 1 && NI >= 1, "wave tile must be at least 32x32");
    static_assert(BK % (WK * 32) == 0 && BT % (WT * 32) == 0, "tile must divide evenly");
};

// kGemmPoint: the 8 input channels c .. c+7 at input element xo of channel 0 (-1: the pixel
// is outside the image), zero for channels >= C. NCHW input: 8 loads one channel plane apart
// (across a wave's lanes they read consecutive pixels); channel-contiguous input: one 16-byte
// load when C % 8 == 0 and it is aligned. Loads are not predicated per lane (a lane without a
// value reads element 0 of the input and discards it), so the compiler's memory-counter waits
// stay per load (with per-lane branches it waits for every load in flight).
__device__ inline half8 point_piece(const mykernelParamType& prm, long long xo, int c, bool nhwc) {
    const half_t* __restrict__ X = (const half_t*)prm.pin;
    const int C = prm.c;
    half8 v;
    if (nhwc && (C & 7) == 0 && (((uintptr_t)X) & 15) == 0) {
        const bool ok = xo >= 0 && c < C;
        const half8 raw = *(const half8*)(X + (ok ? xo + c : 0));
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = ok ? raw[e] : (half_t)0.0f;
        return v;
    }
    const long long es = nhwc ? 1 : (long long)prm.h * prm.w;  // elements between channels
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const bool ok = xo >= 0 && c + e < C;
        const half_t raw = X[ok ? xo + (long long)(c + e) * es : 0];
        v[e] = ok ? raw : (half_t)0.0f;
    }
    return v;
}

}  // namespace

// One GEMM work item (k-block kblk x column block gblk). smem holds 2 stages. Ends after
// the epilogue stores; callers reusing smem for another item must __syncthreads() first.
// Stage layout: substep kk = 0..KS-1, each [UF U fragments][VF V fragments].
// MODE kGemmWino: the Winograd GEMM (16 positions, fold per position, 2x2-tile epilogue).
// MODE kGemmDirect: the implicit-GEMM direct convolution of the general-geometry path: the
// steps run over prm.nsteps chunks of 32 values of the flattened (filter tap, channel)
// reduction index into one accumulator, whose elements are the outputs themselves.
// MODE kGemmPos: positions fpos .. fpos + prm.f43_npos - 1 of the F(4x4) form in one pipeline
// (npos * CB steps), each position's M stored by store_m16 at its last step.
// MODE kGemmPoint: a 1x1 filter (any stride and padding), so the reduction index is the input
// channel: each V piece of a step is gathered from the input (8 channels of one pixel) and
// stored to LDS in the fragment layout the packing pass would have written, so the MFMAs see
// the same operands in the same order as in kGemmDirect (the results are bit-identical).
template <int BK, int BT, int WK, int WT, int D, int MODE = kGemmWino>
__device__ inline void gemm_body(const mykernelParamType& prm, int kblk, int gblk, half_t* smem, int fpos = 0) {
    using S = GemmShape<BK, BT, WK, WT>;
    constexpr int MI = S::MI, NI = S::NI, UF = S::UF, VF = S::VF, NL = S::kLoads;
    static_assert(8 % D == 0 && D % 2 == 0, "ring depth: even, dividing the step count (a multiple of 8)");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int KB = prm.KB32, G32 = prm.G32, CB = prm.CB;
    const int gb0 = gblk * VF;  // first 32-column block of V
    const int kb0 = kblk * UF;  // first 32-row block of U
    // 32-value reduction chunks (a multiple of 8 except for F(4x4) position groups, whose steps
    // past npos * CB issue no MFMAs)
    constexpr bool kDirectSteps = MODE == kGemmDirect || MODE == kGemmPoint;
    const int nsteps = kDirectSteps ? prm.nsteps : (MODE == kGemmPos ? prm.f43_npos * CB : 16 * CB);
    const half_t* __restrict__ U = (const half_t*)prm.transform_filter;
    const half_t* __restrict__ V = (const half_t*)prm.transform_pin;
    if constexpr (MODE == kGemmPos) {
        U += (size_t)fpos * 2 * CB * KB * kFragHalves;
        V += (size_t)fpos * 2 * CB * G32 * kFragHalves;
    }

    // copy slot i of this thread: piece pc = tid + i * threads of every step. Step s covers
    // 16-channel blocks 2s and 2s+1 of the [pos][c16] sequence (CB 32-channel chunks per
    // position), so fragment addresses advance by 2 blocks per step.
    const half_t* src[NL];
    size_t stride[NL];
    // kGemmPoint: whether piece i is a V piece (the same for the 64 lanes of a wave), the first
    // of its 8 channels within a step (xc) and the input element of channel 0 at its pixel
    // (xo; -1 outside the image, past the last column or for a padding slot)
    [[maybe_unused]] bool isv[NL];
    [[maybe_unused]] int xc[NL];
    [[maybe_unused]] long long xo[NL];
    [[maybe_unused]] const bool nhwc = (prm.io & CONV2D_IO_IN_NHWC) != 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int pc = tid + i * S::kThreads;
        const int p = pc < S::kPieces ? pc : 0;  // padding slots copy a dummy piece
        const int kk = p / ((UF + VF) * 64);
        const int q = p % ((UF + VF) * 64);
        if (q < UF * 64) {
            const bool ok = pc < S::kPieces && kb0 + q / 64 < KB;
            src[i] = U + (ok ? ((size_t)kk * KB + kb0) * kFragHalves + (size_t)q * 8 : (size_t)(q % 64) * 8);
            stride[i] = (size_t)2 * KB * kFragHalves;
            if constexpr (MODE == kGemmPoint) isv[i] = false, xc[i] = 0, xo[i] = -1;
        } else {
            const int qv = q - UF * 64;
            const bool ok = pc < S::kPieces && gb0 + qv / 64 < G32;
            src[i] = V + (ok ? ((size_t)kk * G32 + gb0) * kFragHalves + (size_t)qv * 8 : (size_t)(qv % 64) * 8);
            stride[i] = (size_t)2 * G32 * kFragHalves;
            if constexpr (MODE == kGemmPoint) {
                // fragment lane slot qv % 64: pixel (column) g, channels 8 * (slot / 32) of the
                // 16-channel block kk of the step
                const int slot = qv % 64;
                const int g = (gb0 + qv / 64) * 32 + (slot & 31);
                isv[i] = true;
                xc[i] = kk * 16 + (slot >> 5) * 8;
                xo[i] = -1;
                if (pc < S::kPieces && g < prm.NT) {
                    const int img = g / prm.T, pix = g - (g / prm.T) * prm.T;
                    const int oh = pix / prm.Ow, ow = pix - (pix / prm.Ow) * prm.Ow;
                    const int ih = oh * prm.u - prm.p, iw = ow * prm.v - prm.q;
                    const long long H = prm.h, W = prm.w, C = prm.c;
                    if (ih >= 0 && ih < H && iw >= 0 && iw < W)
                        xo[i] = nhwc ? ((long long)img * H * W + ih * W + iw) * C : (long long)img * C * H * W + ih * W + iw;
                }
            }
        }
    }

    half8 ring[D][NL];
    auto gload = [&](half8 (&slot)[NL], int step) {
        const size_t s = (size_t)min(step, nsteps - 1);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if constexpr (MODE == kGemmPoint) {
                if (isv[i]) {
                    slot[i] = point_piece(prm, xo[i], (int)s * 32 + xc[i], nhwc);
                    continue;
                }
            }
            slot[i] = *(const half8*)(src[i] + s * stride[i]);
        }
    };
    auto lstore = [&](const half8 (&slot)[NL], int buf) {
#pragma unroll
        for (int i = 0; i < NL; ++i) *(half8*)(smem + buf * S::kStageHalves + (tid + i * S::kThreads) * 8) = slot[i];
    };

    floatx16 acc[MI][NI];
    floatx16 Y[4][MI][NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) Y[q][mi][ni] = acc[mi][ni];
        }

    // MFMA operands of one step, double-buffered in registers
    half8 ra[2][KS][MI], rb[2][KS][NI];
    auto lread = [&](half8 (&A)[KS][MI], half8 (&B)[KS][NI], int buf) {
        const half_t* sb = smem + buf * S::kStageHalves;
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const half_t* sk = sb + kk * (UF + VF) * kFragHalves;
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) A[kk][mi] = *(const half8*)(sk + ((wk * MI + mi) * 64 + lane) * 8);
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
                B[kk][ni] = *(const half8*)(sk + UF * kFragHalves + ((wt * NI + ni) * 64 + lane) * 8);
        }
    };

    // Pipeline (slot = step % D in the register ring; LDS buffer = step & 1):
    //   prologue: LDS0 <- step 0, regs <- LDS0, LDS1 <- step 1, ring <- steps 2..D+1
    //   iteration s: MFMAs on regs(s) (asynchronous) | regs(s+1) <- LDS[(s+1)&1]
    //                | LDS[s&1] <- step s+2 (that buffer was read into regs(s) before the
    //                previous barrier) | ring refill with step s+2+D | barrier
    gload(ring[0], 0);
    lstore(ring[0], 0);
    gload(ring[0], D);
#pragma unroll
    for (int d = 1; d < D; ++d) gload(ring[d], d);
    __syncthreads();
    lread(ra[0], rb[0], 0);
    lstore(ring[1 % D], 1);
    gload(ring[1 % D], 1 + D);
    __syncthreads();

    [[maybe_unused]] int pos = 0, cb = 0;
    for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = s0 + d;
            const int cur = d & 1;  // D is even, so s & 1 == d & 1
            if (MODE != kGemmPos || s < nsteps) {
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
#pragma unroll
                    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                        for (int ni = 0; ni < NI; ++ni)
                            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ra[cur][kk][mi], rb[cur][kk][ni],
                                                                                 acc[mi][ni], 0, 0, 0);
            }
            lread(ra[cur ^ 1], rb[cur ^ 1], cur ^ 1);
            lstore(ring[(d + 2) % D], cur);
            gload(ring[(d + 2) % D], s + 2 + D);
            __syncthreads();
            if constexpr (MODE == kGemmWino) {
                if (++cb == CB) {
                    fold_position(pos, acc, Y);
#pragma unroll
                    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
                    cb = 0;
                    ++pos;
                }
            } else if constexpr (MODE == kGemmPos) {
                // at the end of each position of the group: M to mid_result, then the next one
                if (s < nsteps && ++cb == CB) {
                    store_m16<MI, NI>(prm, acc, gb0 + wt * NI, kb0 + wk * MI, lane, fpos + pos);
#pragma unroll
                    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
                    cb = 0;
                    ++pos;
                }
            }
        }
    }

    if constexpr (kDirectSteps)
        store_plain32<MI, NI>(prm, acc, gb0 + wt * NI, kb0 + wk * MI, lane);
    else if constexpr (MODE == kGemmWino)
        store_tiles32<MI, NI>(prm, Y, gb0 + wt * NI, kb0 + wk * MI, lane);
    // kGemmPos: every position of the group was stored at its last step
}

// 1x1 filter, stride 1, no padding, N,C,H,W input with H*W % 4 == 0 and an 8-byte aligned
// base: the register-staged pipeline of gemm_body with the input read in 8-byte pieces along
// its rows. Unit (16-channel block kk, channel quad cq, fragment f, pixel quad pq) of a step =
// 4 channels x 4 consecutive pixels (GEMM columns, all in one image since H*W % 4 == 0): four
// 8-byte loads, one per channel, across a wave's lanes 32 consecutive pixel quads of one
// channel row; transposed in registers, it becomes half of the fragment lane slots of its 4
// pixels (8-byte LDS stores). The MFMAs see the operands of the packed direct path in the same
// order, so the results are the same bits.
template <int BK, int BT, int WK, int WT, int D>
__device__ inline void point_quad_body(const mykernelParamType& prm, int kblk, int gblk, half_t* smem) {
    using S = GemmShape<BK, BT, WK, WT>;
    constexpr int MI = S::MI, NI = S::NI, UF = S::UF, VF = S::VF;
    constexpr int kStage = KS * (UF + VF) * kFragHalves;  // halves per LDS stage (<= S::kStageHalves)
    constexpr int kUPieces = KS * UF * 64;                 // 16-byte U pieces per step
    constexpr int NU = (kUPieces + S::kThreads - 1) / S::kThreads;
    constexpr int kVUnits = KS * 4 * VF * 8;               // 4 x 4 input units per step
    static_assert(kVUnits % 64 == 0 && kVUnits <= S::kThreads, "at most one unit per thread, whole waves");
    static_assert(8 % D == 0 && D % 2 == 0, "ring depth: even, dividing the step count (a multiple of 8)");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int KB = prm.KB32, C = prm.c;
    const int gb0 = gblk * VF, kb0 = kblk * UF;
    const int nsteps = prm.nsteps;
    const half_t* __restrict__ U = (const half_t*)prm.transform_filter;
    const half_t* __restrict__ X = (const half_t*)prm.pin;

    // U: 16-byte fragment pieces as in gemm_body (slots past kUPieces, whole waves, store nothing)
    const half_t* usrc[NU];
    int udst[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int pc = tid + i * S::kThreads;
        const int p = pc < kUPieces ? pc : 0;
        const int kk = p / (UF * 64), q = p % (UF * 64);
        const bool ok = pc < kUPieces && kb0 + q / 64 < KB;
        usrc[i] = U + (ok ? ((size_t)kk * KB + kb0) * kFragHalves + (size_t)q * 8 : (size_t)(q % 64) * 8);
        udst[i] = pc < kUPieces ? kk * (UF + VF) * kFragHalves + q * 8 : -1;
    }
    const size_t ustride = (size_t)2 * KB * kFragHalves;

    // V: this thread's unit (pixel quad fastest, then fragment, channel quad, 16-channel block)
    const bool hasv = tid < kVUnits;
    const int pq = tid & 7, f = (tid >> 3) % VF, cq = (tid >> 3) / VF % 4, kkv = (tid >> 3) / VF / 4;
    const int g0 = (gb0 + f) * 32 + pq * 4;
    const bool gok = hasv && g0 < prm.NT;
    const size_t hw = (size_t)prm.T;  // H*W (stride 1, no padding)
    const int img = g0 / prm.T;
    const size_t xbase = (size_t)img * C * hw + (size_t)(g0 - img * prm.T);
    const int c0 = kkv * 16 + cq * 4;
    const int vdst = kkv * (UF + VF) * kFragHalves + (UF + f) * kFragHalves + ((cq >> 1) * 32 + pq * 4) * 8 + (cq & 1) * 4;

    half8 uring[D][NU];
    half4v vring[D][4];
    auto gload = [&](half8 (&us)[NU], half4v (&vs)[4], int step) {
        const int s = min(step, nsteps - 1);
#pragma unroll
        for (int i = 0; i < NU; ++i) us[i] = *(const half8*)(usrc[i] + (size_t)s * ustride);
        if (hasv) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                // not predicated per lane: a lane without a value reads element 0 and drops it
                const int c = s * 32 + c0 + e;
                const bool ok = gok && c < C;
                const half4v raw = *(const half4v*)(X + (ok ? xbase + (size_t)c * hw : 0));
                vs[e] = ok ? raw : half4v{};
            }
        }
    };
    auto lstore = [&](const half8 (&us)[NU], const half4v (&vs)[4], int buf) {
        half_t* b = smem + buf * kStage;
#pragma unroll
        for (int i = 0; i < NU; ++i)
            if (udst[i] >= 0) *(half8*)(b + udst[i]) = us[i];
        if (hasv) {
#pragma unroll
            for (int px = 0; px < 4; ++px) {
                const half4v o = {vs[0][px], vs[1][px], vs[2][px], vs[3][px]};
                *(half4v*)(b + vdst + px * 8) = o;
            }
        }
    };

    floatx16 acc[MI][NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;

    half8 ra[2][KS][MI], rb[2][KS][NI];
    auto lread = [&](half8 (&A)[KS][MI], half8 (&B)[KS][NI], int buf) {
        const half_t* sb = smem + buf * kStage;
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const half_t* sk = sb + kk * (UF + VF) * kFragHalves;
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) A[kk][mi] = *(const half8*)(sk + ((wk * MI + mi) * 64 + lane) * 8);
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
                B[kk][ni] = *(const half8*)(sk + UF * kFragHalves + ((wt * NI + ni) * 64 + lane) * 8);
        }
    };

    // the pipeline of gemm_body: ring slot = step % D, LDS buffer = step & 1
    gload(uring[0], vring[0], 0);
    lstore(uring[0], vring[0], 0);
    gload(uring[0], vring[0], D);
#pragma unroll
    for (int d = 1; d < D; ++d) gload(uring[d], vring[d], d);
    __syncthreads();
    lread(ra[0], rb[0], 0);
    lstore(uring[1 % D], vring[1 % D], 1);
    gload(uring[1 % D], vring[1 % D], 1 + D);
    __syncthreads();

    for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = s0 + d;
            const int cur = d & 1;
#pragma unroll
            for (int kk = 0; kk < KS; ++kk)
#pragma unroll
                for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                    for (int ni = 0; ni < NI; ++ni)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ra[cur][kk][mi], rb[cur][kk][ni],
                                                                             acc[mi][ni], 0, 0, 0);
            lread(ra[cur ^ 1], rb[cur ^ 1], cur ^ 1);
            lstore(uring[(d + 2) % D], vring[(d + 2) % D], cur);
            gload(uring[(d + 2) % D], vring[(d + 2) % D], s + 2 + D);
            __syncthreads();
        }
    }
    store_plain32<MI, NI>(prm, acc, gb0 + wt * NI, kb0 + wk * MI, lane);
}

// ---------------------------------------------------------------------------------------
// LDS-DMA variant. Operand fragments go straight from global memory into an S-stage LDS ring
// with global_load_lds_dwordx4 (each wave moves whole 1 KiB fragments: lane i writes 16 bytes
// at the wave's LDS base in m0 + 16 i), so the prefetch depth costs no registers and no LDS
// store instructions; waves then read their MFMA operands from LDS into a register double
// buffer. Written as inline assembly because the compiler's own waits for LDS DMA drain every
// load in flight (vmcnt(0)) before each LDS read or barrier, which would leave one step of
// prefetch; here the wait for step s is vmcnt((S - 3) * loads per step), counted by hand:
// this wave issues exactly FPW of these loads per step and no other vector-memory loads
// inside the loop.
// ---------------------------------------------------------------------------------------
namespace {

template <int BK, int BT, int WK, int WT, int S>
struct DmaShape {
    static constexpr int kWaves = WK * WT;
    static constexpr int kThreads = kWaves * 64;
    static constexpr int MI = BK / WK / 32, NI = BT / WT / 32;
    static constexpr int UF = BK / 32, VF = BT / 32;
    static constexpr int kFrags = KS * (UF + VF);  // 1 KiB fragments per step
    static constexpr int kFPW = kFrags / kWaves;   // DMA loads per wave per step
    static constexpr int kStageHalves = kFrags * kFragHalves;
    static constexpr int kSmemHalves = S * kStageHalves;
    static_assert(kFrags % kWaves == 0, "fragments per step must split evenly over the waves");
    static_assert(S >= 4 && 8 % S == 0, "stage count: a power of two dividing the step count (a multiple of 8)");
    static_assert((S - 2) * kFPW <= 63, "vmcnt range");
    static_assert(MI >= 1 && NI >= 1 && BK % (WK * 32) == 0 && BT % (WT * 32) == 0, "tile shape");
};

}  // namespace

template <int BK, int BT, int WK, int WT, int S, int MODE = kGemmWino>
__device__ inline void gemm_dma_body(const mykernelParamType& prm, int kblk, int gblk, half_t* smem, int fpos = 0) {
    using Sh = DmaShape<BK, BT, WK, WT, S>;
    constexpr int MI = Sh::MI, NI = Sh::NI, UF = Sh::UF, VF = Sh::VF, FPW = Sh::kFPW;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int KB = prm.KB32, G32 = prm.G32, CB = prm.CB;
    const int gb0 = gblk * VF;
    const int kb0 = kblk * UF;
    const int nsteps = MODE == kGemmDirect ? prm.nsteps : (MODE == kGemmPos ? prm.f43_npos * CB : 16 * CB);
    const half_t* __restrict__ U = (const half_t*)prm.transform_filter;
    const half_t* __restrict__ V = (const half_t*)prm.transform_pin;
    if constexpr (MODE == kGemmPos) {
        U += (size_t)fpos * 2 * CB * KB * kFragHalves;
        V += (size_t)fpos * 2 * CB * G32 * kFragHalves;
    }

    // this wave's DMA slots: fragments wave*FPW .. wave*FPW+FPW-1 of the stage layout
    // [kk][UF U fragments][VF V fragments]; beyond-the-matrix fragments read fragment 0
    const half_t* src[FPW];
    size_t stride[FPW];
#pragma unroll
    for (int j = 0; j < FPW; ++j) {
        const int f = wave * FPW + j;
        const int kk = f / (UF + VF), q = f % (UF + VF);
        if (q < UF) {
            const bool ok = kb0 + q < KB;
            src[j] = U + (ok ? ((size_t)kk * KB + kb0 + q) * kFragHalves : 0) + lane * 8;
            stride[j] = (size_t)2 * KB * kFragHalves;
        } else {
            const bool ok = gb0 + q - UF < G32;
            src[j] = V + (ok ? ((size_t)kk * G32 + gb0 + q - UF) * kFragHalves : 0) + lane * 8;
            stride[j] = (size_t)2 * G32 * kFragHalves;
        }
    }
    const uint32_t lds0 = lds_address(smem);
    auto dma = [&](int step, int stage) {
        const size_t s = (size_t)min(step, nsteps - 1);
#pragma unroll
        for (int j = 0; j < FPW; ++j) {
            const uint32_t dst = __builtin_amdgcn_readfirstlane(
                lds0 + (uint32_t)((stage * Sh::kStageHalves + (wave * FPW + j) * kFragHalves) * 2));
            lds_dma16(src[j] + s * stride[j], dst);
        }
    };

    floatx16 acc[MI][NI];
    floatx16 Y[4][MI][NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) Y[q][mi][ni] = acc[mi][ni];
        }

    half8 ra[2][KS][MI], rb[2][KS][NI];
    auto lread = [&](half8 (&A)[KS][MI], half8 (&B)[KS][NI], int stage) {
        const half_t* sb = smem + stage * Sh::kStageHalves;
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const half_t* sk = sb + kk * (UF + VF) * kFragHalves;
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) A[kk][mi] = *(const half8*)(sk + ((wk * MI + mi) * 64 + lane) * 8);
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
                B[kk][ni] = *(const half8*)(sk + UF * kFragHalves + ((wt * NI + ni) * 64 + lane) * 8);
        }
    };

    // Pipeline (stage = step % S; registers = step & 1):
    //   prologue: DMA steps 0..S-2; wait for step 0; barrier; regs <- stage 0
    //   iteration s: wait for steps <= s+1 (S-3 steps stay in flight) | barrier (stage s+1
    //   complete for every wave; every wave's reads of stage s-1 done) | DMA step s+S-1 into
    //   stage s-1 | regs(s+1) <- stage s+1 | MFMAs on regs(s) | fold at a position's end
#pragma unroll
    for (int p = 0; p < S - 1; ++p) dma(p, p);
    wait_vm<(S - 2) * FPW>();
    __syncthreads();
    lread(ra[0], rb[0], 0);

    [[maybe_unused]] int pos = 0, cb = 0;
    for (int s0 = 0; s0 < nsteps; s0 += S) {
#pragma unroll
        for (int d = 0; d < S; ++d) {
            const int s = s0 + d;
            const int cur = d & 1;  // S is even, so s & 1 == d & 1
            wait_vm<(S - 3) * FPW>();
            __syncthreads();
            dma(s + S - 1, (d + S - 1) % S);
            lread(ra[cur ^ 1], rb[cur ^ 1], (d + 1) % S);
            if (MODE != kGemmPos || s < nsteps) {
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
#pragma unroll
                    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                        for (int ni = 0; ni < NI; ++ni)
                            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ra[cur][kk][mi], rb[cur][kk][ni],
                                                                                 acc[mi][ni], 0, 0, 0);
            }
            if constexpr (MODE == kGemmWino) {
                if (++cb == CB) {
                    fold_position(pos, acc, Y);
#pragma unroll
                    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
                    cb = 0;
                    ++pos;
                }
            } else if constexpr (MODE == kGemmPos) {
                // at the end of each position of the group: M to mid_result, then the next one
                if (s < nsteps && ++cb == CB) {
                    store_m16<MI, NI>(prm, acc, gb0 + wt * NI, kb0 + wk * MI, lane, fpos + pos);
#pragma unroll
                    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                            for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
                    cb = 0;
                    ++pos;
                }
            }
        }
    }
    wait_vm<0>();  // the trailing (re-read) DMAs land before the workgroup's LDS is released

    static_assert(MODE != kGemmPoint, "1x1 gathers need the register-staged body");
    if constexpr (MODE == kGemmDirect)
        store_plain32<MI, NI>(prm, acc, gb0 + wt * NI, kb0 + wk * MI, lane);
    else if constexpr (MODE == kGemmWino)
        store_tiles32<MI, NI>(prm, Y, gb0 + wt * NI, kb0 + wk * MI, lane);
    // kGemmPos: every position of the group was stored at its last step
}

// Waves per SIMD the register allocation must allow: two for 64x32 wave tiles (so two waves
// of a SIMD overlap LDS traffic with MFMAs), one for 64x64 tiles (their output accumulators
// alone take 256 registers).
template <int BK, int BT, int WK, int WT>
constexpr int kWavesPerEU = (BK / WK / 32) * (BT / WT / 32) >= 4 ? 1 : 2;

template <int BK, int BT, int WK, int WT, int S>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_gemm_dma_kernel(mykernelParamType prm) {
    using Sh = DmaShape<BK, BT, WK, WT, S>;
    __shared__ __attribute__((aligned(16))) half_t smem[Sh::kSmemHalves];
    const int nk = (prm.KB32 + Sh::UF - 1) / Sh::UF;
    const int ng = (prm.G32 + Sh::VF - 1) / Sh::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_dma_body<BK, BT, WK, WT, S>(prm, kblk, gblk, smem);
}

template <int BK, int BT, int WK, int WT, int S>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_gemm_dma_grouped_kernel(const BatchLayer* __restrict__ layers, int nlayers) {
    using Sh = DmaShape<BK, BT, WK, WT, S>;
    __shared__ __attribute__((aligned(16))) half_t smem[Sh::kSmemHalves];
    int local;
    const BatchLayer& bl = grouped_layer_item(layers, nlayers, blockIdx.x, &local);
    // the layer's parameters by value, loaded once (read through a reference into global
    // memory, the epilogue's loops grew past the unroller's size limit in the larger tile
    // configurations and their accumulators went to 576 bytes per lane of scratch)
    const mykernelParamType prm = bl.prm;
    const int nk = bl.nk;
    gemm_dma_body<BK, BT, WK, WT, S>(prm, local % nk, local / nk, smem);
}

template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_gemm_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB32 + S::UF - 1) / S::UF;
    const int ng = (prm.G32 + S::VF - 1) / S::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_body<BK, BT, WK, WT, D>(prm, kblk, gblk, smem);
}

// ---------------------------------------------------------------------------------------
// Multi-layer (grouped) GEMM: one launch covers every layer that uses this tile
// configuration. Each layer owns a contiguous range of workgroup ids (layers in decreasing
// cost); a binary search finds the layer, and the id is remapped per XCD within that range (as
// in the single-layer kernel) onto the layer's (k-block, column block) items, so every XCD runs
// an equal share of every layer. One item per workgroup keeps the
// register allocation of the single-layer kernel (a persistent loop around the body spills).
// ---------------------------------------------------------------------------------------
template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_gemm_grouped_kernel(const BatchLayer* __restrict__ layers, int nlayers) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    int local;
    const BatchLayer& bl = grouped_layer_item(layers, nlayers, blockIdx.x, &local);
    // the layer's parameters by value, loaded once (read through a reference into global
    // memory, the epilogue's loops grew past the unroller's size limit in the larger tile
    // configurations and their accumulators went to 576 bytes per lane of scratch)
    const mykernelParamType prm = bl.prm;
    const int nk = bl.nk;
    gemm_body<BK, BT, WK, WT, D>(prm, local % nk, local / nk, smem);
}

// ---- configuration table --------------------------------------------------------------

#define WINO_CFG(BK, BT, WK, WT, D)                                                     \
    {                                                                                   \
        BK, BT, WK * WT * 64, D, (const void*)&wino_gemm_kernel<BK, BT, WK, WT, D>,      \
            (const void*)&wino_gemm_grouped_kernel<BK, BT, WK, WT, D>                      \
    }

// depth field: register-ring depth D for the register-staged kernels, LDS stages S for the
// LDS-DMA kernels
#define WINO_DMA_CFG(BK, BT, WK, WT, S)                                                  \
    {                                                                                    \
        BK, BT, WK * WT * 64, S, (const void*)&wino_gemm_dma_kernel<BK, BT, WK, WT, S>,   \
            (const void*)&wino_gemm_dma_grouped_kernel<BK, BT, WK, WT, S>                  \
    }

const GemmConfig kGemmConfigs[] = {
    WINO_CFG(128, 128, 2, 4, 2),      // 0: wave 64x32, 8 waves
    WINO_CFG(64, 128, 1, 4, 2),       // 1: wave 64x32, 4 waves
    WINO_DMA_CFG(128, 64, 2, 2, 4),   // 2: LDS-DMA, 4-stage ring, wave 64x32, 4 waves
    WINO_CFG(64, 64, 1, 2, 2),        // 3: wave 64x32, 2 waves
    WINO_CFG(32, 128, 1, 4, 2),       // 4: wave 32x32, 4 waves
    WINO_DMA_CFG(128, 128, 2, 4, 4),  // 5: LDS-DMA, 4-stage ring, wave 64x32, 8 waves
    WINO_CFG(32, 256, 1, 8, 2),       // 6: wave 32x32, 8 waves (small K)
    WINO_CFG(64, 256, 1, 8, 2),       // 7: wave 64x32, 8 waves
    WINO_CFG(128, 128, 2, 4, 4),      // 8: wave 64x32, 8 waves, ring depth 4
    WINO_DMA_CFG(128, 128, 2, 4, 8),  // 9: LDS-DMA, 8-stage ring (128 KiB), wave 64x32, 8 waves
    WINO_CFG(32, 64, 1, 2, 2),        // 10: wave 32x32, 2 waves (small problems)
    WINO_DMA_CFG(64, 128, 1, 4, 4),   // 11: LDS-DMA, 4-stage ring, wave 64x32, 4 waves
    WINO_DMA_CFG(64, 64, 1, 2, 4),    // 12: LDS-DMA, 4-stage ring, wave 64x32, 2 waves
    WINO_CFG(64, 128, 1, 4, 4),       // 13: wave 64x32, 4 waves, ring depth 4
    WINO_CFG(128, 64, 2, 2, 4),       // 14: wave 64x32, 4 waves, ring depth 4
    WINO_CFG(32, 128, 1, 4, 4),       // 15: wave 32x32, 4 waves, ring depth 4
};
const int kNumGemmConfigs = (int)(sizeof(kGemmConfigs) / sizeof(kGemmConfigs[0]));

// ---- direct (implicit-GEMM) form of the general-geometry path --------------------------------
// The same two pipeline bodies in kGemmDirect mode: U holds the packed filter taps, V the
// packed im2col input, and the steps run over (tap, channel chunk) pairs.

template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_direct_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB32 + S::UF - 1) / S::UF;
    const int ng = (prm.G32 + S::VF - 1) / S::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_body<BK, BT, WK, WT, D, kGemmDirect>(prm, kblk, gblk, smem);
}

template <int BK, int BT, int WK, int WT, int S>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_direct_dma_kernel(mykernelParamType prm) {
    using Sh = DmaShape<BK, BT, WK, WT, S>;
    __shared__ __attribute__((aligned(16))) half_t smem[Sh::kSmemHalves];
    const int nk = (prm.KB32 + Sh::UF - 1) / Sh::UF;
    const int ng = (prm.G32 + Sh::VF - 1) / Sh::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_dma_body<BK, BT, WK, WT, S, kGemmDirect>(prm, kblk, gblk, smem);
}

#define DIRECT_CFG(BK, BT, WK, WT, D) \
    { BK, BT, WK * WT * 64, D, (const void*)&wino_direct_kernel<BK, BT, WK, WT, D>, nullptr }
#define DIRECT_DMA_CFG(BK, BT, WK, WT, S) \
    { BK, BT, WK * WT * 64, S, (const void*)&wino_direct_dma_kernel<BK, BT, WK, WT, S>, nullptr }

const GemmConfig kDirectConfigs[] = {
    DIRECT_DMA_CFG(128, 128, 2, 4, 8),  // 0: LDS-DMA, 8-stage ring, wave 64x32, 8 waves
    DIRECT_CFG(64, 128, 1, 4, 2),       // 1: wave 64x32, 4 waves
    DIRECT_CFG(32, 64, 1, 2, 2),        // 2: wave 32x32, 2 waves (small problems)
    DIRECT_DMA_CFG(128, 64, 2, 2, 4),   // 3: LDS-DMA, 4-stage ring, wave 64x32, 4 waves
};
const int kNumDirectConfigs = (int)(sizeof(kDirectConfigs) / sizeof(kDirectConfigs[0]));

// 1x1 filters: the register-staged body gathers the input itself (kGemmPoint), so the input
// needs no packing pass. Configuration i has the tile shape of kDirectConfigs[i] (the LDS-DMA
// ones become register-staged: DMA copies whole fragments and cannot gather).
template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_point_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB32 + S::UF - 1) / S::UF;
    const int ng = (prm.G32 + S::VF - 1) / S::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_body<BK, BT, WK, WT, D, kGemmPoint>(prm, kblk, gblk, smem);
}

#define POINT_CFG(BK, BT, WK, WT, D) \
    { BK, BT, WK * WT * 64, D, (const void*)&wino_point_kernel<BK, BT, WK, WT, D>, nullptr }

const GemmConfig kPointConfigs[] = {
    POINT_CFG(128, 128, 2, 4, 2),  // 0: wave 64x32, 8 waves
    POINT_CFG(64, 128, 1, 4, 2),   // 1: wave 64x32, 4 waves
    POINT_CFG(32, 64, 1, 2, 2),    // 2: wave 32x32, 2 waves (small problems)
    POINT_CFG(128, 64, 2, 2, 2),   // 3: wave 64x32, 4 waves
};
static_assert(sizeof(kPointConfigs) == sizeof(kDirectConfigs), "one 1x1 configuration per direct configuration");

// 1x1, stride 1, no padding, N,C,H,W input with H*W % 4 == 0 (8-byte aligned): the input is
// read in 8-byte pieces along its rows and transposed in registers (point_quad_body).
template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<BK, BT, WK, WT>)))
wino_point_quad_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB32 + S::UF - 1) / S::UF;
    const int ng = (prm.G32 + S::VF - 1) / S::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    point_quad_body<BK, BT, WK, WT, D>(prm, kblk, gblk, smem);
}

#define POINT_QUAD_CFG(BK, BT, WK, WT, D) \
    { BK, BT, WK * WT * 64, D, (const void*)&wino_point_quad_kernel<BK, BT, WK, WT, D>, nullptr }

const GemmConfig kPointQuadConfigs[] = {
    POINT_QUAD_CFG(128, 128, 2, 4, 2),  // 0: wave 64x32, 8 waves
    POINT_QUAD_CFG(64, 128, 1, 4, 2),   // 1: wave 64x32, 4 waves
    POINT_QUAD_CFG(32, 64, 1, 2, 2),    // 2: wave 32x32, 2 waves (small problems)
    POINT_QUAD_CFG(128, 64, 2, 2, 2),   // 3: wave 64x32, 4 waves
};
static_assert(sizeof(kPointQuadConfigs) == sizeof(kDirectConfigs), "one configuration per direct configuration");

// ---- per-position GEMM of the F(4x4,3x3) form -----------------------------------------------
// One workgroup per (k-block, column block, group of f43_npos positions): the group's steps in
// one pipeline, each position's M to mid_result. Without the fold's output accumulators a wave
// holds only its MFMA accumulators, so 64 x 64 wave tiles keep two waves per SIMD.

template <int BK, int BT, int WK, int WT, int D>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(2)))
wino_pos_kernel(mykernelParamType prm) {
    using S = GemmShape<BK, BT, WK, WT>;
    __shared__ __attribute__((aligned(16))) half_t smem[2 * S::kStageHalves];
    const int nk = (prm.KB32 + S::UF - 1) / S::UF;
    const int ng = (prm.G32 + S::VF - 1) / S::VF;
    int kblk, rest;
    xcd_remap(blockIdx.x, nk * ng * (36 / prm.f43_npos), nk, &kblk, &rest);
    gemm_body<BK, BT, WK, WT, D, kGemmPos>(prm, kblk, rest % ng, smem, (rest / ng) * prm.f43_npos);
}

template <int BK, int BT, int WK, int WT, int S>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(2)))
wino_pos_dma_kernel(mykernelParamType prm) {
    using Sh = DmaShape<BK, BT, WK, WT, S>;
    __shared__ __attribute__((aligned(16))) half_t smem[Sh::kSmemHalves];
    const int nk = (prm.KB32 + Sh::UF - 1) / Sh::UF;
    const int ng = (prm.G32 + Sh::VF - 1) / Sh::VF;
    int kblk, rest;
    xcd_remap(blockIdx.x, nk * ng * (36 / prm.f43_npos), nk, &kblk, &rest);
    gemm_dma_body<BK, BT, WK, WT, S, kGemmPos>(prm, kblk, rest % ng, smem, (rest / ng) * prm.f43_npos);
}

#define POS_CFG(BK, BT, WK, WT, D) \
    { BK, BT, WK * WT * 64, D, (const void*)&wino_pos_kernel<BK, BT, WK, WT, D>, nullptr }
#define POS_DMA_CFG(BK, BT, WK, WT, S) \
    { BK, BT, WK * WT * 64, S, (const void*)&wino_pos_dma_kernel<BK, BT, WK, WT, S>, nullptr }

const GemmConfig kPosConfigs[] = {
    POS_DMA_CFG(128, 128, 2, 4, 8),  // 0: LDS-DMA, 8-stage ring, wave 64x32, 8 waves
    POS_DMA_CFG(128, 64, 2, 2, 4),   // 1: LDS-DMA, 4-stage ring, wave 64x32, 4 waves
    POS_CFG(64, 128, 1, 4, 2),       // 2: wave 64x32, 4 waves
    POS_CFG(32, 64, 1, 2, 2),        // 3: wave 32x32, 2 waves (small problems)
    POS_DMA_CFG(64, 128, 1, 4, 4),   // 4: LDS-DMA, 4-stage ring, wave 64x32, 4 waves
    POS_DMA_CFG(128, 128, 2, 2, 4),  // 5: LDS-DMA, 4-stage ring, wave 64x64, 4 waves
    POS_DMA_CFG(256, 128, 4, 2, 4),  // 6: LDS-DMA, 4-stage ring, wave 64x64, 8 waves
    POS_DMA_CFG(128, 256, 2, 4, 4),  // 7: LDS-DMA, 4-stage ring, wave 64x64, 8 waves
};
const int kNumPosConfigs = (int)(sizeof(kPosConfigs) / sizeof(kPosConfigs[0]));
