// Fused Winograd F(2x2,3x3) forward: input transform, GEMM and output transform in one
// kernel, so the transformed input V never goes to HBM.
//
// The staged form writes V (16 FP16 values per tile and channel, 4x the input) with the
// input transform and reads it back in the GEMM (reference: BTdB* preprocessing.cuh:273-424,
// matrixMul* matrixmul.cuh:139-556). Here V is computed in registers, directly in the MFMA
// B-operand layout, from an LDS-staged input patch, and U (prepared by the filter
// transform) is read from global memory (L2) straight into A-operand registers.
//
// Workgroup: 8 waves (two per SIMD, up to 256 registers each), 64 output channels x one
// block of FBH x FBW = 64 tiles of one image. Wave w = 4*BH + a owns the Winograd positions
// pos = 4a + 2*BH + bb (bb = 0, 1) of row a of the 4x4 transform domain for the whole
// 64 x 64 tile: 8 FP32 32x32 accumulators (2 positions x 2 channel blocks x 2 tile blocks =
// 128 registers). Two waves per SIMD let one wave's loads, LDS reads and transform overlap
// the other's MFMAs.
//
// Per 32-channel chunk:
//   patch:  input rows 2*th0 - p .. +2*FBH+1 and columns from iw0 = 2*tw0 - q - fsh, all 32
//           channels, zero outside the image, in LDS as [channel group g (8 ch)][row]
//           [column parity][column / 2][8] (one ds_read_b128 = 8 channels of one pixel; the
//           parity split puts the same window column of consecutive tiles 16 bytes apart, so
//           a wave's reads are free of bank conflicts); double-buffered, the next chunk's
//           global loads in flight during this chunk's MFMAs.
//   V:      lane (tile t = tb*32 + lane%32, channels 8*(lane/32) of 16-channel half ks) reads
//           the 2 x 3 window pixels its positions need and computes, in FP16 arithmetic,
//             s_j = B^T row a applied to window column j   (2 rows: d0-d2, d1+d2, d2-d1, d1-d3)
//             V_b = (s B)_b: s0-s2, s1+s2 (BH = 0, columns 0-2); s2-s1, s1-s3 (BH = 1, 1-3)
//           (the reference's input transform is FP16 arithmetic too, preprocessing.cuh:57-72).
//   MFMA:   acc[bb][m][tb] += U[4a+2BH+bb][k block m] . V_{2BH+bb} over the chunk's two
//           16-channel halves.
//
// Epilogue (A^T M A, reference conv2d.cpp:18-47): row a folds into Z_j = sum_b A^T[j][b] M[a][b]
// in the order Z_0 = (M0 + M1) + M2, Z_1 = (M1 - M2) - M3: the BH = 0 wave leaves P = M0 + M1
// and M1 in LDS, the BH = 1 wave adds its M2, M3 and writes Z; then each wave takes half an
// output block (channel block m, tile block tb) and forms Y_0j = (Z0_j + Z1_j) + Z2_j,
// Y_1j = (Z1_j - Z2_j) - Z3_j, rounded once to FP16.
#include "wino_tile_ops.h"
#include "wino_fused.h"

#include <utility>

namespace {

// Raw buffer resource (stride 0, byte offsets); dword 3 = the gfx9 default data format.
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

// Staging of one thread: up to NU units, a unit = (channel group g, patch row, pixel pair) =
// 8 channels x 2 adjacent pixels, one 4-byte buffer load per channel. A pair starts on an even
// patch column. With W even and a 4-byte aligned base (D16 = false) it is entirely inside or
// outside the image and its offset is 4-byte aligned. With W odd or a base that is not 4-byte
// aligned (D16) the offset is only 2-byte aligned (the kernel driver runs compute queues in
// the unaligned access mode, so such a load returns the two pixels), and with W odd the pair
// at column W - 1 has its second pixel past the right edge: that unit loads the dword ending
// with its first pixel (offset - 2, never before the image for W >= 3) and its byte selectors
// take the high half and a zero. Loads are never predicated per lane (a whole wave skips a unit
// set none of its lanes holds): a pixel outside the image, or of a thread slot without a
// unit, gets a byte offset of 2^31 or more, and a channel >= C lands at or beyond the image's
// last byte; both are past the resource's range (the image, C*H*W*2 < 2^31 bytes), where
// buffer loads return zero without accessing memory.
template <int NU>
struct Staging {
    uint32_t off[NU];             // byte offset of the unit's load in channel g*8 of chunk 0
    uint32_t sel0[NU], sel1[NU];  // D16: v_perm selectors of the pair's first and second pixel
    int lds[NU];                  // LDS offset (halves) of the unit's first pixel
    uint32_t w[NU][8];            // loaded pairs (first pixel in the low half)
};

// FP16 x - y as fma(y, m1, x) with m1 = -1: y * -1 is exact, so the single rounding of the
// fma equals the rounding of the subtraction, zero signs included. An 8-wide FP16
// subtraction, or an fma whose -1 the compiler can see, is split into scalar instructions;
// the caller makes m1 opaque (an empty asm statement) so the fma stays packed.
__device__ inline half8 sub8(half8 x, half8 y, half8 m1) { return __builtin_elementwise_fma(y, m1, x); }

__device__ inline uint32_t pack_lo(uint32_t x, uint32_t y) { return __builtin_amdgcn_perm(y, x, 0x05040100u); }
__device__ inline uint32_t pack_hi(uint32_t x, uint32_t y) { return __builtin_amdgcn_perm(y, x, 0x07060302u); }

template <int NU, bool D16, int NT, int PR, int PC>
__device__ inline void stage_setup(const mykernelParamType& prm, int ih0, int iw0, Staging<NU>& st) {
    constexpr int PH = PC >> 1;
    constexpr int nunits = 4 * PR * PH;
    const int H = prm.h, W = prm.w;
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int u = threadIdx.x + NT * i;
        const int g = u / (PR * PH);
        const int rem = u - g * (PR * PH);
        const int pr = rem / PH;
        const int m = rem - pr * PH;
        const int ih = ih0 + pr, iw = iw0 + 2 * m;
        const bool row = u < nunits && ih >= 0 && ih < H;
        const bool pix = row && iw >= 0 && iw < W;
        st.off[i] = pix ? (uint32_t)(((g * 8 * H + ih) * W + iw) * 2) : 0x80000000u;
        if (D16) {
            // iw is even, so only the pair at column W - 1 (W odd) straddles the image edge
            const bool straddle = pix && iw + 1 >= W;
            if (straddle) st.off[i] -= 2u;
            st.sel0[i] = straddle ? 0x07060302u : 0x05040100u;
            st.sel1[i] = straddle ? 0x0C0C0C0Cu : 0x07060302u;  // selector 0x0C: a zero byte
        }
        // pair m: pixel 2m at column parity 0, pixel 2m+1 at parity 1, both at column m; a
        // thread slot without a unit stores to a dummy pixel pair past both patch buffers
        st.lds[i] = u < nunits ? ((g * PR + pr) * PC + m) * 8 : 2 * 4 * PR * PC * 8;
    }
}

// Loads units first .. first + N - 1 of the thread's staging.
template <int N, int NU>
__device__ inline void stage_load(__amdgpu_buffer_rsrc_t in, uint32_t chunk_bytes, uint32_t chan_bytes,
                                  Staging<NU>& st, int first) {
#pragma unroll
    for (int i = first; i < first + N; ++i) {
        const uint32_t base = st.off[i] + chunk_bytes;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            st.w[i][e] = __builtin_amdgcn_raw_buffer_load_b32(in, (int)(base + e * chan_bytes), 0, 0);
    }
}

template <int NU, bool D16>
__device__ inline void stage_store(const Staging<NU>& st, half_t* buf, int half_pc) {
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        u32x4 p0, p1;  // pixel 2m and 2m+1, 8 channels each
        const uint32_t s0 = D16 ? st.sel0[i] : 0x05040100u, s1 = D16 ? st.sel1[i] : 0x07060302u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            p0[k] = __builtin_amdgcn_perm(st.w[i][2 * k + 1], st.w[i][2 * k], s0);
            p1[k] = __builtin_amdgcn_perm(st.w[i][2 * k + 1], st.w[i][2 * k], s1);
        }
        *(u32x4*)(buf + st.lds[i]) = p0;
        *(u32x4*)(buf + st.lds[i] + half_pc * 8) = p1;
    }
}

// 8-byte staging (W a multiple of 4, 8-byte aligned base, patch starting on a column that is
// a multiple of 4): one unit = (channel group g, patch row, pixel quad) = 8 channels x 4
// adjacent pixels, which lie entirely inside or outside the image: one 8-byte buffer load per
// channel, half the load instructions of 4-byte staging. Out-of-image quads and channels >= C
// read zero past the resource's range, as above.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
struct QuadStaging {
    uint32_t off;  // byte offset of the unit's first pixel in channel g*8 of chunk 0
    int lds;       // LDS offset (halves) of the unit's first pixel
    u32x2 w[8];    // loaded quads: pixel pairs 0 and 1 (first pixel of each in the low half)
};

template <int PR, int PC>
__device__ inline void quad_setup(const mykernelParamType& prm, int ih0, int iw0, QuadStaging& st, int set) {
    constexpr int PQ = PC >> 2;
    constexpr int nunits = 4 * PR * PQ;
    static_assert(PC % 4 == 0 && nunits <= 1024, "one or two quad units per thread");
    const int H = prm.h, W = prm.w;
    const int u = threadIdx.x + 512 * set;
    const int g = u / (PR * PQ);
    const int rem = u - g * (PR * PQ);
    const int pr = rem / PQ;
    const int m = rem - pr * PQ;
    const int ih = ih0 + pr, iw = iw0 + 4 * m;
    const bool pix = u < nunits && ih >= 0 && ih < H && iw >= 0 && iw < W;
    st.off = pix ? (uint32_t)(((g * 8 * H + ih) * W + iw) * 2) : 0x80000000u;
    // quad m: pixel pairs 2m and 2m+1; a thread slot without a unit stores to a dummy area past
    // both patch buffers
    st.lds = u < nunits ? ((g * PR + pr) * PC + 2 * m) * 8 : 2 * 4 * PR * PC * 8;
}

__device__ inline void quad_load(__amdgpu_buffer_rsrc_t in, uint32_t chunk_bytes, uint32_t chan_bytes,
                                 QuadStaging& st) {
    const uint32_t base = st.off + chunk_bytes;
#pragma unroll
    for (int e = 0; e < 8; ++e)
        st.w[e] = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(in, (int)(base + e * chan_bytes), 0, 0));
}

template <int PC>
__device__ inline void quad_store(const QuadStaging& st, half_t* buf) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        u32x4 p0, p1;  // pixels 2(2m+d) and 2(2m+d)+1, 8 channels each
        p0[0] = pack_lo(st.w[0][d], st.w[1][d]);
        p0[1] = pack_lo(st.w[2][d], st.w[3][d]);
        p0[2] = pack_lo(st.w[4][d], st.w[5][d]);
        p0[3] = pack_lo(st.w[6][d], st.w[7][d]);
        p1[0] = pack_hi(st.w[0][d], st.w[1][d]);
        p1[1] = pack_hi(st.w[2][d], st.w[3][d]);
        p1[2] = pack_hi(st.w[4][d], st.w[5][d]);
        p1[3] = pack_hi(st.w[6][d], st.w[7][d]);
        *(u32x4*)(buf + st.lds + d * 8) = p0;
        *(u32x4*)(buf + st.lds + d * 8 + (PC / 2) * 8) = p1;
    }
}

// Channel-contiguous (N,H,W,C) input, C % 8 == 0 and a 16-byte aligned base: one unit = (channel
// group g, patch row, pixel) = the 8 channels of one pixel, one 16-byte load stored to LDS as it
// is (the patch layout is already channel-interleaved), no permutes. Thread t holds units t,
// t + 512, ... Pixels outside the image, and channel groups at or beyond C in the last chunk,
// take an offset past the resource's range and read zero.
struct NhwcStaging {
    uint32_t off[kFusedNhwcUnits];  // byte offset of the unit's pixel, channel g*8 of chunk 0
    int g8[kFusedNhwcUnits];        // first channel of the unit within a chunk (g*8)
    int lds[kFusedNhwcUnits];       // LDS offset (halves)
    u32x4 w[kFusedNhwcUnits];
};

template <int PR, int PC>
__device__ inline void nhwc_setup(const mykernelParamType& prm, int ih0, int iw0, NhwcStaging& st) {
    constexpr int nunits = 4 * PR * PC;
    static_assert(nunits <= kFusedNhwcUnits * 512, "NHWC units per thread");
    const int H = prm.h, W = prm.w, C = prm.c;
#pragma unroll
    for (int i = 0; i < kFusedNhwcUnits; ++i) {
        const int u = threadIdx.x + 512 * i;
        const int g = u / (PR * PC);
        const int rem = u - g * (PR * PC);
        const int pr = rem / PC;
        const int col = rem - pr * PC;
        const int ih = ih0 + pr, iw = iw0 + col;
        const bool pix = u < nunits && ih >= 0 && ih < H && iw >= 0 && iw < W;
        st.off[i] = pix ? (uint32_t)(((ih * W + iw) * C + g * 8) * 2) : 0x80000000u;
        st.g8[i] = u < nunits ? g * 8 : 1 << 20;
        // column col at parity col & 1, position col / 2 (a thread slot without a unit stores to a
        // dummy slot past both patch buffers)
        st.lds[i] = u < nunits ? ((g * PR + pr) * PC + (col & 1) * (PC / 2) + (col >> 1)) * 8 : 2 * 4 * PR * PC * 8;
    }
}

// chunk ch: channels ch*32 .. +31; groups at or beyond C (only in the last chunk) read zero
__device__ inline void nhwc_load(__amdgpu_buffer_rsrc_t in, int ch, int C, NhwcStaging& st) {
    const int lim = C - ch * 32;
#pragma unroll
    for (int i = 0; i < kFusedNhwcUnits; ++i) {
        const uint32_t o = st.g8[i] < lim ? st.off[i] + (uint32_t)ch * 64u : 0x80000000u;
        st.w[i] = __builtin_amdgcn_raw_buffer_load_b128(in, (int)o, 0, 0);
    }
}

__device__ inline void nhwc_store(const NhwcStaging& st, half_t* buf) {
#pragma unroll
    for (int i = 0; i < kFusedNhwcUnits; ++i) *(u32x4*)(buf + st.lds[i]) = st.w[i];
}

// The staging of one thread for load width LW: one or two 8-byte quad units, or two or three
// 2- or 4-byte pair units (512 threads x 2 pair units cover the largest 64-tile pair patch, 816
// units; the 128-tile patches of the wide tail items need three pair units, up to 1,360, or two
// quad units, up to 816).
template <int LW, int PR, int PC>
struct PatchStager {
    static constexpr bool quad = LW == kFusedLoad8;
    static constexpr bool d16 = LW == kFusedLoad2;
    static constexpr bool nhwc = LW == kFusedLoadNHWC;
    static constexpr int NU = (4 * PR * (PC / 2) + 511) / 512;
    static constexpr int NQ = quad ? (4 * PR * (PC / 4) + 511) / 512 : 1;
    static_assert(nhwc || quad || (NU >= 2 && NU <= 3), "two or three pair units per thread");
    QuadStaging q[NQ];
    Staging<NU> p;
    NhwcStaging n;
    __device__ void setup(const mykernelParamType& prm, int ih0, int iw0) {
        if constexpr (nhwc) nhwc_setup<PR, PC>(prm, ih0, iw0, n);
        else if constexpr (quad) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) quad_setup<PR, PC>(prm, ih0, iw0, q[i], i);
        }
        else stage_setup<NU, d16, 512, PR, PC>(prm, ih0, iw0, p);
    }
    // NHWC: chunk index ch and the channel count; otherwise the chunk's and one channel's
    // byte strides
    __device__ void load_nhwc(__amdgpu_buffer_rsrc_t in, int ch, int C) { nhwc_load(in, ch, C, n); }
    __device__ void load(__amdgpu_buffer_rsrc_t in, uint32_t chunk_bytes, uint32_t chan_bytes) {
        if constexpr (quad) {
            // 360-432 quad units (720-816 for wide items): the waves past the last unit of a set
            // issue no loads for it (one wave-uniform branch around all eight; their stores go to
            // the dummy area)
            constexpr int nunits = 4 * PR * (PC / 4);
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 64 + 512 * i < nunits)
                    quad_load(in, chunk_bytes, chan_bytes, q[i]);
        } else {
            // 648-1,360 pair units: the last unit set (threads + 512 * (NU - 1)) ends part way
            // through the workgroup; the later waves skip its loads (wave-uniform branch; their
            // stores go to the dummy pair)
            constexpr int nunits = 4 * PR * (PC / 2);
            stage_load<1>(in, chunk_bytes, chan_bytes, p, 0);
#pragma unroll
            for (int i = 1; i < NU; ++i)
                if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 64 + 512 * i < nunits)
                    stage_load<1>(in, chunk_bytes, chan_bytes, p, i);
        }
    }
    __device__ void store(half_t* buf) const {
        if constexpr (nhwc) nhwc_store(n, buf);
        else if constexpr (quad) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) quad_store<PC>(q[i], buf);
        }
        else stage_store<NU, d16>(p, buf, PC / 2);
    }
};

}  // namespace

// One work item: 64 output channels (block kblk) x one FBH x FBW = 64-tile block (gblk,
// image-major) of the problem in prm, by the waves of one half (BH) of a 512-thread
// workgroup, in the launch's dynamic LDS, with staging load width LW. The patch is
// PR = 2*FBH + 2 rows by PC columns starting FSH columns left of 2*tw0 - q
// (fused_patch_cols, fused_col_shift), so every LDS offset except the lane's two bases
// (window rows rp, rm) and the chunk's buffer is a compile-time constant. HALF: the block's
// second 32 output channels lie beyond K (the last block when K % 64 is 1..32); their U loads,
// MFMAs and epilogue work are left out. WIDE (the same last block, the other way): 32 output
// channels x a block of FBH x 2*FBW = 128 tiles (gblk then numbers these wide blocks, half as
// many per block row, rounded up), with the accumulators' channel-block index m standing for
// the block's two 64-tile halves: one item does the MFMA work of a full one, with half its U
// loads, where two HALF items would each pay the per-item prologue and epilogue.
template <int LW, int BH, int FBH, int FBW, int FSH, bool HALF, bool WIDE = false>
__device__ __attribute__((always_inline)) inline void fused_item(const mykernelParamType& prm, int kblk, int gblk) {
    static_assert(!(HALF && WIDE), "one tail form");
    constexpr int BW = WIDE ? 2 * FBW : FBW;  // block width in tiles
    constexpr int PR = fused_patch_rows(FBH), PC = fused_patch_cols(LW, BW, FSH);
    constexpr int PB = 4 * PR * PC * 8;  // halves per patch buffer
    static_assert(FBH * FBW == 64, "64 tiles per block");
    static_assert(2 * PB * 2 + PC * 16 <= 131072, "two patch buffers and the dummy area within the Z area");
    extern __shared__ __attribute__((aligned(16))) half_t lds[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int a = wave & 3;

    const int K = prm.k, C = prm.c;
    const int KB = prm.KB32, CB = prm.CB, C16 = 2 * CB;
    const int nbw = WIDE ? (prm.fnbw + 1) / 2 : prm.fnbw, nbh = prm.fnbh;
    const int img = gblk / (nbh * nbw);
    const int brem = gblk - img * (nbh * nbw);
    const int th0 = (brem / nbw) * FBH;
    const int tw0 = (brem - (brem / nbw) * nbw) * BW;
    const int ih0 = 2 * th0 - prm.p;
    const int iw0 = 2 * tw0 - prm.q - FSH;
    const half_t* in_img = (const half_t*)prm.pin + (size_t)img * C * prm.h * prm.w;
    const uint32_t chan_bytes = (uint32_t)prm.h * prm.w * 2;
    const __amdgpu_buffer_rsrc_t in_rsrc = buffer_rsrc(in_img, (uint32_t)C * chan_bytes);
    constexpr bool kNhwc = LW == kFusedLoadNHWC;

    // U fragments (ks, bb, m) for pos = 4a + 2BH + bb; a block beyond K reads block 0
    const __amdgpu_buffer_rsrc_t u_rsrc =
        buffer_rsrc(prm.transform_filter, (uint32_t)(16u * C16 * KB * kFragHalves * 2u));
    const int ulane = lane * 16;
    auto load_u = [&](half8 (&u)[8], int ch) {
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int m = 0; m < (HALF || WIDE ? 1 : 2); ++m) {
                    const int kb = kblk * 2 + m < KB ? kblk * 2 + m : 0;
                    const uint32_t frag = ((uint32_t)(4 * a + 2 * BH + bb) * C16 + ch * 2 + ks) * KB + kb;
                    const u32x4 raw =
                        __builtin_amdgcn_raw_buffer_load_b128(u_rsrc, ulane, (int)(frag * (kFragHalves * 2u)), 0);
                    u[(ks * 2 + bb) * 2 + m] = __builtin_bit_cast(half8, raw);
                }
    };

    // rows of the window that row a of B^T combines: s = d[rp] + sg * d[rm], sg = +-1 (exact,
    // so the fma rounds once like the addition or subtraction it stands for)
    const int rp = a == 0 ? 0 : (a == 2 ? 2 : 1);
    const int rm = a == 0 ? 2 : (a == 2 ? 1 : (a == 1 ? 2 : 3));
    // LDS offsets (halves) of this lane's window rows rp and rm in channel group lane/32 of
    // buffer 0: tile t = 64*mt + tb*32 + lane%32 (mt the 64-tile half, WIDE only) at patch row
    // 2*(t / BW), column pair t % BW; window column j then sits at parity (FSH + j) & 1, position
    // (FSH + j) / 2 past the pair
    constexpr int NMT = WIDE ? 2 : 1;
    int rbase[NMT][2][2];
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
            const int t = mt * 64 + tb * 32 + (lane & 31);
            const int lb = ((lane >> 5) * PR * PC + 2 * (t / BW) * PC + t % BW) * 8;
            rbase[mt][tb][0] = lb + rp * PC * 8;
            rbase[mt][tb][1] = lb + rm * PC * 8;
        }
    const half_t sgh = a == 1 ? (half_t)1.0f : (half_t)-1.0f;
    const half8 sg = {sgh, sgh, sgh, sgh, sgh, sgh, sgh, sgh};
    half8 m1 = {-1.f16, -1.f16, -1.f16, -1.f16, -1.f16, -1.f16, -1.f16, -1.f16};
    asm volatile("" : "+s"(m1));

    // written first by chunk 0's MFMAs, whose accumulator input is the constant zero (no
    // zeroing instructions per item); with HALF the second channel block is never written or
    // read
    floatx16 acc[2][2][2];

    // the next chunk's patch loads are in flight during this chunk's MFMAs; U has one
    // register set (two waves per SIMD hide its L2 latency), reloaded after its last use
    PatchStager<LW, PR, PC> st;
    st.setup(prm, ih0, iw0);
    auto stage_chunk = [&](int ch) {
        if constexpr (kNhwc) st.load_nhwc(in_rsrc, ch, C);
        else st.load(in_rsrc, (uint32_t)ch * 32u * chan_bytes, chan_bytes);
    };
    stage_chunk(0);
    half8 ucur[8];
    load_u(ucur, 0);  // in flight while the first patch is stored
    st.store(lds);
    __syncthreads();

    auto chunk = [&](int ch, auto first_tag) {
        constexpr bool FIRST = decltype(first_tag)::value;
        const int buf_ofs = (ch & 1) * PB;
        const bool more = ch + 1 < CB;
        // Wait for this chunk's U here, before the next chunk's patch loads are issued: the
        // memory counter drains in issue order, so a wait for U placed after those loads (at the
        // first MFMA) would also wait for them and expose their whole latency every chunk.
        asm volatile("" ::"v"(ucur[0]), "v"(ucur[1]), "v"(ucur[2]), "v"(ucur[3]), "v"(ucur[4]), "v"(ucur[5]),
                     "v"(ucur[6]), "v"(ucur[7]));
        if (more) stage_chunk(ch + 1);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int tb = 0; tb < 2; ++tb) {
                const half_t* wpp = lds + buf_ofs + rbase[mt][tb][0] + ks * 2 * PR * PC * 8;
                const half_t* wpm = lds + buf_ofs + rbase[mt][tb][1] + ks * 2 * PR * PC * 8;
                half8 s3[3];
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    const int j = BH + jj;
                    const int col = (((FSH + j) & 1) * (PC / 2) + ((FSH + j) >> 1)) * 8;
                    const half8 dp = *(const half8*)(wpp + col);
                    const half8 dm = *(const half8*)(wpm + col);
                    s3[jj] = __builtin_elementwise_fma(dm, sg, dp);
                }
                half8 v[2];
                if (BH == 0) {
                    v[0] = sub8(s3[0], s3[2], m1);  // s0 - s2
                    v[1] = s3[1] + s3[2];           // s1 + s2
                } else {
                    v[0] = sub8(s3[1], s3[0], m1);  // s2 - s1
                    v[1] = sub8(s3[0], s3[2], m1);  // s1 - s3
                }
                // issue priority while this wave queues its MFMAs, so the other wave of the SIMD
                // does not interleave its loads and transform between them (about 2% faster)
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int mi = 0; mi < (HALF || WIDE ? 1 : 2); ++mi) {
                        // WIDE: accumulator block m = the tile half, U of the one channel block
                        const int m = WIDE ? mt : mi;
                        acc[bb][m][tb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            ucur[(ks * 2 + bb) * 2 + mi], v[bb], (FIRST && ks == 0) ? floatx16{} : acc[bb][m][tb], 0, 0, 0);
                    }
                __builtin_amdgcn_s_setprio(0);
            }
        }
        if (more) {
            load_u(ucur, ch + 1);
            st.store(lds + (PB - buf_ofs));
        }
        __syncthreads();
    };
    chunk(0, BoolTag<true>{});
    for (int ch = 1; ch < CB; ++ch) chunk(ch, BoolTag<false>{});

    // ---- epilogue; LDS Z layout (floats): [a][j][block = m*2 + tb][e/4][lane][4] ----
    // Row a's fold Z_0 = (M0 + M1) + M2, Z_1 = (M1 - M2) - M3 needs both halves of the row. For
    // channel block m = 0 the BH = 0 wave hands over M0 + M1 and M1 and the BH = 1 wave folds;
    // for m = 1 the BH = 1 wave hands over M2 and M3 and the BH = 0 wave folds, so both halves
    // work in both phases. Z overwrites the handed-over values in place.
    float* Z = (float*)lds;
    constexpr int give = BH;      // channel block whose values this wave hands over
    constexpr int fold = 1 - BH;  // channel block this wave folds
    auto zslot = [&](int j, int m, int tb, int e4) {
        return Z + ((((a * 2 + j) * 4 + m * 2 + tb) * 4 + e4) * 64 + lane) * 4;
    };
    constexpr bool kGive = !(HALF && give == 1), kFold = !(HALF && fold == 1);
#pragma unroll
    for (int tb = 0; tb < (kGive ? 2 : 0); ++tb)
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            floatx4 v0, v1;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = e4 * 4 + q;
                const float x0 = acc[0][give][tb][e], x1 = acc[1][give][tb][e];
                v0[q] = BH == 0 ? x0 + x1 : x0;  // M0 + M1 | M2
                v1[q] = x1;                      // M1      | M3
            }
            *(floatx4*)zslot(0, give, tb, e4) = v0;
            *(floatx4*)zslot(1, give, tb, e4) = v1;
        }
    __syncthreads();
#pragma unroll
    for (int tb = 0; tb < (kFold ? 2 : 0); ++tb)
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            float* z0p = zslot(0, fold, tb, e4);
            float* z1p = zslot(1, fold, tb, e4);
            const floatx4 h0 = *(const floatx4*)z0p, h1 = *(const floatx4*)z1p;
            floatx4 z0, z1;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = e4 * 4 + q;
                const float x0 = acc[0][fold][tb][e], x1 = acc[1][fold][tb][e];
                // BH = 1 holds M2, M3 and received M0 + M1, M1; BH = 0 holds M0, M1 and
                // received M2, M3
                const float P = BH == 1 ? h0[q] : x0 + x1;
                const float M1 = BH == 1 ? h1[q] : x1;
                const float M2 = BH == 1 ? x0 : h0[q];
                const float M3 = BH == 1 ? x1 : h1[q];
                z0[q] = P + M2;          // (M0 + M1) + M2
                z1[q] = (M1 - M2) - M3;  // (M1 - M2) - M3
            }
            *(floatx4*)z0p = z0;
            *(floatx4*)z1p = z1;
        }
    __syncthreads();
    const int blk = wave >> 1;  // output block of this wave, e/4 groups 2*(wave&1) .. +1
    const int mb = blk >> 1, tb = blk & 1;
    if (HALF && mb == 1) return;  // output channels beyond K
    const int m = WIDE ? 0 : mb;  // 32-channel block of the outputs (WIDE: mb is the tile half)
    const int t = (WIDE ? mb * 64 : 0) + tb * 32 + (lane & 31);
    const int th = th0 + t / BW, tw = tw0 + t % BW;
    // Output of this image through a buffer resource: 32-bit byte offsets (the host checks
    // K*Oh*Ow*2 < 2^31), two FP16 values of a row in one 4-byte store when every row starts
    // 4-byte aligned (even Ow, aligned base; then a tile's second column is always inside).
    const int Oh = prm.Oh, Ow = prm.Ow;
    const uint32_t plane = (uint32_t)Oh * (uint32_t)Ow;
    const __amdgpu_buffer_rsrc_t out_rsrc =
        buffer_rsrc((const half_t*)prm.pout + (size_t)img * K * plane, (uint32_t)K * plane * 2u);
    const bool pairs = (Ow & 1) == 0 && ((uintptr_t)prm.pout & 3) == 0;
    // fast path (wave-uniform): NCHW output (FP16 or bfloat16, whatever the input layout),
    // every channel of the block exists, every tile has both rows and columns, no bias or
    // activation: row-pair stores whose channel offset goes in the scalar offset, no
    // per-element checks
    const bool out_bf16 = (prm.io & CONV2D_IO_OUT_BF16) != 0;
    const bool fast = (prm.io & CONV2D_IO_OUT_NHWC) == 0 && pairs && (Oh & 1) == 0 && !prm.bias && !prm.act &&
                      kblk * 64 + (WIDE ? 32 : 64) <= K;
    // one copy of the stores per output type, so that the type test is not repeated per pair
    auto fast_stores = [&](auto bf_tag) {
        constexpr bool BF = decltype(bf_tag)::value;
        if (th < prm.tiles_h && tw < prm.tiles_w) {
            const uint32_t plane2 = plane * 2u;
            const uint32_t row_bytes = (uint32_t)Ow * 2u;
            const uint32_t voff = ((uint32_t)(2 * th) * (uint32_t)Ow + (uint32_t)(2 * tw)) * 2u +
                                  (uint32_t)(4 * (lane >> 5)) * plane2;
#pragma unroll
            for (int eh = 0; eh < 2; ++eh) {
                const int e4 = 2 * (wave & 1) + eh;
                floatx4 z[4][2];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        z[r][j] = *(const floatx4*)(Z + ((((r * 2 + j) * 4 + blk) * 4 + e4) * 64 + lane) * 4);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int e = e4 * 4 + q;
                    const uint32_t kc = (uint32_t)(kblk * 64 + m * 32 + 8 * (e >> 2) + (e & 3));
                    const int soff = (int)__builtin_amdgcn_readfirstlane(kc * plane2);
                    const float y00 = (z[0][0][q] + z[1][0][q]) + z[2][0][q];
                    const float y01 = (z[0][1][q] + z[1][1][q]) + z[2][1][q];
                    const float y10 = (z[1][0][q] - z[2][0][q]) - z[3][0][q];
                    const float y11 = (z[1][1][q] - z[2][1][q]) - z[3][1][q];
                    __builtin_amdgcn_raw_buffer_store_b32(pack_out2(BF, y00, y01), out_rsrc, (int)voff, soff, 0);
                    __builtin_amdgcn_raw_buffer_store_b32(pack_out2(BF, y10, y11), out_rsrc, (int)(voff + row_bytes),
                                                          soff, 0);
                }
            }
        }
    };
    if (fast) {
        if (out_bf16)
            fast_stores(BoolTag<true>{});
        else
            fast_stores(BoolTag<false>{});
    } else if (prm.io != 0) {
        // NHWC and/or bfloat16 output: per pixel of the tile, the 4 consecutive channels of an
        // accumulator group (one 8-byte store in NHWC when K % 4 == 0 and the base is 8-byte
        // aligned, else one 2-byte store per element)
        if (th < prm.tiles_h && tw < prm.tiles_w) {
            const bool nhwc = (prm.io & CONV2D_IO_OUT_NHWC) != 0, bf = (prm.io & CONV2D_IO_OUT_BF16) != 0;
            const bool vec = nhwc && (K & 3) == 0 && ((uintptr_t)prm.pout & 7) == 0;
            const bool rows[2] = {true, 2 * th + 1 < Oh}, cols[2] = {true, 2 * tw + 1 < Ow};
            auto bits = [&](float y) {
                return bf ? (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)y)
                          : (uint32_t)__builtin_bit_cast(uint16_t, (half_t)y);
            };
#pragma unroll
            for (int eh = 0; eh < 2; ++eh) {
                const int e4 = 2 * (wave & 1) + eh;
                floatx4 z[4][2];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        z[r][j] = *(const floatx4*)(Z + ((((r * 2 + j) * 4 + blk) * 4 + e4) * 64 + lane) * 4);
                const int k0 = kblk * 64 + m * 32 + 8 * e4 + 4 * (lane >> 5);
                float yv[4][4];  // [pixel (row, column) of the tile][q]
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float y[4] = {(z[0][0][q] + z[1][0][q]) + z[2][0][q], (z[0][1][q] + z[1][1][q]) + z[2][1][q],
                                  (z[1][0][q] - z[2][0][q]) - z[3][0][q], (z[1][1][q] - z[2][1][q]) - z[3][1][q]};
                    if (prm.bias || prm.act) {
                        const float bk = k0 + q < K ? epilogue_bias(prm, k0 + q) : 0.f;
#pragma unroll
                        for (int i = 0; i < 4; ++i) y[i] = epilogue(prm, y[i], bk);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) yv[i][q] = y[i];
                }
                if (k0 >= K) continue;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (!rows[i >> 1] || !cols[i & 1]) continue;
                    const uint32_t px = (uint32_t)(2 * th + (i >> 1)) * (uint32_t)Ow + (uint32_t)(2 * tw + (i & 1));
                    if (vec) {
                        const u32x2 pk2 = {pack_out2(bf, yv[i][0], yv[i][1]), pack_out2(bf, yv[i][2], yv[i][3])};
                        __builtin_amdgcn_raw_buffer_store_b64(pk2, out_rsrc,
                                                              (int)((px * (uint32_t)K + (uint32_t)k0) * 2u), 0, 0);
                        continue;
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (k0 + q >= K) continue;
                        const uint32_t off = nhwc ? (px * (uint32_t)K + (uint32_t)(k0 + q)) * 2u
                                                  : ((uint32_t)(k0 + q) * plane + px) * 2u;
                        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)bits(yv[i][q]), out_rsrc, (int)off, 0, 0);
                    }
                }
            }
        }
    } else if (th < prm.tiles_h && tw < prm.tiles_w) {
        const bool r1 = 2 * th + 1 < Oh, c1 = 2 * tw + 1 < Ow;
        const uint32_t pix = ((uint32_t)(2 * th) * (uint32_t)Ow + (uint32_t)(2 * tw)) * 2u;
        const uint32_t row_bytes = (uint32_t)Ow * 2u;
        auto put_row = [&](uint32_t off, float ya, float yb) {
            if (pairs) {
                __builtin_amdgcn_raw_buffer_store_b32(pack_f16x2(ya, yb), out_rsrc, (int)off, 0, 0);
            } else {
                const uint16_t ha = __builtin_bit_cast(uint16_t, (half_t)ya);
                const uint16_t hb = __builtin_bit_cast(uint16_t, (half_t)yb);
                __builtin_amdgcn_raw_buffer_store_b16(ha, out_rsrc, (int)off, 0, 0);
                if (c1) __builtin_amdgcn_raw_buffer_store_b16(hb, out_rsrc, (int)(off + 2u), 0, 0);
            }
        };
#pragma unroll
        for (int eh = 0; eh < 2; ++eh) {
            const int e4 = 2 * (wave & 1) + eh;
            floatx4 z[4][2];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    z[r][j] = *(const floatx4*)(Z + ((((r * 2 + j) * 4 + blk) * 4 + e4) * 64 + lane) * 4);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = e4 * 4 + q;
                const int k = kblk * 64 + m * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                if (k >= K) continue;
                float y00 = (z[0][0][q] + z[1][0][q]) + z[2][0][q];
                float y01 = (z[0][1][q] + z[1][1][q]) + z[2][1][q];
                float y10 = (z[1][0][q] - z[2][0][q]) - z[3][0][q];
                float y11 = (z[1][1][q] - z[2][1][q]) - z[3][1][q];
                if (prm.bias || prm.act) {
                    const float bk = epilogue_bias(prm, k);
                    y00 = epilogue(prm, y00, bk);
                    y01 = epilogue(prm, y01, bk);
                    y10 = epilogue(prm, y10, bk);
                    y11 = epilogue(prm, y11, bk);
                }
                const uint32_t off = (uint32_t)k * plane * 2u + pix;
                put_row(off, y00, y01);
                if (r1) put_row(off + row_bytes, y10, y11);
            }
        }
    }
}

// the four 64-tile block shapes, indexed as fused_shape_index numbers them
template <int S> struct BlockShape;
template <> struct BlockShape<0> { static constexpr int h = 8, w = 8; };
template <> struct BlockShape<1> { static constexpr int h = 4, w = 16; };
template <> struct BlockShape<2> { static constexpr int h = 16, w = 4; };
template <> struct BlockShape<3> { static constexpr int h = 2, w = 32; };

// S = 4 * (block shape index) + column shift FSH. With TAIL the item takes the half-block body
// when its second 32 channels lie beyond K (wave-uniform choice per item).
template <int LW, int S, bool TAIL>
__device__ __attribute__((always_inline)) inline void fused_item_dispatch(const mykernelParamType& prm, int kblk, int gblk) {
    constexpr int FBH = BlockShape<S / 4>::h, FBW = BlockShape<S / 4>::w, FSH = S & 3;
    // waves 0-3 take BH = 0, waves 4-7 BH = 1 (wave-uniform; both paths pass the same barriers)
    const bool bh1 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8) != 0;
    if (TAIL && kblk * 64 + 32 >= prm.k) {
        if (!bh1) fused_item<LW, 0, FBH, FBW, FSH, true>(prm, kblk, gblk);
        else fused_item<LW, 1, FBH, FBW, FSH, true>(prm, kblk, gblk);
    } else {
        if (!bh1) fused_item<LW, 0, FBH, FBW, FSH, false>(prm, kblk, gblk);
        else fused_item<LW, 1, FBH, FBW, FSH, false>(prm, kblk, gblk);
    }
}

// The last 64-channel block as wide items (32 channels x 128 tiles), gblk numbering the wide
// blocks; K % 64 is 1..32.
template <int LW, int S>
__device__ __attribute__((always_inline)) inline void fused_wide_dispatch(const mykernelParamType& prm, int gblk) {
    constexpr int FBH = BlockShape<S / 4>::h, FBW = BlockShape<S / 4>::w, FSH = S & 3;
    const bool bh1 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8) != 0;
    if (!bh1) fused_item<LW, 0, FBH, FBW, FSH, false, true>(prm, prm.k >> 6, gblk);
    else fused_item<LW, 1, FBH, FBW, FSH, false, true>(prm, prm.k >> 6, gblk);
}

// Single-layer launch. With TAIL (K % 64 is 1..32) and a wide tail form for this staging and
// block shape, a grid smaller than one item per (64-channel block, 64-tile block) means the host
// chose wide tail items: items [0, nmain) are the K / 64 full 64-channel blocks of every
// 64-tile block and the rest the wide items (fused_wide_items per image), each range remapped
// per XCD on its own. Otherwise the last block's items take the half-block body.
template <int LW, int S, bool TAIL>
__global__ void __launch_bounds__(512, 1) wino_fused_kernel(mykernelParamType prm) {
    int kblk, gblk;
    if constexpr (TAIL && fused_wide_tail(LW, S / 4)) {
        const int nblk = prm.n * prm.fnbh * prm.fnbw;
        const int nfull = prm.k >> 6;
        const int nmain = nfull * nblk;
        const int L = blockIdx.x;
        if ((int)gridDim.x < nmain + nblk) {
            if (L < nmain) {
                xcd_remap(L, nmain, nfull, &kblk, &gblk);
                fused_item_dispatch<LW, S, false>(prm, kblk, gblk);
            } else {
                xcd_remap(L - nmain, (int)gridDim.x - nmain, 1, &kblk, &gblk);
                fused_wide_dispatch<LW, S>(prm, gblk);
            }
            return;
        }
    }
    xcd_remap(blockIdx.x, gridDim.x, (prm.k + 63) / 64, &kblk, &gblk);
    fused_item_dispatch<LW, S, TAIL>(prm, kblk, gblk);
}

// Multi-layer form: one launch over the fused items of several layers (descriptor i owns items
// [s3_first, s3_first + s3_items), nk = 64-channel blocks, remapped per XCD within that range),
// every layer with the same staging load width, block shape and column shift; a layer's last
// block runs as wide or half-block items as in its single-layer launch.
template <int LW, int S>
__global__ void __launch_bounds__(512, 1)
wino_fused_grouped_kernel(const BatchLayer* __restrict__ layers, int nlayers) {
    int local;
    const BatchLayer& bl = grouped_layer_item(layers, nlayers, blockIdx.x, &local);
    // the layer's parameters by value, loaded once, as the single-layer kernel receives them
    const mykernelParamType prm = bl.prm;
    const int nk = bl.nk;
    if constexpr (fused_wide_tail(LW, S / 4)) {
        // fewer items than (64-channel block, 64-tile block) pairs: the host chose wide tail
        // items for this layer, after its K / 64 full blocks of every 64-tile block
        const int nblk = prm.n * prm.fnbh * prm.fnbw;
        const int nmain = (prm.k >> 6) * nblk;
        if (bl.s3_items < nk * nblk) {
            if (local < nmain) fused_item_dispatch<LW, S, false>(prm, local % (prm.k >> 6), local / (prm.k >> 6));
            else fused_wide_dispatch<LW, S>(prm, local - nmain);
            return;
        }
    }
    fused_item_dispatch<LW, S, true>(prm, local % nk, local / nk);
}

int fused_threads() { return kFusedThreads; }

// Instances: 2- and 4-byte staging for every block shape with column shift q & 1; 8-byte
// staging with the shifts (4 - q % 4) % 4 of q = 0, 1, 2 wherever the quad patch fits one unit
// per thread (all but the 16 x 4 block with shift 3); each with and without a tail body
// (layers whose K % 64 is 1..32 take the tail instance: wide tail items where fused_wide_tail
// allows, else the half-block body). The grouped kernels always carry the half-block body
// (their layers have different K). The tables map (lw, shape, fsh, tail)
// to them.
namespace {
template <int LW, int S>
constexpr bool fused_instance() {
    return LW == kFusedLoad8 ? ((S & 3) != 1 && !(S / 4 == 2 && (S & 3) == 3))
                             : (LW == kFusedLoadNHWC ? (S & 3) == 0 : (S & 3) <= 1);
}

// index I = 64 * tail + 16 * lw + S
template <int I, bool Grouped>
const void* fused_ptr_one() {
    constexpr int LW = (I % 64) / 16, S = I % 16;
    constexpr bool TAIL = I >= 64;
    if constexpr (!fused_instance<LW, S>() || (Grouped && !TAIL)) return nullptr;
    else if constexpr (Grouped) return (const void*)&wino_fused_grouped_kernel<LW, S>;
    else return (const void*)&wino_fused_kernel<LW, S, TAIL>;
}

template <bool Grouped, int... I>
const void* fused_ptr_lookup(int idx, std::integer_sequence<int, I...>) {
    static const void* const table[] = {fused_ptr_one<I, Grouped>()...};
    return table[idx];
}
}  // namespace

const void* fused_kernel_ptr(int lw, int shape, int fsh, bool tail) {
    if (lw < 0 || lw >= kFusedLoadWidths || shape < 0 || shape > 3 || fsh < 0 || fsh > 3) return nullptr;
    return fused_ptr_lookup<false>((tail ? 64 : 0) + lw * 16 + shape * 4 + fsh,
                                   std::make_integer_sequence<int, 32 * kFusedLoadWidths>{});
}

const void* fused_grouped_kernel_ptr(int lw, int shape, int fsh) {
    if (lw < 0 || lw >= kFusedLoadWidths || shape < 0 || shape > 3 || fsh < 0 || fsh > 3) return nullptr;
    return fused_ptr_lookup<true>(64 + lw * 16 + shape * 4 + fsh, std::make_integer_sequence<int, 32 * kFusedLoadWidths>{});
}
