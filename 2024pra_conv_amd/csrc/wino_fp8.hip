// FP8 (OCP e4m3fn) compensated form of the transform and GEMM stages.
//
// No counterpart in the reference (it extends the method). Each operand value v is split
// into a high and a low e4m3 part, v ~= hi * 2^E_hi + lo * 2^E_lo, with one E8M0 exponent
// pair per (Winograd position, row, 32-channel block) -- the microscaling block that
// v_mfma_scale_f32_32x32x64_f8f6f4 applies in hardware. E_hi is the smallest exponent with
// max|v| <= 448 * 2^E_hi; E_lo = E_hi - 4. The product is recovered in three FP8 MFMA passes
// accumulated in FP32:  U.V ~ U_hi.V_hi + U_hi.V_lo + U_lo.V_hi  (the dropped U_lo.V_lo term is
// below 2^-8 of the block scales' product). With the split of V alone (two passes) the
// filter's own e4m3 rounding leaves a few percent of error; the third pass removes it.
//
// The transforms keep FP32 arithmetic: U is quantised from the FP16 transformed filter that
// S1 writes, V straight from the FP32 input transform (no FP16 rounding of V). Every step of
// the quantisation is exact in FP32 (power-of-two scalings; the hi residual is v minus v
// rounded to 4 significant bits), then rounded once to e4m3 (nearest, ties to even).
#include "wino_fp8.h"
#include "wino_tile_ops.h"

typedef int int4v __attribute__((ext_vector_type(4)));
typedef int int8v __attribute__((ext_vector_type(8)));

namespace {

// Smallest E with m <= 448 * 2^E (m >= 0), clamped to [-127, 127]; -127 for m = 0.
// 448 = 0.875 * 2^9, and m = f * 2^x with f in [0.5, 1).
__device__ inline int block_exponent(float m) {
    if (!(m > 0.f)) return -127;
    int x;
    const float f = frexpf(m, &x);
    const int E = x - 9 + (f > 0.875f ? 1 : 0);
    return E < -127 ? -127 : (E > 127 ? 127 : E);
}

typedef short short2v __attribute__((ext_vector_type(2)));

// The FP32 scale 2^E of an exponent pair member. Only all-zero blocks get E = -127 (every
// nonzero transformed value is at least 2^-24, so E >= -37); their codes are zero for any
// scale, so the scale is kept at the smallest normal power of two.
__device__ inline float exp_scale(int E) { return ldexpf(1.f, E < -126 ? -126 : E); }

// Four values -> four e4m3 codes of v / s, nearest-even (s a power of two, so the scaling is
// exact; gfx950's scaled conversion gives the same codes as scaling then converting, checked
// on 4M values with ties, subnormals and zeros of both signs).
__device__ inline uint32_t pack4_e4m3(float a, float b, float c, float d, float s) {
    short2v w = {0, 0};
    w = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(w, a, b, s, false);
    w = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(w, c, d, s, true);
    return __builtin_bit_cast(uint32_t, w);
}

// Splits n values (multiple of 4) into hi / lo code words given the block exponents: hi codes
// of v / 2^E_hi, then lo codes of the residual (v - hi * 2^E_hi) / 2^E_lo; the residual is exact
// in FP32, and the scaled decode equals decoding then scaling (same check).
template <int N>
__device__ inline void split_hi_lo(const float (&v)[N], int Ehi, int Elo, uint32_t (&hi)[N / 4],
                                   uint32_t (&lo)[N / 4]) {
    const float shi = exp_scale(Ehi), slo = exp_scale(Elo);
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        hi[q] = pack4_e4m3(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3], shi);
        const int h = (int)hi[q];
        const float r0 = v[4 * q] - __builtin_amdgcn_cvt_scalef32_f32_fp8(h, shi, 0);
        const float r1 = v[4 * q + 1] - __builtin_amdgcn_cvt_scalef32_f32_fp8(h, shi, 1);
        const float r2 = v[4 * q + 2] - __builtin_amdgcn_cvt_scalef32_f32_fp8(h, shi, 2);
        const float r3 = v[4 * q + 3] - __builtin_amdgcn_cvt_scalef32_f32_fp8(h, shi, 3);
        lo[q] = pack4_e4m3(r0, r1, r2, r3, slo);
    }
}

__device__ inline int lo_exponent(int Ehi) { return max(Ehi - kFp8LoOffset, -127); }

}  // namespace

// ---------------------------------------------------------------------------------------
// U (FP16, packed by S1) -> FP8 records. One thread per (position, row, 32-channel block):
// it gathers the block's 32 channels from the four FP16 lane slots that hold them, then
// writes two 16-byte code runs per part and one scale pair. Consecutive threads take
// consecutive rows, so the 16-byte stores of a wave are contiguous.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) wino_filter_fp8(mykernelParamType prm) {
    const int KB = prm.KB32, CB = prm.CB, CQ = fp8_cq(prm.c);
    const int rows = KB * 32;
    const long long total = 16LL * 2 * CQ * rows;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int r = (int)(t % rows);
    const int cb = (int)((t / rows) % (2 * CQ));  // 32-channel block
    const int pos = (int)(t / ((long long)rows * 2 * CQ));
    float v[32];
    if (cb < CB) {
        const half_t* U = (const half_t*)prm.transform_filter;
#pragma unroll
        for (int hc = 0; hc < 2; ++hc)      // 16-channel block 2cb + hc
#pragma unroll
            for (int cg = 0; cg < 2; ++cg) {  // 8-channel group (lane half) inside it
                const half8 h =
                    *(const half8*)(U + frag_offset(pos, 2 * cb + hc, r >> 5, 2 * CB, KB) + (cg * 32 + (r & 31)) * 8);
#pragma unroll
                for (int e = 0; e < 8; ++e) v[hc * 16 + cg * 8 + e] = (float)h[e];
            }
    } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) v[i] = 0.f;
    }
    float m = 0.f;
#pragma unroll
    for (int i = 0; i < 32; ++i) m = fmaxf(m, fabsf(v[i]));
    const int Ehi = block_exponent(m), Elo = lo_exponent(Ehi);
    uint32_t hi[8], lo[8];
    split_hi_lo<32>(v, Ehi, Elo, hi, lo);
    uint8_t* U8 = (uint8_t*)prm.transform_filter + fp8_u8_offset(CB, KB);
    const size_t ri = (size_t)(pos * CQ + (cb >> 1)) * KB + (r >> 5);  // record index
    uint8_t* rec = U8 + ri * kFp8CodeBytes;
    uint8_t* sc = U8 + fp8_scale_offset(prm.c, KB) + ri * kFp8ScaleBytes;
    uint8_t* part = rec + (cb & 1) * 1024;
    const int l0 = r & 31, l1 = 32 + (r & 31);  // lanes holding channels 0-15 / 16-31
    *(u32x4*)(part + l0 * 16) = u32x4{hi[0], hi[1], hi[2], hi[3]};
    *(u32x4*)(part + l1 * 16) = u32x4{hi[4], hi[5], hi[6], hi[7]};
    *(u32x4*)(part + 2048 + l0 * 16) = u32x4{lo[0], lo[1], lo[2], lo[3]};
    *(u32x4*)(part + 2048 + l1 * 16) = u32x4{lo[4], lo[5], lo[6], lo[7]};
    *(uint16_t*)(sc + ((cb & 1) * 32 + l0) * 2) = (uint16_t)((Ehi + 127) | ((Elo + 127) << 8));
}

// ---------------------------------------------------------------------------------------
// Input transform straight to FP8 records. Same workgroup decomposition and LDS staging as
// the FP16 transform (band of tile rows x 32-channel block x image). Work unit = 16
// consecutive tiles x the 32 staged channels; lane = (tile lane % 16, channel quarter
// lane / 16) computes 8 channels x 16 positions in FP32, one row of 4 positions at a time
// (B^T d B row a combines window rows rp(a) and rm(a): the same operations as the full
// transform, so the same bits, in 68 instead of 180 registers); the
// four quarters of a tile agree on the block exponents by a cross-lane maximum, then each
// stores its 8 code bytes per part. A wave's stores cover two contiguous 256-byte runs per
// part and position.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) wino_input_transform_fp8(mykernelParamType prm) {
    extern __shared__ __attribute__((aligned(16))) half_t lds[];
    const int n = blockIdx.z, cb = blockIdx.y, bi = blockIdx.x;
    stage_input_band(prm, n, cb, bi, lds);
    const int TH = prm.band;
    const int th0 = bi * TH;
    const int tiles_w = prm.tiles_w;
    const int th1 = min(th0 + TH, prm.tiles_h);
    const int LR = input_lds_rows(TH);
    const int LWp = input_lds_pitch(tiles_w, prm.w, prm.q);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int T = prm.T, NT = prm.NT, G32 = prm.G32, CB = prm.CB, CQ = fp8_cq(prm.c);
    const int gs = n * T + th0 * tiles_w;
    const int ge = n * T + th1 * tiles_w;
    // the last image's last band also writes the zero columns up to the 32-column padding
    const int gend = ge == NT ? G32 * 32 : ge;
    const int q = lane >> 4;  // channel quarter: staged channels 8q .. 8q+7
    const int part = cb & 1;  // 32-channel block inside the 64-channel record
    // an odd channel-block count leaves the second half of the last record to this block
    const bool pad_half = (cb == CB - 1) && (CB & 1);
    uint8_t* V8 = (uint8_t*)prm.transform_pin;
    const int grp0 = gs >> 4;
    const int nunits = ((gend - 1) >> 4) - grp0 + 1;
    for (int u = wave; u < nunits; u += nwaves) {
        const int g = (grp0 + u) * 16 + (lane & 15);
        const bool valid = g >= gs && g < ge;
        const bool store = g >= gs && g < gend;
        const int t = valid ? g - n * T : 0;
        const int th = t / tiles_w;
        const int tw = t - th * tiles_w;
        const int lcode = (q >> 1) * 32 + (g & 31);  // operand lane holding these channels
        const int lscale = part * 32 + (g & 31);     // lane holding this block's scales
#pragma unroll 1
        for (int a = 0; a < 4; ++a) {
            // row a of B^T d B: s_j = d[rp][j] + sg * d[rm][j] (rows 0-2, 1+2, 2-1, 1-3; the
            // product with sg = +-1 is exact, so the sum rounds once like the subtraction)
            const int rp = a == 0 ? 0 : (a == 2 ? 2 : 1);
            const int rm = a == 0 ? 2 : (a == 2 ? 1 : (a == 1 ? 2 : 3));
            const float sg = a == 1 ? 1.f : -1.f;
            float v[4][8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float sj[4] = {0.f, 0.f, 0.f, 0.f};
                if (valid) {
                    const half_t* base = lds + ((q * 8 + e) * LR + 2 * (th - th0)) * LWp + 2 * tw + 8;
                    const half2_t p0 = *(const half2_t*)(base + rp * LWp), p1 = *(const half2_t*)(base + rp * LWp + 2);
                    const half2_t m0 = *(const half2_t*)(base + rm * LWp), m1 = *(const half2_t*)(base + rm * LWp + 2);
                    const float dp[4] = {(float)p0[0], (float)p0[1], (float)p1[0], (float)p1[1]};
                    const float dm[4] = {(float)m0[0], (float)m0[1], (float)m1[0], (float)m1[1]};
#pragma unroll
                    for (int j = 0; j < 4; ++j) sj[j] = dp[j] + sg * dm[j];
                }
                // columns, as the full transform: s0 - s2, s1 + s2, s2 - s1, s1 - s3
                v[0][e] = sj[0] - sj[2];
                v[1][e] = sj[1] + sj[2];
                v[2][e] = sj[2] - sj[1];
                v[3][e] = sj[1] - sj[3];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int pos = 4 * a + b;
                float m = 0.f;
#pragma unroll
                for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(v[b][e]));
                m = fmaxf(m, __shfl_xor(m, 16));
                m = fmaxf(m, __shfl_xor(m, 32));
                const int Ehi = block_exponent(m), Elo = lo_exponent(Ehi);
                uint32_t hi[2], lo[2];
                split_hi_lo<8>(v[b], Ehi, Elo, hi, lo);
                if (store) {
                    const size_t ri = (size_t)(pos * CQ + (cb >> 1)) * G32 + (g >> 5);  // record index
                    uint8_t* rec = V8 + ri * kFp8CodeBytes;
                    uint8_t* sc = V8 + fp8_scale_offset(prm.c, G32) + ri * kFp8ScaleBytes;
                    uint8_t* dst = rec + part * 1024 + lcode * 16 + (q & 1) * 8;
                    *(uint2*)dst = make_uint2(hi[0], hi[1]);
                    *(uint2*)(dst + 2048) = make_uint2(lo[0], lo[1]);
                    if (q == 0) *(uint16_t*)(sc + lscale * 2) = (uint16_t)((Ehi + 127) | ((Elo + 127) << 8));
                    if (pad_half) {
                        uint8_t* z = rec + 1024 + lcode * 16 + (q & 1) * 8;
                        *(uint2*)z = make_uint2(0u, 0u);
                        *(uint2*)(z + 2048) = make_uint2(0u, 0u);
                        if (q == 0) *(uint16_t*)(sc + (32 + (g & 31)) * 2) = (uint16_t)0;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Three-pass FP8 GEMM with the fused output transform, on the LDS-DMA pipeline of the FP16
// GEMM (workgroup tile BK x BT, waves WK x WT with 32x32 sub-tiles). One step = one 64-channel
// block of one position = UF + VF records. A step's record codes (four 1 KiB units each) go
// global -> LDS with global_load_lds_dwordx4, every wave moving the same number of whole units,
// and its (UF + VF) x 128 scale bytes with one global_load_lds_dword per wave (64 lanes x 4
// bytes, each lane fetching its own word of the step's scale list); S LDS stages, S - 2 steps
// in flight, the waits counted by hand (every wave issues exactly CPW + 1 such loads per step
// and no other vector-memory loads inside the loop).
// ---------------------------------------------------------------------------------------
namespace {

template <int BK, int BT, int WK, int WT, int S>
struct Fp8Shape {
    static constexpr int kWaves = WK * WT;
    static constexpr int kThreads = kWaves * 64;
    static constexpr int MI = BK / WK / 32, NI = BT / WT / 32;
    static constexpr int UF = BK / 32, VF = BT / 32;
    static constexpr int kRecs = UF + VF;                    // records per step: U first, then V
    static constexpr int kCPW = 4 * kRecs / kWaves;          // 1 KiB code units per wave and step
    static constexpr int kVPW = kCPW + 1;                    // vector-memory loads per wave and step
    static constexpr int kScaleOfs = kRecs * kFp8CodeBytes;  // a stage: codes, then the scale words
    static constexpr int kStageBytes = kScaleOfs + kWaves * 256;
    static constexpr int kSmemBytes = S * kStageBytes;
    // operand registers double-buffered (the next step's LDS reads overlap this step's MFMAs)
    // when they fit beside the accumulators and the fold's four output sums
    static constexpr bool kDB = (MI + NI) * 2 * 17 + MI * NI * 80 <= 200 || kWaves <= 4;
    static_assert((4 * kRecs) % kWaves == 0, "code units must split evenly over the waves");
    static_assert(kRecs * kFp8ScaleBytes <= kWaves * 256, "one 256-byte scale load per wave covers the step");
    static_assert(S >= 4 && 16 % S == 0, "stage count: at least 4, dividing the step count (a multiple of 16)");
    static_assert((S - 2) * kVPW <= 63, "vmcnt range");
    static_assert(MI >= 1 && NI >= 1 && BK % (WK * 32) == 0 && BT % (WT * 32) == 0, "tile shape");
    static_assert(kSmemBytes <= 160 * 1024, "LDS of one compute unit");
};

struct Fp8Frag {
    int8v hi, lo;
    int sc;  // byte 0: E_hi + 127, byte 1: E_lo + 127
};

// rec: the record's codes in LDS; sc: its 128 scale bytes
__device__ inline Fp8Frag read_frag(const uint8_t* rec, const uint8_t* sc, int lane) {
    Fp8Frag f;
    const int4v h0 = *(const int4v*)(rec + lane * 16);
    const int4v h1 = *(const int4v*)(rec + 1024 + lane * 16);
    const int4v l0 = *(const int4v*)(rec + 2048 + lane * 16);
    const int4v l1 = *(const int4v*)(rec + 3072 + lane * 16);
    f.hi = __builtin_shufflevector(h0, h1, 0, 1, 2, 3, 4, 5, 6, 7);
    f.lo = __builtin_shufflevector(l0, l1, 0, 1, 2, 3, 4, 5, 6, 7);
    f.sc = *(const uint16_t*)(sc + lane * 2);
    return f;
}

// One global_load_lds_dword: lane i's 4 bytes at gptr land in LDS at lds_byte_addr + 4 i.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ inline void lds_dma4(const void* gptr, uint32_t lds_byte_addr) {
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off"
                 :
                 : "s"(lds_byte_addr), "v"(gptr)
                 : "memory", "m0");
}
#pragma clang diagnostic pop

}  // namespace

template <int BK, int BT, int WK, int WT, int S>
__device__ inline void gemm_fp8_body(const mykernelParamType& prm, int kblk, int gblk, uint8_t* smem) {
    using Sh = Fp8Shape<BK, BT, WK, WT, S>;
    constexpr int MI = Sh::MI, NI = Sh::NI, UF = Sh::UF, VF = Sh::VF, CPW = Sh::kCPW, VPW = Sh::kVPW;
    constexpr bool DB = Sh::kDB;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wk = wave % WK;
    const int wt = wave / WK;
    const int KB = prm.KB32, G32 = prm.G32, CQ = fp8_cq(prm.c);
    const int gb0 = gblk * VF;
    const int kb0 = kblk * UF;
    const int nsteps = 16 * CQ;  // (position, 64-channel block) pairs
    const uint8_t* __restrict__ U8 = (const uint8_t*)prm.transform_filter + fp8_u8_offset(prm.CB, KB);
    const uint8_t* __restrict__ V8 = (const uint8_t*)prm.transform_pin;
    const uint8_t* __restrict__ U8s = U8 + fp8_scale_offset(prm.c, KB);
    const uint8_t* __restrict__ V8s = V8 + fp8_scale_offset(prm.c, G32);

    // code units of this wave: unit u = wave * CPW + j = record u / 4, 1 KiB part u % 4; record
    // (pos, cq, rb) of step s = pos * CQ + cq is record s * RB + rb, so addresses advance by RB
    // records per step; a record beyond the matrix reads record 0 (its rows or columns are never
    // stored)
    const uint8_t* csrc[CPW];
    size_t cstride[CPW];
#pragma unroll
    for (int j = 0; j < CPW; ++j) {
        const int u = wave * CPW + j;
        const int r = u >> 2, part = u & 3;
        if (r < UF) {
            const bool ok = kb0 + r < KB;
            csrc[j] = U8 + (ok ? (size_t)(kb0 + r) : 0) * kFp8CodeBytes + part * 1024 + lane * 16;
            cstride[j] = (size_t)KB * kFp8CodeBytes;
        } else {
            const bool ok = gb0 + r - UF < G32;
            csrc[j] = V8 + (ok ? (size_t)(gb0 + r - UF) : 0) * kFp8CodeBytes + part * 1024 + lane * 16;
            cstride[j] = (size_t)G32 * kFp8CodeBytes;
        }
    }
    // scale word d = wave * 64 + lane of the step's list (record d / 32, word d % 32); words past
    // the step's records re-read record 0 of U (their LDS slots are never read)
    const int sd = wave * 64 + lane, sr = sd >> 5, sw = sd & 31;
    const uint8_t* ssrc;
    size_t sstride;
    if (sr < UF || sr >= UF + VF) {
        const bool ok = sr < UF && kb0 + sr < KB;
        ssrc = U8s + (ok ? (size_t)(kb0 + sr) : 0) * kFp8ScaleBytes + sw * 4;
        sstride = (size_t)KB * kFp8ScaleBytes;
    } else {
        const bool ok = gb0 + sr - UF < G32;
        ssrc = V8s + (ok ? (size_t)(gb0 + sr - UF) : 0) * kFp8ScaleBytes + sw * 4;
        sstride = (size_t)G32 * kFp8ScaleBytes;
    }
    const uint32_t lds0 = lds_address((const half_t*)smem);
    auto dma = [&](int step, int stage) {
        const size_t s = (size_t)min(step, nsteps - 1);
        const uint32_t sb = lds0 + (uint32_t)(stage * Sh::kStageBytes);
#pragma unroll
        for (int j = 0; j < CPW; ++j)
            lds_dma16((const half_t*)(csrc[j] + s * cstride[j]),
                      __builtin_amdgcn_readfirstlane(sb + (uint32_t)((wave * CPW + j) * 1024)));
        lds_dma4(ssrc + s * sstride, __builtin_amdgcn_readfirstlane(sb + (uint32_t)(Sh::kScaleOfs + wave * 256)));
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

    Fp8Frag ra[DB ? 2 : 1][MI], rb[DB ? 2 : 1][NI];
    auto lread = [&](Fp8Frag (&A)[MI], Fp8Frag (&B)[NI], int stage) {
        const uint8_t* sb = smem + stage * Sh::kStageBytes;
        const uint8_t* ss = sb + Sh::kScaleOfs;
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
            const int r = wk * MI + mi;
            A[mi] = read_frag(sb + r * kFp8CodeBytes, ss + r * kFp8ScaleBytes, lane);
        }
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
            const int r = UF + wt * NI + ni;
            B[ni] = read_frag(sb + r * kFp8CodeBytes, ss + r * kFp8ScaleBytes, lane);
        }
    };

    // Pipeline (stage = step % S):
    //   prologue: DMA steps 0..S-2; wait for step 0; barrier; (DB) regs <- stage 0
    //   iteration s: wait for steps <= s+1 (S-3 steps stay in flight) | barrier (stage s+1
    //   complete for every wave; every wave's reads of stage s-1 done) | DMA step s+S-1 into
    //   stage s-1 | DB: regs(s+1) <- stage s+1, MFMAs on regs(s); else regs <- stage s, MFMAs |
    //   fold at a position's end
#pragma unroll
    for (int p = 0; p < S - 1; ++p) dma(p, p);
    wait_vm<(S - 2) * VPW>();
    __syncthreads();
    if constexpr (DB) lread(ra[0], rb[0], 0);

    int pos = 0, cq = 0;
    for (int s0 = 0; s0 < nsteps; s0 += S) {
#pragma unroll
        for (int d = 0; d < S; ++d) {
            const int s = s0 + d;
            const int cur = DB ? (d & 1) : 0;  // S is even, so s & 1 == d & 1
            wait_vm<(S - 3) * VPW>();
            __syncthreads();
            dma(s + S - 1, (d + S - 1) % S);
            if constexpr (DB) lread(ra[cur ^ 1], rb[cur ^ 1], (d + 1) % S);
            else lread(ra[0], rb[0], d);
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) {
                    const Fp8Frag& a = ra[cur][mi];
                    const Fp8Frag& b = rb[cur][ni];
                    floatx16 c = acc[mi][ni];
                    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a.hi, b.hi, c, 0, 0, 0, a.sc, 0, b.sc);
                    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a.hi, b.lo, c, 0, 0, 0, a.sc, 1, b.sc);
                    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a.lo, b.hi, c, 0, 0, 1, a.sc, 0, b.sc);
                    acc[mi][ni] = c;
                }
            if (++cq == CQ) {
                fold_position(pos, acc, Y);
#pragma unroll
                for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                        for (int e = 0; e < 16; ++e) acc[mi][ni][e] = 0.f;
                cq = 0;
                ++pos;
            }
        }
    }
    wait_vm<0>();  // the trailing (re-read) DMAs land before the workgroup's LDS is released

    store_tiles32<MI, NI>(prm, Y, gb0 + wt * NI, kb0 + wk * MI, lane);
}

template <int BK, int BT, int WK, int WT>
constexpr int kFp8WavesPerEU = WK * WT >= 8 ? 2 : 1;

template <int BK, int BT, int WK, int WT, int S>
__global__ void __launch_bounds__(WK * WT * 64) __attribute__((amdgpu_waves_per_eu(kFp8WavesPerEU<BK, BT, WK, WT>)))
wino_gemm_fp8_kernel(mykernelParamType prm) {
    using Sh = Fp8Shape<BK, BT, WK, WT, S>;
    __shared__ __attribute__((aligned(16))) uint8_t smem[Sh::kSmemBytes];
    const int nk = (prm.KB32 + Sh::UF - 1) / Sh::UF;
    const int ng = (prm.G32 + Sh::VF - 1) / Sh::VF;
    int kblk, gblk;
    xcd_remap(blockIdx.x, nk * ng, nk, &kblk, &gblk);
    gemm_fp8_body<BK, BT, WK, WT, S>(prm, kblk, gblk, smem);
}

#define FP8_CFG(BK, BT, WK, WT, S) \
    { BK, BT, WK * WT * 64, S, (const void*)&wino_gemm_fp8_kernel<BK, BT, WK, WT, S>, nullptr }

const GemmConfig kFp8Configs[] = {
    FP8_CFG(128, 128, 2, 4, 4),  // 0: wave 64x32, 8 waves
    FP8_CFG(128, 64, 2, 2, 4),   // 1: wave 64x32, 4 waves
    FP8_CFG(64, 128, 1, 4, 4),   // 2: wave 64x32, 4 waves
    FP8_CFG(32, 128, 1, 4, 4),   // 3: wave 32x32, 4 waves (small K)
    FP8_CFG(64, 64, 1, 2, 4),    // 4: wave 64x32, 2 waves (small problems)
    FP8_CFG(32, 64, 1, 2, 4),    // 5: wave 32x32, 2 waves (small problems)
};
const int kNumFp8Configs = (int)(sizeof(kFp8Configs) / sizeof(kFp8Configs[0]));
