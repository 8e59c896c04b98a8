// FP8 (OCP e4m3fn) compensated form of the GEMM stage.
//
// No counterpart in the reference (it extends the method). Operands are split into a high
// and a low FP8 part, each with an E8M0 scale per (Winograd position, row or column,
// 32-channel block) -- the microscaling granularity v_mfma_scale_f32_16x16x128_f8f6f4
// applies in hardware -- and the product is recovered in three FP8 MFMA passes accumulated
// in FP32:  U.V ~ U_hi.V_hi + U_hi.V_lo + U_lo.V_hi  (the dropped U_lo.V_lo term is ~2^-8 of
// the block scale squared). Two passes (one-sided compensation) measurably miss a 1e-2
// error bound; three meet it. The transforms stay FP16/FP32: U is quantised from the FP16
// transformed filter, V from the FP32 transform before any FP16 rounding.
//
// Packed layouts (bytes): fragment (pos, cq, blk) = 64 lanes x 32 bytes, lane =
// (c % 128 / 32) * 16 + row % 16, byte = c % 32 (the 16x16x128 operand layout: a lane holds
// 32 consecutive channels of one row); scale block (pos, cq, blk) = 64 E8M0 bytes, one per
// lane. Workspace = hi codes | lo codes | hi scales | lo scales.
#include "wino_common.h"
#include "wino_fp8.h"

typedef int int8v __attribute__((ext_vector_type(8)));

namespace {

constexpr int kAT8[2][4] = {{1, 1, 1, 0}, {0, 1, -1, -1}};

// Smallest E with m <= 448 * 2^E (m >= 0), clamped to [-127, 127]; E = -127 for m = 0.
__device__ inline int block_exponent(float m) {
    if (!(m > 0.f)) return -127;
    int x;
    const float f = frexpf(m, &x);  // m = f * 2^x, f in [0.5, 1)
    int E = x - 9 + (f > 0.875f ? 1 : 0);
    return E < -127 ? -127 : (E > 127 ? 127 : E);
}

// Packs 4 floats (already scaled into [-448, 448]) into 4 e4m3 bytes (RNE).
__device__ inline uint32_t pack4_e4m3(float a, float b, float c, float d) {
    uint32_t w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    return __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
}

// Value of an e4m3fn byte as float.
__device__ inline float e4m3_value(uint32_t code) {
    const uint32_t s = code & 0x80u, e = (code >> 3) & 0xFu, m = code & 7u;
    float v = e ? ldexpf((float)(8u + m), (int)e - 10) : ldexpf((float)m, -9);
    return s ? -v : v;
}

__device__ inline float scale_of(int E) { return ldexpf(1.0f, E); }

// Quantises n <= 32 values held by a lane group into hi/lo codes given the group's block
// exponents. vals: this lane's count values; out bytes via pack4.
template <int N>
__device__ inline void quantize_lane(const float (&v)[N], int Ehi, int Elo, uint32_t (&hi)[N / 4],
                                     uint32_t (&lo)[N / 4]) {
    const float ih = scale_of(-Ehi), il = scale_of(-Elo), sh = scale_of(Ehi);
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        float h[4], r[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = v[q * 4 + e] * ih;
        hi[q] = pack4_e4m3(h[0], h[1], h[2], h[3]);
#pragma unroll
        for (int e = 0; e < 4; ++e) r[e] = (v[q * 4 + e] - e4m3_value((hi[q] >> (8 * e)) & 0xFFu) * sh) * il;
        lo[q] = pack4_e4m3(r[0], r[1], r[2], r[3]);
    }
}

template <int N>
__device__ inline float max_abs(const float (&v)[N]) {
    float m = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) m = fmaxf(m, fabsf(v[i]));
    return m;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// U (FP16 packed) -> U8 hi/lo + scales. One lane per (pos, cq, kb, lane): 32 channels of
// one row k, gathered from the four FP16 lane slots that hold them.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) wino_filter_fp8(mykernelParamType prm, uint8_t* __restrict__ ws) {
    const int KB = prm.KB, CB = prm.CB, CQ = (prm.c + 127) / 128;
    const size_t nfrag = (size_t)16 * CQ * KB;
    const size_t gt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= nfrag * 64) return;
    const int lane = (int)(gt & 63);
    const size_t frag = gt >> 6;
    const int kb = (int)(frag % KB);
    const int cq = (int)((frag / KB) % CQ);
    const int pos = (int)(frag / ((size_t)KB * CQ));
    const int cb32 = cq * 4 + (lane >> 4);
    const int r = lane & 15;
    float v[32];
    if (cb32 < CB) {
        const half_t* U = (const half_t*)prm.transform_filter + frag_offset(pos, cb32, kb, CB, KB);
#pragma unroll
        for (int cg = 0; cg < 4; ++cg) {
            const half8 h = *(const half8*)(U + (cg * 16 + r) * 8);
#pragma unroll
            for (int e = 0; e < 8; ++e) v[cg * 8 + e] = (float)h[e];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) v[i] = 0.f;
    }
    const int Ehi = block_exponent(max_abs(v));
    uint32_t hi[8], lo[8];
    // residual maximum needs the hi codes first
    {
        const float ih = scale_of(-Ehi), sh = scale_of(Ehi);
        float rmax = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            hi[q] = pack4_e4m3(v[q * 4] * ih, v[q * 4 + 1] * ih, v[q * 4 + 2] * ih, v[q * 4 + 3] * ih);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                rmax = fmaxf(rmax, fabsf(v[q * 4 + e] - e4m3_value((hi[q] >> (8 * e)) & 0xFFu) * sh));
        }
        const int Elo = block_exponent(rmax);
        quantize_lane<32>(v, Ehi, Elo, hi, lo);
        const size_t codes = nfrag * 2048;
        uint8_t* dst = ws + frag * 2048 + (size_t)lane * 32;
        *(u32x4*)dst = u32x4{hi[0], hi[1], hi[2], hi[3]};
        *(u32x4*)(dst + 16) = u32x4{hi[4], hi[5], hi[6], hi[7]};
        *(u32x4*)(dst + codes) = u32x4{lo[0], lo[1], lo[2], lo[3]};
        *(u32x4*)(dst + codes + 16) = u32x4{lo[4], lo[5], lo[6], lo[7]};
        uint8_t* sc = ws + 2 * codes + frag * 64 + lane;
        sc[0] = (uint8_t)(Ehi + 127);
        sc[nfrag * 64] = (uint8_t)(Elo + 127);
    }
}

// ---------------------------------------------------------------------------------------
// Input transform straight to V8 hi/lo + scales. Same staging as the FP16 transform; a lane
// computes 8 channels x 16 positions of one tile in FP32, the four lanes holding one tile's
// 32-channel block agree on the block exponents by cross-lane max, then each writes its
// 8 bytes of the 32-byte lane slot.
// ---------------------------------------------------------------------------------------
__device__ void input_transform_fp32_tile(const half_t* lds, int LR, int LWp, int lr0, int lc0, int ch,
                                          float v[16]);

__global__ void __launch_bounds__(256) wino_input_transform_fp8(mykernelParamType prm, uint8_t* __restrict__ ws) {
    extern __shared__ __attribute__((aligned(16))) half_t lds[];
    const int n = blockIdx.z, cb = blockIdx.y, bi = blockIdx.x;
    stage_input_band(prm, n, cb, bi, lds);  // zero + load, ends with a barrier
    const int TH = prm.band;
    const int th0 = bi * TH;
    const int tiles_w = prm.tiles_w;
    const int th1 = min(th0 + TH, prm.tiles_h);
    const int LR = input_lds_rows(TH);
    const int LWp = input_lds_pitch(tiles_w, prm.w, prm.q);
    const int T = prm.T, NT = prm.NT, G16 = prm.G16, CQ = (prm.c + 127) / 128;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int cg = lane >> 4;
    const int gs = n * T + th0 * tiles_w;
    const int ge = n * T + th1 * tiles_w;
    const size_t nfrag = (size_t)16 * CQ * G16;
    const size_t codes = nfrag * 2048;
    const int cq = cb >> 2, slot_lane = (cb & 3) * 16;
    for (int grp = (gs >> 4) + wave; grp <= ((ge - 1) >> 4); grp += nwaves) {
        const int g = grp * 16 + (lane & 15);
        const bool valid = g >= gs && g < ge;
        const bool tail = (ge == NT) && g >= NT && g < G16 * 16;
        float v[16][8];
        if (valid) {
            const int t = g - n * T;
            const int th = t / tiles_w;
            const int tw = t - th * tiles_w;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float vv[16];
                input_transform_fp32_tile(lds, LR, LWp, 2 * (th - th0), 2 * tw + 8, cg * 8 + e, vv);
#pragma unroll
                for (int pos = 0; pos < 16; ++pos) v[pos][e] = vv[pos];
            }
        } else {
#pragma unroll
            for (int pos = 0; pos < 16; ++pos)
#pragma unroll
                for (int e = 0; e < 8; ++e) v[pos][e] = 0.f;
        }
        const bool store = valid || tail;
#pragma unroll
        for (int pos = 0; pos < 16; ++pos) {
            float m = max_abs(v[pos]);
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            const int Ehi = block_exponent(m);
            const float ih = scale_of(-Ehi), sh = scale_of(Ehi);
            uint32_t hi[2], lo[2];
            float rm = 0.f;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                hi[q] = pack4_e4m3(v[pos][q * 4] * ih, v[pos][q * 4 + 1] * ih, v[pos][q * 4 + 2] * ih,
                                   v[pos][q * 4 + 3] * ih);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    rm = fmaxf(rm, fabsf(v[pos][q * 4 + e] - e4m3_value((hi[q] >> (8 * e)) & 0xFFu) * sh));
            }
            rm = fmaxf(rm, __shfl_xor(rm, 16));
            rm = fmaxf(rm, __shfl_xor(rm, 32));
            const int Elo = block_exponent(rm);
            const float il = scale_of(-Elo);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                float r[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    r[e] = (v[pos][q * 4 + e] - e4m3_value((hi[q] >> (8 * e)) & 0xFFu) * sh) * il;
                lo[q] = pack4_e4m3(r[0], r[1], r[2], r[3]);
            }
            if (store) {
                const size_t frag = ((size_t)pos * CQ + cq) * G16 + grp;
                const int L = slot_lane + (g & 15);
                uint8_t* dst = ws + frag * 2048 + (size_t)L * 32 + cg * 8;
                *(uint2*)dst = make_uint2(hi[0], hi[1]);
                *(uint2*)(dst + codes) = make_uint2(lo[0], lo[1]);
                if (cg == 0) {
                    uint8_t* sc = ws + 2 * codes + frag * 64 + L;
                    sc[0] = (uint8_t)(Ehi + 127);
                    sc[nfrag * 64] = (uint8_t)(Elo + 127);
                }
            }
        }
    }
}

// Zero fill of the V8 fragments of 32-channel blocks that lie entirely beyond C in the last
// 128-channel block (codes 0, scale bytes 0 = 2^-127): grid.y covers them too, so every
// byte of the workspace is written.

// ---------------------------------------------------------------------------------------
// Three-pass FP8 GEMM with the fused output transform. Workgroup tile BK x BT, waves
// WK x WT, 16x16 sub-tiles; one pipeline step = one 128-channel block of one position.
// LDS stage: U hi | U lo | V hi | V lo codes, then U hi | U lo | V hi | V lo scales.
// ---------------------------------------------------------------------------------------
template <int BK, int BT, int WK, int WT>
__global__ void __launch_bounds__(WK * WT * 64) wino_gemm_fp8_kernel(mykernelParamType prm) {
    constexpr int kThreads = WK * WT * 64;
    constexpr int MI = BK / WK / 16, NI = BT / WT / 16, UF = BK / 16, VF = BT / 16;
    constexpr int kCodeBytes = 2 * (UF + VF) * 2048;
    constexpr int kScaleBytes = 2 * (UF + VF) * 64;
    constexpr int kStage = kCodeBytes + kScaleBytes;
    constexpr int kCodePieces = kCodeBytes / 16, kPieces = kStage / 16;
    constexpr int kLoads = (kPieces + kThreads - 1) / kThreads;
    __shared__ __attribute__((aligned(16))) uint8_t smem[2 * kStage];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wk = wave % WK, wt = wave / WK;
    const int gb0 = blockIdx.x * VF, kb0 = blockIdx.y * UF;
    const int KB = prm.KB, G16 = prm.G16, CQ = (prm.c + 127) / 128;
    const int nsteps = 16 * CQ;
    const uint8_t* U8 = (const uint8_t*)prm.transform_filter;
    const uint8_t* V8 = (const uint8_t*)prm.transform_pin;
    const size_t ufr = (size_t)16 * CQ * KB, vfr = (size_t)16 * CQ * G16;

    // copy slots: piece p of a stage -> (global base, per-step stride, validity)
    const uint8_t* src[kLoads];
    size_t sstride[kLoads];
    bool sval[kLoads];
#pragma unroll
    for (int i = 0; i < kLoads; ++i) {
        const int pc = tid + i * kThreads;
        src[i] = U8;
        sstride[i] = 0;
        sval[i] = false;
        if (pc < kCodePieces) {
            const int part = pc / (128 * (UF + VF));  // 0 hi, 1 lo
            const int rem = pc % (128 * (UF + VF));
            const int f = rem / 128, within = rem % 128;
            if (f < UF) {
                src[i] = U8 + part * ufr * 2048 + ((size_t)kb0 + f) * 2048 + within * 16;
                sstride[i] = (size_t)KB * 2048;
                sval[i] = kb0 + f < KB;
            } else {
                src[i] = V8 + part * vfr * 2048 + ((size_t)gb0 + f - UF) * 2048 + within * 16;
                sstride[i] = (size_t)G16 * 2048;
                sval[i] = gb0 + f - UF < G16;
            }
        } else if (pc < kPieces) {
            const int sp = pc - kCodePieces;  // 4 pieces (64 B) per fragment scale block
            const int part = sp / (4 * (UF + VF));
            const int rem = sp % (4 * (UF + VF));
            const int f = rem / 4, within = rem % 4;
            if (f < UF) {
                src[i] = U8 + 2 * ufr * 2048 + part * ufr * 64 + ((size_t)kb0 + f) * 64 + within * 16;
                sstride[i] = (size_t)KB * 64;
                sval[i] = kb0 + f < KB;
            } else {
                src[i] = V8 + 2 * vfr * 2048 + part * vfr * 64 + ((size_t)gb0 + f - UF) * 64 + within * 16;
                sstride[i] = (size_t)G16 * 64;
                sval[i] = gb0 + f - UF < G16;
            }
        }
    }
    u32x4 pre[kLoads];
    auto gload = [&](int step) {
#pragma unroll
        for (int i = 0; i < kLoads; ++i)
            if (sval[i]) pre[i] = *(const u32x4*)(src[i] + (size_t)step * sstride[i]);
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < kLoads; ++i) {
            const int pc = tid + i * kThreads;
            if (sval[i]) *(u32x4*)(smem + buf * kStage + pc * 16) = pre[i];
        }
    };

    floatx4 acc[MI][NI], Y[4][MI][NI];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) Y[q][mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};

    gload(0);
    lstore(0);
    __syncthreads();
#pragma unroll 1
    for (int pos = 0; pos < 16; ++pos) {
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int cq = 0; cq < CQ; ++cq) {
            const int step = pos * CQ + cq;
            const bool has_next = step + 1 < nsteps;
            if (has_next) gload(step + 1);
            const uint8_t* sb = smem + (step & 1) * kStage;
            const uint8_t* ush = sb;                       // U hi codes
            const uint8_t* usl = sb + UF * 2048;           // U lo codes
            const uint8_t* vsh = sb + 2 * UF * 2048;       // V hi codes
            const uint8_t* vsl = vsh + VF * 2048;          // V lo codes
            const uint8_t* sc = sb + kCodeBytes;           // scales: U hi | V hi | U lo | V lo
            int8v bh[NI], bl[NI];
            int sbh[NI], sbl[NI];
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) {
                const int f = wt * NI + ni;
                bh[ni] = *(const int8v*)(vsh + f * 2048 + lane * 32);
                bl[ni] = *(const int8v*)(vsl + f * 2048 + lane * 32);
                sbh[ni] = sc[(UF + f) * 64 + lane];
                sbl[ni] = sc[(UF + VF) * 64 + (UF + f) * 64 + lane];
            }
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) {
                const int f = wk * MI + mi;
                const int8v ah = *(const int8v*)(ush + f * 2048 + lane * 32);
                const int8v al = *(const int8v*)(usl + f * 2048 + lane * 32);
                const int sah = sc[f * 64 + lane];
                const int sal = sc[(UF + VF) * 64 + f * 64 + lane];
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) {
                    floatx4 c = acc[mi][ni];
                    c = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(ah, bh[ni], c, 0, 0, 0, sah, 0, sbh[ni]);
                    c = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(ah, bl[ni], c, 0, 0, 0, sah, 0, sbl[ni]);
                    c = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(al, bh[ni], c, 0, 0, 0, sal, 0, sbh[ni]);
                    acc[mi][ni] = c;
                }
            }
            if (has_next) lstore((step + 1) & 1);
            __syncthreads();
        }
        const int pa = pos >> 2, pb = pos & 3;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int coef = kAT8[i][pa] * kAT8[j][pb];
                if (coef == 0) continue;
                const float cf = (float)coef;
#pragma unroll
                for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                    for (int ni = 0; ni < NI; ++ni) Y[i * 2 + j][mi][ni] += cf * acc[mi][ni];
            }
    }
    store_tiles_fp8<MI, NI>(prm, Y, gb0 + wt * NI, kb0 + wk * MI, lane);
}

#define FP8_CFG(BK, BT, WK, WT) \
    { BK, BT, WK * WT * 64, 0, (const void*)&wino_gemm_fp8_kernel<BK, BT, WK, WT> }

const GemmConfig kFp8Configs[] = {
    FP8_CFG(64, 64, 2, 2),    // 0: wave 32x32, 4 waves
    FP8_CFG(128, 64, 2, 2),   // 1: wave 64x32, 4 waves
    FP8_CFG(64, 128, 2, 4),   // 2: wave 32x32, 8 waves
};
const int kNumFp8Configs = (int)(sizeof(kFp8Configs) / sizeof(kFp8Configs[0]));

size_t fp8_u_bytes(int c, int k) {
    const size_t CQ = (size_t)(c + 127) / 128, KB = (size_t)(k + 15) / 16;
    return 16 * CQ * KB * (2 * 2048 + 2 * 64);
}
size_t fp8_v_bytes(int c, long long G16) {
    const size_t CQ = (size_t)(c + 127) / 128;
    return 16 * CQ * (size_t)G16 * (2 * 2048 + 2 * 64);
}
