// Winograd F(2x2,3x3) transform stages for gfx950.
//
//   S1  wino_filter_transform : U = G g G^T  (reference preprocessing.cuh:75-270, GgGT*)
//   S2  wino_input_transform  : V = B^T d B  (reference preprocessing.cuh:6-72, 273-424, BTdB*)
//
// Both write the packed MFMA operand layouts described in mykernelParamType.h. All
// arithmetic is FP32 on FP16 inputs, rounded once to FP16 on store; the operation order is
// fixed (and restated by the oracle) so the results are reproducible bit for bit.
#include "wino_common.h"

// ---------------------------------------------------------------------------------------
// S1: filter transform.
//   G = [[1,0,0],[1/2,1/2,1/2],[1/2,-1/2,1/2],[0,0,1]]   (reference preprocessing.cuh:77-81)
// One lane produces one (k, 8-channel group) for all 16 positions, so each of its 16
// stores is one full 16-byte lane slot of a fragment and a wave stores whole 1 KiB
// fragments.
// ---------------------------------------------------------------------------------------
__device__ inline void filter_tile_fp32(const float g[9], float u[16]) {
    float t[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float g0 = g[0 * 3 + j], g1 = g[1 * 3 + j], g2 = g[2 * 3 + j];
        t[0][j] = g0;
        t[1][j] = ((g0 + g1) + g2) * 0.5f;
        t[2][j] = ((g0 - g1) + g2) * 0.5f;
        t[3][j] = g2;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float a0 = t[i][0], a1 = t[i][1], a2 = t[i][2];
        u[i * 4 + 0] = a0;
        u[i * 4 + 1] = ((a0 + a1) + a2) * 0.5f;
        u[i * 4 + 2] = ((a0 - a1) + a2) * 0.5f;
        u[i * 4 + 3] = a2;
    }
}

__global__ void __launch_bounds__(256) wino_filter_transform(mykernelParamType prm) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int frag = gtid >> 6;
    const int lane = gtid & 63;
    const int KB = prm.KB, CB = prm.CB;
    if (frag >= KB * CB) return;
    const int kb = frag % KB;
    const int cb = frag / KB;
    const int k = kb * 16 + (lane & 15);
    const int c0 = cb * 32 + (lane >> 4) * 8;
    const int K = prm.k, C = prm.c;

    half8 outv[16];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int c = c0 + e;
        float u[16];
        if (k < K && c < C) {
            const half_t* src = (const half_t*)prm.pweight + ((size_t)k * C + c) * 9;
            float g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = (float)src[i];
            filter_tile_fp32(g, u);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) u[i] = 0.0f;
        }
#pragma unroll
        for (int pos = 0; pos < 16; ++pos) outv[pos][e] = (half_t)u[pos];
    }
    half_t* U = (half_t*)prm.transform_filter;
#pragma unroll
    for (int pos = 0; pos < 16; ++pos) {
        half8* dst = (half8*)(U + frag_offset(pos, cb, kb, CB, KB) + lane * 8);
        *dst = outv[pos];
    }
}

// ---------------------------------------------------------------------------------------
// S2: input transform.
//   B^T = [[1,0,-1,0],[0,1,1,0],[0,-1,1,0],[0,1,0,-1]]  (reference preprocessing.cuh:50-54)
// Workgroup = (band of prm.band tile rows, 32-channel block, image). The band's input rows
// (with halo and zero padding, ragged right/bottom edges included) are staged in LDS
// with coalesced row loads; then each lane builds the 4x4 windows of one tile for 8
// channels and stores 16 full fragment lane slots. A wave covers 16 consecutive tiles x
// 32 channels, i.e. whole 1 KiB fragments per position.
// ---------------------------------------------------------------------------------------
__device__ inline void input_tile_fp32(const float d[16], float v[16]) {
    float s[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float d0 = d[0 * 4 + j], d1 = d[1 * 4 + j], d2 = d[2 * 4 + j], d3 = d[3 * 4 + j];
        s[0][j] = d0 - d2;
        s[1][j] = d1 + d2;
        s[2][j] = d2 - d1;
        s[3][j] = d1 - d3;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i * 4 + 0] = s[i][0] - s[i][2];
        v[i * 4 + 1] = s[i][1] + s[i][2];
        v[i * 4 + 2] = s[i][2] - s[i][1];
        v[i * 4 + 3] = s[i][1] - s[i][3];
    }
}

__global__ void __launch_bounds__(256) wino_input_transform(mykernelParamType prm) {
    extern __shared__ __attribute__((aligned(16))) half_t lds[];
    input_transform_block(prm, blockIdx.z, blockIdx.y, blockIdx.x, lds);
}

// One input-transform work item: tile-row band bi of image n, channel block cb. Called by
// the per-layer kernel above and by the multi-layer persistent kernel. `lds` must hold
// input_lds_bytes(prm.band, ...) bytes. Ends with no barrier pending; callers that reuse
// `lds` for another item must __syncthreads() first.
__device__ void input_transform_block(const mykernelParamType& prm, int n, int cb, int bi, half_t* lds) {
    const int TH = prm.band;
    const int th0 = bi * TH;
    const int tiles_w = prm.tiles_w, tiles_h = prm.tiles_h;
    const int th1 = min(th0 + TH, tiles_h);
    const int H = prm.h, W = prm.w, C = prm.c;
    const int pp = prm.p, qq = prm.q;
    const int LR = input_lds_rows(TH);
    const int LWp = input_lds_pitch(tiles_w, W, qq);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = blockDim.x >> 6;

    // 1. zero the staging area (padding halo, rows outside the image, channels >= C)
    {
        const int total16 = (32 * LR * LWp) >> 3;
        u32x4* l16 = (u32x4*)lds;
        const u32x4 z = {0u, 0u, 0u, 0u};
        for (int i = tid; i < total16; i += blockDim.x) l16[i] = z;
    }
    __syncthreads();

    // 2. stage the band's input rows
    const half_t* in_n = (const half_t*)prm.pin + (size_t)n * C * H * W;
    const int ih0 = 2 * th0 - pp;
    const int njobs = 32 * LR;
    if (prm.flags & 1) {
        const int rsub = lane >> 3, ck = lane & 7, nck = W >> 3;
        for (int job = wave * 8 + rsub; job < njobs; job += nwaves * 8) {
            const int ch = job / LR;
            const int lr = job - ch * LR;
            const int c = cb * 32 + ch;
            const int ih = ih0 + lr;
            if (c >= C || ih < 0 || ih >= H) continue;
            const half_t* src = in_n + ((size_t)c * H + ih) * W;
            half_t* dst = lds + (ch * LR + lr) * LWp + qq + 8;
            for (int j = ck; j < nck; j += 8) {
                const half8 val = *(const half8*)(src + j * 8);
#pragma unroll
                for (int e = 0; e < 8; ++e) dst[j * 8 + e] = val[e];
            }
        }
    } else {
        for (int job = wave; job < njobs; job += nwaves) {
            const int ch = job / LR;
            const int lr = job - ch * LR;
            const int c = cb * 32 + ch;
            const int ih = ih0 + lr;
            if (c >= C || ih < 0 || ih >= H) continue;
            const half_t* src = in_n + ((size_t)c * H + ih) * W;
            half_t* dst = lds + (ch * LR + lr) * LWp + qq + 8;
            for (int iw = lane; iw < W; iw += 64) dst[iw] = src[iw];
        }
    }
    __syncthreads();

    // 3. transform: lane -> (tile g = grp*16 + lane%16, channels cg*8 .. cg*8+7)
    const int T = prm.T, NT = prm.NT, G16 = prm.G16, CB = prm.CB;
    const int gs = n * T + th0 * tiles_w;
    const int ge = n * T + th1 * tiles_w;
    const int cg = lane >> 4;
    half_t* V = (half_t*)prm.transform_pin;
    for (int grp = (gs >> 4) + wave; grp <= ((ge - 1) >> 4); grp += nwaves) {
        const int g = grp * 16 + (lane & 15);
        const bool valid = g >= gs && g < ge;
        const bool tail = (ge == NT) && g >= NT && g < G16 * 16;
        if (!valid && !tail) continue;
        half8 outv[16];
        if (valid) {
            const int t = g - n * T;
            const int th = t / tiles_w;
            const int tw = t - th * tiles_w;
            const int lr0 = 2 * (th - th0);
            const int lc0 = 2 * tw + 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const half_t* base = lds + ((cg * 8 + e) * LR + lr0) * LWp + lc0;
                float d[16];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const half2_t lo = *(const half2_t*)(base + i * LWp);
                    const half2_t hi = *(const half2_t*)(base + i * LWp + 2);
                    d[i * 4 + 0] = (float)lo[0];
                    d[i * 4 + 1] = (float)lo[1];
                    d[i * 4 + 2] = (float)hi[0];
                    d[i * 4 + 3] = (float)hi[1];
                }
                float v[16];
                input_tile_fp32(d, v);
#pragma unroll
                for (int pos = 0; pos < 16; ++pos) outv[pos][e] = (half_t)v[pos];
            }
        } else {
#pragma unroll
            for (int pos = 0; pos < 16; ++pos)
#pragma unroll
                for (int e = 0; e < 8; ++e) outv[pos][e] = (half_t)0.0f;
        }
        const int slot = (cg * 16 + (lane & 15)) * 8;
#pragma unroll
        for (int pos = 0; pos < 16; ++pos) {
            half8* dst = (half8*)(V + frag_offset(pos, cb, grp, CB, G16) + slot);
            *dst = outv[pos];
        }
    }
}

// Slot-4 placeholder: the output transform runs inside the GEMM epilogue.
__global__ void wino_noop(mykernelParamType) {}

// ---------------------------------------------------------------------------------------
// Multi-layer persistent input transform: workgroups pull (layer, band, channel block,
// image) items from one queue until it is empty. Dynamic LDS = the largest layer's staging.
// ---------------------------------------------------------------------------------------
__global__ void wino_batch_reset(unsigned* counters) {
    if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(256) wino_batch_input_transform(const BatchLayer* __restrict__ layers,
                                                                  int nlayers, int total, unsigned* counter) {
    extern __shared__ __attribute__((aligned(16))) half_t lds[];
    __shared__ int slot;
    for (;;) {
        const int item = batch_next_item(counter, &slot);
        if (item >= total) break;
        const int L = batch_find_layer(layers, nlayers, item, false);
        const BatchLayer& bl = layers[L];
        const int local = item - bl.s2_first;
        const int CB = bl.prm.CB;
        const int bi = local % bl.nbands;
        const int cb = (local / bl.nbands) % CB;
        const int n = local / (bl.nbands * CB);
        input_transform_block(bl.prm, n, cb, bi, lds);
        __syncthreads();
    }
}
