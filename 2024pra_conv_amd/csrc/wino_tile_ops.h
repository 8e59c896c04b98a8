// Device helpers shared by the FP16 and FP8 kernels (header-only: no relocatable device
// code is needed across translation units).
#pragma once

#include "wino_common.h"


// ---- input transform pieces --------------------------------------------------------------

// V = B^T d B for one 4x4 window, FP32, fixed operation order (restated by the oracle).
//   B^T = [[1,0,-1,0],[0,1,1,0],[0,-1,1,0],[0,1,0,-1]]  (reference preprocessing.cuh:50-54)
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

// The 4x4 window of staged channel `ch` whose top-left LDS element is (lr0, lc0), as FP32.
// lc0 is even, so each window row is two 4-byte LDS reads.
__device__ inline void load_window(const half_t* lds, int LR, int LWp, int lr0, int lc0, int ch, float d[16]) {
    const half_t* base = lds + (ch * LR + lr0) * LWp + lc0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const half2_t lo = *(const half2_t*)(base + i * LWp);
        const half2_t hi = *(const half2_t*)(base + i * LWp + 2);
        d[i * 4 + 0] = (float)lo[0];
        d[i * 4 + 1] = (float)lo[1];
        d[i * 4 + 2] = (float)hi[0];
        d[i * 4 + 3] = (float)hi[1];
    }
}

// Row staging for widths that are not a multiple of 8 halves: the band's (channel, row,
// VEC-half chunk) items are spread over all threads of the workgroup (consecutive threads
// take consecutive chunks of a row, and a channel's rows are contiguous in memory), so no
// lane idles on short rows. VEC halves per load: W % VEC == 0 and the input base is
// 2*VEC-byte aligned.
template <int VEC>
__device__ inline void stage_rows_flat(const half_t* in_n, int cb, int C, int H, int W, int ih0, int LR, int LWp,
                                       int qq, half_t* lds) {
    typedef uint16_t chunk_t __attribute__((ext_vector_type(VEC)));
    const int nck = W / VEC;
    const int per_ch = LR * nck;
    const int total = 32 * per_ch;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int ch = i / per_ch;
        const int rem = i - ch * per_ch;
        const int lr = rem / nck;
        const int j = rem - lr * nck;
        const int c = cb * 32 + ch;
        const int ih = ih0 + lr;
        if (c >= C || ih < 0 || ih >= H) continue;
        const chunk_t val = *(const chunk_t*)(in_n + ((size_t)c * H + ih) * W + j * VEC);
        uint16_t* dst = (uint16_t*)lds + (ch * LR + lr) * LWp + qq + 8 + j * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e) dst[e] = val[e];
    }
}

// Row staging of channel-contiguous (N,H,W,C) input: one element per thread and step, with
// consecutive threads on consecutive channels of one pixel (contiguous in memory); the LDS
// layout is the same [channel][row][column] as for N,C,H,W input.
__device__ inline void stage_rows_nhwc(const half_t* in_n, int cb, int C, int H, int W, int ih0, int LR, int LWp,
                                       int qq, half_t* lds) {
    const int total = LR * W * 32;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int ch = i & 31;
        const int rest = i >> 5;
        const int lr = rest / W;
        const int iw = rest - lr * W;
        const int c = cb * 32 + ch;
        const int ih = ih0 + lr;
        if (c >= C || ih < 0 || ih >= H) continue;
        lds[(ch * LR + lr) * LWp + qq + 8 + iw] = in_n[((size_t)ih * W + iw) * C + c];
    }
}

// Stages LR input rows from row ih0 (image n, 32-channel block cb) in LDS as
// [channel][row][column], input column iw at LDS column q + 8 + iw (pitch LWp): zero fill
// (padding halo, rows outside the image, channels >= C), then coalesced row loads. Ends with
// a barrier, so the staging is visible to every wave.
__device__ inline void stage_band(const mykernelParamType& prm, int n, int cb, int ih0, int LR, int LWp,
                                  half_t* lds) {
    const int H = prm.h, W = prm.w, C = prm.c;
    const int qq = prm.q;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = blockDim.x >> 6;
    {
        const int total16 = (32 * LR * LWp) >> 3;
        u32x4* l16 = (u32x4*)lds;
        const u32x4 z = {0u, 0u, 0u, 0u};
        for (int i = tid; i < total16; i += blockDim.x) l16[i] = z;
    }
    __syncthreads();
    const half_t* in_n = (const half_t*)prm.pin + (size_t)n * C * H * W;
    const int njobs = 32 * LR;
    if (prm.io & CONV2D_IO_IN_NHWC) {
        stage_rows_nhwc(in_n, cb, C, H, W, ih0, LR, LWp, qq, lds);
    } else if (prm.flags & 1) {
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
    } else if (prm.flags & 2) {
        stage_rows_flat<4>(in_n, cb, C, H, W, ih0, LR, LWp, qq, lds);
    } else if (prm.flags & 4) {
        stage_rows_flat<2>(in_n, cb, C, H, W, ih0, LR, LWp, qq, lds);
    } else {
        stage_rows_flat<1>(in_n, cb, C, H, W, ih0, LR, LWp, qq, lds);
    }
    __syncthreads();
}

// The input rows of 2x2-tile row band bi (tile rows bi*band ..): rows 2*th0 - p .. + 2*band + 1.
__device__ inline void stage_input_band(const mykernelParamType& prm, int n, int cb, int bi, half_t* lds) {
    const int th0 = bi * prm.band;
    stage_band(prm, n, cb, 2 * th0 - prm.p, input_lds_rows(prm.band), input_lds_pitch(prm.tiles_w, prm.w, prm.q),
               lds);
}

// ---- GEMM epilogue pieces -----------------------------------------------------------------

// Folds the FP32 accumulators of position pos into the four output accumulators:
// Y[i][j] += A^T[i][a] A^T[j][b] M(a,b), pos = 4a + b. Coefficients are wave-uniform values
// in {-1, 0, 1}; zero ones are skipped with a uniform branch.
// A^T[i][a] of F(2,3), A^T = [[1,1,1,0],[0,1,-1,-1]] (reference conv2d.cpp:18-22), computed
// arithmetically (scalar ALU): reading the table from memory would issue
// scalar loads whose s_waitcnt lgkmcnt(0) also waits for every LDS operation in flight.
__device__ inline int wino_at(int i, int a) { return i == 0 ? (a != 3 ? 1 : 0) : (a == 0 ? 0 : (a == 1 ? 1 : -1)); }

template <typename Acc, int MI, int NI>
__device__ inline void fold_position(int pos, const Acc (&acc)[MI][NI], Acc (&Y)[4][MI][NI]) {
    const int pa = pos >> 2, pb = pos & 3;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int coef = wino_at(i, pa) * wino_at(j, pb);
            if (coef == 0) continue;
            const float cf = (float)coef;
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) Y[i * 2 + j][mi][ni] += cf * acc[mi][ni];
        }
}

// Output position of GEMM column g: image, and top-left output row/column of its 2x2 tile.
struct TileCoord {
    int img, oh, ow;
    bool r1, c1;  // second row / column inside the image (ragged bottom / right edge)
};
__device__ inline TileCoord tile_coord(const mykernelParamType& prm, int g) {
    TileCoord tc;
    tc.img = g / prm.T;
    const int t = g - tc.img * prm.T;
    const int th = t / prm.tiles_w;
    const int tw = t - th * prm.tiles_w;
    tc.oh = 2 * th;
    tc.ow = 2 * tw;
    tc.r1 = tc.oh + 1 < prm.Oh;
    tc.c1 = tc.ow + 1 < prm.Ow;
    return tc;
}

// Output epilogue in FP32: y + bias[k], then the activation (1: ReLU; a NaN stays NaN).
// prm.bias and prm.act are kernel arguments, so the branches are wave-uniform.
__device__ inline float epilogue(const mykernelParamType& prm, float y, float b) {
    y += b;
    return (prm.act == 1 && y < 0.f) ? 0.f : y;
}
// Output and bias live in global memory: the address-space casts keep the accesses on the
// global (not flat) path also when prm was read from memory (multi-layer descriptors).
typedef __attribute__((address_space(1))) half_t global_half_t;

__device__ inline float epilogue_bias(const mykernelParamType& prm, int k) {
    return prm.bias ? (float)((const global_half_t*)prm.bias)[k] : 0.f;
}

typedef __attribute__((address_space(1))) uint16_t global_u16_t;

// One output value as the bits of its output type (FP16, or bfloat16 with CONV2D_IO_OUT_BF16),
// rounded once to nearest even from FP32.
__device__ inline uint16_t out_bits(const mykernelParamType& prm, float y) {
    return (prm.io & CONV2D_IO_OUT_BF16) ? __builtin_bit_cast(uint16_t, (__bf16)y)
                                         : __builtin_bit_cast(uint16_t, (half_t)y);
}

// Element offset of output (image, channel k, row oh, column ow) in the output layout.
__device__ inline size_t out_index(const mykernelParamType& prm, int img, int k, int oh, int ow) {
    return (prm.io & CONV2D_IO_OUT_NHWC) ? (((size_t)img * prm.Oh + oh) * prm.Ow + ow) * prm.k + k
                                         : (((size_t)img * prm.k + k) * prm.Oh + oh) * prm.Ow + ow;
}

__device__ inline void store_tile(const mykernelParamType& prm, const TileCoord& tc, int k, float y00, float y01,
                                  float y10, float y11) {
    if (prm.bias || prm.act) {
        const float b = epilogue_bias(prm, k);
        y00 = epilogue(prm, y00, b);
        y01 = epilogue(prm, y01, b);
        y10 = epilogue(prm, y10, b);
        y11 = epilogue(prm, y11, b);
    }
    const int Oh = prm.Oh, Ow = prm.Ow;
    const bool out_bf16 = (prm.io & CONV2D_IO_OUT_BF16) != 0;
    if ((prm.io & CONV2D_IO_OUT_NHWC) == 0 && (Ow & 1) == 0 && ((uintptr_t)prm.pout & 3) == 0) {
        // NCHW output whose rows all start 4-byte aligned (a tile's second column is then
        // inside): one 4-byte store per row
        typedef __attribute__((address_space(1))) uint32_t global_u32_t;
        global_u16_t* o = (global_u16_t*)prm.pout + (((size_t)tc.img * prm.k + k) * Oh + tc.oh) * Ow + tc.ow;
        *(global_u32_t*)o = pack_out2(out_bf16, y00, y01);
        if (tc.r1) *(global_u32_t*)(o + Ow) = pack_out2(out_bf16, y10, y11);
        return;
    }
    if (prm.io != 0) {
        // NHWC and/or bfloat16 output: one 2-byte store per element
        global_u16_t* o = (global_u16_t*)prm.pout;
        o[out_index(prm, tc.img, k, tc.oh, tc.ow)] = out_bits(prm, y00);
        if (tc.c1) o[out_index(prm, tc.img, k, tc.oh, tc.ow + 1)] = out_bits(prm, y01);
        if (tc.r1) {
            o[out_index(prm, tc.img, k, tc.oh + 1, tc.ow)] = out_bits(prm, y10);
            if (tc.c1) o[out_index(prm, tc.img, k, tc.oh + 1, tc.ow + 1)] = out_bits(prm, y11);
        }
        return;
    }
    global_half_t* o = (global_half_t*)prm.pout + (((size_t)tc.img * prm.k + k) * Oh + tc.oh) * Ow + tc.ow;
    o[0] = (half_t)y00;
    if (tc.c1) o[1] = (half_t)y01;
    if (tc.r1) {
        o[Ow] = (half_t)y10;
        if (tc.c1) o[Ow + 1] = (half_t)y11;
    }
}

// Epilogue for 32x32 MFMA accumulators: the 2x2 output tile of each (k, g) this lane
// holds, FP32 -> FP16, ragged edges guarded. gbw / kbw: the wave's first 32-column / 32-row
// block. v_mfma_f32_32x32x16 accumulator layout: element j of lane l is row
// 8*(j/4) + 4*(l/32) + j%4, column l%32.
template <int MI, int NI>
__device__ inline void store_tiles32(const mykernelParamType& prm, const floatx16 (&Y)[4][MI][NI], int gbw,
                                     int kbw, int lane) {
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
        const int g = (gbw + ni) * 32 + (lane & 31);
        if (g >= prm.NT) continue;
        const TileCoord tc = tile_coord(prm, g);
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int k = (kbw + mi) * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                if (k >= prm.k) continue;
                store_tile(prm, tc, k, Y[0][mi][ni][j], Y[1][mi][ni][j], Y[2][mi][ni][j], Y[3][mi][ni][j]);
            }
    }
}

// Epilogue of the direct (implicit-GEMM) form: GEMM column g is output pixel g of the batch
// (image g / T, pixel g % T with T = Oh*Ow), so each accumulator element is one output
// value, rounded once to FP16. Consecutive lanes hold consecutive pixels of one output
// channel, so the 2-byte stores of a row coalesce.
template <int MI, int NI>
__device__ inline void store_plain32(const mykernelParamType& prm, const floatx16 (&acc)[MI][NI], int gbw, int kbw,
                                     int lane) {
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
        const int g = (gbw + ni) * 32 + (lane & 31);
        if (g >= prm.NT) continue;
        const int img = g / prm.T;
        const int pix = g - img * prm.T;
        global_half_t* o = (global_half_t*)prm.pout + (size_t)img * prm.k * prm.T + pix;
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int k = (kbw + mi) * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                if (k >= prm.k) continue;
                float y = acc[mi][ni][j];
                if (prm.bias || prm.act) y = epilogue(prm, y, epilogue_bias(prm, k));
                if (prm.io != 0) {
                    // NHWC and/or bfloat16 output: pixel pix of the image is (pix / Ow, pix % Ow)
                    const size_t off = (prm.io & CONV2D_IO_OUT_NHWC) ? ((size_t)img * prm.T + pix) * prm.k + k
                                                                     : ((size_t)img * prm.k + k) * prm.T + pix;
                    ((global_u16_t*)prm.pout)[off] = out_bits(prm, y);
                    continue;
                }
                o[(size_t)k * prm.T] = (half_t)y;
            }
    }
}

// Epilogue of the F(4x4) per-position GEMM (position pos): M / 4 rounded to FP16 into
// mid_result as [pos][k / 4][column][k % 4] (rows padded to KB32*32, columns to G32*32). The
// four consecutive rows a lane holds per accumulator quarter are one 8-byte store, and lanes
// 0-31 (consecutive columns) write 256 contiguous bytes.
template <int MI, int NI>
__device__ inline void store_m16(const mykernelParamType& prm, const floatx16 (&acc)[MI][NI], int gbw, int kbw,
                                 int lane, int pos) {
    typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) u32x2_t global_u32x2_t;
    const size_t KQp = (size_t)prm.KB32 * 8, NTp = (size_t)prm.G32 * 32;
    global_u32x2_t* M = (global_u32x2_t*)prm.mid_result;
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
        if (gbw + ni >= prm.G32) continue;
        const size_t t = (size_t)(gbw + ni) * 32 + (lane & 31);
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
            if (kbw + mi >= prm.KB32) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const size_t kq = (size_t)(kbw + mi) * 8 + 2 * i + (lane >> 5);
                const u32x2_t pk = {pack_f16x2(0.25f * acc[mi][ni][4 * i], 0.25f * acc[mi][ni][4 * i + 1]),
                                    pack_f16x2(0.25f * acc[mi][ni][4 * i + 2], 0.25f * acc[mi][ni][4 * i + 3])};
                M[((size_t)pos * KQp + kq) * NTp + t] = pk;
            }
        }
    }
}

// ---- LDS DMA ------------------------------------------------------------------------------

// One global_load_lds_dwordx4: lane i's 16 bytes at gptr land in LDS at lds_byte_addr + 16 i
// (the address goes through m0, so it must be wave-uniform). No destination registers; the
// load counts in vmcnt until its bytes are in LDS.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ inline void lds_dma16(const half_t* gptr, uint32_t lds_byte_addr) {
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                 :
                 : "s"(lds_byte_addr), "v"(gptr)
                 : "memory", "m0");
}
#pragma clang diagnostic pop

// Waits until at most N vector-memory operations of this wave are outstanding.
template <int N>
__device__ inline void wait_vm() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

__device__ inline uint32_t lds_address(const half_t* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) half_t*)p;
}

// Hardware dispatches workgroup ids round-robin over the 8 XCDs; this remaps id L so every
// XCD gets a contiguous run of the (k-block fastest) work list, keeping the k-blocks that
// share one column block of V adjacent on one XCD (same L2).
__device__ inline void xcd_remap(int L, int total, int nk, int* kblk, int* gblk) {
    const int xcd = L & 7, j = L >> 3;
    const int base = total >> 3, rem = total & 7;
    const int m = xcd * base + min(xcd, rem) + j;
    *kblk = m % nk;
    *gblk = m / nk;
}

// Multi-layer (grouped) launches: the workgroup ids of each layer form one contiguous range
// (first id s3_first, s3_items ids), layers in decreasing per-item cost. Id L is remapped within
// its layer's range as above, so every XCD runs an equal share of every layer (a contiguous run
// of its k-block-fastest items) and the layers follow one another in launch order. (Remapped
// over the whole group instead, the first XCDs would get the costliest layers' items and the
// last ones the cheapest, while each XCD still runs the same number of workgroups.) Returns the
// layer; *local is the item within it.
__device__ inline const BatchLayer& grouped_layer_item(const BatchLayer* __restrict__ layers, int nlayers, int L,
                                                       int* local) {
    int lo = 0, hi = nlayers - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (layers[mid].s3_first <= L) lo = mid; else hi = mid - 1;
    }
    const BatchLayer& bl = layers[lo];
    int unused;
    xcd_remap(L - bl.s3_first, bl.s3_items, bl.s3_items, local, &unused);
    return bl;
}
