// Host side of the plugin: geometry validation, tile-configuration choice, launch-slot
// filling and the additive C-ABI entry points.
//
// Reference counterpart: getParamsize / getkernelInfo (src/conv2d.cpp:9-14, 254-517), which
// pick one of six hard-coded kernel quadruples by (n, c) alone and never check the other
// dimensions. Here every dimension is validated, the Winograd geometry is derived once and
// appended to the kernel argument, and the GEMM tile configuration comes from the tuned
// table (see tuner.hip) or, failing that, from the shape.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "conv2d.h"
#include "mykernelParamType.h"
#include "wino_common.h"
#include "wino_fp8.h"
#include "wino_fused.h"
#include "wino_gemm.h"
#include "plugin_internal.h"

namespace wino {

// Input-transform band: the first of 4/2/1/8 tile rows (8/4/2/1 with prefer8) whose LDS
// staging fits in 64 KiB (several workgroups per CU); for wide images one tile row in up to
// 160 KiB (the whole LDS of a gfx950 compute unit, one workgroup per unit: w up to 630).
// Eight-row bands halve the staged halo of small images; measured on MI355X they speed up the
// FP16 transform when w % 4 == 0 (N=64 C=512 28x28: 63.8 -> 58.8 us; odd 27x27: 53 -> 72 us)
// and the FP8 transform at any width (28x28: 115 -> 91 us, 27x27: 112 -> 87 us).
static int input_band(const problem_t* pb, const Geometry& g, bool prefer8) {
    const int c4[4] = {4, 2, 1, 8}, c8[4] = {8, 4, 2, 1};
    const int* cands = prefer8 ? c8 : c4;
    for (int i = 0; i < 4; ++i) {
        const int b = cands[i];
        if (b > 1 && b > g.tiles_h) continue;
        // the same number of bands, evenly sized (14 tile rows: 7 + 7 rather than 8 + 6, 58 ->
        // 54 us for N=64 C=512 28x28), which never needs more LDS than b
        if (input_lds_bytes(b, g.tiles_w, (int)pb->w, (int)pb->q) <= 65536) {
            const int nb = (g.tiles_h + b - 1) / b;
            return (g.tiles_h + nb - 1) / nb;
        }
    }
    return input_lds_bytes(1, g.tiles_w, (int)pb->w, (int)pb->q) <= 163840 ? 1 : 0;
}

static int compute_units();

// Finish time of a list schedule on P identical compute units: n1 workgroups of duration d1,
// then n2 of duration d2, each started on the earliest free unit (units free at the same time
// are kept as one group, so the loop runs about twice per round).
static double list_schedule_time(long long n1, double d1, long long n2, double d2, long long P) {
    typedef std::pair<double, long long> Group;  // (free at, units)
    std::priority_queue<Group, std::vector<Group>, std::greater<Group>> free_at;
    free_at.push(Group(0.0, P));
    double end = 0.0;
    const long long n[2] = {n1, n2};
    const double d[2] = {d1, d2};
    for (int i = 0; i < 2; ++i) {
        long long left = n[i];
        while (left > 0) {
            Group g = free_at.top();
            free_at.pop();
            while (!free_at.empty() && free_at.top().first == g.first) {
                g.second += free_at.top().second;
                free_at.pop();
            }
            const long long take = std::min(g.second, left);
            left -= take;
            free_at.push(Group(g.first + d[i], take));
            end = std::max(end, g.first + d[i]);
            if (g.second > take) free_at.push(Group(g.first, g.second - take));
        }
    }
    return end;
}

int derive_geometry(const problem_t* pb, Geometry* g) { return derive_geometry_io(pb, CONV2D_IO_NCHW, g); }

int derive_geometry_io(const problem_t* pb, int io, Geometry* g) {
    if (!pb || !g) return CONV2D_ERR_NULL;
    memset(g, 0, sizeof(*g));
    if (io & ~(CONV2D_IO_IN_NHWC | CONV2D_IO_OUT_NHWC | CONV2D_IO_OUT_BF16)) return CONV2D_ERR_INVALID;
    g->io = io;
    if (pb->n == 0 || pb->c == 0 || pb->h == 0 || pb->w == 0 || pb->k == 0 || pb->r == 0 ||
        pb->s == 0 || pb->u == 0 || pb->v == 0)
        return CONV2D_ERR_INVALID;
    const long long hp = (long long)pb->h + 2LL * pb->p, wp = (long long)pb->w + 2LL * pb->q;
    if (hp < pb->r || wp < pb->s) return CONV2D_ERR_INVALID;
    const long long Oh = (hp - pb->r) / pb->u + 1;
    const long long Ow = (wp - pb->s) / pb->v + 1;
    const unsigned lim = 32767u;
    if (pb->n > lim || pb->c > lim || pb->h > lim || pb->w > lim || pb->k > lim || pb->r > lim ||
        pb->s > lim || pb->u > lim || pb->v > lim || pb->p > lim || pb->q > lim || Oh > lim ||
        Ow > lim)
        return CONV2D_ERR_OVERFLOW;
    g->Oh = (int)Oh;
    g->Ow = (int)Ow;
    g->winograd_ok = pb->r == 3 && pb->s == 3 && pb->u == 1 && pb->v == 1 && pb->p <= 2 &&
                     pb->q <= 2;
    g->tiles_h = (g->Oh + 1) / 2;
    g->tiles_w = (g->Ow + 1) / 2;
    g->T = (long long)g->tiles_h * g->tiles_w;
    g->NT = g->T * pb->n;
    if (g->NT > (1LL << 30)) return CONV2D_ERR_OVERFLOW;
    g->G32 = (g->NT + 31) / 32;
    g->CB = ((int)pb->c + 31) / 32;
    g->KB32 = ((int)pb->k + 31) / 32;
    g->v_bytes = (size_t)16 * 2 * g->CB * (size_t)g->G32 * kFragHalves * 2;
    g->u_bytes = (size_t)16 * 2 * g->CB * (size_t)g->KB32 * kFragHalves * 2;
    g->band = input_band(pb, *g, pb->w % 4 == 0);
    g->nsteps = 16 * g->CB;
    g->fused_ok = false;
    if (g->winograd_ok) {
        // fused form: the 64-tile block shape that pads the tile grid least (ties: the
        // squarer block, whose input patch is smaller)
        long long best = -1;
        for (int i = 0; i < 4; ++i) {
            const int bh = kFusedShapes[i][0], bw = kFusedShapes[i][1];
            const long long padded = (long long)((g->tiles_h + bh - 1) / bh) * bh * ((g->tiles_w + bw - 1) / bw) * bw;
            if (best < 0 || padded < best) {
                best = padded;
                g->fbh = bh;
                g->fbw = bw;
            }
        }
        g->fnbh = (g->tiles_h + g->fbh - 1) / g->fbh;
        g->fnbw = (g->tiles_w + g->fbw - 1) / g->fbw;
        // staging load width: 8-byte pixel quads when the width and base allow them and the
        // quad patch fits one unit per thread, else 4-byte aligned pairs, else 4-byte pairs at
        // 2-byte aligned offsets (odd width or a base that is not 4-byte aligned)
        const uintptr_t in_addr = (uintptr_t)pb->in;
        g->fpr = fused_patch_rows(g->fbh);
        g->flw = (pb->w % 2) != 0 || (in_addr & 3) != 0 ? kFusedLoad2 : kFusedLoad4;
        if (g->flw == kFusedLoad4 && pb->w % 4 == 0 && (in_addr & 7) == 0) {
            const int pc8 = fused_patch_cols(kFusedLoad8, g->fbw, fused_col_shift(kFusedLoad8, pb->q));
            if (4 * g->fpr * (pc8 / 4) <= kFusedThreads) g->flw = kFusedLoad8;
        }
        // channel-contiguous input: 16-byte loads of 8 channels per pixel (whole channel groups
        // and 16-byte aligned pixels needed; otherwise the layer has no fused form)
        const bool nhwc_in = (io & CONV2D_IO_IN_NHWC) != 0;
        const bool nhwc_ok = pb->c % 8 == 0 && (in_addr & 15) == 0;
        if (nhwc_in) g->flw = kFusedLoadNHWC;
        g->fsh = fused_col_shift(g->flw, pb->q);
        g->fpc = fused_patch_cols(g->flw, g->fbw, g->fsh);
        const int units = 4 * g->fpr * (g->fpc / fused_unit_pixels(g->flw));
        g->fnu = (units + kFusedThreads - 1) / kFusedThreads;
        g->fgroup_items = (long long)pb->n * g->fnbh * g->fnbw * ((pb->k + 63) / 64);
        // a last block of 1..32 channels as wide items (32 channels x 128 tiles, where the
        // single-layer kernel has them) when a list schedule over the compute units finishes
        // sooner, counting a full or half-block item as 1 and a wide item as 1.5 (measured
        // 1.3-1.55 half-block items): N=32 C=K=96 55x55 896 -> 672 items, 42 -> 35 us;
        // N=32 C=K=32 55x55 13.4 -> 9.9 us; 64x256x27x160 (3 vs 3.5 rounds), 16x128x30x96 and
        // 64x16x33x32 (3 rounds of half-block items vs 2 of wide ones) keep half-block items,
        // with which they measured 13%, 33% and 4% faster
        const long long full = (long long)pb->n * g->fnbh * g->fnbw * (pb->k / 64);
        const long long wide_items = full + (long long)pb->n * fused_wide_items(g->fnbh, g->fnbw);
        const long long cus = compute_units();
        const bool wide = fused_tail(g->KB32) && fused_wide_tail(g->flw, fused_shape_index(g->fbh, g->fbw)) &&
                          list_schedule_time(full, 1.0, wide_items - full, 1.5, cus) <
                              list_schedule_time(full, 1.0, g->fgroup_items - full, 1.0, cus);
        g->fitems = wide ? wide_items : g->fgroup_items;
        // buffer offsets are 32-bit: one input image with 32 channels of slack under 2^31
        // bytes (out-of-image offsets are placed at 2^31 and above), one output image and U
        // under 2^31 bytes
        const long long img_bytes = ((long long)pb->c + 32) * pb->h * pb->w * 2;
        const long long out_img_bytes = (long long)pb->k * g->Oh * g->Ow * 2;
        // (a one-pixel-wide image has no fused form: its straddling pixel pair would load the
        // dword before the image's first pixel)
        g->fused_ok = (!nhwc_in || nhwc_ok) && g->fnu <= (nhwc_in ? kFusedNhwcUnits : 2) &&
                      (g->flw != kFusedLoad2 || pb->w != 1) &&
                      fused_kernel_ptr(g->flw, fused_shape_index(g->fbh, g->fbw), g->fsh, fused_tail(g->KB32)) &&
                      g->fgroup_items < (1LL << 31) && img_bytes < (1LL << 31) &&
                      out_img_bytes < (1LL << 31) && (long long)g->u_bytes < (1LL << 31);
    }
    g->f43_ok = false;
    if (g->winograd_ok) {
        // F(4x4,3x3) form: 4x4 tiles; the first of 4/2/1 tile rows whose input staging fits in
        // 64 KiB, else one tile row in up to 160 KiB
        g->t4h = (g->Oh + 3) / 4;
        g->t4w = (g->Ow + 3) / 4;
        g->T4 = (long long)g->t4h * g->t4w;
        g->NT4 = g->T4 * pb->n;
        g->G32_4 = (g->NT4 + 31) / 32;
        g->band4 = 0;
        const int c4[3] = {4, 2, 1};
        for (int i = 0; i < 3 && !g->band4; ++i) {
            const int b = c4[i];
            if (b > 1 && b > g->t4h) continue;
            if (input43_lds_bytes(b, g->t4w, (int)pb->w, (int)pb->q) <= 65536) g->band4 = b;
        }
        if (!g->band4 && input43_lds_bytes(1, g->t4w, (int)pb->w, (int)pb->q) <= 163840) g->band4 = 1;
        g->v43_bytes = (size_t)36 * 2 * g->CB * (size_t)g->G32_4 * kFragHalves * 2;
        g->u43_bytes = (size_t)36 * 2 * g->CB * (size_t)g->KB32 * kFragHalves * 2;
        g->m43_bytes = (size_t)36 * g->KB32 * 32 * (size_t)g->G32_4 * 32 * 2;
        g->f43_ok = g->band4 > 0 && g->NT4 <= (1LL << 30);
    }
    // the Winograd path needs the staged form's input-transform band or the fused form
    g->direct = !(g->winograd_ok && (g->band || g->fused_ok));
    if (g->direct) g->f43_ok = false;
    g->staged_ok = !g->direct && g->band > 0;
    if (g->direct) {
        // general-geometry path: GEMM columns are output pixels, steps run over 32 values
        // of the flattened reduction index j = tap * C + c, padded with zero steps to a
        // multiple of 8 (the deepest ring of the direct configurations)
        g->tiles_h = g->Oh;
        g->tiles_w = g->Ow;
        g->T = (long long)g->Oh * g->Ow;
        g->NT = g->T * pb->n;
        if (g->NT > (1LL << 30)) return CONV2D_ERR_OVERFLOW;
        g->G32 = (g->NT + 31) / 32;
        const long long jn = (long long)pb->r * pb->s * pb->c;
        const long long steps = ((jn + 31) / 32 + 7) / 8 * 8;
        // the packing kernels index j in 32 bits and launch one thread per 16 bytes of an
        // operand (a 1-D grid of fewer than 2^32 threads): at most 64 GiB per operand
        const long long rows = g->G32 > g->KB32 ? g->G32 : g->KB32;
        if (jn > (1LL << 30) || steps * 2 * rows * 64 > (1LL << 32) - 256) return CONV2D_ERR_OVERFLOW;
        g->nsteps = (int)steps;
        g->band = 0;
        g->fused_ok = false;
        g->u_bytes = (size_t)steps * 2 * (size_t)g->KB32 * kFragHalves * 2;
        // a 1x1 filter makes the reduction index the input channel: the GEMM gathers its B
        // operand from the input (no packed V, no packing pass), in every input layout, so the
        // workspace size does not depend on the layout. (Gathering larger filters' taps the same
        // way measured slower than packing: N=32 C=128 56x56 K=256 3x3/2 90 us in the GEMM
        // against 30 us GEMM + 25 us packing on N,C,H,W input, 63-88 us against 22-25 us for
        // N=16 C=256 14x14 on channel-contiguous input.)
        g->pointwise = pb->r == 1 && pb->s == 1;
        g->point_quads = g->pointwise && pb->u == 1 && pb->v == 1 && pb->p == 0 && pb->q == 0 &&
                         (io & CONV2D_IO_IN_NHWC) == 0 && ((long long)pb->h * pb->w) % 4 == 0 &&
                         ((uintptr_t)pb->in & 7) == 0;
        g->v_bytes = g->pointwise ? 0 : (size_t)steps * 2 * (size_t)g->G32 * kFragHalves * 2;
    }
    return CONV2D_OK;
}

int require_winograd(const problem_t* pb, Geometry* g) {
    int rc = derive_geometry(pb, g);
    if (rc) return rc;
    if (g->direct) return CONV2D_ERR_UNSUPPORTED;
    return CONV2D_OK;
}

int require_staged(const problem_t* pb, Geometry* g) {
    int rc = require_winograd(pb, g);
    if (rc) return rc;
    return g->staged_ok ? CONV2D_OK : CONV2D_ERR_UNSUPPORTED;
}

static int compute_units() {
    static int cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            cus = 256;
    }
    return cus;
}

// Shape heuristic (used when the tuned table has no entry): the smallest tile (32x64) unless
// the 64x128 tile gives at least one workgroup per two compute units; then the 128x128
// LDS-DMA tile for deep reductions (C >= 256) that give at least one workgroup per two
// compute units, else the 64x128 LDS-DMA tile. Fitted to autotuner measurements of 2,000
// seeded shapes on MI355X (held-out: 1.1% mean regret against exhaustive search); the
// 128x128 threshold was lowered from one workgroup per compute unit after a re-measurement
// of 400 seeded shapes with the current kernels (mean regret 1.14% -> 0.99%; 32x512x28x512:
// 71 vs 81 us).
int heuristic_config(const problem_t* pb, const Geometry& g) {
    const long long cus = compute_units();
    auto items = [&](int bk, int bt) {
        return (long long)((g.KB32 + bk / 32 - 1) / (bk / 32)) * ((g.G32 + bt / 32 - 1) / (bt / 32));
    };
    if (2 * items(64, 128) < cus) return 10;
    if (pb->c >= 256 && 2 * items(128, 128) >= cus) return 9;
    return 11;
}

// General-geometry path: the 128x128 LDS-DMA tile for K >= 128 when it gives every compute
// unit a workgroup, 64x128 when that gives one per two, else 32x64.
static int direct_heuristic_config(const Geometry& g) {
    const long long cus = compute_units();
    auto items = [&](int bk, int bt) {
        return (long long)((g.KB32 + bk / 32 - 1) / (bk / 32)) * ((g.G32 + bt / 32 - 1) / (bt / 32));
    };
    if (g.KB32 >= 4 && items(128, 128) >= cus) return 0;
    if (2 * items(64, 128) >= cus) return 1;
    return 2;
}

int select_config(const problem_t* pb, const Geometry& g) {
    if (g.direct) return direct_heuristic_config(g);
    int cfg = tuned_config_lookup(pb);
    if (cfg >= 0 && cfg < kNumGemmConfigs) return cfg;
    return heuristic_config(pb, g);
}

// Form choice on the Winograd path. The fused kernel recomputes V once per 64 output
// channels instead of writing and reading it through HBM; the staged GEMM's 128 x 128 tiles
// reuse each loaded operand more. A shape in the tuned table takes the form the autotuner
// measured faster.
bool use_fused(const problem_t* pb, const Geometry& g) {
    if (g.direct || !g.fused_ok) return false;
    if (!g.staged_ok) return true;  // too wide for the input transform's LDS staging
    const int tuned = tuned_form_lookup(pb);
    if (tuned == CONV2D_FORM_STAGED || tuned == CONV2D_FORM_FUSED)
        return tuned == CONV2D_FORM_FUSED;  // measured by the autotuner
    // Measured on 45 shapes (both forms timed after each had run 50 ms, so at sustained
    // clocks): with K <= 384 the fused form was faster on all but two (at most 6% slower,
    // up to 3.2x faster); with K = 512 the staged GEMM's 128-row tiles win from
    // N*T*K*C ~ 1.6e9 (up to 1.15x), and from 8e8 when C < K.
    const double work = (double)g.NT * (double)pb->k * (double)pb->c;
    return pb->k <= 384 || work < 8e8 || (work < 1.2e9 && pb->c >= pb->k);
}

void fill_param(const problem_t* pb, const Geometry& g, int cfg, mykernelParamType* a) {
    memset(a, 0, sizeof(*a));
    a->pin = pb->in;
    a->pweight = pb->weight;
    a->pout = pb->out;
    a->transform_pin = pb->transform_pin;
    a->transform_filter = pb->transform_filter;
    a->mid_result = pb->mid_result;
    a->n = (short)pb->n;
    a->c = (short)pb->c;
    a->h = (short)pb->h;
    a->w = (short)pb->w;
    a->k = (short)pb->k;
    a->r = (short)pb->r;
    a->s = (short)pb->s;
    a->u = (short)pb->u;
    a->v = (short)pb->v;
    a->p = (short)pb->p;
    a->q = (short)pb->q;
    a->Oh = (short)g.Oh;
    a->Ow = (short)g.Ow;
    a->tiles_h = g.tiles_h;
    a->tiles_w = g.tiles_w;
    a->T = (int)g.T;
    a->NT = (int)g.NT;
    a->G32 = (int)g.G32;
    a->CB = g.CB;
    a->KB32 = g.KB32;
    a->band = g.band;
    a->cfg = cfg;
    a->nsteps = g.nsteps;
    a->fbh = g.fbh;
    a->fbw = g.fbw;
    a->fnbh = g.fnbh;
    a->fnbw = g.fnbw;
    a->fpr = g.fpr;
    a->fpc = g.fpc;
    a->fsh = g.fsh;
    a->fnu = g.fnu;
    a->io = g.io;
    // widest input-row load the staging may use (rows start at multiples of w halves)
    const uintptr_t in_addr = (uintptr_t)pb->in;
    if (pb->w % 8 == 0 && (in_addr & 15) == 0)
        a->flags = 1;
    else if (pb->w % 4 == 0 && (in_addr & 7) == 0)
        a->flags = 2;
    else if (pb->w % 2 == 0 && (in_addr & 3) == 0)
        a->flags = 4;
    else
        a->flags = 0;
}

int select_form(const problem_t* pb, const Geometry& g) {
    if (g.direct) return CONV2D_FORM_STAGED;
    if (g.f43_ok && pb->mid_result && tuned_form_lookup(pb) == CONV2D_FORM_F43) return CONV2D_FORM_F43;
    return use_fused(pb, g) ? CONV2D_FORM_FUSED : CONV2D_FORM_STAGED;
}

// Positions per F(4x4) GEMM workgroup: the most (a divisor of 36) that still gives two
// workgroups per compute unit, so each pipeline runs over several positions' steps instead of
// filling and draining its ring once per position.
static int f43_positions_per_item(const Geometry& g, int cfg) {
    const GemmConfig& c = kPosConfigs[cfg];
    const long long base = (long long)((g.KB32 + c.bk / 32 - 1) / (c.bk / 32)) * ((g.G32_4 + c.bt / 32 - 1) / (c.bt / 32));
    const int divs[9] = {36, 18, 12, 9, 6, 4, 3, 2, 1};
    for (int i = 0; i < 9; ++i)
        if (base * (36 / divs[i]) >= 2LL * compute_units()) return divs[i];
    return 1;
}

void fill_param43(const problem_t* pb, const Geometry& g, int cfg, mykernelParamType* a) {
    fill_param(pb, g, cfg, a);
    a->f43_npos = f43_positions_per_item(g, cfg);
    a->tiles_h = g.t4h;
    a->tiles_w = g.t4w;
    a->T = (int)g.T4;
    a->NT = (int)g.NT4;
    a->G32 = (int)g.G32_4;
    a->band = g.band4;
    a->nsteps = g.CB;
    a->form = CONV2D_FORM_F43;
}

// F(4x4) GEMM heuristic (used when the tuned table has no entry), from measurements on MI355X:
// the 128x128 tile of 64x64 waves for deep reductions (CB >= 8) with four 128x128 items per
// compute unit, else the 128x64 tile when it gives one item per unit, else 32x64.
int select_config43(const problem_t* pb, const Geometry& g) {
    if (tuned_form_lookup(pb) == CONV2D_FORM_F43) {
        const int cfg = tuned_config_lookup(pb);
        if (cfg >= 0 && cfg < kNumPosConfigs) return cfg;
    }
    const long long cus = compute_units();
    auto items = [&](int bk, int bt) {
        return 36LL * ((g.KB32 + bk / 32 - 1) / (bk / 32)) * ((g.G32_4 + bt / 32 - 1) / (bt / 32));
    };
    if (g.CB >= 8 && items(128, 128) >= 4 * cus) return 5;
    if (items(128, 64) >= cus) return 1;
    return 3;
}

Launch slot_filter43(const Geometry& g) {
    Launch L;
    L.kernel = (const void*)&wino43_filter_transform;
    // one thread per element of one position's packed U
    L.grid = dim3((unsigned)(((long long)g.KB32 * 2 * g.CB * kFragHalves + 255) / 256), 1, 1);
    L.block = dim3(256, 1, 1);
    L.lds = 0;
    return L;
}

Launch slot_input43(const problem_t* pb, const Geometry& g) {
    Launch L;
    L.kernel = (const void*)&wino43_input_transform;
    L.grid = dim3((unsigned)((g.t4h + g.band4 - 1) / g.band4), (unsigned)g.CB, pb->n);
    L.block = dim3(256, 1, 1);
    L.lds = (unsigned)input43_lds_bytes(g.band4, g.t4w, (int)pb->w, (int)pb->q);
    return L;
}

Launch slot_gemm43(const Geometry& g, int cfg) {
    const GemmConfig& c = kPosConfigs[cfg];
    Launch L;
    L.kernel = c.kernel;
    const int vf = c.bt / 32, uf = c.bk / 32;
    const unsigned ng = (unsigned)((g.G32_4 + vf - 1) / vf), nk = (unsigned)((g.KB32 + uf - 1) / uf);
    L.grid = dim3(ng * nk * (unsigned)(36 / f43_positions_per_item(g, cfg)), 1, 1);
    L.block = dim3((unsigned)c.threads, 1, 1);
    L.lds = 0;
    return L;
}

Launch slot_output43(const Geometry& g, int k) {
    Launch L;
    L.kernel = (const void*)&wino43_output_transform;
    L.grid = dim3((unsigned)((g.NT4 + 255) / 256), (unsigned)((k + 3) / 4), 1);
    L.block = dim3(256, 1, 1);
    L.lds = 0;
    return L;
}

int prepare_weights_form(const problem_t* pb, const Geometry& g, int form, hipStream_t st) {
    mykernelParamType a;
    if (form == CONV2D_FORM_F43) {
        fill_param43(pb, g, 0, &a);
        return launch(slot_filter43(g), &a, st);
    }
    fill_param(pb, g, 0, &a);
    return launch(slot_filter(g), &a, st);
}

// S2 + S3 + S4 of the F(4x4) form with the epilogue in a (a filled by fill_param43).
static int forward43(const problem_t* pb, const Geometry& g, int cfg, mykernelParamType* a, hipStream_t st) {
    int rc;
    if ((rc = launch(slot_input43(pb, g), a, st))) return rc;
    if ((rc = launch(slot_gemm43(g, cfg), a, st))) return rc;
    return launch(slot_output43(g, (int)pb->k), a, st);
}

Launch slot_filter(const Geometry& g) {
    Launch L;
    L.lds = 0;
    if (g.direct) {
        L.kernel = (const void*)&direct_pack_weights;
        const long long threads = (long long)g.KB32 * 2 * g.nsteps * 64;
        L.grid = dim3((unsigned)((threads + 255) / 256), 1, 1);
        L.block = dim3(256, 1, 1);
        return L;
    }
    // one one-wave workgroup per (U fragment column of 32 rows x 16 channels, transform row)
    L.kernel = (const void*)&wino_filter_transform;
    L.grid = dim3((unsigned)(g.KB32 * 2 * g.CB * 4), 1, 1);
    L.block = dim3(64, 1, 1);
    return L;
}

// Fused form: one workgroup of 8 waves per (64 output channels, 64-tile block), or per
// (32 channels, 128-tile block) for wide tail items (g.fitems counts both); LDS holds two input
// patches during the channel loop and the 128 KiB Z exchange after it.
Launch slot_fused(const Geometry& g) {
    Launch L;
    L.kernel = fused_kernel_ptr(g.flw, fused_shape_index(g.fbh, g.fbw), g.fsh, fused_tail(g.KB32));
    L.grid = dim3((unsigned)g.fitems, 1, 1);
    L.block = dim3((unsigned)fused_threads(), 1, 1);
    const unsigned patches = 2u * 4u * (unsigned)g.fpr * (unsigned)g.fpc * 16u;
    L.lds = patches > 131072u ? patches : 131072u;
    return L;
}

Launch slot_input(const problem_t* pb, const Geometry& g) {
    Launch L;
    if (g.direct && g.pointwise) return slot_noop();  // the GEMM reads the input itself
    if (g.direct && (g.io & CONV2D_IO_IN_NHWC) == 0 && pb->c % 8 != 0) {
        // few channels: a wave per 32-pixel block walks every 16-value block
        L.kernel = (const void*)&direct_pack_input_walk;
        L.grid = dim3((unsigned)((g.G32 + 3) / 4), 1, 1);
        L.block = dim3(256, 1, 1);
        L.lds = 0;
        return L;
    }
    if (g.direct) {
        L.kernel = (const void*)&direct_pack_input;
        const long long threads = (long long)g.G32 * 2 * g.nsteps * 64;
        L.grid = dim3((unsigned)((threads + 255) / 256), 1, 1);
        L.block = dim3(256, 1, 1);
        L.lds = 0;
        return L;
    }
    L.kernel = (const void*)&wino_input_transform;
    L.grid = dim3((unsigned)((g.tiles_h + g.band - 1) / g.band), (unsigned)g.CB, pb->n);
    L.block = dim3(256, 1, 1);
    L.lds = (unsigned)input_lds_bytes(g.band, g.tiles_w, (int)pb->w, (int)pb->q);
    return L;
}

Launch slot_gemm(const Geometry& g, int cfg) {
    const GemmConfig* table = !g.direct ? kGemmConfigs
                              : (g.point_quads ? kPointQuadConfigs : (g.pointwise ? kPointConfigs : kDirectConfigs));
    const GemmConfig& c = table[cfg];
    Launch L;
    L.kernel = c.kernel;
    const int vf = c.bt / 32, uf = c.bk / 32;
    const unsigned ng = (unsigned)((g.G32 + vf - 1) / vf), nk = (unsigned)((g.KB32 + uf - 1) / uf);
    L.grid = c.depth > 0 ? dim3(ng * nk, 1, 1) : dim3(ng, nk, 1);
    L.block = dim3((unsigned)c.threads, 1, 1);
    L.lds = 0;
    return L;
}

Launch slot_noop() {
    Launch L;
    L.kernel = (const void*)&wino_noop;
    L.grid = dim3(1, 1, 1);
    L.block = dim3(64, 1, 1);
    L.lds = 0;
    return L;
}

int launch(const Launch& L, mykernelParamType* a, hipStream_t stream) {
    void* args[1] = {(void*)a};
    hipError_t e = hipExtLaunchKernel(L.kernel, L.grid, L.block, args, L.lds, stream, nullptr,
                                      nullptr, 0);
    return e == hipSuccess ? CONV2D_OK : CONV2D_ERR_HIP;
}

}  // namespace wino

using namespace wino;

static void put_slot(const Launch& L, unsigned* bx, unsigned* by, unsigned* bz, unsigned* tx,
                     unsigned* ty, unsigned* tz, unsigned* lds, void** ptr) {
    *bx = L.grid.x;
    *by = L.grid.y;
    *bz = L.grid.z;
    *tx = L.block.x;
    *ty = L.block.y;
    *tz = L.block.z;
    *lds = L.lds;
    *ptr = (void*)L.kernel;
}

extern "C" {

int getParamsize(problem_t* problem, int* paramSize) {
    if (!paramSize) return CONV2D_ERR_NULL;
    (void)problem;
    *paramSize = (int)sizeof(mykernelParamType);
    return CONV2D_OK;
}

int getkernelInfo(problem_t* problem, kernelInfo_t* ki, void* param) {
    return getkernelInfo_io(problem, CONV2D_IO_NCHW, ki, param);
}

int getkernelInfo_io(problem_t* problem, int io, kernelInfo_t* ki, void* param) {
    if (!problem || !ki || !param) return CONV2D_ERR_NULL;
    if (!problem->in || !problem->weight || !problem->out || !problem->transform_filter)
        return CONV2D_ERR_NULL;
    Geometry g;
    int rc = derive_geometry_io(problem, io, &g);
    if (rc) return rc;
    if (!problem->transform_pin && g.v_bytes > 0) return CONV2D_ERR_NULL;
    const int form = select_form(problem, g);
    if (form == CONV2D_FORM_F43) {
        const int cfg = select_config43(problem, g);
        fill_param43(problem, g, cfg, (mykernelParamType*)param);
        put_slot(slot_filter43(g), &ki->blockx1, &ki->blocky1, &ki->blockz1, &ki->threadx1, &ki->thready1,
                 &ki->threadz1, &ki->dynmicLdsSize1, &ki->kernelPtr1);
        put_slot(slot_input43(problem, g), &ki->blockx2, &ki->blocky2, &ki->blockz2, &ki->threadx2, &ki->thready2,
                 &ki->threadz2, &ki->dynmicLdsSize2, &ki->kernelPtr2);
        put_slot(slot_gemm43(g, cfg), &ki->blockx3, &ki->blocky3, &ki->blockz3, &ki->threadx3, &ki->thready3,
                 &ki->threadz3, &ki->dynmicLdsSize3, &ki->kernelPtr3);
        put_slot(slot_output43(g, (int)problem->k), &ki->blockx4, &ki->blocky4, &ki->blockz4, &ki->threadx4,
                 &ki->thready4, &ki->threadz4, &ki->dynmicLdsSize4, &ki->kernelPtr4);
        return CONV2D_OK;
    }
    const int cfg = select_config(problem, g);
    const bool fused = form == CONV2D_FORM_FUSED;
    fill_param(problem, g, cfg, (mykernelParamType*)param);
    ((mykernelParamType*)param)->form = fused ? CONV2D_FORM_FUSED : CONV2D_FORM_STAGED;
    put_slot(slot_filter(g), &ki->blockx1, &ki->blocky1, &ki->blockz1, &ki->threadx1,
             &ki->thready1, &ki->threadz1, &ki->dynmicLdsSize1, &ki->kernelPtr1);
    put_slot(fused ? slot_fused(g) : slot_input(problem, g), &ki->blockx2, &ki->blocky2, &ki->blockz2,
             &ki->threadx2, &ki->thready2, &ki->threadz2, &ki->dynmicLdsSize2, &ki->kernelPtr2);
    put_slot(fused ? slot_noop() : slot_gemm(g, cfg), &ki->blockx3, &ki->blocky3, &ki->blockz3,
             &ki->threadx3, &ki->thready3, &ki->threadz3, &ki->dynmicLdsSize3, &ki->kernelPtr3);
    put_slot(slot_noop(), &ki->blockx4, &ki->blocky4, &ki->blockz4, &ki->threadx4,
             &ki->thready4, &ki->threadz4, &ki->dynmicLdsSize4, &ki->kernelPtr4);
    return CONV2D_OK;
}

int conv2d_workspace_size(const problem_t* problem, size_t* pin_bytes, size_t* filter_bytes,
                          size_t* mid_bytes) {
    if (!problem) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = derive_geometry(problem, &g);
    if (rc) return rc;
    // every form the library may choose for the shape: the F(4x4) form's U and V where it
    // applies (V is smaller than F(2x2)'s, U 2.25x larger) and its M in mid_result
    if (pin_bytes) *pin_bytes = g.f43_ok && g.v43_bytes > g.v_bytes ? g.v43_bytes : g.v_bytes;
    if (filter_bytes) *filter_bytes = g.f43_ok && g.u43_bytes > g.u_bytes ? g.u43_bytes : g.u_bytes;
    if (mid_bytes) *mid_bytes = g.f43_ok ? g.m43_bytes : 0;
    return CONV2D_OK;
}

int conv2d_form_workspace_size(const problem_t* problem, int form, size_t* pin_bytes, size_t* filter_bytes,
                               size_t* mid_bytes) {
    if (!problem) return CONV2D_ERR_NULL;
    if (form != CONV2D_FORM_STAGED && form != CONV2D_FORM_FUSED && form != CONV2D_FORM_F43)
        return CONV2D_ERR_INVALID;
    Geometry g;
    int rc = derive_geometry(problem, &g);
    if (rc) return rc;
    const bool f43 = form == CONV2D_FORM_F43;
    if (f43 && !g.f43_ok) return CONV2D_ERR_UNSUPPORTED;
    if (pin_bytes) *pin_bytes = f43 ? g.v43_bytes : g.v_bytes;
    if (filter_bytes) *filter_bytes = f43 ? g.u43_bytes : g.u_bytes;
    if (mid_bytes) *mid_bytes = f43 ? g.m43_bytes : 0;
    return CONV2D_OK;
}

int conv2d_launch(const kernelInfo_t* ki, void* param, void* stream) {
    if (!ki || !param) return CONV2D_ERR_NULL;
    if (!ki->kernelPtr1 || !ki->kernelPtr2 || !ki->kernelPtr3 || !ki->kernelPtr4)
        return CONV2D_ERR_NULL;
    hipStream_t st = (hipStream_t)stream;
    void* args[1] = {param};
    const void* ptrs[4] = {ki->kernelPtr1, ki->kernelPtr2, ki->kernelPtr3, ki->kernelPtr4};
    const dim3 grids[4] = {dim3(ki->blockx1, ki->blocky1, ki->blockz1),
                           dim3(ki->blockx2, ki->blocky2, ki->blockz2),
                           dim3(ki->blockx3, ki->blocky3, ki->blockz3),
                           dim3(ki->blockx4, ki->blocky4, ki->blockz4)};
    const dim3 blocks[4] = {dim3(ki->threadx1, ki->thready1, ki->threadz1),
                            dim3(ki->threadx2, ki->thready2, ki->threadz2),
                            dim3(ki->threadx3, ki->thready3, ki->threadz3),
                            dim3(ki->threadx4, ki->thready4, ki->threadz4)};
    const unsigned lds[4] = {ki->dynmicLdsSize1, ki->dynmicLdsSize2, ki->dynmicLdsSize3,
                             ki->dynmicLdsSize4};
    for (int i = 0; i < 4; ++i) {
        hipError_t e = hipExtLaunchKernel(ptrs[i], grids[i], blocks[i], args, lds[i], st,
                                          nullptr, nullptr, 0);
        if (e != hipSuccess) return CONV2D_ERR_HIP;
    }
    return CONV2D_OK;
}

// transform_pin may be NULL where the input workspace is 0 bytes (the general-geometry path
// gathering its input inside the GEMM): hipMalloc(0) gives NULL to callers that allocate the
// reported size.
static bool v_missing(const problem_t* pb) {
    if (pb->transform_pin) return false;
    Geometry g;
    return derive_geometry(pb, &g) != CONV2D_OK || g.v_bytes > 0;
}

static int check_ptrs(const problem_t* pb, bool need_in, bool need_w, bool need_out,
                      bool need_v, bool need_u) {
    if (!pb) return CONV2D_ERR_NULL;
    if ((need_in && !pb->in) || (need_w && !pb->weight) || (need_out && !pb->out) ||
        (need_v && v_missing(pb)) || (need_u && !pb->transform_filter))
        return CONV2D_ERR_NULL;
    return CONV2D_OK;
}

int conv2d_prepare_weights(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, false, true, false, false, true);
    if (rc) return rc;
    Geometry g;
    if ((rc = derive_geometry(pb, &g))) return rc;
    return prepare_weights_form(pb, g, select_form(pb, g), (hipStream_t)stream);
}

int conv2d_input_transform(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, true, false, false, true, false);
    if (rc) return rc;
    Geometry g;
    if ((rc = derive_geometry(pb, &g))) return rc;
    if (!g.direct && !g.staged_ok) return CONV2D_ERR_UNSUPPORTED;
    mykernelParamType a;
    fill_param(pb, g, 0, &a);
    return launch(slot_input(pb, g), &a, (hipStream_t)stream);
}

int conv2d_gemm(const problem_t* pb, int cfg, void* stream) {
    int rc = check_ptrs(pb, false, false, true, true, true);
    if (rc) return rc;
    Geometry g;
    if ((rc = derive_geometry(pb, &g))) return rc;
    if (!g.direct && !g.staged_ok) return CONV2D_ERR_UNSUPPORTED;
    if (cfg < 0) cfg = select_config(pb, g);
    if (cfg >= (g.direct ? kNumDirectConfigs : kNumGemmConfigs)) return CONV2D_ERR_INVALID;
    mykernelParamType a;
    fill_param(pb, g, cfg, &a);
    return launch(slot_gemm(g, cfg), &a, (hipStream_t)stream);
}

int conv2d_forward(const problem_t* pb, int cfg, void* stream) { return conv2d_forward_io(pb, cfg, CONV2D_IO_NCHW, stream); }

int conv2d_forward_io(const problem_t* pb, int cfg, int io, void* stream) {
    int rc = check_ptrs(pb, true, false, true, true, true);
    if (rc) return rc;
    Geometry g;
    if ((rc = derive_geometry_io(pb, io, &g))) return rc;
    const int form = cfg < 0 ? select_form(pb, g) : CONV2D_FORM_STAGED;
    if (form == CONV2D_FORM_F43) {
        mykernelParamType a;
        const int c43 = select_config43(pb, g);
        fill_param43(pb, g, c43, &a);
        return forward43(pb, g, c43, &a, (hipStream_t)stream);
    }
    if (form == CONV2D_FORM_FUSED) {
        mykernelParamType a;
        fill_param(pb, g, select_config(pb, g), &a);
        a.form = CONV2D_FORM_FUSED;
        return launch(slot_fused(g), &a, (hipStream_t)stream);
    }
    if (!g.direct && !g.staged_ok) return CONV2D_ERR_UNSUPPORTED;
    if (cfg < 0) cfg = select_config(pb, g);
    if (cfg >= (g.direct ? kNumDirectConfigs : kNumGemmConfigs)) return CONV2D_ERR_INVALID;
    mykernelParamType a;
    fill_param(pb, g, cfg, &a);
    if ((rc = launch(slot_input(pb, g), &a, (hipStream_t)stream))) return rc;
    return launch(slot_gemm(g, cfg), &a, (hipStream_t)stream);
}

int conv2d_set_epilogue(void* param, const void* bias, int act) {
    if (!param) return CONV2D_ERR_NULL;
    if (act != CONV2D_ACT_NONE && act != CONV2D_ACT_RELU) return CONV2D_ERR_INVALID;
    mykernelParamType* a = (mykernelParamType*)param;
    a->bias = bias;
    a->act = act;
    return CONV2D_OK;
}

int conv2d_forward_epilogue(const problem_t* pb, int cfg, const void* bias, int act, void* stream) {
    int rc = check_ptrs(pb, true, false, true, true, true);
    if (rc) return rc;
    if (act != CONV2D_ACT_NONE && act != CONV2D_ACT_RELU) return CONV2D_ERR_INVALID;
    Geometry g;
    if ((rc = derive_geometry(pb, &g))) return rc;
    const int form = cfg < 0 ? select_form(pb, g) : CONV2D_FORM_STAGED;
    if (form == CONV2D_FORM_F43) {
        mykernelParamType a;
        const int c43 = select_config43(pb, g);
        fill_param43(pb, g, c43, &a);
        a.bias = bias;
        a.act = act;
        return forward43(pb, g, c43, &a, (hipStream_t)stream);
    }
    if (form == CONV2D_FORM_FUSED) {
        mykernelParamType a;
        fill_param(pb, g, select_config(pb, g), &a);
        a.form = CONV2D_FORM_FUSED;
        a.bias = bias;
        a.act = act;
        return launch(slot_fused(g), &a, (hipStream_t)stream);
    }
    if (!g.direct && !g.staged_ok) return CONV2D_ERR_UNSUPPORTED;
    if (cfg < 0) cfg = select_config(pb, g);
    if (cfg >= (g.direct ? kNumDirectConfigs : kNumGemmConfigs)) return CONV2D_ERR_INVALID;
    mykernelParamType a;
    fill_param(pb, g, cfg, &a);
    a.bias = bias;
    a.act = act;
    if ((rc = launch(slot_input(pb, g), &a, (hipStream_t)stream))) return rc;
    return launch(slot_gemm(g, cfg), &a, (hipStream_t)stream);
}

int conv2d_form(const problem_t* pb, int* form) {
    if (!pb || !form) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = derive_geometry(pb, &g);
    if (rc) return rc;
    *form = select_form(pb, g);
    return CONV2D_OK;
}

int conv2d_fused_gemm(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, true, false, true, false, true);
    if (rc) return rc;
    Geometry g;
    if ((rc = require_winograd(pb, &g))) return rc;
    if (!g.fused_ok) return CONV2D_ERR_UNSUPPORTED;
    mykernelParamType a;
    fill_param(pb, g, 0, &a);
    return launch(slot_fused(g), &a, (hipStream_t)stream);
}

int conv2d_num_tile_configs(void) { return kNumGemmConfigs; }

int conv2d_tile_config(int cfg, int* bk, int* bt, int* threads) {
    if (cfg < 0 || cfg >= kNumGemmConfigs) return CONV2D_ERR_INVALID;
    if (bk) *bk = kGemmConfigs[cfg].bk;
    if (bt) *bt = kGemmConfigs[cfg].bt;
    if (threads) *threads = kGemmConfigs[cfg].threads;
    return CONV2D_OK;
}

int conv2d_select_tile_config(const problem_t* pb, int* cfg) {
    if (!pb || !cfg) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = derive_geometry(pb, &g);
    if (rc) return rc;
    *cfg = select_config(pb, g);
    return CONV2D_OK;
}

int conv2d_method(const problem_t* pb, int* method) {
    if (!pb || !method) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = derive_geometry(pb, &g);
    if (rc) return rc;
    *method = g.direct ? CONV2D_METHOD_DIRECT : CONV2D_METHOD_WINOGRAD;
    return CONV2D_OK;
}

int conv2d_direct_num_tile_configs(void) { return kNumDirectConfigs; }

int conv2d_direct_tile_config(int cfg, int* bk, int* bt, int* threads) {
    if (cfg < 0 || cfg >= kNumDirectConfigs) return CONV2D_ERR_INVALID;
    if (bk) *bk = kDirectConfigs[cfg].bk;
    if (bt) *bt = kDirectConfigs[cfg].bt;
    if (threads) *threads = kDirectConfigs[cfg].threads;
    return CONV2D_OK;
}

int conv2d_direct_f64(const problem_t* pb, double* out_f64, void* stream) {
    return conv2d_direct_f64_io(pb, CONV2D_IO_NCHW, out_f64, stream);
}

int conv2d_direct_f64_io(const problem_t* pb, int io, double* out_f64, void* stream) {
    if (!pb || !out_f64 || !pb->in || !pb->weight) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = derive_geometry_io(pb, io, &g);
    if (rc) return rc;
    mykernelParamType a;
    fill_param(pb, g, 0, &a);
    const size_t total = (size_t)pb->n * pb->k * g.Oh * g.Ow;
    size_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(conv_direct_f64_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, a, out_f64);
    return hipGetLastError() == hipSuccess ? CONV2D_OK : CONV2D_ERR_HIP;
}

// ---- F(4x4,3x3) form stage entry points ----------------------------------------------------

static int f43_prepare(const problem_t* pb, Geometry* g) {
    if (!pb) return CONV2D_ERR_NULL;
    int rc = derive_geometry(pb, g);
    if (rc) return rc;
    return g->f43_ok ? CONV2D_OK : CONV2D_ERR_UNSUPPORTED;
}

int conv2d_f43_prepare_weights(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, false, true, false, false, true);
    if (rc) return rc;
    Geometry g;
    if ((rc = f43_prepare(pb, &g))) return rc;
    return prepare_weights_form(pb, g, CONV2D_FORM_F43, (hipStream_t)stream);
}

int conv2d_f43_input_transform(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, true, false, false, true, false);
    if (rc) return rc;
    Geometry g;
    if ((rc = f43_prepare(pb, &g))) return rc;
    mykernelParamType a;
    fill_param43(pb, g, 0, &a);
    return launch(slot_input43(pb, g), &a, (hipStream_t)stream);
}

int conv2d_f43_gemm(const problem_t* pb, int cfg, void* stream) {
    int rc = check_ptrs(pb, false, false, false, true, true);
    if (rc) return rc;
    if (!pb->mid_result) return CONV2D_ERR_NULL;
    Geometry g;
    if ((rc = f43_prepare(pb, &g))) return rc;
    if (cfg < 0) cfg = select_config43(pb, g);
    if (cfg >= kNumPosConfigs) return CONV2D_ERR_INVALID;
    mykernelParamType a;
    fill_param43(pb, g, cfg, &a);
    return launch(slot_gemm43(g, cfg), &a, (hipStream_t)stream);
}

int conv2d_f43_output_transform(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, false, false, true, false, false);
    if (rc) return rc;
    if (!pb->mid_result) return CONV2D_ERR_NULL;
    Geometry g;
    if ((rc = f43_prepare(pb, &g))) return rc;
    mykernelParamType a;
    fill_param43(pb, g, 0, &a);
    return launch(slot_output43(g, (int)pb->k), &a, (hipStream_t)stream);
}

int conv2d_f43_forward(const problem_t* pb, int cfg, void* stream) {
    int rc = check_ptrs(pb, true, false, true, true, true);
    if (rc) return rc;
    if (!pb->mid_result) return CONV2D_ERR_NULL;
    Geometry g;
    if ((rc = f43_prepare(pb, &g))) return rc;
    if (cfg < 0) cfg = select_config43(pb, g);
    if (cfg >= kNumPosConfigs) return CONV2D_ERR_INVALID;
    mykernelParamType a;
    fill_param43(pb, g, cfg, &a);
    return forward43(pb, g, cfg, &a, (hipStream_t)stream);
}

int conv2d_f43_num_tile_configs(void) { return kNumPosConfigs; }

int conv2d_f43_tile_config(int cfg, int* bk, int* bt, int* threads) {
    if (cfg < 0 || cfg >= kNumPosConfigs) return CONV2D_ERR_INVALID;
    if (bk) *bk = kPosConfigs[cfg].bk;
    if (bt) *bt = kPosConfigs[cfg].bt;
    if (threads) *threads = kPosConfigs[cfg].threads;
    return CONV2D_OK;
}

int conv2d_f43_select_tile_config(const problem_t* pb, int* cfg) {
    if (!cfg) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = f43_prepare(pb, &g);
    if (rc) return rc;
    *cfg = select_config43(pb, g);
    return CONV2D_OK;
}

// ---- FP8 compensated form ----------------------------------------------------------------

// FP8 tile heuristic (LDS-DMA pipeline, measured on MI355X): the 128x128 tile of eight waves
// once it gives every compute unit a workgroup (N=64 C=K=512 28x28: 184 us against 238-347 us
// for the smaller tiles), else 64x64 with one workgroup per unit, else 32x64.
static int fp8_heuristic_config(const problem_t* pb) {
    Geometry g;
    if (derive_geometry(pb, &g)) return 0;
    const long long cus = compute_units();
    auto items = [&](int bk, int bt) { return (long long)((g.KB32 + bk / 32 - 1) / (bk / 32)) * ((g.G32 + bt / 32 - 1) / (bt / 32)); };
    if (items(128, 128) >= cus) return 0;
    return items(64, 64) >= cus ? 4 : 5;
}

static int fp8_prepare(const problem_t* pb, int cfg, Geometry* g, mykernelParamType* a, int* out_cfg) {
    int rc = require_staged(pb, g);
    if (rc) return rc;
    g->band = input_band(pb, *g, true);  // the FP8 input transform prefers eight-row bands
    if (cfg < 0) cfg = fp8_heuristic_config(pb);
    if (cfg >= kNumFp8Configs) return CONV2D_ERR_INVALID;
    fill_param(pb, *g, cfg, a);
    if (out_cfg) *out_cfg = cfg;
    return CONV2D_OK;
}

static Launch fp8_slot_filter(const problem_t* pb, const Geometry& g) {
    Launch L;
    L.kernel = (const void*)&wino_filter_fp8;
    const long long threads = 16LL * 2 * fp8_cq((int)pb->c) * g.KB32 * 32;
    L.grid = dim3((unsigned)((threads + 255) / 256), 1, 1);
    L.block = dim3(256, 1, 1);
    L.lds = 0;
    return L;
}

static Launch fp8_slot_input(const problem_t* pb, const Geometry& g) {
    Launch L = slot_input(pb, g);
    L.kernel = (const void*)&wino_input_transform_fp8;
    return L;
}

static Launch fp8_slot_gemm(const Geometry& g, int cfg) {
    const GemmConfig& c = kFp8Configs[cfg];
    Launch L;
    L.kernel = c.kernel;
    const int vf = c.bt / 32, uf = c.bk / 32;
    const unsigned ng = (unsigned)((g.G32 + vf - 1) / vf), nk = (unsigned)((g.KB32 + uf - 1) / uf);
    L.grid = dim3(ng * nk, 1, 1);
    L.block = dim3((unsigned)c.threads, 1, 1);
    L.lds = 0;
    return L;
}

int conv2d_fp8_workspace_size(const problem_t* pb, size_t* pin_bytes, size_t* filter_bytes) {
    if (!pb) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = require_staged(pb, &g);
    if (rc) return rc;
    if (pin_bytes) *pin_bytes = fp8_operand_bytes((int)pb->c, g.G32);
    if (filter_bytes) *filter_bytes = fp8_u8_offset(g.CB, g.KB32) + fp8_operand_bytes((int)pb->c, g.KB32);
    return CONV2D_OK;
}

int conv2d_fp8_prepare_weights(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, false, true, false, false, true);
    if (rc) return rc;
    Geometry g;
    mykernelParamType a;
    if ((rc = fp8_prepare(pb, 0, &g, &a, nullptr))) return rc;
    if ((rc = launch(slot_filter(g), &a, (hipStream_t)stream))) return rc;
    return launch(fp8_slot_filter(pb, g), &a, (hipStream_t)stream);
}

int conv2d_fp8_input_transform(const problem_t* pb, void* stream) {
    int rc = check_ptrs(pb, true, false, false, true, false);
    if (rc) return rc;
    Geometry g;
    mykernelParamType a;
    if ((rc = fp8_prepare(pb, 0, &g, &a, nullptr))) return rc;
    return launch(fp8_slot_input(pb, g), &a, (hipStream_t)stream);
}

int conv2d_fp8_gemm(const problem_t* pb, int cfg, void* stream) {
    int rc = check_ptrs(pb, false, false, true, true, true);
    if (rc) return rc;
    Geometry g;
    mykernelParamType a;
    if ((rc = fp8_prepare(pb, cfg, &g, &a, &cfg))) return rc;
    return launch(fp8_slot_gemm(g, cfg), &a, (hipStream_t)stream);
}

int conv2d_fp8_forward(const problem_t* pb, int cfg, void* stream) {
    int rc = check_ptrs(pb, true, false, true, true, true);
    if (rc) return rc;
    Geometry g;
    mykernelParamType a;
    if ((rc = fp8_prepare(pb, cfg, &g, &a, &cfg))) return rc;
    if ((rc = launch(fp8_slot_input(pb, g), &a, (hipStream_t)stream))) return rc;
    return launch(fp8_slot_gemm(g, cfg), &a, (hipStream_t)stream);
}

int conv2d_fp8_num_tile_configs(void) { return kNumFp8Configs; }

int conv2d_fp8_tile_config(int cfg, int* bk, int* bt, int* threads) {
    if (cfg < 0 || cfg >= kNumFp8Configs) return CONV2D_ERR_INVALID;
    if (bk) *bk = kFp8Configs[cfg].bk;
    if (bt) *bt = kFp8Configs[cfg].bt;
    if (threads) *threads = kFp8Configs[cfg].threads;
    return CONV2D_OK;
}

int conv2d_fp8_select_tile_config(const problem_t* pb, int* cfg) {
    if (!pb || !cfg) return CONV2D_ERR_NULL;
    Geometry g;
    int rc = require_staged(pb, &g);
    if (rc) return rc;
    *cfg = fp8_heuristic_config(pb);
    return CONV2D_OK;
}

const char* conv2d_version(void) { return "winoconv-mi355x 0.1 (gfx950)"; }

}  // extern "C"
