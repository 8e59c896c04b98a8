// Multi-layer (batched) form: many conv layers of uneven shape. Staged-form layers run as one
// persistent input transform over a (layer, band, channel block, image) work queue and one
// grouped GEMM launch per tile configuration in use over (layer, k-block, column block)
// items; fused-form layers as one grouped fused launch per (staging width, tile-block shape)
// over (layer, 64-channel block, tile block) items. The whole forward pass can be captured once into a
// hipGraph and replayed with a single graph launch.
//
// No counterpart in the reference, whose harness runs one shape per process
// (src/main.cpp:8-166) with a hand-picked kernel quadruple per (n, c) (conv2d.cpp:273-494).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "plugin_internal.h"
#include "wino_common.h"
#include "wino_fused.h"
#include "wino_gemm.h"

struct conv2d_batch {
    struct Segment {
        int cfg;      // staged: GEMM tile configuration of every layer in the segment
        int fused;    // 1 for a fused segment, 0 for a staged one
        int lw;       // fused: staging load width (FusedLoad)
        int shape;    // fused: 4 * tile-block shape index (fused_shape_index) + column shift
        int first;    // first descriptor of the segment
        int count;    // layers in the segment
        int items;    // work items of the segment's queue
        unsigned lds; // fused: dynamic LDS of the launch
    };
    std::vector<problem_t> layers;
    std::vector<int> cfg;                // GEMM tile configuration per layer (staged form)
    std::vector<int> form;               // CONV2D_FORM_STAGED or CONV2D_FORM_FUSED per layer
    int n_staged = 0;                    // staged descriptors come first
    std::vector<BatchLayer> host;        // descriptors, grouped by configuration
    std::vector<Segment> segs;
    BatchLayer* d_layers = nullptr;
    unsigned* d_counters = nullptr;      // input-transform queue
    int s2_total = 0;
    unsigned s2_lds = 0;
    int cus = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Rebuilds the device descriptors from layers + cfg + form. Staged layers come first, grouped
// by tile configuration (one grouped GEMM launch per group); fused layers follow, grouped by
// staging width and tile-block shape (one grouped fused launch per group). Inside a group layers are queued
// by decreasing per-item cost so large items start first (the order changes no result).
int upload(conv2d_batch* b) {
    const int nl = (int)b->layers.size();
    std::vector<wino::Geometry> geo(nl);
    for (int i = 0; i < nl; ++i) {
        int rc = wino::require_winograd(&b->layers[i], &geo[i]);
        if (rc) return rc;
    }
    auto key = [&](int i) {
        return b->form[i] == CONV2D_FORM_FUSED ? 1000 + 16 * geo[i].flw + 4 * fused_shape_index(geo[i].fbh, geo[i].fbw) + geo[i].fsh : b->cfg[i];
    };
    auto cost = [&](int i) {
        if (b->form[i] == CONV2D_FORM_FUSED) return (double)b->layers[i].c;
        const GemmConfig& c = kGemmConfigs[b->cfg[i]];
        return (double)c.bk * c.bt * b->layers[i].c;
    };
    std::vector<int> order(nl);
    for (int i = 0; i < nl; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        if (key(a) != key(c)) return key(a) < key(c);
        return cost(a) > cost(c);
    });
    b->host.assign(nl, BatchLayer{});
    b->segs.clear();
    b->n_staged = 0;
    int s2 = 0;
    unsigned lds = 0;
    for (int qi = 0; qi < nl; ++qi) {
        const int i = order[qi];
        const problem_t& pb = b->layers[i];
        const wino::Geometry& g = geo[i];
        const int cfg = b->cfg[i];
        const bool fused = b->form[i] == CONV2D_FORM_FUSED;
        const int fz = fused ? 1 : 0;
        const int lw = fused ? g.flw : 0;
        const int shape = fused ? 4 * fused_shape_index(g.fbh, g.fbw) + g.fsh : 0;
        if (b->segs.empty() || b->segs.back().fused != fz || b->segs.back().lw != lw ||
            b->segs.back().shape != shape || (!fused && b->segs.back().cfg != cfg))
            b->segs.push_back(conv2d_batch::Segment{cfg, fz, lw, shape, qi, 0, 0, 0u});
        conv2d_batch::Segment& seg = b->segs.back();
        BatchLayer& bl = b->host[qi];
        wino::fill_param(&pb, g, cfg, &bl.prm);
        bl.cfg = cfg;
        bl.s3_first = seg.items;
        if (fused) {
            bl.prm.form = CONV2D_FORM_FUSED;
            bl.nbands = 0;
            bl.s2_first = s2;
            bl.s2_items = 0;
            bl.nk = ((int)pb.k + 63) / 64;
            bl.s3_items = (int)g.fitems;  // wide tail items where the single-layer launch has them
            const wino::Launch L = wino::slot_fused(g);
            seg.lds = std::max(seg.lds, L.lds);
        } else {
            b->n_staged = qi + 1;
            bl.nbands = (g.tiles_h + g.band - 1) / g.band;
            bl.s2_first = s2;
            bl.s2_items = bl.nbands * g.CB * (int)pb.n;
            s2 += bl.s2_items;
            const GemmConfig& c = kGemmConfigs[cfg];
            const int uf = c.bk / 32, vf = c.bt / 32;
            bl.nk = (g.KB32 + uf - 1) / uf;
            const int ng = (int)((g.G32 + vf - 1) / vf);
            bl.s3_items = bl.nk * ng;
            const unsigned l = (unsigned)input_lds_bytes(g.band, g.tiles_w, (int)pb.w, (int)pb.q);
            lds = std::max(lds, l);
        }
        seg.items += bl.s3_items;
        seg.count += 1;
    }
    b->s2_total = s2;
    b->s2_lds = lds;
    if (!b->d_layers && hipMalloc(&b->d_layers, sizeof(BatchLayer) * nl) != hipSuccess) return CONV2D_ERR_HIP;
    if (hipMemcpy(b->d_layers, b->host.data(), sizeof(BatchLayer) * nl, hipMemcpyHostToDevice) != hipSuccess)
        return CONV2D_ERR_HIP;
    return CONV2D_OK;
}

void drop_graph(conv2d_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Staged layers: resets the input-transform queue with a kernel (not a memset node: it is
// captured into graphs), one persistent input-transform launch over the staged layers and one
// grouped GEMM launch per tile configuration in use. Fused layers: one grouped fused launch per
// (staging width, tile-block shape).
int launch_forward(conv2d_batch* b, hipStream_t st) {
    const int nseg = (int)b->segs.size();
    if (b->s2_total > 0) {
        hipLaunchKernelGGL(wino_batch_reset, dim3(1), dim3(64), 0, st, b->d_counters, 1);
        const int g2 = std::min(b->s2_total, b->cus * 4);
        hipLaunchKernelGGL(wino_batch_input_transform, dim3(g2), dim3(256), b->s2_lds, st,
                           (const BatchLayer*)b->d_layers, b->n_staged, b->s2_total, b->d_counters);
    }
    for (int i = 0; i < nseg; ++i) {
        const conv2d_batch::Segment& seg = b->segs[i];
        const BatchLayer* layers = b->d_layers + seg.first;
        int count = seg.count;
        void* args[2] = {(void*)&layers, (void*)&count};
        const hipError_t e =
            seg.fused ? hipLaunchKernel(fused_grouped_kernel_ptr(seg.lw, seg.shape >> 2, seg.shape & 3), dim3(seg.items), dim3(fused_threads()), args, seg.lds, st)
                   : hipLaunchKernel(kGemmConfigs[seg.cfg].grouped_kernel, dim3(seg.items),
                                     dim3(kGemmConfigs[seg.cfg].threads), args, 0, st);
        if (e != hipSuccess) return CONV2D_ERR_HIP;
    }
    return hipGetLastError() == hipSuccess ? CONV2D_OK : CONV2D_ERR_HIP;
}

}  // namespace

extern "C" {

int conv2d_batch_create(const problem_t* layers, int nlayers, conv2d_batch** out) {
    if (!layers || !out) return CONV2D_ERR_NULL;
    if (nlayers <= 0) return CONV2D_ERR_INVALID;
    *out = nullptr;
    for (int i = 0; i < nlayers; ++i) {
        const problem_t& pb = layers[i];
        if (!pb.in || !pb.weight || !pb.out || !pb.transform_pin || !pb.transform_filter) return CONV2D_ERR_NULL;
        wino::Geometry g;
        int rc = wino::require_winograd(&pb, &g);
        if (rc) return rc;
    }
    conv2d_batch* b = new (std::nothrow) conv2d_batch;
    if (!b) return CONV2D_ERR_INVALID;
    b->layers.assign(layers, layers + nlayers);
    b->cfg.resize(nlayers);
    b->form.resize(nlayers);
    for (int i = 0; i < nlayers; ++i) {
        wino::Geometry g;
        (void)wino::require_winograd(&layers[i], &g);
        b->cfg[i] = wino::select_config(&layers[i], g);
        b->form[i] = wino::use_fused(&layers[i], g) ? CONV2D_FORM_FUSED : CONV2D_FORM_STAGED;
    }
    int rc = CONV2D_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        rc = CONV2D_ERR_HIP;
    if (!rc && hipMalloc(&b->d_counters, sizeof(unsigned)) != hipSuccess) rc = CONV2D_ERR_HIP;
    if (!rc) rc = upload(b);
    if (rc) {
        if (b->d_layers) (void)hipFree(b->d_layers);
        if (b->d_counters) (void)hipFree(b->d_counters);
        delete b;
        return rc;
    }
    *out = b;
    return CONV2D_OK;
}

int conv2d_batch_destroy(conv2d_batch* b) {
    if (!b) return CONV2D_ERR_NULL;
    drop_graph(b);
    if (b->d_layers) (void)hipFree(b->d_layers);
    if (b->d_counters) (void)hipFree(b->d_counters);
    delete b;
    return CONV2D_OK;
}

// Filter transforms of every layer (once per weight set): the F(2x2) U, which both batched
// forms use (a shape tuned to the F(4x4) form still runs fused or staged here).
int conv2d_batch_prepare_weights(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    for (const problem_t& pb : b->layers) {
        if (!pb.weight || !pb.transform_filter) return CONV2D_ERR_NULL;
        wino::Geometry g;
        int rc = wino::derive_geometry(&pb, &g);
        if (!rc) rc = wino::prepare_weights_form(&pb, g, CONV2D_FORM_STAGED, (hipStream_t)stream);
        if (rc) return rc;
    }
    return CONV2D_OK;
}

// Input transforms and GEMMs of every layer: one persistent launch each (plus a queue reset).
int conv2d_batch_forward(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    return launch_forward(b, (hipStream_t)stream);
}

// The same work as separate per-layer launches (each layer in its form: S2 then S3 with the
// same tile configuration, or the fused kernel): the baseline the batched form is compared
// against.
int conv2d_batch_forward_separate(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    for (size_t i = 0; i < b->layers.size(); ++i) {
        int rc = b->form[i] == CONV2D_FORM_FUSED ? conv2d_fused_gemm(&b->layers[i], stream)
                                                 : conv2d_forward(&b->layers[i], b->cfg[i], stream);
        if (rc) return rc;
    }
    return CONV2D_OK;
}

// Captures conv2d_batch_forward into a hipGraph on `stream` (must not be the null stream).
int conv2d_batch_capture_graph(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    if (!stream) return CONV2D_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    drop_graph(b);
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return CONV2D_ERR_HIP;
    int rc = launch_forward(b, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : CONV2D_ERR_HIP;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(b);
        return CONV2D_ERR_HIP;
    }
    return CONV2D_OK;
}

int conv2d_batch_launch_graph(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    if (!b->exec) return CONV2D_ERR_INVALID;
    return hipGraphLaunch(b->exec, (hipStream_t)stream) == hipSuccess ? CONV2D_OK : CONV2D_ERR_HIP;
}

// Per-layer tile configurations from the autotuner (every configuration) and, where the fused
// form applies, the faster form: the fused kernel against the input transform plus the best
// GEMM configuration. Each layer's U and V must be prepared; every layer's output is
// overwritten.
int conv2d_batch_autotune(conv2d_batch* b, int reps, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> all(kNumGemmConfigs);
    for (int c = 0; c < kNumGemmConfigs; ++c) all[c] = c;
    std::vector<float> times(kNumGemmConfigs);
    for (size_t i = 0; i < b->layers.size(); ++i) {
        const problem_t* pb = &b->layers[i];
        wino::Geometry g0;
        if (int rc0 = wino::require_winograd(pb, &g0)) return rc0;
        if (!g0.staged_ok) {  // too wide for the input transform: the fused form only
            b->form[i] = CONV2D_FORM_FUSED;
            continue;
        }
        int best = -1;
        int rc = wino::autotune_subset(pb, reps, stream, all.data(), kNumGemmConfigs, &best, times.data());
        if (rc) return rc;
        b->cfg[i] = best;
        b->form[i] = CONV2D_FORM_STAGED;
        wino::Geometry g;
        if ((rc = wino::require_winograd(pb, &g))) return rc;
        if (!g.fused_ok) continue;
        mykernelParamType a;
        wino::fill_param(pb, g, best, &a);
        float t_s2 = 0.f, t_fused = 0.f;
        if ((rc = wino::time_launch_us(wino::slot_input(pb, g), &a, reps, st, &t_s2))) return rc;
        if ((rc = wino::time_launch_us(wino::slot_fused(g), &a, reps, st, &t_fused))) return rc;
        if (t_fused < t_s2 + times[best]) b->form[i] = CONV2D_FORM_FUSED;
    }
    drop_graph(b);
    return upload(b);
}

// Form of layer i (CONV2D_FORM_STAGED or CONV2D_FORM_FUSED).
int conv2d_batch_layer_form(const conv2d_batch* b, int i, int* form) {
    if (!b || !form) return CONV2D_ERR_NULL;
    if (i < 0 || i >= (int)b->layers.size()) return CONV2D_ERR_INVALID;
    *form = b->form[i];
    return CONV2D_OK;
}

// GEMM tile configuration of layer i.
int conv2d_batch_layer_config(const conv2d_batch* b, int i, int* cfg) {
    if (!b || !cfg) return CONV2D_ERR_NULL;
    if (i < 0 || i >= (int)b->layers.size()) return CONV2D_ERR_INVALID;
    *cfg = b->cfg[i];
    return CONV2D_OK;
}

}  // extern "C"
