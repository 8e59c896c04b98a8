// Multi-layer (batched) form: many conv layers of uneven shape run as one persistent input
// transform and one persistent GEMM over (layer, tile-block) work queues, optionally
// captured once into a hipGraph and replayed with a single launch.
//
// No counterpart in the reference, whose harness runs one shape per process
// (src/main.cpp:8-166) with a hand-picked kernel quadruple per (n, c) (conv2d.cpp:273-494).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "plugin_internal.h"
#include "wino_common.h"
#include "wino_gemm.h"

struct conv2d_batch {
    std::vector<problem_t> layers;
    std::vector<int> bcfg;               // batched GEMM configuration per layer
    std::vector<BatchLayer> host;        // descriptors in queue order
    BatchLayer* d_layers = nullptr;
    unsigned* d_counters = nullptr;      // [0] input-transform queue, [1] GEMM queue
    int s2_total = 0, s3_total = 0;
    unsigned s2_lds = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Global GEMM configuration index of each batched configuration (same tile shapes).
const int kBatchToGlobal[kBatchCfgs] = {1, 2, 4};

int global_to_batch(int cfg) {
    for (int b = 0; b < kBatchCfgs; ++b)
        if (kBatchToGlobal[b] == cfg) return b;
    return -1;
}

int batch_heuristic(const problem_t& pb) {
    const int kp = ((int)pb.k + 15) / 16 * 16;
    if (kp >= 128) return 1;
    if (kp >= 64) return 0;
    return 2;
}

// Rebuilds the device descriptors from layers + bcfg. Layers are queued by decreasing
// per-item GEMM cost so large items start first (the order changes no result).
int upload(conv2d_batch* b) {
    const int nl = (int)b->layers.size();
    std::vector<int> order(nl);
    for (int i = 0; i < nl; ++i) order[i] = i;
    auto cost = [&](int i) {
        const problem_t& p = b->layers[i];
        return (double)batch_cfg_bk(b->bcfg[i]) * batch_cfg_bt(b->bcfg[i]) * p.c;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost(a) > cost(c); });
    b->host.assign(nl, BatchLayer{});
    int s2 = 0, s3 = 0;
    unsigned lds = 0;
    for (int qi = 0; qi < nl; ++qi) {
        const int i = order[qi];
        const problem_t& pb = b->layers[i];
        wino::Geometry g;
        int rc = wino::require_winograd(&pb, &g);
        if (rc) return rc;
        BatchLayer& bl = b->host[qi];
        wino::fill_param(&pb, g, kBatchToGlobal[b->bcfg[i]], &bl.prm);
        bl.nbands = (g.tiles_h + g.band - 1) / g.band;
        bl.s2_first = s2;
        bl.s2_items = bl.nbands * g.CB * (int)pb.n;
        s2 += bl.s2_items;
        const int uf = batch_cfg_bk(b->bcfg[i]) / 32, vf = batch_cfg_bt(b->bcfg[i]) / 32;
        bl.nk = (g.KB32 + uf - 1) / uf;
        const int ng = (int)((g.G32 + vf - 1) / vf);
        bl.s3_first = s3;
        bl.s3_items = bl.nk * ng;
        s3 += bl.s3_items;
        bl.bcfg = b->bcfg[i];
        const unsigned l = (unsigned)input_lds_bytes(g.band, g.tiles_w, (int)pb.w, (int)pb.q);
        lds = std::max(lds, l);
    }
    b->s2_total = s2;
    b->s3_total = s3;
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

// Resets the queues with a kernel (not a memset node: it is captured into graphs).
int launch_forward(conv2d_batch* b, hipStream_t st) {
    const int nl = (int)b->layers.size();
    hipLaunchKernelGGL(wino_batch_reset, dim3(1), dim3(64), 0, st, b->d_counters);
    const int g2 = std::min(b->s2_total, 256 * 4);
    hipLaunchKernelGGL(wino_batch_input_transform, dim3(g2), dim3(256), b->s2_lds, st,
                       (const BatchLayer*)b->d_layers, nl, b->s2_total, b->d_counters);
    const int g3 = std::min(b->s3_total, 256 * 2);
    hipLaunchKernelGGL(wino_batch_gemm, dim3(g3), dim3(256), 0, st, (const BatchLayer*)b->d_layers, nl,
                       b->s3_total, b->d_counters + 1);
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
    b->bcfg.resize(nlayers);
    for (int i = 0; i < nlayers; ++i) {
        const int t = global_to_batch(wino::tuned_config_lookup(&layers[i]));
        b->bcfg[i] = t >= 0 ? t : batch_heuristic(layers[i]);
    }
    int rc = CONV2D_OK;
    if (hipMalloc(&b->d_counters, 2 * sizeof(unsigned)) != hipSuccess) rc = CONV2D_ERR_HIP;
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

// Filter transforms of every layer (once per weight set).
int conv2d_batch_prepare_weights(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    for (const problem_t& pb : b->layers) {
        int rc = conv2d_prepare_weights(&pb, stream);
        if (rc) return rc;
    }
    return CONV2D_OK;
}

// Input transforms and GEMMs of every layer: one persistent launch each (plus a queue reset).
int conv2d_batch_forward(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    return launch_forward(b, (hipStream_t)stream);
}

// The same work as separate per-layer launches (S2 then S3 for each layer, same tile
// configurations): the baseline the batched form is compared against.
int conv2d_batch_forward_separate(conv2d_batch* b, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    for (size_t i = 0; i < b->layers.size(); ++i) {
        int rc = conv2d_forward(&b->layers[i], kBatchToGlobal[b->bcfg[i]], stream);
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

// Per-layer tile configurations from the autotuner, restricted to the batched set. Each
// layer's U and V must be prepared; every layer's output is overwritten.
int conv2d_batch_autotune(conv2d_batch* b, int reps, void* stream) {
    if (!b) return CONV2D_ERR_NULL;
    for (size_t i = 0; i < b->layers.size(); ++i) {
        int best = -1;
        int rc = wino::autotune_subset(&b->layers[i], reps, stream, kBatchToGlobal, kBatchCfgs, &best, nullptr);
        if (rc) return rc;
        b->bcfg[i] = global_to_batch(best);
    }
    drop_graph(b);
    return upload(b);
}

// Batched GEMM configuration of layer i (as a global configuration index).
int conv2d_batch_layer_config(const conv2d_batch* b, int i, int* cfg) {
    if (!b || !cfg) return CONV2D_ERR_NULL;
    if (i < 0 || i >= (int)b->layers.size()) return CONV2D_ERR_INVALID;
    *cfg = kBatchToGlobal[b->bcfg[i]];
    return CONV2D_OK;
}

}  // extern "C"
