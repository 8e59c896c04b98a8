// conv2dfp16demo — command-line harness for the plugin.
//
// Counterpart of the reference harness (src/main.cpp:8-166), same positional arguments
// "n c h w k r s u v p q" (main.cpp:10-20). Like the reference it is the plugin's caller: it
// queries getParamsize / getkernelInfo and launches the four slots itself, each with
// hipExtLaunchKernel and the parameter blob as the only argument (main.cpp:88-108, 121-124).
// Changes:
//   * inputs come from a seeded SplitMix64 generator instead of unseeded rand()
//     (main.cpp:40-48); --dist positive reproduces the reference's (rand()%255)/255 range;
//   * workspaces (mid_result included, for the F(4x4) form) are sized from
//     conv2d_workspace_size, not from 4*h*h*n*c etc. (main.cpp:33-38), which is too small for
//     ragged shapes;
//   * every HIP call is checked; the plugin's return codes are checked;
//   * timing runs on one created stream: the reference protocol (all four slots, filter
//     transform included, main.cpp:112-131) and the steady state (S2 + S3, weights
//     prepared once);
//   * verification uses the FP64 direct-convolution GPU kernel, reporting normwise max and
//     RMS relative error, plus the reference's own elementwise criterion
//     (main.cpp:140-151) on positive data.
// Prints one JSON line.
#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "conv2d.h"

#define HIPCHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                        \
            exit(2);                                                                  \
        }                                                                             \
    } while (0)

static uint64_t splitmix64(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unit(uint64_t seed, uint64_t i) { return (double)(splitmix64(seed, i) >> 11) * 0x1p-53; }

// The four slots in order, launched by the caller as the reference harness does
// (main.cpp:105-108): grid and block from the slot, dynamic LDS from the slot, the blob as the
// kernel's only (by-value) argument, on `st` (the reference used the null stream).
static void launch_slots(const kernelInfo_t& ki, void* param, hipStream_t st) {
    const dim3 groups[4] = {dim3(ki.blockx1, ki.blocky1, ki.blockz1), dim3(ki.blockx2, ki.blocky2, ki.blockz2),
                            dim3(ki.blockx3, ki.blocky3, ki.blockz3), dim3(ki.blockx4, ki.blocky4, ki.blockz4)};
    const dim3 threads[4] = {dim3(ki.threadx1, ki.thready1, ki.threadz1), dim3(ki.threadx2, ki.thready2, ki.threadz2),
                             dim3(ki.threadx3, ki.thready3, ki.threadz3), dim3(ki.threadx4, ki.thready4, ki.threadz4)};
    const unsigned lds[4] = {ki.dynmicLdsSize1, ki.dynmicLdsSize2, ki.dynmicLdsSize3, ki.dynmicLdsSize4};
    void* const ptrs[4] = {ki.kernelPtr1, ki.kernelPtr2, ki.kernelPtr3, ki.kernelPtr4};
    for (int i = 0; i < 4; ++i)
        HIPCHECK(hipExtLaunchKernel(ptrs[i], groups[i], threads[i], (void**)&param, lds[i], st, nullptr, nullptr, 0));
}

int main(int argc, char** argv) {
    if (argc < 12) {
        fprintf(stderr,
                "usage: %s n c h w k r s u v p q [--seed S] [--iters I] [--dist uniform|positive]\n",
                argv[0]);
        return 1;
    }
    unsigned d[11];
    for (int i = 0; i < 11; ++i) d[i] = (unsigned)atoi(argv[1 + i]);
    uint64_t seed = 1;
    int iters = 100;
    bool positive = false;
    for (int i = 12; i < argc; ++i) {
        if (!strcmp(argv[i], "--seed") && i + 1 < argc) seed = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--iters") && i + 1 < argc) iters = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--dist") && i + 1 < argc) positive = !strcmp(argv[++i], "positive");
    }
    problem_t pb;
    memset(&pb, 0, sizeof pb);
    pb.n = d[0]; pb.c = d[1]; pb.h = d[2]; pb.w = d[3]; pb.k = d[4];
    pb.r = d[5]; pb.s = d[6]; pb.u = d[7]; pb.v = d[8]; pb.p = d[9]; pb.q = d[10];
    const int oh = ((int)pb.h - (int)pb.r + 2 * (int)pb.p) / (int)pb.u + 1;
    const int ow = ((int)pb.w - (int)pb.s + 2 * (int)pb.q) / (int)pb.v + 1;
    const size_t n_in = (size_t)pb.n * pb.c * pb.h * pb.w;
    const size_t n_w = (size_t)pb.k * pb.c * pb.r * pb.s;
    const size_t n_out = (size_t)pb.n * pb.k * oh * ow;

    size_t v_bytes = 0, u_bytes = 0, m_bytes = 0;
    int rc = conv2d_workspace_size(&pb, &v_bytes, &u_bytes, &m_bytes);
    if (rc) {
        printf("{\"error\": %d, \"where\": \"conv2d_workspace_size\"}\n", rc);
        return 3;
    }

    std::vector<_Float16> h_in(n_in), h_w(n_w), h_out(n_out);
    for (size_t i = 0; i < n_in; ++i)
        h_in[i] = (_Float16)(positive ? (double)(splitmix64(seed, i) % 255) / 255.0
                                      : 2.0 * unit(seed, i) - 1.0);
    for (size_t i = 0; i < n_w; ++i)
        h_w[i] = (_Float16)(positive ? (double)(splitmix64(seed + 1, i) % 255) / 255.0
                                     : 2.0 * unit(seed + 1, i) - 1.0);

    _Float16 *d_in, *d_w, *d_out, *d_v, *d_u, *d_m = nullptr;
    double* d_ref;
    HIPCHECK(hipMalloc(&d_in, n_in * 2));
    HIPCHECK(hipMalloc(&d_w, n_w * 2));
    HIPCHECK(hipMalloc(&d_out, n_out * 2));
    HIPCHECK(hipMalloc(&d_v, v_bytes));
    HIPCHECK(hipMalloc(&d_u, u_bytes));
    if (m_bytes) HIPCHECK(hipMalloc(&d_m, m_bytes));  // the F(4x4) form's M, where that form applies
    HIPCHECK(hipMalloc(&d_ref, n_out * sizeof(double)));
    HIPCHECK(hipMemcpy(d_in, h_in.data(), n_in * 2, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_w, h_w.data(), n_w * 2, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(d_out, 0, n_out * 2));
    pb.in = d_in; pb.weight = d_w; pb.out = d_out;
    pb.transform_pin = d_v; pb.transform_filter = d_u; pb.mid_result = d_m;

    int psize = 0;
    if ((rc = getParamsize(&pb, &psize))) return 3;
    void* param = malloc(psize);
    kernelInfo_t ki;
    if ((rc = getkernelInfo(&pb, &ki, param))) {
        printf("{\"error\": %d, \"where\": \"getkernelInfo\"}\n", rc);
        return 3;
    }
    hipStream_t st;
    HIPCHECK(hipStreamCreate(&st));
    launch_slots(ki, param, st);  // warm-up, and the result that is checked
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipMemcpy(h_out.data(), d_out, n_out * 2, hipMemcpyDeviceToHost));

    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    float ms_ref = 0.f, ms_steady = 0.f;
    HIPCHECK(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i) launch_slots(ki, param, st);
    HIPCHECK(hipEventRecord(e1, st));
    HIPCHECK(hipEventSynchronize(e1));
    HIPCHECK(hipEventElapsedTime(&ms_ref, e0, e1));
    HIPCHECK(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i)
        if ((rc = conv2d_forward(&pb, -1, st))) return 3;
    HIPCHECK(hipEventRecord(e1, st));
    HIPCHECK(hipEventSynchronize(e1));
    HIPCHECK(hipEventElapsedTime(&ms_steady, e0, e1));

    // FP64 validation
    if ((rc = conv2d_direct_f64(&pb, d_ref, st))) return 3;
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<double> ref(n_out);
    HIPCHECK(hipMemcpy(ref.data(), d_ref, n_out * sizeof(double), hipMemcpyDeviceToHost));
    double max_err = 0, max_ref = 0, se = 0, sr = 0;
    long ref_criterion_errors = 0;
    for (size_t i = 0; i < n_out; ++i) {
        const double y = (double)h_out[i];
        const double e = fabs(y - ref[i]);
        if (!(e <= max_err)) max_err = e;  // NaN propagates
        if (fabs(ref[i]) > max_ref) max_ref = fabs(ref[i]);
        se += e * e;
        sr += ref[i] * ref[i];
        // reference acceptance test (main.cpp:144): |gpu - cpu| / gpu > 1% or NaN/Inf
        if (fabs(y - ref[i]) / y > 0.01 || std::isnan(y) || std::isinf(y)) ++ref_criterion_errors;
    }
    const double flops = 2.0 * pb.n * pb.k * pb.c * pb.r * pb.s * (double)oh * ow;
    const double us_ref = ms_ref * 1000.0 / iters, us_steady = ms_steady * 1000.0 / iters;
    int cfg = -1, form = -1, method = -1;
    conv2d_select_tile_config(&pb, &cfg);
    conv2d_form(&pb, &form);
    conv2d_method(&pb, &method);
    printf("{\"shape\": [%u,%u,%u,%u,%u,%u,%u,%u,%u,%u,%u], \"seed\": %llu, \"dist\": \"%s\", "
           "\"method\": \"%s\", \"form\": \"%s\", \"cfg\": %d, \"param_bytes\": %d, "
           "\"us_reference_protocol\": %.3f, \"us_steady\": %.3f, "
           "\"tflops_reference_protocol\": %.3f, \"tflops_steady\": %.3f, "
           "\"max_rel\": %.3e, \"rms_rel\": %.3e, \"reference_criterion_errors\": %ld}\n",
           pb.n, pb.c, pb.h, pb.w, pb.k, pb.r, pb.s, pb.u, pb.v, pb.p, pb.q,
           (unsigned long long)seed, positive ? "positive" : "uniform",
           method == CONV2D_METHOD_DIRECT ? "direct" : "winograd", form == CONV2D_FORM_FUSED ? "fused" : "staged",
           cfg, psize, us_ref, us_steady,
           flops / us_ref * 1e-6, flops / us_steady * 1e-6, max_err / (max_ref > 0 ? max_ref : 1),
           sqrt(se / (sr > 0 ? sr : 1)), ref_criterion_errors);
    free(param);
    HIPCHECK(hipFree(d_in)); HIPCHECK(hipFree(d_w)); HIPCHECK(hipFree(d_out));
    HIPCHECK(hipFree(d_v)); HIPCHECK(hipFree(d_u)); HIPCHECK(hipFree(d_ref));
    if (d_m) HIPCHECK(hipFree(d_m));
    return 0;
}
