// Tuned tile-configuration table consulted by getkernelInfo.
//
// Replaces the reference's hand-written (n, c) if/else dispatch (src/conv2d.cpp:273-494):
// a table of exact shapes -> GEMM tile configuration, filled by the autotuner (or loaded
// from its saved file). Shapes not in the table fall back to the shape heuristic.
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "plugin_internal.h"
#include "wino_gemm.h"

namespace {

struct TunedEntry {
    unsigned n, c, h, w, k, p, q;
    int cfg;
    int form;  // CONV2D_FORM_STAGED, CONV2D_FORM_FUSED or CONV2D_FORM_F43 (cfg then indexes its table)
};

std::mutex g_mu;
std::vector<TunedEntry> g_table;

bool same_shape(const TunedEntry& e, const problem_t* pb) {
    return e.n == pb->n && e.c == pb->c && e.h == pb->h && e.w == pb->w && e.k == pb->k &&
           e.p == pb->p && e.q == pb->q;
}

}  // namespace

namespace wino {

int tuned_config_lookup(const problem_t* pb) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& e : g_table)
        if (same_shape(e, pb)) return e.cfg;
    return -1;
}

int tuned_form_lookup(const problem_t* pb) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& e : g_table)
        if (same_shape(e, pb)) return e.form;
    return -1;
}

}  // namespace wino

extern "C" {

// Whether cfg indexes the GEMM table of `form` (the F(4x4) form has its own).
static bool valid_entry(int cfg, int form) {
    if (form == CONV2D_FORM_F43) return cfg >= 0 && cfg < kNumPosConfigs;
    return (form == CONV2D_FORM_STAGED || form == CONV2D_FORM_FUSED) && cfg >= 0 && cfg < kNumGemmConfigs;
}

// Records (or replaces) the form and its GEMM configuration for one exact shape.
int conv2d_tuned_set_form(const problem_t* pb, int cfg, int form) {
    if (!pb) return CONV2D_ERR_NULL;
    if (!valid_entry(cfg, form)) return CONV2D_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& e : g_table)
        if (same_shape(e, pb)) {
            e.cfg = cfg;
            e.form = form;
            return CONV2D_OK;
        }
    g_table.push_back({pb->n, pb->c, pb->h, pb->w, pb->k, pb->p, pb->q, cfg, form});
    return CONV2D_OK;
}

// Records (or replaces) the configuration for one exact shape, in the staged form.
int conv2d_tuned_set(const problem_t* pb, int cfg) { return conv2d_tuned_set_form(pb, cfg, CONV2D_FORM_STAGED); }

// Empties the table (every shape falls back to the heuristic).
int conv2d_tuned_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    return CONV2D_OK;
}

int conv2d_tuned_count(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (int)g_table.size();
}

// Text format, one shape per line: "n c h w k p q cfg form" (form 0 staged, 1 fused, 2 F(4x4); files
// without the form column load as staged). Lines starting with '#' are comments.
int conv2d_tuned_save(const char* path) {
    if (!path) return CONV2D_ERR_NULL;
    FILE* f = fopen(path, "w");
    if (!f) return CONV2D_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    fprintf(f, "# winoconv tuned tile configurations: n c h w k p q cfg form\n");
    for (const auto& e : g_table)
        fprintf(f, "%u %u %u %u %u %u %u %d %d\n", e.n, e.c, e.h, e.w, e.k, e.p, e.q, e.cfg, e.form);
    fclose(f);
    return CONV2D_OK;
}

int conv2d_tuned_load(const char* path) {
    if (!path) return CONV2D_ERR_NULL;
    FILE* f = fopen(path, "r");
    if (!f) return CONV2D_ERR_INVALID;
    char line[256];
    std::vector<TunedEntry> rows;
    int rc = CONV2D_OK;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        TunedEntry e;
        e.form = CONV2D_FORM_STAGED;
        const int got = sscanf(line, "%u %u %u %u %u %u %u %d %d", &e.n, &e.c, &e.h, &e.w, &e.k, &e.p,
                               &e.q, &e.cfg, &e.form);
        if ((got != 8 && got != 9) || !valid_entry(e.cfg, e.form)) {
            rc = CONV2D_ERR_INVALID;
            break;
        }
        rows.push_back(e);
    }
    fclose(f);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& r : rows) {
        bool found = false;
        for (auto& e : g_table)
            if (e.n == r.n && e.c == r.c && e.h == r.h && e.w == r.w && e.k == r.k &&
                e.p == r.p && e.q == r.q) {
                e.cfg = r.cfg;
                e.form = r.form;
                found = true;
            }
        if (!found) g_table.push_back(r);
    }
    return CONV2D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Autotuner: exhaustive, deterministic best-of-reps timing of every GEMM configuration,
// behind a parity gate.
// ---------------------------------------------------------------------------------------
__global__ void tuner_count_mismatch(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                                     size_t n, unsigned* __restrict__ count) {
    unsigned local = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        local += a[i] != b[i];
    if (local) atomicAdd(count, local);
}

// max |a - b| and max |b| over n FP16 values, as the bit patterns of non-negative floats (which
// order like the floats; a NaN difference orders above every finite one): the fused form's gate.
__global__ void tuner_max_abs_diff(const _Float16* __restrict__ a, const _Float16* __restrict__ b, size_t n,
                                   unsigned* __restrict__ out2) {
    float d = 0.f, m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = (float)a[i], y = (float)b[i];
        const float e = fabsf(x - y);
        d = (e > d || e != e) ? e : d;
        m = fmaxf(m, fabsf(y));
    }
    if (d != 0.f) atomicMax(out2, __float_as_uint(d));
    if (m != 0.f) atomicMax(out2 + 1, __float_as_uint(m));
}

namespace {

struct TunerBuffers {
    uint16_t* golden = nullptr;
    unsigned* count = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TunerBuffers() {
        if (golden) (void)hipFree(golden);
        if (count) (void)hipFree(count);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

}  // namespace

namespace wino {

int time_launch_us(const Launch& L, mykernelParamType* a, int reps, hipStream_t st, float* us) {
    if (reps < 1 || !us) return CONV2D_ERR_INVALID;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CONV2D_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = CONV2D_ERR_HIP;
    if (!rc) rc = launch(L, a, st);
    float best = -1.f;
    for (int r = 0; r < reps && !rc; ++r) {
        float ms = 0.f;
        if (hipEventRecord(e0, st) != hipSuccess || (rc = launch(L, a, st)) != CONV2D_OK ||
            hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = rc ? rc : CONV2D_ERR_HIP;
            break;
        }
        if (best < 0.f || ms < best) best = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!rc) *us = best * 1e3f;
    return rc;
}

int autotune_subset(const problem_t* pb, int reps, void* stream, const int* cfgs, int ncfg, int* best_cfg,
                    float* times_us) {
    if (!pb || !pb->out || !pb->transform_pin || !pb->transform_filter) return CONV2D_ERR_NULL;
    if (reps < 1 || ncfg < 1 || !cfgs) return CONV2D_ERR_INVALID;
    for (int i = 0; i < ncfg; ++i)
        if (cfgs[i] < 0 || cfgs[i] >= kNumGemmConfigs) return CONV2D_ERR_INVALID;
    Geometry g;
    int rc = require_staged(pb, &g);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t nout = (size_t)pb->n * pb->k * g.Oh * g.Ow;
    TunerBuffers B;
    if (hipMalloc(&B.golden, nout * 2) != hipSuccess || hipMalloc(&B.count, sizeof(unsigned)) != hipSuccess ||
        hipEventCreate(&B.e0) != hipSuccess || hipEventCreate(&B.e1) != hipSuccess)
        return CONV2D_ERR_HIP;
    mykernelParamType a;
    int best = -1;
    float best_t = 0.f;
    for (int ci = 0; ci < ncfg; ++ci) {
        const int cfg = cfgs[ci];
        fill_param(pb, g, cfg, &a);
        const Launch L = slot_gemm(g, cfg);
        if ((rc = launch(L, &a, st))) return rc;  // warm-up, and the output to check
        bool pass = true;
        if (ci == 0) {
            if (hipMemcpyAsync(B.golden, pb->out, nout * 2, hipMemcpyDeviceToDevice, st) != hipSuccess)
                return CONV2D_ERR_HIP;
        } else {
            unsigned h = 0;
            if (hipMemsetAsync(B.count, 0, sizeof(unsigned), st) != hipSuccess) return CONV2D_ERR_HIP;
            hipLaunchKernelGGL(tuner_count_mismatch, dim3(1024), dim3(256), 0, st,
                               (const uint16_t*)B.golden, (const uint16_t*)pb->out, nout, B.count);
            if (hipMemcpyAsync(&h, B.count, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return CONV2D_ERR_HIP;
            pass = h == 0;
        }
        float t_min = -1.f;
        if (pass) {
            for (int r = 0; r < reps; ++r) {
                float ms = 0.f;
                if (hipEventRecord(B.e0, st) != hipSuccess) return CONV2D_ERR_HIP;
                if ((rc = launch(L, &a, st))) return rc;
                if (hipEventRecord(B.e1, st) != hipSuccess || hipEventSynchronize(B.e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, B.e0, B.e1) != hipSuccess)
                    return CONV2D_ERR_HIP;
                if (t_min < 0.f || ms < t_min) t_min = ms;
            }
            if (best < 0 || t_min < best_t) {
                best = cfg;
                best_t = t_min;
            }
        }
        if (times_us) times_us[ci] = pass ? t_min * 1000.f : -1.f;
    }
    if (best < 0) return CONV2D_ERR_INVALID;
    if (best_cfg) *best_cfg = best;
    return CONV2D_OK;
}

}  // namespace wino

// Whether the output the problem's out holds is within the library's accuracy bound of the
// staged output in `golden`: max |out - staged| <= 1e-2 * max |staged| (the fused and F(4x4)
// forms round differently from the staged form and are not bit-identical to it).
static int gate_against(const problem_t* pb, const wino::Geometry& g, hipStream_t st, const uint16_t* golden,
                        bool* pass) {
    *pass = false;
    const size_t nout = (size_t)pb->n * pb->k * g.Oh * g.Ow;
    unsigned* d2 = nullptr;
    unsigned h2[2] = {0u, 0u};
    if (hipMalloc(&d2, 2 * sizeof(unsigned)) != hipSuccess) return CONV2D_ERR_HIP;
    int rc = CONV2D_OK;
    if (hipMemsetAsync(d2, 0, 2 * sizeof(unsigned), st) != hipSuccess) rc = CONV2D_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(tuner_max_abs_diff, dim3(1024), dim3(256), 0, st, (const _Float16*)pb->out,
                           (const _Float16*)golden, nout, d2);
        if (hipMemcpyAsync(h2, d2, sizeof h2, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = CONV2D_ERR_HIP;
    }
    (void)hipFree(d2);
    if (rc) return rc;
    float dmax, ymax;
    memcpy(&dmax, &h2[0], 4);
    memcpy(&ymax, &h2[1], 4);
    *pass = dmax <= 1e-2f * ymax;
    return CONV2D_OK;
}

// Times the fused kernel on the problem, gated against the staged output in `golden`.
// *us = best of reps, or -1 when the fused form does not apply or fails the gate.
static int time_fused(const problem_t* pb, const wino::Geometry& g, int reps, hipStream_t st,
                      const uint16_t* golden, float* us) {
    *us = -1.f;
    if (!g.fused_ok) return CONV2D_OK;
    mykernelParamType a;
    wino::fill_param(pb, g, 0, &a);
    a.form = CONV2D_FORM_FUSED;
    float t = 0.f;
    int rc = wino::time_launch_us(wino::slot_fused(g), &a, reps, st, &t);
    if (rc) return rc;
    bool pass = false;
    if ((rc = gate_against(pb, g, st, golden, &pass))) return rc;
    if (pass) *us = t;
    return CONV2D_OK;
}

// Times the F(4x4) form: its filter transform (into transform_filter), input transform, every
// GEMM configuration (each bit-identical to the first, as all configurations must be) and
// output transform; the output is gated against the staged output in `golden`. *us = input
// transform + best GEMM + output transform, or -1 when the form does not apply (or mid_result
// is not set) or fails the gate; *best_cfg its fastest GEMM configuration.
static int time_f43(const problem_t* pb, const wino::Geometry& g, int reps, hipStream_t st, const uint16_t* golden,
                    float* us, int* best_cfg) {
    *us = -1.f;
    *best_cfg = -1;
    if (!g.f43_ok || !pb->mid_result) return CONV2D_OK;
    mykernelParamType a;
    wino::fill_param43(pb, g, 0, &a);
    int rc = wino::prepare_weights_form(pb, g, CONV2D_FORM_F43, st);
    float t_in = 0.f, t_out = 0.f;
    if (!rc) rc = wino::time_launch_us(wino::slot_input43(pb, g), &a, reps, st, &t_in);
    if (rc) return rc;
    const size_t mhalves = g.m43_bytes / 2;
    uint16_t* first = nullptr;
    unsigned* count = nullptr;
    if (hipMalloc(&first, g.m43_bytes) != hipSuccess || hipMalloc(&count, sizeof(unsigned)) != hipSuccess) {
        (void)hipFree(first);
        return CONV2D_ERR_HIP;
    }
    float best_t = -1.f;
    for (int cfg = 0; cfg < kNumPosConfigs && !rc; ++cfg) {
        wino::fill_param43(pb, g, cfg, &a);
        float t = 0.f;
        if ((rc = wino::time_launch_us(wino::slot_gemm43(g, cfg), &a, reps, st, &t))) break;
        bool same = true;
        if (cfg == 0) {
            if (hipMemcpyAsync(first, pb->mid_result, g.m43_bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
                rc = CONV2D_ERR_HIP;
        } else {
            unsigned h = 0;
            if (hipMemsetAsync(count, 0, sizeof(unsigned), st) != hipSuccess) rc = CONV2D_ERR_HIP;
            if (!rc) {
                hipLaunchKernelGGL(tuner_count_mismatch, dim3(1024), dim3(256), 0, st, (const uint16_t*)first,
                                   (const uint16_t*)pb->mid_result, mhalves, count);
                if (hipMemcpyAsync(&h, count, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess)
                    rc = CONV2D_ERR_HIP;
            }
            same = h == 0;
        }
        if (!rc && same && (best_t < 0.f || t < best_t)) {
            best_t = t;
            *best_cfg = cfg;
        }
    }
    (void)hipFree(first);
    (void)hipFree(count);
    if (rc) return rc;
    if (*best_cfg < 0) return CONV2D_OK;
    // M of the best configuration, then the output transform (timed) and the gate
    wino::fill_param43(pb, g, *best_cfg, &a);
    if ((rc = wino::launch(wino::slot_gemm43(g, *best_cfg), &a, st))) return rc;
    if ((rc = wino::time_launch_us(wino::slot_output43(g, (int)pb->k), &a, reps, st, &t_out))) return rc;
    bool pass = false;
    if ((rc = gate_against(pb, g, st, golden, &pass))) return rc;
    if (pass) *us = t_in + best_t + t_out;
    return CONV2D_OK;
}

extern "C" int conv2d_autotune(const problem_t* pb, int reps, void* stream, int* best_cfg,
                               float* times_us, int ntimes) {
    std::vector<int> all(kNumGemmConfigs);
    for (int i = 0; i < kNumGemmConfigs; ++i) all[i] = i;
    std::vector<float> t(kNumGemmConfigs, -1.f);
    int best = -1;
    int rc = wino::autotune_subset(pb, reps, stream, all.data(), kNumGemmConfigs, &best, t.data());
    if (rc) return rc;
    wino::Geometry g;
    if ((rc = wino::require_staged(pb, &g))) return rc;
    hipStream_t st = (hipStream_t)stream;
    // the staged output of the winning configuration is the fused form's reference
    const size_t nout = (size_t)pb->n * pb->k * g.Oh * g.Ow;
    uint16_t* golden = nullptr;
    if (hipMalloc(&golden, nout * 2) != hipSuccess) return CONV2D_ERR_HIP;
    mykernelParamType a;
    wino::fill_param(pb, g, best, &a);
    float t_s2 = -1.f, t_fused = -1.f, t_f43 = -1.f;
    int cfg43 = -1;
    rc = wino::launch(wino::slot_gemm(g, best), &a, st);
    if (!rc && hipMemcpyAsync(golden, pb->out, nout * 2, hipMemcpyDeviceToDevice, st) != hipSuccess) rc = CONV2D_ERR_HIP;
    if (!rc) rc = wino::time_launch_us(wino::slot_input(pb, g), &a, reps, st, &t_s2);
    if (!rc) rc = time_fused(pb, g, reps, st, golden, &t_fused);
    if (!rc) rc = time_f43(pb, g, reps, st, golden, &t_f43, &cfg43);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = CONV2D_ERR_HIP;
    (void)hipFree(golden);
    if (rc) return rc;
    if (times_us) {
        for (int i = 0; i < ntimes && i < kNumGemmConfigs; ++i) times_us[i] = t[i];
        if (ntimes > kNumGemmConfigs) times_us[kNumGemmConfigs] = t_s2;
        if (ntimes > kNumGemmConfigs + 1) times_us[kNumGemmConfigs + 1] = t_fused;
        if (ntimes > kNumGemmConfigs + 2) times_us[kNumGemmConfigs + 2] = t_f43;
    }
    if (best_cfg) *best_cfg = best;
    // the fastest of: input transform + best GEMM (staged), the fused kernel, the F(4x4) form
    int form = CONV2D_FORM_STAGED, cfg = best;
    float t_best = t_s2 + t[best];
    if (t_fused > 0.f && t_fused < t_best) {
        form = CONV2D_FORM_FUSED;
        t_best = t_fused;
    }
    if (t_f43 > 0.f && t_f43 < t_best) {
        form = CONV2D_FORM_F43;
        cfg = cfg43;
    }
    if ((rc = conv2d_tuned_set_form(pb, cfg, form))) return rc;
    // transform_filter holds the recorded form's U (the F(4x4) timing overwrote it)
    if (g.f43_ok && pb->mid_result) {
        if ((rc = wino::prepare_weights_form(pb, g, form, st))) return rc;
        if (hipStreamSynchronize(st) != hipSuccess) return CONV2D_ERR_HIP;
    }
    return CONV2D_OK;
}
