// Internal host-side declarations shared by the plugin translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "conv2d.h"
#include "mykernelParamType.h"

namespace wino {

struct Geometry {
    int Oh, Ow;
    bool winograd_ok;      // 3x3, stride 1, pad <= 2
    bool direct;           // runs on the general-geometry (implicit-GEMM direct) path
    bool pointwise;        // direct path with a 1x1 filter: the GEMM reads the input itself
    bool point_quads;      // ... in 8-byte pieces along its rows (stride 1, no padding, N,C,H,W,
                           // H*W % 4 == 0, 8-byte aligned input)
    int tiles_h, tiles_w;  // Winograd: 2x2 output tiles; direct: Oh, Ow
    long long T, NT, G32;  // GEMM columns per image / in total, 32-column blocks of V
    int CB, KB32;          // CB: 32-channel blocks; KB32: 32-row blocks of U
    int band;
    int nsteps;            // GEMM steps of 32 reduction values: 16 * CB, or ceil(r*s*c/32) padded to 8
    size_t v_bytes, u_bytes;
    // fused form: block of fbh x fbw tiles, fnbh x fnbw blocks per image (fused_ok: even w)
    bool fused_ok;
    bool staged_ok;  // Winograd path with an input-transform band (the staged form applies)
    int flw;    // staging load width (FusedLoad: 2-, 4- or 8-byte loads per channel)
    int fbh, fbw, fnbh, fnbw, fpr, fpc, fsh, fnu;
    // workgroups of the layer's launch, single-layer or its share of a grouped one (wide tail
    // items, when chosen, counted once per 128-tile block), and the count with half-block tail
    // items (one per 64-channel block and 64-tile block)
    long long fitems, fgroup_items;
    int io;  // CONV2D_IO_* tensor layouts and output type
    // F(4x4,3x3) form: 4x4 output tiles (t4h x t4w per image), input-transform band of band4
    // tile rows (4*band4 + 2 input rows in LDS); f43_ok: the band fits
    bool f43_ok;
    int t4h, t4w, band4;
    long long T4, NT4, G32_4;
    size_t v43_bytes, u43_bytes, m43_bytes;
};

struct Launch {
    const void* kernel;
    dim3 grid;
    dim3 block;
    unsigned lds;
};

// Validates a problem and derives its geometry (Winograd or general-geometry path) for the
// tensor layouts and output type of io (CONV2D_IO_*; derive_geometry: N,C,H,W FP16).
int derive_geometry_io(const problem_t* pb, int io, Geometry* g);
int derive_geometry(const problem_t* pb, Geometry* g);
// As derive_geometry, but only for problems on the Winograd path (else CONV2D_ERR_UNSUPPORTED).
int require_winograd(const problem_t* pb, Geometry* g);
// As require_winograd, and the staged form applies (an input-transform band fits in LDS).
int require_staged(const problem_t* pb, Geometry* g);
int heuristic_config(const problem_t* pb, const Geometry& g);
int select_config(const problem_t* pb, const Geometry& g);
void fill_param(const problem_t* pb, const Geometry& g, int cfg, mykernelParamType* a);
Launch slot_filter(const Geometry& g);
Launch slot_fused(const Geometry& g);
Launch slot_input(const problem_t* pb, const Geometry& g);
Launch slot_gemm(const Geometry& g, int cfg);
Launch slot_noop();
int launch(const Launch& L, mykernelParamType* a, hipStream_t stream);

// Tuned-table lookup (tuner.hip): configuration index / form (CONV2D_FORM_*) recorded for this
// exact shape, or -1.
int tuned_config_lookup(const problem_t* pb);
int tuned_form_lookup(const problem_t* pb);

// Autotuner over a subset of GEMM configurations (tuner.hip): best-of-reps timing of each,
// gated on bit-identity with the first one's output; does not touch the tuned table.
// times_us[i] = best time of cfgs[i] (us) or -1 if gated out.
int autotune_subset(const problem_t* pb, int reps, void* stream, const int* cfgs, int ncfg, int* best_cfg,
                    float* times_us);
// Whether getkernelInfo runs this Winograd-path problem in the fused form.
bool use_fused(const problem_t* pb, const Geometry& g);
// The form getkernelInfo and conv2d_forward use: CONV2D_FORM_F43 for a shape the tuned table
// records as such (when the F(4x4) form applies and mid_result is set), else fused or staged.
int select_form(const problem_t* pb, const Geometry& g);
// F(4x4,3x3) form: kernel argument, configuration choice (kPosConfigs) and launch slots.
void fill_param43(const problem_t* pb, const Geometry& g, int cfg, mykernelParamType* a);
int select_config43(const problem_t* pb, const Geometry& g);
Launch slot_filter43(const Geometry& g);
Launch slot_input43(const problem_t* pb, const Geometry& g);
Launch slot_gemm43(const Geometry& g, int cfg);
Launch slot_output43(const Geometry& g, int k);
// The filter transform of `form` (F(2x2) for staged and fused, F(4x4) for CONV2D_FORM_F43).
int prepare_weights_form(const problem_t* pb, const Geometry& g, int form, hipStream_t st);
// Fastest of reps launches of L with argument a on stream st, in microseconds (one warm-up).
int time_launch_us(const Launch& L, mykernelParamType* a, int reps, hipStream_t st, float* us);

}  // namespace wino
