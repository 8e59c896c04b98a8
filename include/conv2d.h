/*
 * conv2d.h — C-ABI of the MI355X (gfx950) FP16 Winograd F(2x2,3x3) convolution plugin.
 *
 * Drop-in for the contest plugin interface of the reference
 * (include/conv2d.h:8-70 in Yhorm26/2024PRA_CONV): the same problem_t and kernelInfo_t
 * layouts (field order and types unchanged) and the same getParamsize / getkernelInfo
 * signatures, now with C linkage. Every function below returns 0 on success and a
 * CONV2D_ERR_* code otherwise; nothing is launched on error.
 *
 * Tensor layouts follow the reference's spec (readme.txt:21-24): input NCHW, weight KCRS,
 * output N,K,Oh,Ow, all FP16, with Oh = (h - r + 2p)/u + 1 and Ow = (w - s + 2q)/v + 1
 * (readme.txt:17-19).
 *
 * The Winograd path takes r = s = 3, u = v = 1 and any pad p, q in [0, 2]. Every other valid
 * geometry runs on the general-geometry path, an implicit-GEMM direct convolution behind the
 * same four slots (the reference computed garbage for them, conv2d.cpp:254-494);
 * conv2d_method reports which path a problem takes. The FP8, multi-layer and autotuning forms
 * are Winograd-only and return CONV2D_ERR_UNSUPPORTED for other geometry.
 */
#ifndef CONV2D_MI355X_H_
#define CONV2D_MI355X_H_

#include <stddef.h>
#include <stdint.h>

#define __in__
#define __out__
#define __in_out__

#if defined(__clang__) || (defined(__GNUC__) && __GNUC__ >= 12)
typedef _Float16 conv2d_half_t;
#else
typedef uint16_t conv2d_half_t; /* same size and alignment as _Float16 */
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
enum {
    CONV2D_OK = 0,
    CONV2D_ERR_NULL = 1,        /* a required pointer is NULL */
    CONV2D_ERR_UNSUPPORTED = 2, /* geometry the requested form does not implement (the FP8,
                                   multi-layer and autotuning forms are Winograd-only) */
    CONV2D_ERR_OVERFLOW = 3,    /* a dimension does not fit the kernel argument (short) */
    CONV2D_ERR_WORKSPACE = 4,   /* caller-declared workspace smaller than required */
    CONV2D_ERR_HIP = 5,         /* a HIP runtime call failed */
    CONV2D_ERR_INVALID = 6,     /* bad argument (zero dimension, unknown config, ...) */
    CONV2D_ERR_IO = 7,          /* a file could not be opened, read or written in full */
    CONV2D_ERR_MISMATCH = 8     /* a persisted U is not for this problem, these weights or
                                   this layout version, or its payload is corrupt */
};

/* Method a problem runs on (conv2d_method). */
enum {
    CONV2D_METHOD_WINOGRAD = 0, /* 3x3, stride 1, pad <= 2: the four-stage Winograd path */
    CONV2D_METHOD_DIRECT = 1    /* any other geometry: implicit-GEMM direct convolution */
};

/* Problem descriptor. Same layout as the reference's problem_t (conv2d.h:8-27). The three
 * workspaces are caller-owned device buffers:
 *   transform_pin    : transformed input V (size from conv2d_workspace_size)
 *   transform_filter : transformed filter U (size from conv2d_workspace_size)
 *   mid_result       : the F(4x4,3x3) form's per-position products M (size from
 *                      conv2d_workspace_size); the F(2x2) forms fuse the output transform into
 *                      the GEMM and never use it. May be NULL (the F(4x4) form is then never
 *                      chosen). */
typedef struct {
    conv2d_half_t* in;
    conv2d_half_t* weight;
    conv2d_half_t* out;
    conv2d_half_t* transform_pin;
    conv2d_half_t* transform_filter;
    conv2d_half_t* mid_result;
    unsigned int n;
    unsigned int c;
    unsigned int h;
    unsigned int w;
    unsigned int k;
    unsigned int r;
    unsigned int s;
    unsigned int u;
    unsigned int v;
    unsigned int p;
    unsigned int q;
} problem_t;

/* Four launch slots, as in the reference (conv2d.h:29-66). This implementation fills, for the
 * form conv2d_form reports:
 *   staged form (Winograd path):
 *     slot 1: filter transform  U = G g G^T        (wino_filter_transform)
 *     slot 2: input transform   V = B^T d B        (wino_input_transform)
 *     slot 3: 16-position FP16 MFMA GEMM with the output transform A^T M A fused into its
 *             epilogue (wino_gemm_*)
 *     slot 4: an empty one-workgroup kernel (the fourth reference stage, conv2d.cpp:17-250,
 *             no longer exists as a separate pass)
 *   fused form (Winograd path):
 *     slot 1: filter transform; slot 2: wino_fused_kernel (input transform, GEMM and output
 *     transform in one launch, V never written); slots 3 and 4: empty kernels
 *   general-geometry path:
 *     slot 1: filter-tap packing; slot 2: im2col input packing; slot 3: the MFMA GEMM;
 *     slot 4: an empty kernel
 * Every slot holds a valid kernel, so callers that launch all four unconditionally (as the
 * reference harness does, main.cpp:105-108) stay valid. */
typedef struct {
    unsigned int blockx1, blocky1, blockz1;
    unsigned int threadx1, thready1, threadz1;
    unsigned int dynmicLdsSize1;
    void* kernelPtr1;

    unsigned int blockx2, blocky2, blockz2;
    unsigned int threadx2, thready2, threadz2;
    unsigned int dynmicLdsSize2;
    void* kernelPtr2;

    unsigned int blockx3, blocky3, blockz3;
    unsigned int threadx3, thready3, threadz3;
    unsigned int dynmicLdsSize3;
    void* kernelPtr3;

    unsigned int blockx4, blocky4, blockz4;
    unsigned int threadx4, thready4, threadz4;
    unsigned int dynmicLdsSize4;
    void* kernelPtr4;
} kernelInfo_t;

/* ---- Reference entry points (drop-in) ------------------------------------------------ */

/* Replaces getParamsize (reference conv2d.h:69, conv2d.cpp:9-14): writes
 * sizeof(mykernelParamType) of this implementation (the reference struct plus appended
 * fields). */
int getParamsize(__in__ problem_t* problem, __out__ int* paramSize);

/* Replaces getkernelInfo (reference conv2d.h:70, conv2d.cpp:254-517): validates the
 * geometry, selects a GEMM tile configuration (the tuned table if one is loaded, else a
 * shape heuristic), fills the four launch slots and packs the kernel argument blob.
 * The caller must have allocated the workspaces at least as large as
 * conv2d_workspace_size reports. Each slot's kernel takes the blob by value as its only
 * argument, exactly as the reference harness launches them (main.cpp:105-108). */
int getkernelInfo(__in__ problem_t* problem, __out__ kernelInfo_t* kernelInfo,
                  __in_out__ void* param);

/* ---- Additive entry points (no counterpart in the reference) ------------------------- */

/* Workspace sizes in bytes for transform_pin (V), transform_filter (U) and mid_result: enough
 * for every form the library may run the problem in. mid_result is 0 bytes unless the
 * F(4x4,3x3) form (CONV2D_FORM_F43) applies; it is then that form's M, and transform_filter
 * its U (2.25x the F(2x2) U). The reference harness sized them as 4*h*h*n*c, 16*k*c and
 * 4*k*h*w*n halves (main.cpp:33-38); for the F(2x2) forms those sizes suffice when h == w is
 * even, c % 32 == 0, k % 32 == 0 and n*(h/2)*(w/2) % 32 == 0, and are too small otherwise
 * (ragged or padded shapes). */
int conv2d_workspace_size(const problem_t* problem, size_t* pin_bytes,
                          size_t* filter_bytes, size_t* mid_bytes);
/* The bytes one form uses (CONV2D_FORM_STAGED and CONV2D_FORM_FUSED: the F(2x2) V and U, or the
 * general-geometry path's packed operands; CONV2D_FORM_F43: its V, U and M;
 * CONV2D_ERR_UNSUPPORTED where that form does not apply). */
int conv2d_form_workspace_size(const problem_t* problem, int form, size_t* pin_bytes, size_t* filter_bytes,
                               size_t* mid_bytes);

/* Launches the four slots in order on a HIP stream (NULL = null stream) with
 * hipExtLaunchKernel, as the reference harness does (main.cpp:105-108). */
int conv2d_launch(const kernelInfo_t* kernelInfo, void* param, void* stream);

/* S1 only: U = G g G^T into problem->transform_filter. Run once per weight set. */
int conv2d_prepare_weights(const problem_t* problem, void* stream);

/* S2 only: V = B^T d B into problem->transform_pin. */
int conv2d_input_transform(const problem_t* problem, void* stream);

/* S3 only (GEMM + fused output transform) with tile configuration cfg
 * (-1 = the same choice getkernelInfo makes). Needs U and V prepared. */
int conv2d_gemm(const problem_t* problem, int cfg, void* stream);

/* S2 + S3 (steady state: U prepared once). cfg as for conv2d_gemm. */
int conv2d_forward(const problem_t* problem, int cfg, void* stream);

/* Allocation-free forward pass. No counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same as conv2d_forward, but the transformed input (V of the staged
 * Winograd form, the im2col input of the general-geometry path) lives in a library-owned
 * device array of conv2d_static_scratch_bytes() bytes (64 MiB; one copy per GPU, held until the
 * process ends) instead of problem->transform_pin, which is not used and may be NULL. U must be
 * prepared in transform_filter (conv2d_prepare_weights). Calls no allocating routine. A problem
 * whose transformed input exceeds the array returns CONV2D_ERR_WORKSPACE before launching
 * anything (the fused form needs no array). Every call on one GPU shares its array: calls that
 * overlap on one GPU (concurrent streams or threads) race. */
int conv2d_forward_static(const problem_t* problem, int cfg, void* stream);
size_t conv2d_static_scratch_bytes(void);

/* Output epilogue (no counterpart in the reference, whose stage 4 stores the plain
 * convolution, conv2d.cpp:37-47): every output becomes act(y + bias[k]), computed in FP32
 * before the output's one FP16 rounding. bias: K FP16 values in device memory, or NULL;
 * act: CONV2D_ACT_NONE or CONV2D_ACT_RELU (a NaN stays NaN). */
#define CONV2D_ACT_NONE 0
#define CONV2D_ACT_RELU 1
/* Sets the epilogue in a kernel-argument blob filled by getkernelInfo, so that the four
 * slots launched with it (conv2d_launch, or the harness's own launches) apply it. */
int conv2d_set_epilogue(void* param, const void* bias, int act);
/* S2 + S3 with the epilogue; cfg as for conv2d_forward (-1: getkernelInfo's choice). Works on
 * both the Winograd and the general-geometry path. */
int conv2d_forward_epilogue(const problem_t* problem, int cfg, const void* bias, int act, void* stream);

/* Tensor layouts and output type (no counterpart in the reference, whose tensors are NCHW
 * FP16, readme.txt:21-24): an `io` word of these bits. The weights stay KCRS FP16 and the
 * workspaces keep their layouts and sizes. */
#define CONV2D_IO_NCHW 0     /* input N,C,H,W and output N,K,Oh,Ow, FP16 (the reference's layout) */
#define CONV2D_IO_IN_NHWC 1  /* input N,H,W,C */
#define CONV2D_IO_OUT_NHWC 2 /* output N,Oh,Ow,K */
#define CONV2D_IO_OUT_BF16 4 /* output bfloat16, rounded to nearest even from the FP32 result */
/* getkernelInfo for an io word: the same slots and blob, with kernels for those layouts (the
 * fused form takes NHWC input in 16-byte loads of 8 channels when C % 8 == 0 and the input
 * base is 16-byte aligned; otherwise such a layer runs staged). CONV2D_ERR_INVALID for
 * unknown bits. Every path (fused, staged, general geometry) and the epilogue
 * (conv2d_set_epilogue) work with every io word. */
int getkernelInfo_io(problem_t* problem, int io, kernelInfo_t* kernelInfo, void* param);
/* conv2d_forward for an io word (cfg -1: getkernelInfo_io's form and configuration). */
int conv2d_forward_io(const problem_t* problem, int cfg, int io, void* stream);
/* conv2d_direct_f64 reading the input in the io word's input layout (the FP64 result is always
 * N,K,Oh,Ow). */
int conv2d_direct_f64_io(const problem_t* problem, int io, double* out_f64, void* stream);

/* Fused form of S2 + S3 (replaces BTdB*, preprocessing.cuh:273-424, and matrixMul*,
 * matrixmul.cuh:139-556, in one launch): the input transform runs inside the GEMM in FP16
 * arithmetic, as the reference's does, so transform_pin is not used. Needs U prepared
 * (conv2d_prepare_weights). Winograd geometry with even w and a 4-byte aligned input only;
 * CONV2D_ERR_UNSUPPORTED otherwise. */
int conv2d_fused_gemm(const problem_t* problem, void* stream);

/* Persisted transformed weights (the reference keeps U in its own transform_filter buffer,
 * conv2d.h:14, main.cpp:38, and recomputes it every run). File: a 64-byte header (magic
 * "WINOU16\0", layout version, method, k, c, r, s, FNV-1a 64 hash of the K*C*R*S FP16
 * weights, U bytes, FNV-1a 64 hash of the U bytes) followed by U exactly as
 * problem->transform_filter holds it after conv2d_prepare_weights.
 * save: U must be prepared; writes the file (synchronises the stream).
 * load: checks the header against the problem (CONV2D_ERR_MISMATCH before any device access
 * for a foreign file), hashes the problem's current weights and the payload, and only when
 * everything matches copies U into problem->transform_filter (synchronises the stream). A
 * stale cache (the weights changed) is CONV2D_ERR_MISMATCH, never silently used. */
#define CONV2D_U_LAYOUT_VERSION 1
int conv2d_weights_save(const problem_t* problem, const char* path, void* stream);
int conv2d_weights_load(const problem_t* problem, const char* path, void* stream);
/* FNV-1a 64 of host bytes, the hash the weight cache uses. */
int conv2d_weight_hash(const void* host_bytes, size_t nbytes, unsigned long long* hash);

/* Which form getkernelInfo and conv2d_forward(-1) use for a Winograd-path problem: the form
 * recorded for the shape in the tuned table (conv2d_autotune records the fastest one it
 * measured; CONV2D_FORM_F43 only where that form applies and problem->mid_result is set),
 * else CONV2D_FORM_FUSED when the fused form applies and K <= 384, or N*T*K*C < 8e8, or
 * N*T*K*C < 1.2e9 with C >= K (T = 2x2 tiles per image; where the fused form measured faster
 * on MI355X, by up to 3.2x), else CONV2D_FORM_STAGED (also for the direct path). The shape
 * heuristic never picks CONV2D_FORM_F43, whose U does not fit the reference harness's
 * transform_filter (16*k*c halves, main.cpp:38). */
#define CONV2D_FORM_STAGED 0
#define CONV2D_FORM_FUSED 1
#define CONV2D_FORM_F43 2
int conv2d_form(const problem_t* problem, int* form);

/* ---- Winograd F(4x4,3x3) form: no counterpart in the reference, which implements F(2x2,3x3)
 * only. 36 positions per 4x4 output tile (2.25 multiplications per output instead of 4) on the
 * interpolation points {0, 1, -1, 2, -1/2} and infinity; four slots as the reference's four
 * stages: S1 U = G' g G'^T and S2 V = B^T d B exact in FP64 and rounded once to FP16, S3 the
 * 36 per-position MFMA GEMMs storing FP16(M / 4) in mid_result, S4 the output transform in FP32
 * with the epilogue and io options. Workspaces from conv2d_workspace_size (which covers this
 * form wherever it applies); CONV2D_ERR_UNSUPPORTED for problems it does not take (the direct
 * path, input rows too wide for its staging). Normwise error against the exact convolution
 * modelled at 3.4e-3 (uniform data) and 3.9e-3 (normal data) on a C = 256 layer, against 4.1e-4
 * and 5.9e-4 for the F(2x2) forms on the same data. ---- */
/* S1 into transform_filter (what conv2d_prepare_weights runs when this form is selected). */
int conv2d_f43_prepare_weights(const problem_t* problem, void* stream);
/* S2 into transform_pin. */
int conv2d_f43_input_transform(const problem_t* problem, void* stream);
/* S3 with GEMM configuration cfg (-1: the tuned or heuristic one) into mid_result. */
int conv2d_f43_gemm(const problem_t* problem, int cfg, void* stream);
/* S4 from mid_result into out (N,K,Oh,Ow FP16). */
int conv2d_f43_output_transform(const problem_t* problem, void* stream);
/* S2 + S3 + S4 (weights prepared once). */
int conv2d_f43_forward(const problem_t* problem, int cfg, void* stream);
int conv2d_f43_num_tile_configs(void);
int conv2d_f43_tile_config(int cfg, int* bk, int* bt, int* threads);
int conv2d_f43_select_tile_config(const problem_t* problem, int* cfg);

/* Tile configurations of the GEMM stage. */
int conv2d_num_tile_configs(void);
/* Describes configuration cfg: rows of k (bk) and columns of tiles (bt) per workgroup,
 * threads per workgroup. */
int conv2d_tile_config(int cfg, int* bk, int* bt, int* threads);
/* The configuration getkernelInfo would select for this problem (an index into the
 * Winograd table, or into the direct table for problems on the direct path). */
int conv2d_select_tile_config(const problem_t* problem, int* cfg);

/* General-geometry path (no counterpart in the reference, which accepts r, s, u, v, p, q
 * in problem_t, conv2d.h:21-26, but ignores them, conv2d.cpp:254-271). Every valid problem
 * that the Winograd path does not take (filter not 3x3, stride != 1, pad > 2, rows too wide
 * for its LDS staging) runs as an implicit-GEMM direct convolution on the same MFMA
 * pipeline: slot 1 packs the filter taps, slot 2 the im2col input, slot 3 is the GEMM,
 * slot 4 the empty kernel; conv2d_workspace_size, conv2d_prepare_weights,
 * conv2d_input_transform, conv2d_gemm and conv2d_forward serve both paths. A 1x1 filter (any
 * stride and padding) has no im2col: its GEMM reads the input itself, slot 2 is the empty
 * kernel too, conv2d_input_transform launches nothing but that kernel, the input workspace
 * (transform_pin) needs 0 bytes and may be NULL, and conv2d_gemm reads the input as N,C,H,W
 * (other layouts through conv2d_forward_io or getkernelInfo_io). */
int conv2d_method(const problem_t* problem, int* method);
int conv2d_direct_num_tile_configs(void);
int conv2d_direct_tile_config(int cfg, int* bk, int* bt, int* threads);

/* Validation-only FP64 direct convolution on the GPU (any r, s, u, v, p, q): every
 * product of two FP16 values is formed exactly in FP64 and summed in FP64, written as
 * N,K,Oh,Ow doubles to out_f64 (device memory). The only FP64 arithmetic on the GPU;
 * no production stage uses it. */
int conv2d_direct_f64(const problem_t* problem, double* out_f64, void* stream);

/* ---- FP8 (OCP e4m3fn) compensated form: no counterpart in the reference. The GEMM stage
 * runs on the scaled FP8 MFMA (v_mfma_scale_f32_32x32x64_f8f6f4): every transformed value is
 * split into hi + lo e4m3 parts with one E8M0 exponent pair per (position, row, 32-channel
 * block), and U.V is recovered as U_hi.V_hi + U_hi.V_lo + U_lo.V_hi in FP32. Same problem_t,
 * different workspace layouts (sizes from conv2d_fp8_workspace_size); mid_result unused. ---- */
/* Bytes of transform_pin (V records) and transform_filter (the FP16 U, then the U records). */
int conv2d_fp8_workspace_size(const problem_t* problem, size_t* pin_bytes, size_t* filter_bytes);
/* Filter transform (FP16, as conv2d_prepare_weights) followed by its FP8 quantisation. */
int conv2d_fp8_prepare_weights(const problem_t* problem, void* stream);
/* Input transform in FP32, quantised straight into V records. */
int conv2d_fp8_input_transform(const problem_t* problem, void* stream);
/* Three-pass FP8 GEMM with the fused output transform, tile configuration cfg
 * (-1 = shape heuristic). */
int conv2d_fp8_gemm(const problem_t* problem, int cfg, void* stream);
/* Input transform + GEMM (weights prepared once). */
int conv2d_fp8_forward(const problem_t* problem, int cfg, void* stream);
int conv2d_fp8_num_tile_configs(void);
int conv2d_fp8_tile_config(int cfg, int* bk, int* bt, int* threads);
int conv2d_fp8_select_tile_config(const problem_t* problem, int* cfg);

/* ---- Tuned tile-configuration table (replaces the reference's hard-coded (n, c) dispatch,
 * conv2d.cpp:273-494). getkernelInfo consults it before the shape heuristic. Each entry holds
 * the form (CONV2D_FORM_STAGED, CONV2D_FORM_FUSED or CONV2D_FORM_F43) and its GEMM
 * configuration. ---- */
/* Staged form with configuration cfg. */
int conv2d_tuned_set(const problem_t* problem, int cfg);
/* Form `form` with GEMM configuration cfg (an index into conv2d_tile_config's table, or into
 * conv2d_f43_tile_config's for CONV2D_FORM_F43). */
int conv2d_tuned_set_form(const problem_t* problem, int cfg, int form);
int conv2d_tuned_clear(void);
int conv2d_tuned_count(void);
/* Text file, one line per shape: "n c h w k p q cfg form" (files without the form column load
 * as the staged form). */
int conv2d_tuned_save(const char* path);
int conv2d_tuned_load(const char* path);

/* Autotuner (replaces the reference's hand-picked per-shape kernels, conv2d.cpp:273-494):
 * times every GEMM tile configuration on this problem (U and V must already be prepared;
 * problem->out is overwritten) with a deterministic best-of-`reps` protocol on `stream`,
 * gates each configuration on its output being bit-identical to configuration 0's, then times
 * the input transform and, where the fused form applies, the fused kernel (gated on
 * max |fused - staged| <= 1e-2 max |staged|: its FP16 input transform rounds differently), and,
 * where the F(4x4,3x3) form applies and problem->mid_result is set, that form (its filter and
 * input transforms, every one of its GEMM configurations, each gated on bit-identity with the
 * first, and its output transform; the output gated as the fused form's), then records the
 * fastest form with its fastest configuration in the tuned table. With mid_result set, every
 * workspace must have conv2d_workspace_size's size. On return transform_filter holds the U of
 * the recorded form. times_us[i] (i < ntimes) receives: for i < 16 the best time of
 * configuration i in microseconds, or -1 if it failed the gate; times_us[16] the input
 * transform; times_us[17] the fused kernel (-1 when the fused form does not apply or failed
 * its gate); times_us[18] the F(4x4) form's input transform + best GEMM + output transform
 * (-1 when it was not timed or failed its gate). Allocates temporary buffers of the output's
 * and M's size. */
int conv2d_autotune(const problem_t* problem, int reps, void* stream, int* best_cfg,
                    float* times_us, int ntimes);

/* ---- Multi-layer (batched) form: no counterpart in the reference (its harness runs one
 * shape per process, main.cpp:8-166). Many layers of uneven shape run as one persistent
 * input-transform launch over a (layer, band, channel block, image) work queue and one
 * grouped GEMM launch per tile configuration in use over (layer, tile-block) items; the
 * sequence can be captured once into a hipGraph and replayed with one launch. ---- */
typedef struct conv2d_batch conv2d_batch;
/* Validates every layer (all pointers set, Winograd geometry) and uploads descriptors.
 * Per-layer tile configurations come from the tuned table or the shape heuristic. */
int conv2d_batch_create(const problem_t* layers, int nlayers, conv2d_batch** batch);
int conv2d_batch_destroy(conv2d_batch* batch);
/* Filter transforms of every layer (once per weight set). */
int conv2d_batch_prepare_weights(conv2d_batch* batch, void* stream);
/* Input transforms + GEMMs of all layers (the launches described above). */
int conv2d_batch_forward(conv2d_batch* batch, void* stream);
/* The same work as separate per-layer launches with the same tile configurations. */
int conv2d_batch_forward_separate(conv2d_batch* batch, void* stream);
/* Captures conv2d_batch_forward into a hipGraph on a non-null stream / replays it. */
int conv2d_batch_capture_graph(conv2d_batch* batch, void* stream);
int conv2d_batch_launch_graph(conv2d_batch* batch, void* stream);
/* Autotunes each layer over every GEMM tile configuration (U and V prepared; outputs
 * overwritten). Invalidates a captured graph. */
int conv2d_batch_autotune(conv2d_batch* batch, int reps, void* stream);
/* Tile configuration (global index) used for layer i. */
int conv2d_batch_layer_config(const conv2d_batch* batch, int i, int* cfg);
/* Form of layer i in the batch (CONV2D_FORM_STAGED or CONV2D_FORM_FUSED): chosen as
 * conv2d_form does, or by conv2d_batch_autotune from measured times. */
int conv2d_batch_layer_form(const conv2d_batch* batch, int i, int* form);

/* ---- Learned cost model (no counterpart in the reference): a 21 -> 64 -> 64 -> 1 FP32 MLP
 * predicting log(GEMM time, us) from log2(n, c, h, w, k) and a one-hot of the tile
 * configuration. Parameters: float[conv2d_mlp_param_count()] laid out
 * W1[64][21] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3. All array arguments of the
 * forward / training calls are device pointers. ---- */
int conv2d_mlp_param_count(void);
/* Host: the 21 features of (shape, cfg), cfg in [0, 16). */
int conv2d_mlp_features(const problem_t* problem, int cfg, float* out21);
/* y[b] = MLP(X[b]), X: B x 21. */
int conv2d_mlp_forward(const float* params, const float* X, int B, float* y, void* stream);
/* Mean-squared-error loss over B samples with targets T and its gradient (5633 floats). */
int conv2d_mlp_loss_grad(const float* params, const float* X, const float* T, int B, float* grad,
                         float* loss, void* stream);
/* One bias-corrected Adam step t >= 1 (beta1 0.9, beta2 0.999, eps 1e-8). */
int conv2d_mlp_adam_step(float* params, const float* grad, float* m, float* v, int t, float lr,
                         void* stream);
/* `steps` full-batch Adam steps continuing from step t0; final loss to *loss_out (host). */
int conv2d_mlp_train(float* params, float* m, float* v, int t0, const float* X, const float* T,
                     int B, int steps, float lr, float* loss_out, void* stream);
/* The configuration with the smallest predicted time (GPU inference over all
 * configurations); pred_us (host, optional) receives each prediction. */
int conv2d_mlp_select(const float* params, const problem_t* problem, int* best_cfg,
                      float* pred_us, void* stream);
/* Form and configuration from two networks of this architecture: params_staged predicts
 * log(input transform + GEMM time) per configuration, params_fused log(fused kernel time) from
 * the feature row of configuration 0. *form = CONV2D_FORM_FUSED when the fused form applies and
 * its prediction is the smallest, else CONV2D_FORM_STAGED with *best_cfg the configuration of
 * smallest prediction. pred_us (host, optional, 17 floats): the 16 staged predictions, then the
 * fused one (-1 when the fused form does not apply). Staged-form (Winograd) problems only. */
int conv2d_mlp_select_form(const float* params_staged, const float* params_fused, const problem_t* problem,
                           int* form, int* best_cfg, float* pred_us, void* stream);

/* Library version string. */
const char* conv2d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CONV2D_MI355X_H_ */
