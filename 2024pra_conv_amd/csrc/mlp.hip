// Learned tile-configuration cost model: a 21 -> 64 -> 64 -> 1 FP32 MLP (ReLU hidden
// layers) predicting log(GEMM time in us) from (log2 N, C, H, W, K; one-hot of 16 tile
// configurations). Training (MSE loss, full-batch Adam) and inference run in the kernels
// below; the host only sequences them.
//
// No counterpart in the reference (its dispatch is a hand-written if/else, conv2d.cpp:273-494).
//
// Parameter vector layout (float, kMlpParams = 5633):
//   W1[64][21] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "plugin_internal.h"
#include "wino_gemm.h"

namespace {

constexpr int kIn = 21, kH = 64;
constexpr int oW1 = 0, ob1 = oW1 + kH * kIn, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, ow3 = ob2 + kH,
              ob3 = ow3 + kH;
constexpr int kMlpParams = ob3 + 1;
static_assert(kMlpParams == 5633, "parameter count");

// Forward: one 64-thread block per sample. Stores pre-activations z1, z2 (for backward)
// and the output y.
__global__ void __launch_bounds__(64) mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                          int B, float* __restrict__ Z1, float* __restrict__ Z2,
                                                          float* __restrict__ Y) {
    __shared__ float x[kIn];
    __shared__ float a1[kH];
    __shared__ float a2[kH];
    const int b = blockIdx.x, j = threadIdx.x;
    if (b >= B) return;
    if (j < kIn) x[j] = X[(size_t)b * kIn + j];
    __syncthreads();
    float z = P[ob1 + j];
    for (int i = 0; i < kIn; ++i) z += P[oW1 + j * kIn + i] * x[i];
    if (Z1) Z1[(size_t)b * kH + j] = z;
    a1[j] = z > 0.f ? z : 0.f;
    __syncthreads();
    float z2 = P[ob2 + j];
    for (int i = 0; i < kH; ++i) z2 += P[oW2 + j * kH + i] * a1[i];
    if (Z2) Z2[(size_t)b * kH + j] = z2;
    a2[j] = z2 > 0.f ? z2 : 0.f;
    __syncthreads();
    if (j == 0) {
        float y = P[ob3];
        for (int i = 0; i < kH; ++i) y += P[ow3 + i] * a2[i];
        Y[b] = y;
    }
}

// Backward through the activations for one sample (MSE loss mean over B):
// dy = 2 (y - t) / B, dz2 = dy w3 . [z2 > 0], dz1 = (W2^T dz2) . [z1 > 0].
__global__ void __launch_bounds__(64) mlp_backward_kernel(const float* __restrict__ P, const float* __restrict__ Z1,
                                                           const float* __restrict__ Z2, const float* __restrict__ Y,
                                                           const float* __restrict__ Tgt, int B,
                                                           float* __restrict__ DY, float* __restrict__ DZ1,
                                                           float* __restrict__ DZ2) {
    __shared__ float dz2[kH];
    const int b = blockIdx.x, j = threadIdx.x;
    if (b >= B) return;
    const float dy = 2.f * (Y[b] - Tgt[b]) / (float)B;
    if (j == 0) DY[b] = dy;
    const float z2 = Z2[(size_t)b * kH + j];
    const float d2 = z2 > 0.f ? dy * P[ow3 + j] : 0.f;
    dz2[j] = d2;
    DZ2[(size_t)b * kH + j] = d2;
    __syncthreads();
    float da1 = 0.f;
    for (int i = 0; i < kH; ++i) da1 += P[oW2 + i * kH + j] * dz2[i];
    const float z1 = Z1[(size_t)b * kH + j];
    DZ1[(size_t)b * kH + j] = z1 > 0.f ? da1 : 0.f;
}

// Parameter gradients: one thread per parameter, summing over samples in order
// (deterministic).
__global__ void __launch_bounds__(256) mlp_param_grad_kernel(const float* __restrict__ X, const float* __restrict__ Z1,
                                                              const float* __restrict__ Z2, const float* __restrict__ DY,
                                                              const float* __restrict__ DZ1,
                                                              const float* __restrict__ DZ2, int B,
                                                              float* __restrict__ G) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float s = 0.f;
    if (p < ob1) {  // W1[j][i]
        const int j = p / kIn, i = p % kIn;
        for (int b = 0; b < B; ++b) s += DZ1[(size_t)b * kH + j] * X[(size_t)b * kIn + i];
    } else if (p < oW2) {  // b1[j]
        const int j = p - ob1;
        for (int b = 0; b < B; ++b) s += DZ1[(size_t)b * kH + j];
    } else if (p < ob2) {  // W2[j][i]
        const int j = (p - oW2) / kH, i = (p - oW2) % kH;
        for (int b = 0; b < B; ++b) {
            const float z1 = Z1[(size_t)b * kH + i];
            s += DZ2[(size_t)b * kH + j] * (z1 > 0.f ? z1 : 0.f);
        }
    } else if (p < ow3) {  // b2[j]
        const int j = p - ob2;
        for (int b = 0; b < B; ++b) s += DZ2[(size_t)b * kH + j];
    } else if (p < ob3) {  // w3[i]
        const int i = p - ow3;
        for (int b = 0; b < B; ++b) {
            const float z2 = Z2[(size_t)b * kH + i];
            s += DY[b] * (z2 > 0.f ? z2 : 0.f);
        }
    } else {  // b3
        for (int b = 0; b < B; ++b) s += DY[b];
    }
    G[p] = s;
}

// MSE loss, one block, deterministic tree reduction.
__global__ void __launch_bounds__(256) mlp_loss_kernel(const float* __restrict__ Y, const float* __restrict__ Tgt,
                                                        int B, float* __restrict__ loss) {
    __shared__ float red[256];
    float s = 0.f;
    for (int b = threadIdx.x; b < B; b += 256) {
        const float d = Y[b] - Tgt[b];
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / (float)B;
}

// Adam (bias-corrected), elementwise.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                                                        float* __restrict__ M, float* __restrict__ V, int n, float lr,
                                                        float beta1, float beta2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float g = G[p];
    const float m = beta1 * M[p] + (1.f - beta1) * g;
    const float v = beta2 * V[p] + (1.f - beta2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Scratch for one batch size (activations and their gradients).
struct MlpScratch {
    float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *DY = nullptr, *DZ1 = nullptr, *DZ2 = nullptr;
    ~MlpScratch() {
        for (float* p : {Z1, Z2, Y, DY, DZ1, DZ2})
            if (p) (void)hipFree(p);
    }
    bool alloc(int B) {
        const size_t h = (size_t)B * kH * sizeof(float), s = (size_t)B * sizeof(float);
        return hipMalloc(&Z1, h) == hipSuccess && hipMalloc(&Z2, h) == hipSuccess && hipMalloc(&Y, s) == hipSuccess &&
               hipMalloc(&DY, s) == hipSuccess && hipMalloc(&DZ1, h) == hipSuccess && hipMalloc(&DZ2, h) == hipSuccess;
    }
};

int launch_ok() { return hipGetLastError() == hipSuccess ? CONV2D_OK : CONV2D_ERR_HIP; }

void loss_grad(const float* P, const float* X, const float* T, int B, float* G, const MlpScratch& s,
               hipStream_t st) {
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(B), dim3(64), 0, st, P, X, B, s.Z1, s.Z2, s.Y);
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(B), dim3(64), 0, st, P, (const float*)s.Z1, (const float*)s.Z2,
                       (const float*)s.Y, T, B, s.DY, s.DZ1, s.DZ2);
    hipLaunchKernelGGL(mlp_param_grad_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, X,
                       (const float*)s.Z1, (const float*)s.Z2, (const float*)s.DY, (const float*)s.DZ1,
                       (const float*)s.DZ2, B, G);
}

}  // namespace

extern "C" {

int conv2d_mlp_param_count(void) { return kMlpParams; }

// Feature row of (shape, configuration): log2 of n, c, h, w, k, then a one-hot of cfg
// over 16 slots.
int conv2d_mlp_features(const problem_t* pb, int cfg, float* out21) {
    if (!pb || !out21) return CONV2D_ERR_NULL;
    if (cfg < 0 || cfg >= 16) return CONV2D_ERR_INVALID;
    if (!pb->n || !pb->c || !pb->h || !pb->w || !pb->k) return CONV2D_ERR_INVALID;
    out21[0] = log2f((float)pb->n);
    out21[1] = log2f((float)pb->c);
    out21[2] = log2f((float)pb->h);
    out21[3] = log2f((float)pb->w);
    out21[4] = log2f((float)pb->k);
    for (int i = 0; i < 16; ++i) out21[5 + i] = i == cfg ? 1.f : 0.f;
    return CONV2D_OK;
}

// y[b] = MLP(X[b]) for B feature rows (device pointers).
int conv2d_mlp_forward(const float* params, const float* X, int B, float* y, void* stream) {
    if (!params || !X || !y) return CONV2D_ERR_NULL;
    if (B <= 0) return CONV2D_ERR_INVALID;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(B), dim3(64), 0, (hipStream_t)stream, params, X, B, (float*)nullptr,
                       (float*)nullptr, y);
    return launch_ok();
}

// MSE loss and its gradient with respect to every parameter (device pointers; loss may be
// NULL). Allocates activation scratch for B samples.
int conv2d_mlp_loss_grad(const float* params, const float* X, const float* T, int B, float* grad, float* loss,
                         void* stream) {
    if (!params || !X || !T || !grad) return CONV2D_ERR_NULL;
    if (B <= 0) return CONV2D_ERR_INVALID;
    MlpScratch s;
    if (!s.alloc(B)) return CONV2D_ERR_HIP;
    hipStream_t st = (hipStream_t)stream;
    loss_grad(params, X, T, B, grad, s, st);
    if (loss) hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, st, (const float*)s.Y, T, B, loss);
    int rc = launch_ok();
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = CONV2D_ERR_HIP;
    return rc;
}

// One Adam update (step t >= 1) of params from grad with moment buffers m, v.
int conv2d_mlp_adam_step(float* params, const float* grad, float* m, float* v, int t, float lr, void* stream) {
    if (!params || !grad || !m || !v) return CONV2D_ERR_NULL;
    if (t < 1) return CONV2D_ERR_INVALID;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, params,
                       grad, m, v, kMlpParams, lr, b1, b2, 1e-8f, c1, c2);
    return launch_ok();
}

// Full-batch Adam training for `steps` steps on device data; moments m, v persist across
// calls (t0 = steps already taken). Writes the final MSE loss to *loss_out (host) if set.
int conv2d_mlp_train(float* params, float* m, float* v, int t0, const float* X, const float* T, int B, int steps,
                     float lr, float* loss_out, void* stream) {
    if (!params || !m || !v || !X || !T) return CONV2D_ERR_NULL;
    if (B <= 0 || steps < 0 || t0 < 0) return CONV2D_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    MlpScratch s;
    float* G = nullptr;
    float* L = nullptr;
    if (!s.alloc(B) || hipMalloc(&G, kMlpParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&L, sizeof(float)) != hipSuccess) {
        if (G) (void)hipFree(G);
        return CONV2D_ERR_HIP;
    }
    int rc = CONV2D_OK;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int i = 1; i <= steps && !rc; ++i) {
        const int t = t0 + i;
        loss_grad(params, X, T, B, G, s, st);
        const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, params,
                           (const float*)G, m, v, kMlpParams, lr, b1, b2, 1e-8f, c1, c2);
        rc = launch_ok();
    }
    if (!rc) {
        hipLaunchKernelGGL(mlp_forward_kernel, dim3(B), dim3(64), 0, st, (const float*)params, X, B, s.Z1, s.Z2, s.Y);
        hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, st, (const float*)s.Y, T, B, L);
        rc = launch_ok();
    }
    if (!rc && loss_out &&
        (hipMemcpyAsync(loss_out, L, sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
        rc = CONV2D_ERR_HIP;
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = CONV2D_ERR_HIP;
    (void)hipFree(G);
    (void)hipFree(L);
    return rc;
}

// Predictions of the MLP with parameters `params` for nrows feature rows (host in, host out):
// one GPU inference launch.
static int mlp_predict(const float* params, const std::vector<float>& feats, int nrows, float* y, hipStream_t st) {
    float *dX = nullptr, *dY = nullptr;
    int rc = CONV2D_OK;
    if (hipMalloc(&dX, feats.size() * sizeof(float)) != hipSuccess || hipMalloc(&dY, nrows * sizeof(float)) != hipSuccess)
        rc = CONV2D_ERR_HIP;
    if (!rc && hipMemcpyAsync(dX, feats.data(), feats.size() * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess)
        rc = CONV2D_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(mlp_forward_kernel, dim3(nrows), dim3(64), 0, st, params, (const float*)dX, nrows,
                           (float*)nullptr, (float*)nullptr, dY);
        rc = launch_ok();
    }
    if (!rc && (hipMemcpyAsync(y, dY, nrows * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess))
        rc = CONV2D_ERR_HIP;
    if (dX) (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    return rc;
}

// Predicts the time of every tile configuration for this shape with the MLP (inference on
// the GPU) and returns the configuration with the smallest prediction. pred_us (optional)
// receives kNumGemmConfigs predicted times in microseconds.
int conv2d_mlp_select(const float* params, const problem_t* pb, int* best_cfg, float* pred_us, void* stream) {
    if (!params || !pb || !best_cfg) return CONV2D_ERR_NULL;
    const int nc = kNumGemmConfigs < 16 ? kNumGemmConfigs : 16;
    std::vector<float> feats((size_t)nc * kIn), y(nc);
    for (int c = 0; c < nc; ++c) {
        int rc = conv2d_mlp_features(pb, c, &feats[(size_t)c * kIn]);
        if (rc) return rc;
    }
    int rc = mlp_predict(params, feats, nc, y.data(), (hipStream_t)stream);
    if (rc) return rc;
    int best = 0;
    for (int c = 1; c < nc; ++c)
        if (y[c] < y[best]) best = c;
    *best_cfg = best;
    if (pred_us)
        for (int c = 0; c < nc; ++c) pred_us[c] = expf(y[c]);
    return CONV2D_OK;
}

// Form and configuration choice from two networks of the same architecture: params_staged
// predicts the staged form's time (input transform + GEMM) per configuration, params_fused the
// fused kernel's time (feature row of configuration 0). The fused form is chosen when the
// problem admits it and its prediction is the smallest. pred_us (optional) receives the 16
// staged predictions then the fused one (-1 when the fused form does not apply).
int conv2d_mlp_select_form(const float* params_staged, const float* params_fused, const problem_t* pb, int* form,
                           int* best_cfg, float* pred_us, void* stream) {
    if (!params_staged || !params_fused || !pb || !form || !best_cfg) return CONV2D_ERR_NULL;
    wino::Geometry g;
    int rc = wino::require_staged(pb, &g);
    if (rc) return rc;
    const int nc = kNumGemmConfigs < 16 ? kNumGemmConfigs : 16;
    std::vector<float> feats((size_t)nc * kIn), y(nc);
    for (int c = 0; c < nc; ++c)
        if ((rc = conv2d_mlp_features(pb, c, &feats[(size_t)c * kIn]))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = mlp_predict(params_staged, feats, nc, y.data(), st))) return rc;
    int best = 0;
    for (int c = 1; c < nc; ++c)
        if (y[c] < y[best]) best = c;
    float yf = 0.f;
    if (g.fused_ok) {
        std::vector<float> f0(feats.begin(), feats.begin() + kIn);
        if ((rc = mlp_predict(params_fused, f0, 1, &yf, st))) return rc;
    }
    *best_cfg = best;
    *form = g.fused_ok && yf < y[best] ? CONV2D_FORM_FUSED : CONV2D_FORM_STAGED;
    if (pred_us) {
        for (int c = 0; c < nc; ++c) pred_us[c] = expf(y[c]);
        pred_us[nc] = g.fused_ok ? expf(yf) : -1.f;
    }
    return CONV2D_OK;
}

}  // extern "C"
