// Learned CMF2 step-size controller: a 12 -> 64 -> 64 -> 1 multilayer
// perceptron with SiLU activations, BF16 weights and FP32 accumulation, and
// its training (mean-squared error, Adam) -- all on the device.
//
// Parameter vector (FP32, kMlpNParam entries):
//   [0, 12)      feature means (input standardisation, fixed during training)
//   [12, 24)     feature standard deviations
//   W1 [12 x 64] (input i -> unit j at W1 + 64 i + j), b1 [64],
//   W2 [64 x 64] (unit j -> unit k at W2 + 64 j + k), b2 [64],
//   w3 [64], b3 [1].
// Weights are used as BF16 (round to nearest even of the FP32 master copy);
// biases, activations, sums and gradients are FP32.
#include "mg_internal.h"

#include <math.h>

namespace mg {

namespace {

constexpr int F = kMlpNFeat, H = kMlpHidden;
constexpr int oW1 = 2 * F, oB1 = oW1 + F * H, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H;
static_assert(oB3 + 1 == kMlpNParam, "parameter layout");

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);      // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// BF16 copies of W1, W2, w3 (stored contiguously in that order).
__global__ void k_mlp_bf16(const float *p, uint16_t *wb) {
    const int n1 = F * H, n2 = H * H;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n1 + n2 + H; i += gridDim.x * blockDim.x) {
        float v;
        if (i < n1) v = p[oW1 + i];
        else if (i < n1 + n2) v = p[oW2 + i - n1];
        else v = p[oW3 + i - n1 - n2];
        wb[i] = to_bf16(v);
    }
}

// Forward pass, one workgroup (64 threads = hidden units) per sample.
// z1, h1, z2, h2: [B x 64]; y: [B]; with a target t: dy = 2 (y - t) / B and
// err2 = (y - t)^2.
__global__ void __launch_bounds__(64) k_mlp_forward(const float *X, const float *p, const uint16_t *wb, float *z1,
                                                    float *h1, float *z2, float *h2, float *y, const float *t,
                                                    float *dy, float *err2, int B) {
    const int b = blockIdx.x, j = threadIdx.x;
    if (b >= B) return;
    __shared__ float xs[F], a1[H], red[H];
    if (j < F) xs[j] = (X[(size_t)F * b + j] - p[j]) / p[F + j];
    __syncthreads();
    const uint16_t *w1 = wb, *w2 = wb + F * H, *w3 = wb + F * H + H * H;
    float z = p[oB1 + j];
    for (int i = 0; i < F; i++) z = fmaf(xs[i], from_bf16(w1[H * i + j]), z);
    const float a = silu(z);
    z1[(size_t)H * b + j] = z;
    h1[(size_t)H * b + j] = a;
    a1[j] = a;
    __syncthreads();
    float zz = p[oB2 + j];
    for (int i = 0; i < H; i++) zz = fmaf(a1[i], from_bf16(w2[H * i + j]), zz);
    const float a2 = silu(zz);
    z2[(size_t)H * b + j] = zz;
    h2[(size_t)H * b + j] = a2;
    red[j] = a2 * from_bf16(w3[j]);
    __syncthreads();
    if (j == 0) {
        float s = p[oB3];
        for (int i = 0; i < H; i++) s += red[i];
        y[b] = s;
        if (t) {
            const float e = s - t[b];
            dy[b] = 2.0f * e / (float)B;
            err2[b] = e * e;
        }
    }
}

// Backward through the two hidden layers, one workgroup per sample.
__global__ void __launch_bounds__(64) k_mlp_backward(const float *dy, const float *z1, const float *z2,
                                                     const uint16_t *wb, float *dz1, float *dz2, int B) {
    const int b = blockIdx.x, k = threadIdx.x;
    if (b >= B) return;
    __shared__ float d2[H];
    const uint16_t *w2 = wb + F * H, *w3 = wb + F * H + H * H;
    const float g2 = dy[b] * from_bf16(w3[k]) * dsilu(z2[(size_t)H * b + k]);
    d2[k] = g2;
    dz2[(size_t)H * b + k] = g2;
    __syncthreads();
    const int j = k;
    float g = 0.0f;
    for (int q = 0; q < H; q++) g = fmaf(from_bf16(w2[H * j + q]), d2[q], g);
    dz1[(size_t)H * b + j] = g * dsilu(z1[(size_t)H * b + j]);
}

// Gradient of every trainable parameter, summed over the batch in sample
// order (one thread per parameter; deterministic).
__global__ void k_mlp_grad(const float *X, const float *p, const float *h1, const float *h2, const float *dz1,
                           const float *dz2, const float *dy, float *grad, int B) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = kMlpNParam - oW1;
    if (q >= n) return;
    const int o = q + oW1;
    float s = 0.0f;
    if (o < oB1) {
        const int i = (o - oW1) / H, j = (o - oW1) % H;
        for (int b = 0; b < B; b++) s = fmaf((X[(size_t)F * b + i] - p[i]) / p[F + i], dz1[(size_t)H * b + j], s);
    } else if (o < oW2) {
        const int j = o - oB1;
        for (int b = 0; b < B; b++) s += dz1[(size_t)H * b + j];
    } else if (o < oB2) {
        const int j = (o - oW2) / H, k = (o - oW2) % H;
        for (int b = 0; b < B; b++) s = fmaf(h1[(size_t)H * b + j], dz2[(size_t)H * b + k], s);
    } else if (o < oW3) {
        const int k = o - oB2;
        for (int b = 0; b < B; b++) s += dz2[(size_t)H * b + k];
    } else if (o < oB3) {
        const int k = o - oW3;
        for (int b = 0; b < B; b++) s = fmaf(h2[(size_t)H * b + k], dy[b], s);
    } else {
        for (int b = 0; b < B; b++) s += dy[b];
    }
    grad[q] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_mlp_adam(float *p, const float *grad, float *m, float *v, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = kMlpNParam - oW1;
    if (q >= n) return;
    const float g = grad[q];
    const float mq = b1 * m[q] + (1.0f - b1) * g;
    const float vq = b2 * v[q] + (1.0f - b2) * g * g;
    m[q] = mq;
    v[q] = vq;
    p[q + oW1] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace

void launch_mlp_bf16(const float *p, uint16_t *wb, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_bf16, dim3(8), dim3(256), 0, s, p, wb);
}

void launch_mlp_forward(const float *X, const float *p, const uint16_t *wb, MlpWork w, float *y, const float *t,
                        int B, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_forward, dim3(B), dim3(H), 0, s, X, p, wb, w.z1, w.h1, w.z2, w.h2, y, t, w.dy, w.err2,
                       B);
}

void launch_mlp_backward_grad(const float *X, const float *p, const uint16_t *wb, MlpWork w, float *grad, int B,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_backward, dim3(B), dim3(H), 0, s, w.dy, w.z1, w.z2, wb, w.dz1, w.dz2, B);
    const int n = kMlpNParam - oW1;
    hipLaunchKernelGGL(k_mlp_grad, dim3((n + 255) / 256), dim3(256), 0, s, X, p, w.h1, w.h2, w.dz1, w.dz2, w.dy,
                       grad, B);
}

void launch_mlp_adam(float *p, const float *grad, float *m, float *v, float lr, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    const int n = kMlpNParam - oW1;
    hipLaunchKernelGGL(k_mlp_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, grad, m, v, lr, b1, b2, eps, c1, c2);
}

}  // namespace mg
