// Learned refinement predictor: a 2 x 64 tanh MLP (FP32) with hand-written forward,
// backward and Adam kernels.  Input: NF per-node features (standardised on the
// device with the training set's mean and deviation); output: probability that an
// output node will be split by the wavelet-norm test.
//
// Parameter vector layout: W1[64][NF] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mwg_predict.h"

namespace mwg {

constexpr int H = MLP_HIDDEN;
constexpr int NF = MLP_FEATURES;
constexpr int OW1 = 0, OB1 = OW1 + H * NF, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
static_assert(OB3 + 1 == MLP_PARAMS, "parameter layout");

// forward for the samples idx[0..n) (idx null: 0..n); stores normalised inputs and
// both hidden activations when the buffers are given (training).  One wave per
// sample: lane o computes hidden unit o of both layers (the layer inputs through the
// wave's LDS row), lane 0 the output; every sum runs over its inputs in index order.
constexpr int FWD_WAVES = 4;
__global__ void __launch_bounds__(64 * FWD_WAVES) k_mlp_forward(int n, const int *idx, const float *X,
                                                                const float *mean, const float *sdev,
                                                                const float *P, float *prob, float *Xn, float *H1,
                                                                float *H2) {
    static_assert(H == 64, "one lane per hidden unit");
    // parameters with the weight matrices transposed (input-major: the lanes of a wave
    // read consecutive words)
    __shared__ float sp[MLP_PARAMS];
    __shared__ float sx[FWD_WAVES][NF], sh[FWD_WAVES][H];
    for (int e = threadIdx.x; e < MLP_PARAMS; e += blockDim.x) {
        int src = e;
        if (e < OB1) src = OW1 + (e % H) * NF + e / H;                   // W1T[i][o] = W1[o][i]
        else if (e >= OW2 && e < OB2) src = OW2 + ((e - OW2) % H) * H + (e - OW2) / H;  // W2T[i][o]
        sp[e] = P[src];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, o = threadIdx.x & 63;
    for (int s = blockIdx.x * FWD_WAVES + wave; s < n; s += gridDim.x * FWD_WAVES) {
        const int r = idx ? idx[s] : s;
        if (o < NF) {
            const float xv = (X[(size_t)r * NF + o] - mean[o]) / sdev[o];
            sx[wave][o] = xv;
            if (Xn) Xn[(size_t)s * NF + o] = xv;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        float z = sp[OB1 + o];
        for (int i = 0; i < NF; i++) z += sp[OW1 + i * H + o] * sx[wave][i];
        const float h1 = tanhf(z);
        sh[wave][o] = h1;
        if (H1) H1[(size_t)s * H + o] = h1;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        z = sp[OB2 + o];
        for (int i = 0; i < H; i++) z += sp[OW2 + i * H + o] * sh[wave][i];
        const float h2 = tanhf(z);
        if (H2) H2[(size_t)s * H + o] = h2;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        sh[wave][o] = h2;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (o == 0) {
            float z3 = sp[OB3];
            for (int i = 0; i < H; i++) z3 += sp[OW3 + i] * sh[wave][i];
            prob[s] = 1.0f / (1.0f + expf(-z3));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// per-sample backward of the mean binary cross-entropy: dZ3, dZ2, dZ1
__global__ void k_mlp_backward(int n, const int *idx, const float *Y, const float *P, const float *prob,
                               const float *H1, const float *H2, float *dZ1, float *dZ2, float *dZ3) {
    __shared__ float sp[MLP_PARAMS];
    for (int e = threadIdx.x; e < MLP_PARAMS; e += blockDim.x) sp[e] = P[e];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int r = idx ? idx[s] : s;
    const float d3 = (prob[s] - Y[r]) / (float)n;
    dZ3[s] = d3;
    float d2[H];
    for (int i = 0; i < H; i++) {
        const float h = H2[(size_t)s * H + i];
        d2[i] = d3 * sp[OW3 + i] * (1.0f - h * h);
        dZ2[(size_t)s * H + i] = d2[i];
    }
    for (int i = 0; i < H; i++) {
        float g = 0.0f;
        for (int o = 0; o < H; o++) g += sp[OW2 + o * H + i] * d2[o];
        const float h = H1[(size_t)s * H + i];
        dZ1[(size_t)s * H + i] = g * (1.0f - h * h);
    }
}

// gradient: one thread per parameter, summing over the n samples in order
__global__ void k_mlp_grad(int n, const float *Xn, const float *H1, const float *H2, const float *dZ1,
                           const float *dZ2, const float *dZ3, float *G) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    float g = 0.0f;
    if (p < OB1) {
        const int o = p / NF, i = p % NF;
        for (int s = 0; s < n; s++) g += dZ1[(size_t)s * H + o] * Xn[(size_t)s * NF + i];
    } else if (p < OW2) {
        const int o = p - OB1;
        for (int s = 0; s < n; s++) g += dZ1[(size_t)s * H + o];
    } else if (p < OB2) {
        const int o = (p - OW2) / H, i = (p - OW2) % H;
        for (int s = 0; s < n; s++) g += dZ2[(size_t)s * H + o] * H1[(size_t)s * H + i];
    } else if (p < OW3) {
        const int o = p - OB2;
        for (int s = 0; s < n; s++) g += dZ2[(size_t)s * H + o];
    } else if (p < OB3) {
        const int i = p - OW3;
        for (int s = 0; s < n; s++) g += dZ3[s] * H2[(size_t)s * H + i];
    } else {
        for (int s = 0; s < n; s++) g += dZ3[s];
    }
    G[p] = g;
}

__global__ void k_adam(float *P, const float *G, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    const float g = G[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

hipError_t launch_mlp_forward(hipStream_t s, int n, const int *idx, const float *X, const float *mean,
                              const float *sdev, const float *P, float *prob, float *Xn, float *H1, float *H2) {
    if (n <= 0) return hipSuccess;
    const int blocks = std::min((n + FWD_WAVES - 1) / FWD_WAVES, 512);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(64 * FWD_WAVES), 0, s, n, idx, X, mean, sdev, P, prob, Xn,
                       H1, H2);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(hipStream_t s, int n, const int *idx, const float *Y, const float *P,
                               const float *prob, const float *H1, const float *H2, float *dZ1, float *dZ2,
                               float *dZ3) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_backward, dim3((n + 127) / 128), dim3(128), 0, s, n, idx, Y, P, prob, H1, H2, dZ1, dZ2,
                       dZ3);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(hipStream_t s, int n, const float *Xn, const float *H1, const float *H2,
                           const float *dZ1, const float *dZ2, const float *dZ3, float *G) {
    hipLaunchKernelGGL(k_mlp_grad, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, n, Xn, H1, H2, dZ1, dZ2, dZ3, G);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t s, float *P, const float *G, float *m, float *v, float lr, float b1, float b2,
                       float eps, int t) {
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, P, G, m, v, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace mwg
