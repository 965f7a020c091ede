// Learned refinement predictor: MLP sizes and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

namespace mwg {

// per-node features: 10 values computed on the device for the children of every node of
// a refinement iteration (k_child_features in mwg_kernels.hip lists them)
constexpr int MLP_FEATURES = 10;
constexpr int MLP_HIDDEN = 64;
constexpr int MLP_PARAMS = MLP_HIDDEN * MLP_FEATURES + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1;

hipError_t launch_mlp_forward(hipStream_t s, int n, const int *idx, const float *X, const float *mean,
                              const float *sdev, const float *P, float *prob, float *Xn, float *H1, float *H2);
hipError_t launch_mlp_backward(hipStream_t s, int n, const int *idx, const float *Y, const float *P,
                               const float *prob, const float *H1, const float *H2, float *dZ1, float *dZ2,
                               float *dZ3);
hipError_t launch_mlp_grad(hipStream_t s, int n, const float *Xn, const float *H1, const float *H2,
                           const float *dZ1, const float *dZ2, const float *dZ3, float *G);
hipError_t launch_adam(hipStream_t s, float *P, const float *G, float *m, float *v, float lr, float b1, float b2,
                       float eps, int t);

}  // namespace mwg
