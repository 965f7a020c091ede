// Learned refinement predictor: MLP sizes and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

namespace mwg {

// per-node features (computed on the host when a node is created by a split); norms are
// taken relative to split thresholds so that one network serves several precisions:
//  0 log10(||w_parent|| / split threshold of the parent)
//  1 log10(||s_parent|| / split threshold of the node)
//  2 log10(||w_parent|| / ||w_grandparent||) (0 below the roots)
//  3 log10(||w_parent|| / ||s_parent||)  4 input tree has the node (0/1)
//  5 log10(||w_input(node)|| ||g|| / ||input|| / split threshold of the node), -20 absent
//  6 input tree is refined below the node (0/1)
//  7 as 5 for the largest of the 26 neighbours, -20 when none
//  8 log10(mu + 1e-3)                    9 log10(cell width of the node)
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
