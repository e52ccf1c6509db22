// Launch declarations of the preconditioner and network kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "geo.hpp"

namespace mlfma {

// ---- near-field polynomial preconditioner (k_precond.hip); blocks RHS-fastest
void launch_near_diag_inv(hipStream_t st, const NearDev& nd, const double2* Z, double2* dinv);
void launch_poly_init(hipStream_t st, int n, int R, const double2* dinv, const double2* x, double2 a0, double2* t,
                      double2* y);
void launch_poly_step(hipStream_t st, int n, int R, const double2* dinv, const double2* u, double2 a, double2* t,
                      double2* y);
void launch_diag_scale(hipStream_t st, int n, int R, const double2* dinv, const double2* u, double2* y);
void launch_probe(hipStream_t st, int n, unsigned seed, double2* z);

// ---- MLP (k_mlp.hip), FP32, row-major [sample][feature]
// Z = X W^T + b (W: nout x nin); H = tanh(Z) when act, else H = Z
void launch_dense_fwd(hipStream_t st, int B, int nin, int nout, const float* X, const float* W, const float* b,
                      float* Z, float* H, int act);
// dZ = dH * (1 - H^2) when act, else dZ = dH
void launch_act_bwd(hipStream_t st, int count, const float* H, const float* dH, float* dZ, int act);
// dX = dZ W
void launch_dense_bwd_input(hipStream_t st, int B, int nin, int nout, const float* dZ, const float* W, float* dX);
// dW = dZ^T X, db = column sums of dZ (sample order fixed)
void launch_dense_bwd_weights(hipStream_t st, int B, int nin, int nout, const float* dZ, const float* X, float* dW,
                              float* db);
// Quadratic loss of complex coefficients a (Y row = [re a_0..a_d, im a_0..a_d]):
//   l_s = (a^H G_s a - 2 Re a^H h_s + c_s) / c_s,  loss = mean_s l_s
// G_s (m x m), h_s (m) complex FP64; writes l_s and dY = d loss / dY.
void launch_quad_loss(hipStream_t st, int B, int m, const float* Y, const double2* G, const double2* h,
                      const double* c, float* loss, float* dY);
// Adam step on n parameters (bias-corrected, step t >= 1)
void launch_adam(hipStream_t st, int n, float* p, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                 float eps, int t);

}  // namespace mlfma
