// FP32 kernels of the preconditioner network: dense layers (forward, backward
// to inputs and to weights), tanh, the quadratic loss of the polynomial
// coefficients and Adam.  No reference counterpart.  Every reduction runs in a
// fixed order, so training is reproducible.  Sizes are small (a few hundred
// samples, width 256): one thread per output element.
#include <cmath>

#include "cplx.hpp"
#include "precond.hpp"

namespace mlfma {

__global__ void k_dense_fwd(int B, int nin, int nout, const float* __restrict__ X, const float* __restrict__ W,
                            const float* __restrict__ bias, float* __restrict__ Z, float* __restrict__ H, int act) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (o >= nout || s >= B) return;
    const float* x = X + (size_t)s * nin;
    const float* w = W + (size_t)o * nin;
    float z = bias[o];
    for (int i = 0; i < nin; ++i) z = fmaf(x[i], w[i], z);
    Z[(size_t)s * nout + o] = z;
    H[(size_t)s * nout + o] = act ? tanhf(z) : z;
}

void launch_dense_fwd(hipStream_t st, int B, int nin, int nout, const float* X, const float* W, const float* b,
                      float* Z, float* H, int act) {
    k_dense_fwd<<<dim3((nout + 63) / 64, B), 64, 0, st>>>(B, nin, nout, X, W, b, Z, H, act);
}

__global__ void k_act_bwd(int count, const float* __restrict__ H, const float* __restrict__ dH,
                          float* __restrict__ dZ, int act) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    dZ[e] = act ? dH[e] * (1.0f - H[e] * H[e]) : dH[e];
}

void launch_act_bwd(hipStream_t st, int count, const float* H, const float* dH, float* dZ, int act) {
    k_act_bwd<<<(count + 255) / 256, 256, 0, st>>>(count, H, dH, dZ, act);
}

__global__ void k_dense_bwd_input(int B, int nin, int nout, const float* __restrict__ dZ, const float* __restrict__ W,
                                  float* __restrict__ dX) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= nin || s >= B) return;
    const float* dz = dZ + (size_t)s * nout;
    float a = 0.0f;
    for (int o = 0; o < nout; ++o) a = fmaf(dz[o], W[(size_t)o * nin + i], a);
    dX[(size_t)s * nin + i] = a;
}

void launch_dense_bwd_input(hipStream_t st, int B, int nin, int nout, const float* dZ, const float* W, float* dX) {
    k_dense_bwd_input<<<dim3((nin + 63) / 64, B), 64, 0, st>>>(B, nin, nout, dZ, W, dX);
}

// dW[o][i] = sum_s dZ[s][o] X[s][i]; the extra column i = nin holds db[o].
__global__ void k_dense_bwd_weights(int B, int nin, int nout, const float* __restrict__ dZ,
                                    const float* __restrict__ X, float* __restrict__ dW, float* __restrict__ db) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int o = blockIdx.y;
    if (i > nin || o >= nout) return;
    float a = 0.0f;
    if (i < nin) {
        for (int s = 0; s < B; ++s) a = fmaf(dZ[(size_t)s * nout + o], X[(size_t)s * nin + i], a);
        dW[(size_t)o * nin + i] = a;
    } else {
        for (int s = 0; s < B; ++s) a += dZ[(size_t)s * nout + o];
        db[o] = a;
    }
}

void launch_dense_bwd_weights(hipStream_t st, int B, int nin, int nout, const float* dZ, const float* X, float* dW,
                              float* db) {
    k_dense_bwd_weights<<<dim3((nin + 1 + 63) / 64, nout), 64, 0, st>>>(B, nin, nout, dZ, X, dW, db);
}

// One thread per sample; m <= 16 coefficients.
__global__ void k_quad_loss(int B, int m, const float* __restrict__ Y, const double2* __restrict__ G,
                            const double2* __restrict__ h, const double* __restrict__ c, float* __restrict__ loss,
                            float* __restrict__ dY) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    double2 a[16];
    for (int j = 0; j < m; ++j) a[j] = make_double2(Y[(size_t)s * 2 * m + j], Y[(size_t)s * 2 * m + m + j]);
    const double2* Gs = G + (size_t)s * m * m;
    const double2* hs = h + (size_t)s * m;
    const double cs = c[s] > 0.0 ? c[s] : 1.0;
    double q = 0.0;
    for (int j = 0; j < m; ++j) {
        double2 g = make_double2(-hs[j].x, -hs[j].y);   // (G a - h)_j, G row-major
        for (int l = 0; l < m; ++l) cfma(g, Gs[(size_t)j * m + l], a[l]);
        // a^H G a - 2 Re a^H h = Re sum_j conj(a_j) ((G a)_j - 2 h_j)
        const double2 ga = make_double2(g.x + hs[j].x, g.y + hs[j].y);
        q += a[j].x * (ga.x - 2.0 * hs[j].x) + a[j].y * (ga.y - 2.0 * hs[j].y);
        dY[(size_t)s * 2 * m + j] = (float)(2.0 * g.x / cs / B);
        dY[(size_t)s * 2 * m + m + j] = (float)(2.0 * g.y / cs / B);
    }
    loss[s] = (float)((q + c[s]) / cs);
}

void launch_quad_loss(hipStream_t st, int B, int m, const float* Y, const double2* G, const double2* h,
                      const double* c, float* loss, float* dY) {
    k_quad_loss<<<(B + 63) / 64, 64, 0, st>>>(B, m, Y, G, h, c, loss, dY);
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float a = b1 * m1[e] + (1.0f - b1) * gg;
    const float v = b2 * m2[e] + (1.0f - b2) * gg * gg;
    m1[e] = a;
    m2[e] = v;
    p[e] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

void launch_adam(hipStream_t st, int n, float* p, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                 float eps, int t) {
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    k_adam<<<(n + 255) / 256, 256, 0, st>>>(n, p, g, m1, m2, lr, b1, b2, eps, c1, c2);
}

}  // namespace mlfma
