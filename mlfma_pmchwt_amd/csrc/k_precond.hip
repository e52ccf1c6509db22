// Near-field polynomial preconditioner kernels (FP64 vectors, RHS-fastest
// blocks: element (u, r) at u * R + r).  No reference counterpart: the
// reference solves unpreconditioned (Code/MLFMM.cpp:370-397).
//
//   M^-1 r = sum_{k=0..d} alpha_k t_k,  t_0 = D^-1 r,  t_{k+1} = t_k - D^-1 Z_near t_k
//
// with D the diagonal of the operator (every diagonal entry is a near-field
// self entry) and Z_near the stored near field.
#include "cplx.hpp"
#include "precond.hpp"

namespace mlfma {

// Diagonal of the operator from the self block of every leaf (the last pair of
// its block row, columns C_t - nt .. C_t - 1).  PMCHWT: z11 for J rows, z22 for
// M rows.  Written in caller order.
__global__ void k_near_diag(NearDev nd, const double2* __restrict__ Z, double2* __restrict__ dinv) {
    const int t = blockIdx.x;
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    const int C = nd.tgt_cols[t];
    const double2* Zt = Z + nd.tgt_off[t];
    const size_t plane = (size_t)C * nt;
    for (int r = threadIdx.x; r < nt; r += blockDim.x) {
        const size_t e = (size_t)(C - nt + r) * nt + r;
        const int u = nd.perm[i0 + r];
        dinv[u] = cdiv(make_double2(1.0, 0.0), Zt[e]);
        if (nd.planes == 3) dinv[nd.N + u] = cdiv(make_double2(1.0, 0.0), Zt[2 * plane + e]);
    }
}

void launch_near_diag_inv(hipStream_t st, const NearDev& nd, const double2* Z, double2* dinv) {
    if (nd.nleaf == 0) return;
    k_near_diag<<<nd.nleaf, 256, 0, st>>>(nd, Z, dinv);
}

// t = D^-1 x ; y = a0 t
__global__ void k_poly_init(size_t nR, int R, const double2* __restrict__ dinv, const double2* __restrict__ x,
                            double2 a0, double2* __restrict__ t, double2* __restrict__ y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nR) return;
    const double2 v = cmul(dinv[e / R], x[e]);
    t[e] = v;
    y[e] = cmul(a0, v);
}

void launch_poly_init(hipStream_t st, int n, int R, const double2* dinv, const double2* x, double2 a0, double2* t,
                      double2* y) {
    const size_t nR = (size_t)n * R;
    k_poly_init<<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, dinv, x, a0, t, y);
}

// t -= D^-1 u ; y += a t   (y may be null: basis only)
__global__ void k_poly_step(size_t nR, int R, const double2* __restrict__ dinv, const double2* __restrict__ u,
                            double2 a, double2* __restrict__ t, double2* __restrict__ y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nR) return;
    double2 v = t[e];
    cfms(v, dinv[e / R], u[e]);
    t[e] = v;
    if (y) {
        double2 w = y[e];
        cfma(w, a, v);
        y[e] = w;
    }
}

void launch_poly_step(hipStream_t st, int n, int R, const double2* dinv, const double2* u, double2 a, double2* t,
                      double2* y) {
    const size_t nR = (size_t)n * R;
    k_poly_step<<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, dinv, u, a, t, y);
}

// y = D^-1 u (elementwise over a block)
__global__ void k_diag_scale(size_t nR, int R, const double2* __restrict__ dinv, const double2* __restrict__ u,
                             double2* __restrict__ y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nR) y[e] = cmul(dinv[e / R], u[e]);
}

void launch_diag_scale(hipStream_t st, int n, int R, const double2* dinv, const double2* u, double2* y) {
    const size_t nR = (size_t)n * R;
    k_diag_scale<<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, dinv, u, y);
}

// Unit-modulus probe vector with hashed phases (seeded, reproducible).
__global__ void k_probe(int n, unsigned seed, double2* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned h = (unsigned)i * 2654435761u ^ (seed * 0x9e3779b9u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    const double ph = 6.283185307179586 * (h * (1.0 / 4294967296.0));
    double s, c;
    sincos(ph, &s, &c);
    z[i] = make_double2(c, s);
}

void launch_probe(hipStream_t st, int n, unsigned seed, double2* z) {
    k_probe<<<(n + 255) / 256, 256, 0, st>>>(n, seed, z);
}

}  // namespace mlfma
