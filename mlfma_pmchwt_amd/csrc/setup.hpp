// Launch declarations of the FP64 setup / excitation / far-field kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "geo.hpp"

namespace mlfma {

struct NearConsts {
    double2 k1, k2, eta1, etai, inv_etai;
    double alpha;
};

struct ExcitationArgs {
    double kinc[3], einc[3], hinc[3];
    double2 k;
    double alpha;
};

// mode 0 EFIE, 1 CFIE, 2 PMCHWT; one workgroup per leaf-pair block.
void launch_near_fill(hipStream_t st, int mode, const NearDev& nd, const RwgGeo* geo, const GaussRule& g,
                      const NearConsts& c, double2* Z);
void launch_pair_probe(hipStream_t st, const RwgGeo* geo, int f, int s, const GaussRule& g, const NearConsts& c,
                       double2* out);
void launch_patterns(hipStream_t st, bool cfie, int N, int K, const RwgGeo* geo, const int* leaf_of,
                     const double* centre, const double* kp, double2 k, const GaussRule& g, double alpha, double2* Vs,
                     double2* Vf);
void launch_excitation(hipStream_t st, int N, int mode, const RwgGeo* geo, const int* perm, const GaussRule& g,
                       const ExcitationArgs& a, double2* v);
// istride: offset between the current vectors of consecutive angles (0: one vector for all angles)
void launch_farfield(hipStream_t st, int N, int pmchwt, const RwgGeo* geo, const int* perm, const GaussRule& g,
                     double2 k, int nang, const double* dirs, const double2* I, size_t istride, double2* partial,
                     int nblocks);

// Allocation-free RCS (k_rcs_noalloc.hip): directions and partial sums in a
// file-scope device array, looked up once per device by rcs_scratch_init.
void rcs_scratch_init(int device);
size_t rcs_scratch_capacity();
void launch_rcs_noalloc(hipStream_t st, int device, int N, int pmchwt, const RwgGeo* geo, const int* perm,
                        const GaussRule& g, double2 k1, double2 coeff, int nang, const double* th_dev,
                        const double* ph_dev, const double2* I_dev, double* out_dev);

// Krylov vector kernels (k_solve.hip)
// partial[j*nblk + b] = sum over block b of conj(V_j) . w, j < m
void launch_multidot(hipStream_t st, int n, int m, const double2* V, size_t ldv, const double2* w, double2* partial,
                     int nblk);
// out[j] = sum_b partial[j*nblk + b] (fixed order), j < m
void launch_reduce_partials(hipStream_t st, int m, int nblk, const double2* partial, double2* out);
// w -= sum_j h[j] V_j
void launch_sub_combination(hipStream_t st, int n, int m, const double2* V, size_t ldv, const double2* h, double2* w);
// x += sum_j y[j] V_j
void launch_add_combination(hipStream_t st, int n, int m, const double2* V, size_t ldv, const double2* y, double2* x);
// y = a * x (a real)
void launch_scale_copy(hipStream_t st, int n, double a, const double2* x, double2* y);
// r = b - r
void launch_residual(hipStream_t st, int n, const double2* b, double2* r);
// partial[b] = sum |x|^2 over block b
void launch_norm2_partial(hipStream_t st, int n, const double2* x, double* partial, int nblk);
// y = x + a * z   (complex a)
void launch_axpy(hipStream_t st, int n, double2 a, const double2* z, const double2* x, double2* y);
// sum_i a_i b_i (bilinear) partials
void launch_bidot_partial(hipStream_t st, int n, const double2* a, const double2* b, double2* partial, int nblk);
// Batched Krylov kernels: R <= 256 right-hand sides, vectors RHS-fastest
// (element (u, r) at u * R + r); per-r scalars indexed j * R + r.
void launch_multidot_b(hipStream_t st, int n, int R, int m, const double2* V, size_t ldv, const double2* w,
                       double2* partial, int nblk);
void launch_reduce_partials_b(hipStream_t st, int m, int R, int nblk, const double2* partial, double2* out);
void launch_norm2_partial_b(hipStream_t st, int n, int R, const double2* x, double2* partial, int nblk);
void launch_sub_combination_b(hipStream_t st, int n, int R, int m, const double2* V, size_t ldv, const double2* h,
                              double2* w);
void launch_add_combination_b(hipStream_t st, int n, int R, int m, const double2* V, size_t ldv, const double2* y,
                              double2* x);
void launch_scale_copy_b(hipStream_t st, int n, int R, const double* a, const double2* x, double2* y);
// the same with an FP32 basis V (FP32-mode batched solves): V is widened exactly
// and the sums stay in FP64; scale_copy rounds its output to FP32
void launch_multidot_b(hipStream_t st, int n, int R, int m, const float2* V, size_t ldv, const double2* w,
                       double2* partial, int nblk);
void launch_sub_combination_b(hipStream_t st, int n, int R, int m, const float2* V, size_t ldv, const double2* h,
                              double2* w);
void launch_add_combination_b(hipStream_t st, int n, int R, int m, const float2* V, size_t ldv, const double2* y,
                              double2* x);
void launch_scale_copy_b(hipStream_t st, int n, int R, const double* a, const double2* x, float2* y);
void launch_widen(hipStream_t st, size_t n, const float2* in, double2* out);
// out (cols x rows) = transpose of in (rows x cols), both row-major
void launch_transpose(hipStream_t st, int rows, int cols, const double2* in, double2* out);
void launch_gather_cols(hipStream_t st, int n, int R, int Rc, const int* idx, const double2* in, double2* out);
void launch_scatter_cols(hipStream_t st, int n, int R, int Rc, const int* idx, const double2* in, double2* out);

// block-diagonal preconditioner: y_t = D_t^{-1} x_t per leaf (dense inverse blocks)
void launch_block_diag(hipStream_t st, int planes, int nleaf, int N, const int* leaf_ptr, const int* perm,
                       const long long* inv_off, const double2* Dinv, int max_leaf, const double2* x, double2* y);
// the same on RHS-fastest blocks of R right-hand sides (element (u, r) at u * R + r)
void launch_block_diag_b(hipStream_t st, int planes, int nleaf, int N, int R, const int* leaf_ptr, const int* perm,
                         const long long* inv_off, const double2* Dinv, int max_leaf, const double2* x, double2* y);

}  // namespace mlfma
