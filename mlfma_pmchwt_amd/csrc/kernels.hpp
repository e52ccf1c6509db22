// Device-side helpers and kernel launch declarations of the matvec.
// Solver vectors are FP64 complex (double2); the far field (patterns, tables,
// workspaces) is stored and accumulated in R = double or float.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "geo.hpp"

namespace mlfma {

template <typename R>
struct cx_of;
template <>
struct cx_of<double> {
    using type = double2;
};
template <>
struct cx_of<float> {
    using type = float2;
};
template <typename R>
using cplx_t = typename cx_of<R>::type;

// Launch wrappers; every one enqueues on `st` and returns immediately.
// Workspace layout of one tree level: [box][cur][comp][K] complex, cur < ncur.

// Sm_leaf[b][cur][c][k] = sum_{i in leaf b} Vs[i][c][k] * x[cur*N + perm[i]]
// (order: the leaves in launch order, largest first)
template <typename R>
void launch_aggregate(hipStream_t st, int ncur, int nleaf, int K, int N, const int* order, const int* leaf_ptr,
                      const int* perm, const cplx_t<R>* Vs, const double2* x, cplx_t<R>* Sm);

// Sm_parent[p][..][q] = sum_sons shift_up[oct][q] * sum_9 w * Sm_child[son][..][tap]
template <typename R>
void launch_m2m(hipStream_t st, int ncur, int nparent, int Kp, int Kc, const int* son_ptr, const int* son_idx,
                const int* octant_child, const int* tap, const R* w, const cplx_t<R>* shift_up, const cplx_t<R>* Sm_child,
                cplx_t<R>* Sm_parent);

// Separable (theta, then phi) form of launch_m2m from the direction tables of
// the stencil; returns false when it does not apply (the caller then runs
// launch_m2m).
template <typename R>
bool launch_m2m_sep(hipStream_t st, int ncur, int nparent, int Lp, int Lc, const int* son_ptr, const int* son_idx,
                    const int* octant_child, const int* th_tap, const R* th_w, const int* ph_tap, const R* ph_w,
                    const cplx_t<R>* shift_up, const cplx_t<R>* Sm_child, cplx_t<R>* Sm_parent);

// Separable anterpolation, after the translation of a level: for every box of
// the level, Bm += the parent's local expansion shifted to the box and
// anterpolated through the transposed direction stencils (theta rows into
// LDS, then phi).  l2l_sep_rows gives the child rows per workgroup that fit
// (0 when the pass does not apply; the translation then fuses the
// anterpolation).
int l2l_sep_rows(int ncur, int Lp, int Lc, size_t csize);
template <typename R>
void launch_l2l_sep(hipStream_t st, int ncur, int nbox, int Lp, int Lc, int rw, const int* parent, const int* oct,
                    const int* at_ptr, const int* at_idx, const R* at_w, const int* ap_ptr, const int* ap_idx,
                    const R* ap_w, const cplx_t<R>* shift_dn, const cplx_t<R>* Bm_parent, cplx_t<R>* Bm);

// Bm[b][..][q] = sum_dr TFw[vec][q] * Sm[dr][..][q]
//              + (parent present) sum_{(kp,w) in T[q]} w * shift_down[oct][kp] * Bm_parent[par][..][kp]
template <typename R>
void launch_m2l_l2l(hipStream_t st, int ncur, int nbox, int K, const int* dr_ptr, const int* dr_idx,
                    const int* dr_vec, const cplx_t<R>* TFw, const cplx_t<R>* Sm, int Kp, const int* parent,
                    const int* octant, const int* t_ptr, const int* t_kp, const R* t_w, const cplx_t<R>* shift_down,
                    const cplx_t<R>* Bm_parent, cplx_t<R>* Bm);

// Same product by sibling clusters: up to 4 children of one parent share every
// source box of the union of their distant-relative lists, read once per column
// group.  variant 1 (FP32, 4 or more currents): 12 workspace columns per
// thread instead of 6, so each translation value serves twice the columns.
template <typename R>
void launch_m2l_cl(hipStream_t st, int variant, int ncur, int ncl, int K, const int* cl_child, const int* cl_ptr,
                   const int* cl_src, const int* cl_vec, const cplx_t<R>* TFw, const cplx_t<R>* Sm, int Kp,
                   const int* parent, const int* octant, const int* t_ptr, const int* t_kp, const R* t_w,
                   const cplx_t<R>* shift_down, const cplx_t<R>* Bm_parent, cplx_t<R>* Bm);

// Same product by all-sibling clusters (up to 8 children of one parent;
// cl_child and cl_vec have 8 slots per cluster) with the members' translation
// samples staged in LDS once per workgroup (FP32, 4, 8 or 16 currents).
template <typename R>
void launch_m2l_cls(hipStream_t st, int ncur, int ncl, int K, const int* cl_child, const int* cl_ptr,
                    const int* cl_src, const int* cl_vec, const cplx_t<R>* TFw, const cplx_t<R>* Sm, int Kp,
                    const int* parent, const int* octant, const int* t_ptr, const int* t_kp, const R* t_w,
                    const cplx_t<R>* shift_down, const cplx_t<R>* Bm_parent, cplx_t<R>* Bm);

// PEC: y[perm[i]] += c * sum_k Vf[i].B_leaf[k], leaf = leaf_of[i]
template <typename R>
void launch_disagg_pec(hipStream_t st, int N, int K, const int* leaf_of, const int* perm, const cplx_t<R>* Vf,
                       const cplx_t<R>* Bm, double2 c, double2* y);

// PMCHWT, both wavenumbers: L and K forms against B_J, B_M (Code/MLFMM.cpp:314-351)
template <typename R>
struct PmchwtFarArgs {
    int K1, K2, N;
    const cplx_t<R>* Vf1;
    const cplx_t<R>* Vf2;
    const cplx_t<R>* B1;   // finest Bm for k1: [leaf][2][3][K1]
    const cplx_t<R>* B2;
    const R* kp1;      // [3][K1] directions
    const R* kp2;
    double2 c1, c2, etai, inv_etai;
};
template <typename R>
void launch_disagg_pmchwt(hipStream_t st, const int* leaf_of, const int* perm, const PmchwtFarArgs<R>& a,
                          double2* y);

// Near-field product y = Z_near x (layout in geo.hpp NearDev), entries stored
// in S = double or float, accumulated in FP64.
template <typename S>
void launch_near(hipStream_t st, const NearDev& nd, const cplx_t<S>* Z, const double2* x, double2* y);

// Batched forms: rb right-hand sides per pass (rb in 1, 2, 4, 8 with
// rb * ncur <= 8); solver vectors RHS-fastest, element (unknown u, rhs r) at
// u * ldr + r; the pass covers r0 .. r0 + rb - 1; workspace current
// c = rr * ncur + cur (m2m / m2l_l2l take nc = rb * ncur as their ncur).
template <typename R>
void launch_aggregate_b(hipStream_t st, int ncur, int rb, int nleaf, int K, int N, int ldr, int r0, const int* order,
                        const int* leaf_ptr, const int* perm, const cplx_t<R>* Vs, const double2* x, cplx_t<R>* Sm);
template <typename R>
void launch_disagg_pec_b(hipStream_t st, int rb, int N, int K, int ldr, int r0, const int* leaf_of, const int* perm,
                         const cplx_t<R>* Vf, const cplx_t<R>* Bm, double2 c, double2* y);
template <typename R>
void launch_disagg_pmchwt_b(hipStream_t st, int rb, int ldr, int r0, const int* leaf_of, const int* perm,
                            const PmchwtFarArgs<R>& a, double2* y);
template <typename S>
void launch_near_b(hipStream_t st, int rb, int ldr, int r0, const NearDev& nd, const cplx_t<S>* Z, const double2* x,
                   double2* y);

// Near product of R right-hand sides at once on the matrix cores (FP32 near
// entries, FP32 accumulation); vectors RHS-fastest with stride ldr from r0;
// xf: workspace of (unknowns x R) float2 for the FP32 copy of x.
void launch_near_mfma(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const float2* Z, const double2* x,
                      float2* xf, double2* y);
void launch_near_mfma64(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const double2* Z,
                        const double2* x, double2* y);

// Batched FP32 disaggregation on the matrix cores: effective receive-side
// matrices E[leaf][rr][o][3][K] from the finest Bm of a pass (constants and
// kp x B folded in), then y += V E per leaf (o = 0: rows of J / PEC, 1: M).
template <typename T>
void launch_beff(hipStream_t st, int nleaf, int rb, int K, bool pmchwt, const cplx_t<T>* Bm, const T* kp,
                 double2 ca, double2 cb, double2 cc, double2 cd, cplx_t<T>* E);
template <typename T>
void launch_disagg_mfma(hipStream_t st, int nleaf, int N, int ldr, int r0, int rb, const int* order,
                        const int* leaf_ptr, const int* perm, bool pmchwt, int K1, const cplx_t<T>* V1,
                        const cplx_t<T>* E1, int K2, const cplx_t<T>* V2, const cplx_t<T>* E2, double2* y);
// All right-hand sides at once (FP32): the passes' effective matrices go into
// Eall[leaf][o][3K][R] at columns r0 .. r0 + rb - 1, then one product per
// matvec over the near product's work items (nd.items, leaf rows).
// (split: each element stored as its BF16 pairs, uint2 .x = hi Re | hi Im << 16,
// .y = lo Re | lo Im << 16, for launch_disagg_bf3)
void launch_beff_all(hipStream_t st, int nleaf, int rb, int K, bool pmchwt, const float2* Bm, const float* kp,
                     double2 ca, double2 cb, double2 cc, double2 cd, int R, int r0, bool split, float2* Eall);
void launch_disagg_all(hipStream_t st, int R, const NearDev& nd, bool pmchwt, int K1, const float2* V1,
                       const float2* E1, int K2, const float2* V2, const float2* E2, double2* y);
// Split-pair BF16 disaggregation of all R: pattern planes [Re hi, Im hi,
// Re lo, Im lo][N][Qp] (Qp = 3K rounded up to 32) from the FP32 patterns,
// and the product from Eall stored as BF16 pairs (launch_beff_all, split).
void launch_split_rows(hipStream_t st, int N, int Q, int Qp, const float2* V, unsigned short* planes);
void launch_disagg_bf3(hipStream_t st, int R, const NearDev& nd, bool pmchwt, int K1, const unsigned short* P1,
                       const uint2* E1, int K2, const unsigned short* P2, const uint2* E2, double2* y);
// Aggregation of all R right-hand sides (FP32 patterns, x RHS-fastest with
// stride R) into the leaf workspaces of every pass of the batched matvec
// (passes of rb = largest power of two <= cap, pass (r0, rb) at
// Sall + nleaf ncur r0 3K).
void launch_aggregate_all(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, const int* leaf_ptr,
                          const int* perm, const float2* Vs, const double2* x, float2* Sall);
// The same product on split-pair BF16 from transposed pattern planes
// [hi Re, hi Im, lo Re, lo Im][Qp][Np] (launch_split_cols; Qp a multiple of
// 64, leaf t's RWGs from the 32-aligned column joff[t] on, jmap[column] = RWG
// or -1).
void launch_split_cols(hipStream_t st, int Q, int Qp, int Np, const int* jmap, const float2* V,
                       unsigned short* planes);
void launch_aggregate_bf3(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, int Qp, int Np,
                          const int* leaf_ptr, const int* joff, const int* perm, const unsigned short* planes,
                          const double2* x, float2* Sall);
// (x already split into BF16 pairs, RHS-fastest: .x = hi Re | hi Im << 16,
// .y = lo Re | lo Im << 16, as the split-pair near product stages it)
void launch_aggregate_bf3(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, int Qp, int Np,
                          const int* leaf_ptr, const int* joff, const int* perm, const unsigned short* planes,
                          const uint2* xsp, float2* Sall);

// FP32 copy of an FP64 device array (set-up: patterns computed in FP64)
void launch_narrow(hipStream_t st, size_t n, const double2* in, float2* out);

// BF16 near-field copies: per-target record offsets (in 16-byte units;
// returns the total), packing from the FP64 entries, and the batched products.
// Single BF16 (the polynomial preconditioner): x rounded to BF16 into xb,
// nu x R.  Split pairs (split = true; the operator's batched products with
// near_precision = MLFMA_BF16): hi + lo BF16 of entries and of x (xb nu x R
// pairs), three partial products.
long long bf16_near_layout(const std::vector<int>& leaf_ptr, const std::vector<int>& tgt_cols, int planes,
                           bool split, std::vector<long long>& bf_off);
void launch_pack_bf16(hipStream_t st, const NearDev& nd, const double2* Z, const long long* bf_off, bool split,
                      uint4* out);
void launch_near_bf16(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                      const long long* bf_off, const double2* x, unsigned* xb, double2* y);
void launch_near_bf16x3(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                        const long long* bf_off, const double2* x, uint2* xb, double2* y);
void launch_near_bf16x3(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                        const long long* bf_off, const float2* x, uint2* xb, double2* y);

// The polynomial preconditioner with FP32 vectors around the single-BF16
// products (all unknowns, R right-hand sides, stride R):
//   t = D^-1 x, y = a0 t, xb = BF16(t)                       (launch_poly_init_bf16)
//   t -= D^-1 Z xb, y += pe.a t, pe.xbn = BF16(t)            (launch_near_bf16_poly)
// pe.xbn may be null (the last step) and must not be xb.
struct PolyEpi {
    const float2* dinv;   // 1 / diagonal, per unknown
    float2 a;             // coefficient of this step
    float2* t;            // current polynomial basis vector (updated)
    float2* y;            // accumulated result (updated)
    unsigned* xbn;        // BF16 of the new t, the next product's input
};
void launch_poly_init_bf16(hipStream_t st, int R, const NearDev& nd, const float2* dinv, const float2* x, float2 a0,
                           float2* t, float2* y, unsigned* xb);
void launch_near_bf16_poly(hipStream_t st, int R, const NearDev& nd, const uint4* Zb, const long long* bf_off,
                           const unsigned* xb, const PolyEpi& pe);

}  // namespace mlfma
