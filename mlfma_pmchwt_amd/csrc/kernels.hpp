// Device-side helpers and kernel launch declarations (FP64 complex = double2).
#pragma once
#include <hip/hip_runtime.h>

#include "geo.hpp"

namespace mlfma {

// Launch wrappers; every one enqueues on `st` and returns immediately.
// Workspace layout of one tree level: [box][cur][comp][K] complex, cur < ncur.

// Sm_leaf[b][cur][c][k] = sum_{i in leaf b} Vs[i][c][k] * x[cur*N + perm[i]]
// (order: the leaves in launch order, largest first)
void launch_aggregate(hipStream_t st, int ncur, int nleaf, int K, int N, const int* order, const int* leaf_ptr,
                      const int* perm, const double2* Vs, const double2* x, double2* Sm);

// Sm_parent[p][..][q] = sum_sons shift_up[oct][q] * sum_9 w * Sm_child[son][..][tap]
void launch_m2m(hipStream_t st, int ncur, int nparent, int Kp, int Kc, const int* son_ptr, const int* son_idx,
                const int* octant_child, const int* tap, const double* w, const double2* shift_up,
                const double2* Sm_child, double2* Sm_parent);

// Bm[b][..][q] = sum_dr TFw[vec][q] * Sm[dr][..][q]
//              + (parent present) sum_{(kp,w) in T[q]} w * shift_down[oct][kp] * Bm_parent[par][..][kp]
void launch_m2l_l2l(hipStream_t st, int ncur, int nbox, int K, const int* dr_ptr, const int* dr_idx,
                    const int* dr_vec, const double2* TFw, const double2* Sm, int Kp, const int* parent,
                    const int* octant, const int* t_ptr, const int* t_kp, const double* t_w,
                    const double2* shift_down, const double2* Bm_parent, double2* Bm);

// PEC: y[perm[i]] += c * sum_k Vf[i].B_leaf[k], leaf = leaf_of[i]
void launch_disagg_pec(hipStream_t st, int N, int K, const int* leaf_of, const int* perm, const double2* Vf,
                       const double2* Bm, double2 c, double2* y);

// PMCHWT, both wavenumbers: L and K forms against B_J, B_M (Code/MLFMM.cpp:314-351)
struct PmchwtFarArgs {
    int K1, K2, N;
    const double2* Vf1;
    const double2* Vf2;
    const double2* B1;   // finest Bm for k1: [leaf][2][3][K1]
    const double2* B2;
    const double* kp1;   // [3][K1] directions
    const double* kp2;
    double2 c1, c2, etai, inv_etai;
};
void launch_disagg_pmchwt(hipStream_t st, const int* leaf_of, const int* perm, const PmchwtFarArgs& a, double2* y);

// Near-field product y = Z_near x (layout in geo.hpp NearDev).
void launch_near(hipStream_t st, const NearDev& nd, const double2* Z, const double2* x, double2* y);

}  // namespace mlfma
