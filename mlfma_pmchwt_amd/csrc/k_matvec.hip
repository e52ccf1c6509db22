// FP64 MLFMA matvec kernels for gfx950 (64-wide wavefronts).
//
// Workspaces are laid out [box][cur][comp][K] so that consecutive threads
// (consecutive k-space samples) touch consecutive addresses.  RWG functions are
// stored in leaf-major order (internal index i); perm[i] is the reference index
// used by the caller's x and y vectors.
#include "kernels.hpp"
#include "cplx.hpp"

namespace mlfma {

// Blocks are dispatched round-robin over the 8 XCDs of the MI355X; remap the
// linear block id so that every XCD works on one contiguous (Morton-ordered)
// range of boxes and its L2 sees the shared source boxes.
__device__ __forceinline__ int xcd_remap(int lin, int total) {
    const int per = total >> 3, rem = total & 7;
    const int x = lin & 7, s = lin >> 3;
    return x * per + min(x, rem) + s;
}

// ---------------------------------------------------------------- aggregation
// Code/ProcessMLFMM.h:33-56 (mexp), J and M currents of PMCHWT in one pass.
template <int NC>
__global__ __launch_bounds__(256) void k_aggregate(int K, int N, const int* __restrict__ leaf_ptr,
                                                   const int* __restrict__ perm, const double2* __restrict__ Vs,
                                                   const double2* __restrict__ x, double2* __restrict__ Sm) {
    const int leaf = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= K) return;
    double2 acc[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[c][d] = make_double2(0.0, 0.0);
    const int i0 = leaf_ptr[leaf], i1 = leaf_ptr[leaf + 1];
    for (int i = i0; i < i1; ++i) {
        const int r = perm[i];
        const double2* v = Vs + (size_t)i * 3 * K + q;
        const double2 v0 = v[0], v1 = v[K], v2 = v[2 * K];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double2 xv = x[(size_t)c * N + r];
            cfma(acc[c][0], v0, xv);
            cfma(acc[c][1], v1, xv);
            cfma(acc[c][2], v2, xv);
        }
    }
    double2* s = Sm + (size_t)leaf * NC * 3 * K + q;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) s[(size_t)(c * 3 + d) * K] = acc[c][d];
}

void launch_aggregate(hipStream_t st, int ncur, int nleaf, int K, int N, const int* leaf_ptr, const int* perm,
                      const double2* Vs, const double2* x, double2* Sm) {
    dim3 grid((K + 255) / 256, nleaf);
    if (ncur == 1) k_aggregate<1><<<grid, 256, 0, st>>>(K, N, leaf_ptr, perm, Vs, x, Sm);
    else k_aggregate<2><<<grid, 256, 0, st>>>(K, N, leaf_ptr, perm, Vs, x, Sm);
}

// ---------------------------------------------------------------- interpolation
// Code/ProcessMLFMM.h:61-107 (m2m): 3x3 Lagrange taps into the child grid, then
// the octant phase shift, summed over the children of each parent.
template <int NC>
__global__ __launch_bounds__(256) void k_m2m(int Kp, int Kc, const int* __restrict__ son_ptr,
                                             const int* __restrict__ son_idx, const int* __restrict__ oct_child,
                                             const int* __restrict__ tap, const double* __restrict__ w,
                                             const double2* __restrict__ shup, const double2* __restrict__ Smc,
                                             double2* __restrict__ Smp) {
    const int p = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Kp) return;
    int t[9];
    double ww[9];
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        t[s] = tap[(size_t)s * Kp + q];
        ww[s] = w[(size_t)s * Kp + q];
    }
    double2 acc[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[c][d] = make_double2(0.0, 0.0);
    for (int j = son_ptr[p]; j < son_ptr[p + 1]; ++j) {
        const int c = son_idx[j];
        const double2 e = shup[(size_t)oct_child[c] * Kp + q];
        const double2* S = Smc + (size_t)c * NC * 3 * Kc;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double2* s = S + (size_t)(cc * 3 + d) * Kc;
                double2 r = make_double2(0.0, 0.0);
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const double2 f = s[t[k]];
                    r.x += f.x * ww[k];
                    r.y += f.y * ww[k];
                }
                cfma(acc[cc][d], e, r);
            }
    }
    double2* o = Smp + (size_t)p * NC * 3 * Kp + q;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int d = 0; d < 3; ++d) o[(size_t)(cc * 3 + d) * Kp] = acc[cc][d];
}

void launch_m2m(hipStream_t st, int ncur, int nparent, int Kp, int Kc, const int* son_ptr, const int* son_idx,
                const int* octant_child, const int* tap, const double* w, const double2* shift_up,
                const double2* Sm_child, double2* Sm_parent) {
    dim3 grid((Kp + 255) / 256, nparent);
    if (ncur == 1)
        k_m2m<1><<<grid, 256, 0, st>>>(Kp, Kc, son_ptr, son_idx, octant_child, tap, w, shift_up, Sm_child, Sm_parent);
    else
        k_m2m<2><<<grid, 256, 0, st>>>(Kp, Kc, son_ptr, son_idx, octant_child, tap, w, shift_up, Sm_child, Sm_parent);
}

// ---------------------------------------------------------------- translation + anterpolation
// Code/ProcessMLFMM.h:112-147 (m2l, quadrature weight folded into TFw) and
// :152-202 (l2l) in gather form: each child sample sums the parent samples whose
// 9-tap stencil contains it (transposed table, ascending parent sample), so no
// atomics and a fixed summation order.
template <int NC>
__global__ __launch_bounds__(256) void k_m2l_l2l(int nbox, int K, int ktiles, const int* __restrict__ dr_ptr,
                                                 const int* __restrict__ dr_idx, const int* __restrict__ dr_vec,
                                                 const double2* __restrict__ TFw, const double2* __restrict__ Sm,
                                                 int Kp, const int* __restrict__ parent, const int* __restrict__ oct,
                                                 const int* __restrict__ t_ptr, const int* __restrict__ t_kp,
                                                 const double* __restrict__ t_w, const double2* __restrict__ shdn,
                                                 const double2* __restrict__ Bmp, double2* __restrict__ Bm) {
    const int lin = xcd_remap(blockIdx.x, nbox * ktiles);
    const int b = lin / ktiles;
    const int q = (lin % ktiles) * blockDim.x + threadIdx.x;
    if (q >= K) return;
    double2 acc[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[c][d] = make_double2(0.0, 0.0);
    for (int j = dr_ptr[b]; j < dr_ptr[b + 1]; ++j) {
        const double2 T = TFw[(size_t)dr_vec[j] * K + q];
        const double2* S = Sm + (size_t)dr_idx[j] * NC * 3 * K + q;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int d = 0; d < 3; ++d) cfma(acc[cc][d], T, S[(size_t)(cc * 3 + d) * K]);
    }
    if (Bmp != nullptr) {
        const int pb = parent[b];
        const double2* sh = shdn + (size_t)oct[b] * Kp;
        const double2* P = Bmp + (size_t)pb * NC * 3 * Kp;
        for (int j = t_ptr[q]; j < t_ptr[q + 1]; ++j) {
            const int kp = t_kp[j];
            double2 e = sh[kp];
            e.x *= t_w[j];
            e.y *= t_w[j];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int d = 0; d < 3; ++d) cfma(acc[cc][d], e, P[(size_t)(cc * 3 + d) * Kp + kp]);
        }
    }
    double2* o = Bm + (size_t)b * NC * 3 * K + q;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int d = 0; d < 3; ++d) o[(size_t)(cc * 3 + d) * K] = acc[cc][d];
}

void launch_m2l_l2l(hipStream_t st, int ncur, int nbox, int K, const int* dr_ptr, const int* dr_idx,
                    const int* dr_vec, const double2* TFw, const double2* Sm, int Kp, const int* parent,
                    const int* octant, const int* t_ptr, const int* t_kp, const double* t_w,
                    const double2* shift_down, const double2* Bm_parent, double2* Bm) {
    const int kt = (K + 255) / 256;
    dim3 grid(kt * nbox);
    if (ncur == 1)
        k_m2l_l2l<1><<<grid, 256, 0, st>>>(nbox, K, kt, dr_ptr, dr_idx, dr_vec, TFw, Sm, Kp, parent, octant, t_ptr,
                                           t_kp, t_w, shift_down, Bm_parent, Bm);
    else
        k_m2l_l2l<2><<<grid, 256, 0, st>>>(nbox, K, kt, dr_ptr, dr_idx, dr_vec, TFw, Sm, Kp, parent, octant, t_ptr,
                                           t_kp, t_w, shift_down, Bm_parent, Bm);
}

// ---------------------------------------------------------------- disaggregation
// Code/MLFMM.cpp:211-227: w_m += (k eta/4pi)(k/4pi) sum_k Vfmj_m . Bm_leaf.
// One wavefront per RWG row, lanes over k, then a wavefront reduction.
__global__ __launch_bounds__(256) void k_disagg_pec(int K, const int* __restrict__ leaf_ptr,
                                                    const int* __restrict__ perm, const double2* __restrict__ Vf,
                                                    const double2* __restrict__ Bm, double2 c,
                                                    double2* __restrict__ y) {
    const int leaf = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double2* B = Bm + (size_t)leaf * 3 * K;
    for (int i = leaf_ptr[leaf] + wave; i < leaf_ptr[leaf + 1]; i += 4) {
        const double2* V = Vf + (size_t)i * 3 * K;
        double2 s = make_double2(0.0, 0.0);
        for (int q = lane; q < K; q += 64) {
            cfma(s, V[q], B[q]);
            cfma(s, V[K + q], B[K + q]);
            cfma(s, V[2 * K + q], B[2 * K + q]);
        }
        s = wave_sum(s);
        if (lane == 0) {
            const int r = perm[i];
            double2 t = y[r];
            cfma(t, c, s);
            y[r] = t;
        }
    }
}

void launch_disagg_pec(hipStream_t st, int nleaf, int K, const int* leaf_ptr, const int* perm, const double2* Vf,
                       const double2* Bm, double2 c, double2* y) {
    k_disagg_pec<<<nleaf, 256, 0, st>>>(K, leaf_ptr, perm, Vf, Bm, c, y);
}

// Code/MLFMM.cpp:285-351.  The K integral sum (-khat x V).B with khat = -kp is
// evaluated as sum V.(kp x B).
__device__ __forceinline__ void far_sums(int K, const double2* __restrict__ V, const double2* __restrict__ BJ,
                                         const double2* __restrict__ BM, const double* __restrict__ kp, int lane,
                                         double2& LJ, double2& KJ, double2& LM, double2& KM) {
    for (int q = lane; q < K; q += 64) {
        const double kx = kp[q], ky = kp[K + q], kz = kp[2 * K + q];
        const double2 v0 = V[q], v1 = V[K + q], v2 = V[2 * K + q];
        const double2 j0 = BJ[q], j1 = BJ[K + q], j2 = BJ[2 * K + q];
        const double2 m0 = BM[q], m1 = BM[K + q], m2 = BM[2 * K + q];
        cfma(LJ, v0, j0);
        cfma(LJ, v1, j1);
        cfma(LJ, v2, j2);
        cfma(LM, v0, m0);
        cfma(LM, v1, m1);
        cfma(LM, v2, m2);
        // kp x B
        const double2 cj0 = csub(cscale(j2, ky), cscale(j1, kz));
        const double2 cj1 = csub(cscale(j0, kz), cscale(j2, kx));
        const double2 cj2 = csub(cscale(j1, kx), cscale(j0, ky));
        const double2 cm0 = csub(cscale(m2, ky), cscale(m1, kz));
        const double2 cm1 = csub(cscale(m0, kz), cscale(m2, kx));
        const double2 cm2 = csub(cscale(m1, kx), cscale(m0, ky));
        cfma(KJ, v0, cj0);
        cfma(KJ, v1, cj1);
        cfma(KJ, v2, cj2);
        cfma(KM, v0, cm0);
        cfma(KM, v1, cm1);
        cfma(KM, v2, cm2);
    }
}

__global__ __launch_bounds__(256) void k_disagg_pmchwt(const int* __restrict__ leaf_ptr, const int* __restrict__ perm,
                                                       PmchwtFarArgs a, double2* __restrict__ y) {
    const int leaf = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double2* BJ1 = a.B1 + (size_t)leaf * 6 * a.K1;
    const double2* BM1 = BJ1 + 3 * (size_t)a.K1;
    const double2* BJ2 = a.B2 + (size_t)leaf * 6 * a.K2;
    const double2* BM2 = BJ2 + 3 * (size_t)a.K2;
    for (int i = leaf_ptr[leaf] + wave; i < leaf_ptr[leaf + 1]; i += 4) {
        double2 L1J = {0, 0}, K1J = {0, 0}, L1M = {0, 0}, K1M = {0, 0};
        double2 L2J = {0, 0}, K2J = {0, 0}, L2M = {0, 0}, K2M = {0, 0};
        far_sums(a.K1, a.Vf1 + (size_t)i * 3 * a.K1, BJ1, BM1, a.kp1, lane, L1J, K1J, L1M, K1M);
        far_sums(a.K2, a.Vf2 + (size_t)i * 3 * a.K2, BJ2, BM2, a.kp2, lane, L2J, K2J, L2M, K2M);
        L1J = wave_sum(L1J);
        K1J = wave_sum(K1J);
        L1M = wave_sum(L1M);
        K1M = wave_sum(K1M);
        L2J = wave_sum(L2J);
        K2J = wave_sum(K2J);
        L2M = wave_sum(L2M);
        K2M = wave_sum(K2M);
        if (lane == 0) {
            const int r = perm[i];
            double2 wj = y[r], wm = y[a.N + r];
            // Code/MLFMM.cpp:339-349
            wj = cadd(wj, cmul(a.c1, L1J));
            wj = csub(wj, cmul(a.c1, K1M));
            wm = cadd(wm, cmul(a.c1, K1J));
            wm = cadd(wm, cmul(a.c1, L1M));
            wj = cadd(wj, cmul(cmul(a.c2, a.etai), L2J));
            wj = csub(wj, cmul(a.c2, K2M));
            wm = cadd(wm, cmul(a.c2, K2J));
            wm = cadd(wm, cmul(cmul(a.c2, L2M), a.inv_etai));
            y[r] = wj;
            y[a.N + r] = wm;
        }
    }
}

void launch_disagg_pmchwt(hipStream_t st, int nleaf, const int* leaf_ptr, const int* perm, const PmchwtFarArgs& a,
                          double2* y) {
    k_disagg_pmchwt<<<nleaf, 256, 0, st>>>(leaf_ptr, perm, a, y);
}

// ---------------------------------------------------------------- near field
// Code/MLFMM.cpp:228-231 (PEC) and :354-366 (PMCHWT) on leaf-pair blocks.  A
// block is stored column-major (consecutive rows of one source column are
// contiguous) so that a wavefront reads 64 consecutive entries.  PMCHWT blocks
// hold three planes z11, zK, z22 with z12 = -zK, z21 = zK (Code/Zmartix.h:611-614).
template <int P>
__global__ __launch_bounds__(256) void k_near(int N, int max_leaf, const int* __restrict__ leaf_ptr,
                                              const int* __restrict__ perm, const int* __restrict__ pair_ptr,
                                              const int* __restrict__ pair_src,
                                              const long long* __restrict__ pair_off, const double2* __restrict__ Z,
                                              const double2* __restrict__ x, double2* __restrict__ y) {
    extern __shared__ double2 xs[];   // [P == 3 ? 2 : 1][max_leaf]
    const int t = blockIdx.x;
    const int i0 = leaf_ptr[t], nt = leaf_ptr[t + 1] - i0;
    for (int r0 = 0; r0 < nt; r0 += blockDim.x) {
        const int row = r0 + threadIdx.x;
        double2 aJ = {0, 0}, aM = {0, 0};
        for (int p = pair_ptr[t]; p < pair_ptr[t + 1]; ++p) {
            const int s = pair_src[p];
            const int j0 = leaf_ptr[s], ns = leaf_ptr[s + 1] - j0;
            __syncthreads();
            for (int j = threadIdx.x; j < ns; j += blockDim.x) {
                const int r = perm[j0 + j];
                xs[j] = x[r];
                if (P == 3) xs[max_leaf + j] = x[N + r];
            }
            __syncthreads();
            if (row < nt) {
                const double2* z = Z + pair_off[p] + row;
                if (P == 1) {
                    int j = 0;
                    for (; j + 3 < ns; j += 4) {
                        const double2 z0 = z[(size_t)j * nt], z1 = z[(size_t)(j + 1) * nt];
                        const double2 z2 = z[(size_t)(j + 2) * nt], z3 = z[(size_t)(j + 3) * nt];
                        cfma(aJ, z0, xs[j]);
                        cfma(aJ, z1, xs[j + 1]);
                        cfma(aJ, z2, xs[j + 2]);
                        cfma(aJ, z3, xs[j + 3]);
                    }
                    for (; j < ns; ++j) cfma(aJ, z[(size_t)j * nt], xs[j]);
                } else {
                    const double2* zk = z + (size_t)ns * nt;
                    const double2* zm = z + (size_t)2 * ns * nt;
                    for (int j = 0; j < ns; ++j) {
                        const double2 a11 = z[(size_t)j * nt], ak = zk[(size_t)j * nt], a22 = zm[(size_t)j * nt];
                        const double2 xJ = xs[j], xM = xs[max_leaf + j];
                        cfma(aJ, a11, xJ);
                        cfms(aJ, ak, xM);
                        cfma(aM, ak, xJ);
                        cfma(aM, a22, xM);
                    }
                }
            }
        }
        if (row < nt) {
            const int r = perm[i0 + row];
            y[r] = aJ;
            if (P == 3) y[N + r] = aM;
        }
    }
}

void launch_near(hipStream_t st, int planes, int nleaf, int N, const int* leaf_ptr, const int* perm,
                 const int* pair_ptr, const int* pair_src, const long long* pair_off, const double2* Z, int max_leaf,
                 const double2* x, double2* y) {
    if (planes == 1) {
        size_t sh = (size_t)max_leaf * sizeof(double2);
        k_near<1><<<nleaf, 256, sh, st>>>(N, max_leaf, leaf_ptr, perm, pair_ptr, pair_src, pair_off, Z, x, y);
    } else {
        size_t sh = (size_t)2 * max_leaf * sizeof(double2);
        k_near<3><<<nleaf, 256, sh, st>>>(N, max_leaf, leaf_ptr, perm, pair_ptr, pair_src, pair_off, Z, x, y);
    }
}

}  // namespace mlfma
