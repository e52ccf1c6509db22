// FP64 / FP32 MLFMA matvec kernels for gfx950 (64-wide wavefronts).
//
// Workspaces are laid out [box][cur][comp][K] so that consecutive threads
// (consecutive k-space samples) touch consecutive addresses.  RWG functions are
// stored in leaf-major order (internal index i); perm[i] is the reference index
// used by the caller's x and y vectors.  The far-field kernels are templates on
// the real type R of the patterns, tables and workspaces (double or float);
// they accumulate in R.  Solver vectors x, y and the near field stay FP64.
#include <algorithm>

#include "kernels.hpp"
#include <stdexcept>

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

// Window order of n items x nt tiles: items go in windows of wc, and inside
// a window the tile index is the slow one, so the workgroups that run at
// once on one XCD work on the same tile of wc neighbouring items (which
// share most of their inputs in that XCD's L2); wc = 1 is item-major order.
__device__ __forceinline__ void window_order(int lin, int n, int nt, int wc, int& item, int& tile) {
    const int span = wc * nt;
    const int win = lin / span, r = lin - win * span;
    const int w = min(wc, n - win * wc);
    tile = r / w;
    item = win * wc + (r - tile * w);
}

// ---------------------------------------------------------------- aggregation
// Code/ProcessMLFMM.h:33-56 (mexp), J and M currents of PMCHWT in one pass.
// One workgroup per leaf, threads over the K samples (blockDim = K rounded up
// to a wavefront, at most 512), so that a workgroup streams whole pattern rows
// [i][3][K] of its leaf from HBM.  Leaves are taken largest first.  The leaf's
// x values are gathered into LDS in chunks, which keeps the dependent
// perm -> x loads out of the pattern stream.  RB right-hand sides share each
// pattern load: solver vectors are RHS-fastest (unknown u, rhs r at
// u * ldr + r) and workspace current c = rr * NCUR + cur.
template <typename R, int NCUR, int RB>
__global__ __launch_bounds__(512) void k_aggregate(int K, int N, int ldr, int r0, const int* __restrict__ order,
                                                   const int* __restrict__ leaf_ptr, const int* __restrict__ perm,
                                                   const cplx_t<R>* __restrict__ Vs, const double2* __restrict__ x,
                                                   cplx_t<R>* __restrict__ Sm) {
    using C = cplx_t<R>;
    constexpr int NC = NCUR * RB;
    constexpr int CHUNK = NC <= 2 ? 128 : 64;
    __shared__ C xl[NC][CHUNK];
    const int leaf = order[blockIdx.x];
    const int i0 = leaf_ptr[leaf], i1 = leaf_ptr[leaf + 1];
    for (int q0 = 0; q0 < K; q0 += blockDim.x) {
        const int q = q0 + threadIdx.x;
        const bool act = q < K;
        C acc[NC][3];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) acc[c][d] = C{0, 0};
        for (int b = i0; b < i1; b += CHUNK) {
            const int n = min(CHUNK, i1 - b);
            __syncthreads();
            for (int e = threadIdx.x; e < n * NC; e += blockDim.x) {
                const int j = e / NC, c = e - j * NC;
                const int rr = c / NCUR, cur = c - rr * NCUR;
                narrow(x[((size_t)cur * N + perm[b + j]) * ldr + r0 + rr], xl[c][j]);
            }
            __syncthreads();
            if (act) {
                const C* v = Vs + (size_t)b * 3 * K + q;
                const size_t s3 = (size_t)3 * K;
                int j = 0;
                for (; j + 1 < n; j += 2) {
                    const C a0 = v[0], a1 = v[K], a2 = v[2 * K];
                    const C b0 = v[s3], b1 = v[s3 + K], b2 = v[s3 + 2 * K];
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const C xa = xl[c][j], xb = xl[c][j + 1];
                        cfma(acc[c][0], a0, xa);
                        cfma(acc[c][1], a1, xa);
                        cfma(acc[c][2], a2, xa);
                        cfma(acc[c][0], b0, xb);
                        cfma(acc[c][1], b1, xb);
                        cfma(acc[c][2], b2, xb);
                    }
                    v += 2 * s3;
                }
                if (j < n) {
                    const C a0 = v[0], a1 = v[K], a2 = v[2 * K];
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const C xa = xl[c][j];
                        cfma(acc[c][0], a0, xa);
                        cfma(acc[c][1], a1, xa);
                        cfma(acc[c][2], a2, xa);
                    }
                }
            }
        }
        if (act) {
            C* s = Sm + (size_t)leaf * NC * 3 * K + q;
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) s[(size_t)(c * 3 + d) * K] = acc[c][d];
        }
    }
}

template <typename R>
void launch_aggregate_b(hipStream_t st, int ncur, int rb, int nleaf, int K, int N, int ldr, int r0, const int* order,
                        const int* leaf_ptr, const int* perm, const cplx_t<R>* Vs, const double2* x,
                        cplx_t<R>* Sm) {
    if (nleaf == 0) return;
    const int threads = std::min(512, (K + 63) / 64 * 64);
#define AGG(NCUR, RB) k_aggregate<R, NCUR, RB><<<nleaf, threads, 0, st>>>(K, N, ldr, r0, order, leaf_ptr, perm, Vs, x, Sm)
    if (ncur == 1) {
        if (rb == 1) AGG(1, 1);
        else if (rb == 2) AGG(1, 2);
        else if (rb == 4) AGG(1, 4);
        else AGG(1, 8);
    } else {
        if (rb == 1) AGG(2, 1);
        else if (rb == 2) AGG(2, 2);
        else if (rb == 4) AGG(2, 4);
        else if constexpr (sizeof(R) == 4) AGG(2, 8);
        else throw std::invalid_argument("FP64 aggregation: at most 4 right-hand sides per pass");
    }
#undef AGG
}

template <typename R>
void launch_aggregate(hipStream_t st, int ncur, int nleaf, int K, int N, const int* order, const int* leaf_ptr,
                      const int* perm, const cplx_t<R>* Vs, const double2* x, cplx_t<R>* Sm) {
    launch_aggregate_b<R>(st, ncur, 1, nleaf, K, N, 1, 0, order, leaf_ptr, perm, Vs, x, Sm);
}

// ---------------------------------------------------------------- interpolation
// Code/ProcessMLFMM.h:61-107 (m2m): 3x3 Lagrange taps into the child grid, then
// the octant phase shift, summed over the children of each parent.  A thread
// owns one parent sample and CPT of the nc x 3 workspace columns (one
// wavefront per column group); the q tiles of one parent
// are consecutive after the XCD remap, so they share that XCD's L2 copy of the
// children's workspaces.
template <typename R, int NC, int CPT>
__global__ __launch_bounds__(64 * NC * 3 / CPT) void k_m2m(int nparent, int qtiles, int Kp, int Kc,
                                                 const int* __restrict__ son_ptr, const int* __restrict__ son_idx,
                                                 const int* __restrict__ oct_child, const int* __restrict__ tap,
                                                 const R* __restrict__ w, const cplx_t<R>* __restrict__ shup,
                                                 const cplx_t<R>* __restrict__ Smc, cplx_t<R>* __restrict__ Smp) {
    using C = cplx_t<R>;
    const int lin = xcd_remap(blockIdx.x, nparent * qtiles);
    const int p = lin / qtiles;
    const int q = (lin % qtiles) * 64 + (threadIdx.x & 63);
    const int c0 = (threadIdx.x >> 6) * CPT;   // first workspace column of this thread
    if (q >= Kp) return;
    int t[9];
    R ww[9];
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        t[s] = tap[(size_t)s * Kp + q];
        ww[s] = w[(size_t)s * Kp + q];
    }
    C acc[CPT];
#pragma unroll
    for (int d = 0; d < CPT; ++d) acc[d] = C{0, 0};
    for (int j = son_ptr[p]; j < son_ptr[p + 1]; ++j) {
        const int c = son_idx[j];
        const C e = shup[(size_t)oct_child[c] * Kp + q];
        const C* S = Smc + ((size_t)c * NC * 3 + c0) * Kc;
#pragma unroll
        for (int d = 0; d < CPT; ++d) {
            const C* sd = S + (size_t)d * Kc;
            C r = C{0, 0};
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const C f = sd[t[k]];
                r.x += f.x * ww[k];
                r.y += f.y * ww[k];
            }
            cfma(acc[d], e, r);
        }
    }
    C* o = Smp + ((size_t)p * NC * 3 + c0) * Kp + q;
#pragma unroll
    for (int d = 0; d < CPT; ++d) o[(size_t)d * Kp] = acc[d];
}

template <typename R>
void launch_m2m(hipStream_t st, int ncur, int nparent, int Kp, int Kc, const int* son_ptr, const int* son_idx,
                const int* octant_child, const int* tap, const R* w, const cplx_t<R>* shift_up, const cplx_t<R>* Sm_child,
                cplx_t<R>* Sm_parent) {
    if (nparent == 0) return;
    const int qt = (Kp + 63) / 64;
#define M2M(NC, CPT)                                                                                                  \
    k_m2m<R, NC, CPT><<<nparent * qt, 64 * NC * 3 / CPT, 0, st>>>(nparent, qt, Kp, Kc, son_ptr, son_idx,             \
                                                                  octant_child, tap, w, shift_up, Sm_child, Sm_parent)
    if (ncur == 1) M2M(1, 3);
    else if (ncur == 2) M2M(2, 3);
    else if (ncur == 4) M2M(4, 3);
    else if (ncur == 8) M2M(8, 3);
    else if constexpr (sizeof(R) == 4) M2M(16, 6);
    else throw std::invalid_argument("FP64 far field: at most 8 workspace currents per pass");
#undef M2M
}

// Separable form of the same interpolation: the 3 x 3 stencil of a parent
// sample (i, j) is child rows th_tap[i][a] x child columns ph_tap[j][b] with
// weights th_w[i][a] ph_w[j][b], so a workgroup owning RW consecutive rows of
// parent p first interpolates in theta, U(i, m) = sum_a th_w[i][a]
// S_c(th_tap[i][a], m), for all 2 Lc child columns m of each of its CPT
// columns into LDS (contiguous rows of the child workspace), then in phi,
// r(i, j) = sum_b ph_w[j][b] U(i, ph_tap[j][b]), and adds shift x r, child
// after child.  About half the loads and multiply-adds of the 9-tap form; the
// sums are regrouped (theta inside), which changes only rounding.  The lanes
// of a wavefront run over (row, column) pairs of its rows, so short rows
// still fill the wavefront and a workgroup's barriers serve RW rows; a thread
// owns outputs t = lane + 64 jj (jj < JJ) of the RW x 2 Lp block.  G = nc x 3 /
// CPT wavefronts, one per column group, each with its own LDS rows.
template <typename R, int NC, int CPT, int JJ>
__global__ __launch_bounds__(64 * NC * 3 / CPT) void k_m2m_sep(int Lp, int Lc, int RW, const int* __restrict__ son_ptr,
                                                        const int* __restrict__ son_idx,
                                                        const int* __restrict__ oct_child,
                                                        const int* __restrict__ th_tap, const R* __restrict__ th_w,
                                                        const int* __restrict__ ph_tap, const R* __restrict__ ph_w,
                                                        const cplx_t<R>* __restrict__ shup,
                                                        const cplx_t<R>* __restrict__ Smc,
                                                        cplx_t<R>* __restrict__ Smp) {
    using C = cplx_t<R>;
    extern __shared__ __align__(16) unsigned char m2m_lds[];
    const int pfn = 2 * Lp, psn = 2 * Lc, Kp = Lp * pfn, Kc = Lc * psn;
    const int nrg = (Lp + RW - 1) / RW;   // row groups per parent
    const int lin = xcd_remap(blockIdx.x, (int)gridDim.x);
    const int p = lin / nrg, i0 = (lin % nrg) * RW;   // the row groups of one parent are consecutive
    const int nr = min(RW, Lp - i0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c0 = wave * CPT;
    const int us = RW * psn;                                            // one column's U block
    C* U = reinterpret_cast<C*>(m2m_lds) + (size_t)wave * CPT * us;   // this wavefront's [CPT][RW][2 Lc]
    // phi-pass outputs of this thread: row ir[jj] (relative), column jc[jj]
    int ir[JJ], jc[JJ], pj[JJ][3];
    R wj[JJ][3];
    bool ok[JJ];
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj) {
        const int t = lane + 64 * jj;
        ok[jj] = t < nr * pfn;
        const int tc = min(t, nr * pfn - 1);
        ir[jj] = tc / pfn;
        jc[jj] = tc - ir[jj] * pfn;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            pj[jj][b] = ir[jj] * psn + ph_tap[3 * jc[jj] + b];
            wj[jj][b] = ph_w[3 * jc[jj] + b];
        }
    }
    C acc[JJ][CPT];
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj)
#pragma unroll
        for (int d = 0; d < CPT; ++d) acc[jj][d] = C{0, 0};
    // theta-pass items of this thread (t = lane + 64 k < RW 2 Lc <= 192 with the
    // launch's LDS bound): child offsets and weights, the same for every child
    constexpr int KT = 3;
    int ti[KT][3];
    R lt[KT][3];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int t = min(lane + 64 * k, nr * psn - 1);
        const int r = t / psn, m = t - r * psn;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ti[k][a] = th_tap[3 * (i0 + r) + a] * psn + m;
            lt[k][a] = th_w[3 * (i0 + r) + a];
        }
    }
    for (int s = son_ptr[p]; s < son_ptr[p + 1]; ++s) {
        const int c = son_idx[s];
        const C* S = Smc + ((size_t)c * NC * 3 + c0) * Kc;
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int t = lane + 64 * k;
            if (t < nr * psn) {
#pragma unroll
                for (int d = 0; d < CPT; ++d) {
                    const C* sd = S + (size_t)d * Kc;
                    const C f0 = sd[ti[k][0]], f1 = sd[ti[k][1]], f2 = sd[ti[k][2]];
                    C u;
                    u.x = f0.x * lt[k][0] + f1.x * lt[k][1] + f2.x * lt[k][2];
                    u.y = f0.y * lt[k][0] + f1.y * lt[k][1] + f2.y * lt[k][2];
                    U[d * us + t] = u;
                }
            }
        }
        __syncthreads();
        const C* sh = shup + (size_t)oct_child[c] * Kp + (size_t)i0 * pfn;
#pragma unroll
        for (int jj = 0; jj < JJ; ++jj) {
            if (ok[jj]) {
                const C e = sh[ir[jj] * pfn + jc[jj]];
#pragma unroll
                for (int d = 0; d < CPT; ++d) {
                    const C u0 = U[d * us + pj[jj][0]], u1 = U[d * us + pj[jj][1]], u2 = U[d * us + pj[jj][2]];
                    C r;
                    r.x = u0.x * wj[jj][0] + u1.x * wj[jj][1] + u2.x * wj[jj][2];
                    r.y = u0.y * wj[jj][0] + u1.y * wj[jj][1] + u2.y * wj[jj][2];
                    cfma(acc[jj][d], e, r);
                }
            }
        }
        __syncthreads();
    }
    C* o = Smp + ((size_t)p * NC * 3 + c0) * Kp + (size_t)i0 * pfn;
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj)
        if (ok[jj])
#pragma unroll
            for (int d = 0; d < CPT; ++d) o[(size_t)d * Kp + ir[jj] * pfn + jc[jj]] = acc[jj][d];
}

// The separable interpolation when a parent row has at most 256 samples and
// one row of every column fits in LDS (64 KB per workgroup); false otherwise
// (the caller takes the 9-tap kernel).  A workgroup takes as many rows as fit
// both (RW 2 Lp <= 256 outputs per lane group, RW rows of LDS).
template <typename R>
bool launch_m2m_sep(hipStream_t st, int ncur, int nparent, int Lp, int Lc, const int* son_ptr, const int* son_idx,
                    const int* octant_child, const int* th_tap, const R* th_w, const int* ph_tap, const R* ph_w,
                    const cplx_t<R>* shift_up, const cplx_t<R>* Sm_child, cplx_t<R>* Sm_parent) {
    if (nparent == 0) return true;
    const size_t row_lds = (size_t)ncur * 3 * 2 * Lc * sizeof(cplx_t<R>);   // one row of all column groups
    if (2 * Lp > 256 || row_lds > 64 * 1024) return false;
    if (2 * Lc > 192) return false;   // theta-pass items per lane (KT = 3)
    const int RW = std::max(1, std::min({Lp, (int)(64 * 1024 / row_lds), 256 / (2 * Lp), 192 / (2 * Lc)}));
    const int JJ = (RW * 2 * Lp + 63) / 64;
    const size_t lds = row_lds * RW;
    const unsigned grid = (unsigned)nparent * ((Lp + RW - 1) / RW);
#define M2MS(NC, CPT, J)                                                                                              \
    k_m2m_sep<R, NC, CPT, J><<<grid, 64 * NC * 3 / CPT, lds, st>>>(Lp, Lc, RW, son_ptr, son_idx, octant_child,      \
                                                                    th_tap, th_w, ph_tap, ph_w, shift_up, Sm_child, \
                                                                    Sm_parent)
#define M2MS_J(NC, CPT)                                                                                               \
    if (JJ == 1) M2MS(NC, CPT, 1);                                                                                    \
    else if (JJ == 2) M2MS(NC, CPT, 2);                                                                               \
    else if (JJ == 3) M2MS(NC, CPT, 3);                                                                               \
    else M2MS(NC, CPT, 4);
    if (ncur == 1) {
        M2MS_J(1, 3)
    } else if (ncur == 2) {
        M2MS_J(2, 3)
    } else if (ncur == 4) {
        M2MS_J(4, 3)
    } else if (ncur == 8) {
        M2MS_J(8, 3)
    } else if constexpr (sizeof(R) == 4) {
        M2MS_J(16, 6)
    } else {
        throw std::invalid_argument("FP64 far field: at most 8 workspace currents per pass");
    }
#undef M2MS_J
#undef M2MS
    return true;
}

// Anterpolation from the parent level as a separable pass (the transpose of
// k_m2m_sep): for box b with parent P and octant o, child sample (i, j) gets
//   sum over (parent row ip, wt) with th_tap[ip][a] = i, (parent column jp, wp)
//   with ph_tap[jp][b] = j of  wt wp shift_o(ip, jp) P(ip, jp),
// the same products as the fused gather (Code/ProcessMLFMM.h:152-202) with the
// sums regrouped (theta first, into LDS), which changes only rounding.  A
// workgroup takes RW consecutive child rows of one box; the theta pass goes
// row by row with lanes over parent columns, the phi pass has lanes over
// (row, child column) pairs and adds into Bm (the translation wrote it before).
template <typename R, int NC, int CPT, int JJ>
__global__ __launch_bounds__(64 * NC * 3 / CPT) void k_l2l_sep(int Lp, int Lc, int RW, const int* __restrict__ parent,
                                                        const int* __restrict__ oct, const int* __restrict__ at_ptr,
                                                        const int* __restrict__ at_idx, const R* __restrict__ at_w,
                                                        const int* __restrict__ ap_ptr, const int* __restrict__ ap_idx,
                                                        const R* __restrict__ ap_w, const cplx_t<R>* __restrict__ shdn,
                                                        const cplx_t<R>* __restrict__ Bmp, cplx_t<R>* __restrict__ Bm) {
    using C = cplx_t<R>;
    extern __shared__ __align__(16) unsigned char l2l_lds[];
    const int pfn = 2 * Lp, psn = 2 * Lc, Kp = Lp * pfn, Kc = Lc * psn;
    const int nrg = (Lc + RW - 1) / RW;   // row groups per box
    const int lin = xcd_remap(blockIdx.x, (int)gridDim.x);
    const int b = lin / nrg, r0 = (lin % nrg) * RW;   // the row groups of one box are consecutive
    const int nr = min(RW, Lc - r0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c0 = wave * CPT;
    const int us = RW * pfn;                                             // one column's block
    C* V = reinterpret_cast<C*>(l2l_lds) + (size_t)wave * CPT * us;    // this wavefront's [CPT][RW][2 Lp]
    const C* P = Bmp + ((size_t)parent[b] * NC * 3 + c0) * Kp;
    const C* sh = shdn + (size_t)oct[b] * Kp;
    // rows one after another, so a row's taps are the same for every lane
    // (scalar loads) and each parent sample load depends on nothing per lane
    for (int r = 0; r < nr; ++r) {
        const int k0 = at_ptr[r0 + r], k1 = at_ptr[r0 + r + 1];
        for (int jp = lane; jp < pfn; jp += 64) {
            C acc[CPT];
#pragma unroll
            for (int d = 0; d < CPT; ++d) acc[d] = C{0, 0};
            for (int k = k0; k < k1; ++k) {
                const int kp = at_idx[k] * pfn + jp;
                const R w = at_w[k];
                C e = sh[kp];
                e.x *= w;
                e.y *= w;
#pragma unroll
                for (int d = 0; d < CPT; ++d) cfma(acc[d], e, P[(size_t)d * Kp + kp]);
            }
#pragma unroll
            for (int d = 0; d < CPT; ++d) V[d * us + r * pfn + jp] = acc[d];
        }
    }
    __syncthreads();
    C* O = Bm + ((size_t)b * NC * 3 + c0) * Kc + (size_t)r0 * psn;
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj) {
        const int t = lane + 64 * jj;
        if (t < nr * psn) {
            const int r = t / psn, jc = t - r * psn;
            C s[CPT];
#pragma unroll
            for (int d = 0; d < CPT; ++d) s[d] = C{0, 0};
            for (int k = ap_ptr[jc]; k < ap_ptr[jc + 1]; ++k) {
                const int vi = r * pfn + ap_idx[k];
                const R w = ap_w[k];
#pragma unroll
                for (int d = 0; d < CPT; ++d) {
                    const C v = V[d * us + vi];
                    s[d].x += w * v.x;
                    s[d].y += w * v.y;
                }
            }
#pragma unroll
            for (int d = 0; d < CPT; ++d) {
                C o = O[(size_t)d * Kc + t];
                o.x += s[d].x;
                o.y += s[d].y;
                O[(size_t)d * Kc + t] = o;
            }
        }
    }
}

// child rows per workgroup: as many as fit 64 KB of LDS (the theta pass's rows
// of every column) and 256 outputs per lane group (JJ <= 4); 0 if none fits
int l2l_sep_rows(int ncur, int Lp, int Lc, size_t csize) {
    const size_t row_lds = (size_t)ncur * 3 * 2 * Lp * csize;
    if (row_lds > 64 * 1024 || 2 * Lc > 256) return 0;
    return std::max(1, std::min({Lc, (int)(64 * 1024 / row_lds), 256 / (2 * Lc)}));
}

template <typename R>
void launch_l2l_sep(hipStream_t st, int ncur, int nbox, int Lp, int Lc, int rw, const int* parent, const int* oct,
                    const int* at_ptr, const int* at_idx, const R* at_w, const int* ap_ptr, const int* ap_idx,
                    const R* ap_w, const cplx_t<R>* shift_dn, const cplx_t<R>* Bm_parent, cplx_t<R>* Bm) {
    if (nbox == 0) return;
    const int JJ = (rw * 2 * Lc + 63) / 64;
    if (rw < 1 || rw > Lc || JJ > 4 || (size_t)ncur * 3 * 2 * Lp * sizeof(cplx_t<R>) * rw > 64 * 1024)
        throw std::invalid_argument("separable anterpolation: rows per workgroup out of range");
    const size_t lds = (size_t)ncur * 3 * 2 * Lp * sizeof(cplx_t<R>) * rw;
    const unsigned grid = (unsigned)nbox * ((Lc + rw - 1) / rw);
#define L2LS(NC, CPT, J)                                                                                              \
    k_l2l_sep<R, NC, CPT, J><<<grid, 64 * NC * 3 / CPT, lds, st>>>(Lp, Lc, rw, parent, oct, at_ptr, at_idx, at_w,    \
                                                                    ap_ptr, ap_idx, ap_w, shift_dn, Bm_parent, Bm)
#define L2LS_J(NC, CPT)                                                                                               \
    if (JJ == 1) L2LS(NC, CPT, 1);                                                                                    \
    else if (JJ == 2) L2LS(NC, CPT, 2);                                                                               \
    else if (JJ == 3) L2LS(NC, CPT, 3);                                                                               \
    else L2LS(NC, CPT, 4);
    if (ncur == 1) {
        L2LS_J(1, 3)
    } else if (ncur == 2) {
        L2LS_J(2, 3)
    } else if (ncur == 4) {
        L2LS_J(4, 3)
    } else if (ncur == 8) {
        L2LS_J(8, 3)
    } else if constexpr (sizeof(R) == 4) {
        L2LS_J(16, 6)
    } else {
        throw std::invalid_argument("FP64 far field: at most 8 workspace currents per pass");
    }
#undef L2LS_J
#undef L2LS
}

// ---------------------------------------------------------------- translation + anterpolation
// Code/ProcessMLFMM.h:112-147 (m2l, quadrature weight folded into TFw) and
// :152-202 (l2l) in gather form: each child sample sums the parent samples whose
// 9-tap stencil contains it (transposed table, ascending parent sample), so no
// atomics and a fixed summation order.
template <typename R, int NC>
__global__ __launch_bounds__(256) void k_m2l_l2l(int nbox, int K, int ktiles, const int* __restrict__ dr_ptr,
                                                 const int* __restrict__ dr_idx, const int* __restrict__ dr_vec,
                                                 const cplx_t<R>* __restrict__ TFw, const cplx_t<R>* __restrict__ Sm,
                                                 int Kp, const int* __restrict__ parent, const int* __restrict__ oct,
                                                 const int* __restrict__ t_ptr, const int* __restrict__ t_kp,
                                                 const R* __restrict__ t_w, const cplx_t<R>* __restrict__ shdn,
                                                 const cplx_t<R>* __restrict__ Bmp, cplx_t<R>* __restrict__ Bm) {
    using C = cplx_t<R>;
    const int lin = xcd_remap(blockIdx.x, nbox * ktiles);
    const int b = lin / ktiles;
    const int q = (lin % ktiles) * blockDim.x + threadIdx.x;
    if (q >= K) return;
    C acc[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[c][d] = C{0, 0};
    for (int j = dr_ptr[b]; j < dr_ptr[b + 1]; ++j) {
        const C T = TFw[(size_t)dr_vec[j] * K + q];
        const C* S = Sm + (size_t)dr_idx[j] * NC * 3 * K + q;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int d = 0; d < 3; ++d) cfma(acc[cc][d], T, S[(size_t)(cc * 3 + d) * K]);
    }
    if (Bmp != nullptr) {
        const int pb = parent[b];
        const C* sh = shdn + (size_t)oct[b] * Kp;
        const C* P = Bmp + (size_t)pb * NC * 3 * Kp;
        for (int j = t_ptr[q]; j < t_ptr[q + 1]; ++j) {
            const int kp = t_kp[j];
            C e = sh[kp];
            e.x *= t_w[j];
            e.y *= t_w[j];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int d = 0; d < 3; ++d) cfma(acc[cc][d], e, P[(size_t)(cc * 3 + d) * Kp + kp]);
        }
    }
    C* o = Bm + (size_t)b * NC * 3 * K + q;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int d = 0; d < 3; ++d) o[(size_t)(cc * 3 + d) * K] = acc[cc][d];
}

template <typename R>
void launch_m2l_l2l(hipStream_t st, int ncur, int nbox, int K, const int* dr_ptr, const int* dr_idx,
                    const int* dr_vec, const cplx_t<R>* TFw, const cplx_t<R>* Sm, int Kp, const int* parent,
                    const int* octant, const int* t_ptr, const int* t_kp, const R* t_w, const cplx_t<R>* shift_down,
                    const cplx_t<R>* Bm_parent, cplx_t<R>* Bm) {
    const int kt = (K + 255) / 256;
    dim3 grid(kt * nbox);
#define M2L(NC) k_m2l_l2l<R, NC><<<grid, 256, 0, st>>>(nbox, K, kt, dr_ptr, dr_idx, dr_vec, TFw, Sm, Kp, parent, octant, t_ptr, t_kp, t_w, shift_down, Bm_parent, Bm)
    if (ncur == 1) M2L(1);
    else if (ncur == 2) M2L(2);
    else if (ncur == 4) M2L(4);
    else if (ncur == 8) M2L(8);
    else if constexpr (sizeof(R) == 4) M2L(16);
    else throw std::invalid_argument("FP64 far field: at most 8 workspace currents per pass");
#undef M2L
}

// Sibling-cluster form of the same product: a thread owns one sample q, CPT of
// the nc x 3 workspace columns and all NM (1..4) children of one parent.  It
// walks the union of their distant-relative lists in ascending box order,
// reading each source box's CPT values once and applying them to every member
// with that member's translation vector; a member without the source in its
// list reads the all-zero row the table keeps after its last vector, so the
// member loop has no branches (a member's sum thus runs over its own list in
// ascending box order, plus exact zeros).  The interpolation from the parent
// reads each parent value once for all members.  G = nc x 3 / CPT wavefronts
// per workgroup, one per column group; no LDS and no barriers.
template <typename R, int NC, int G, int NM>
__device__ __forceinline__ void m2l_cl_body(int cl, int K, int q, int c0, const int* __restrict__ cl_child,
                                            const int* __restrict__ cl_ptr, const int* __restrict__ cl_src,
                                            const int* __restrict__ cl_vec, const cplx_t<R>* __restrict__ TFw,
                                            const cplx_t<R>* __restrict__ Sm, int Kp, const int* __restrict__ parent,
                                            const int* __restrict__ oct, const int* __restrict__ t_ptr,
                                            const int* __restrict__ t_kp, const R* __restrict__ t_w,
                                            const cplx_t<R>* __restrict__ shdn, const cplx_t<R>* __restrict__ Bmp,
                                            cplx_t<R>* __restrict__ Bm) {
    using C = cplx_t<R>;
    constexpr int CPT = NC * 3 / G;
    int child[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) child[k] = cl_child[cl * 4 + k];
    C acc[NM][CPT];
#pragma unroll
    for (int k = 0; k < NM; ++k)
#pragma unroll
        for (int j = 0; j < CPT; ++j) acc[k][j] = C{0, 0};
    const int u0 = cl_ptr[cl], u1 = cl_ptr[cl + 1];
    // software pipeline over two register sets: the next source's values and
    // its members' translation samples are loaded into one set while the
    // other is applied
    C sa[CPT], sb[CPT], ta[NM], tb[NM];
    auto fetch = [&](C (&s)[CPT], C (&t)[NM], int u) {
        const int4 v = *reinterpret_cast<const int4*>(cl_vec + (size_t)u * 4);
        const int vk[4] = {v.x, v.y, v.z, v.w};
        const C* S = Sm + ((size_t)cl_src[u] * NC * 3 + c0) * K + q;
#pragma unroll
        for (int j = 0; j < CPT; ++j) s[j] = S[(size_t)j * K];
#pragma unroll
        for (int k = 0; k < NM; ++k) t[k] = TFw[(size_t)vk[k] * K + q];
    };
    auto apply = [&](const C (&s)[CPT], const C (&t)[NM]) {
#pragma unroll
        for (int k = 0; k < NM; ++k)
#pragma unroll
            for (int j = 0; j < CPT; ++j) cmac(acc[k][j], t[k], s[j]);
    };
    // Sources go in pairs, so that every apply has the other set's loads in
    // flight; the prefetch after a pair is unconditional (the last source is
    // read again past the end), so every path reaches an apply with the same
    // loads outstanding, and the scheduling barriers keep the compiler from
    // sinking the loads to their uses.
    int u = u0;
    if (u < u1) fetch(sa, ta, u);
    for (; u + 1 < u1; u += 2) {
        fetch(sb, tb, u + 1);
        __builtin_amdgcn_sched_barrier(0);
        apply(sa, ta);
        fetch(sa, ta, min(u + 2, u1 - 1));
        __builtin_amdgcn_sched_barrier(0);
        apply(sb, tb);
    }
    if (u < u1) apply(sa, ta);
    if (Bmp != nullptr) {
        const C* P = Bmp + ((size_t)parent[child[0]] * NC * 3 + c0) * Kp;   // siblings share the parent
        const C* sh[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k) sh[k] = shdn + (size_t)oct[child[k]] * Kp;
        for (int j = t_ptr[q]; j < t_ptr[q + 1]; ++j) {
            const int kp = t_kp[j];
            const R w = t_w[j];
            C pv[CPT];
#pragma unroll
            for (int jj = 0; jj < CPT; ++jj) pv[jj] = P[(size_t)jj * Kp + kp];
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                C e = sh[k][kp];
                e.x *= w;
                e.y *= w;
#pragma unroll
                for (int jj = 0; jj < CPT; ++jj) cfma(acc[k][jj], e, pv[jj]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        C* o = Bm + ((size_t)child[k] * NC * 3 + c0) * K + q;
#pragma unroll
        for (int j = 0; j < CPT; ++j) o[(size_t)j * K] = acc[k][j];
    }
}

// All-sibling clusters with LDS-staged translation samples (FP32): a cluster
// is every child of one parent (NM <= 8 members; with up to 4 a parent of 5-8
// children needs two clusters, which read the union's sources twice), a
// thread owns one sample q and CPT = 6 columns.  The translation samples of
// the members for kS sources at a time are staged in LDS by the whole
// workgroup, so each is read from memory once per workgroup instead of once
// per column-group wavefront, and a source costs a wavefront CPT vector loads
// (its workspace values) and NM LDS reads.  The staging is double-buffered
// (the next chunk's samples are loaded into registers during the current
// chunk's products, one barrier per chunk); the last chunk is padded with
// zero samples, and members without a source read the table's zero row, so
// the product loop has no branches.  Per sample and member the sums run over
// the same sources in the same order as the other cluster forms (plus exact
// zeros), so the results are identical.
constexpr int kClsS = 4;   // sources per staged chunk (8 spills registers)

template <int NC, int G, int NM>
__device__ __forceinline__ void m2l_cls_body(int cl, int K, int kt, int c0, float2 (*tfs)[kClsS][8][64],
                                             const int* __restrict__ cl_child, const int* __restrict__ cl_ptr,
                                             const int* __restrict__ cl_src, const int* __restrict__ cl_vec,
                                             const float2* __restrict__ TFw, const float2* __restrict__ Sm, int Kp,
                                             const int* __restrict__ parent, const int* __restrict__ oct,
                                             const int* __restrict__ t_ptr, const int* __restrict__ t_kp,
                                             const float* __restrict__ t_w, const float2* __restrict__ shdn,
                                             const float2* __restrict__ Bmp, float2* __restrict__ Bm) {
    using C = float2;
    constexpr int CPT = NC * 3 / G;
    constexpr int NTH = 64 * G;
    constexpr int NE = kClsS * NM * 64, EPT = (NE + NTH - 1) / NTH;   // staged samples per chunk, per thread
    const int lane = threadIdx.x & 63;
    const int q = kt * 64 + lane, qc = min(q, K - 1);   // lanes past K compute on clamped data, store nothing
    int child[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) child[k] = cl_child[cl * 8 + k];
    C acc[NM][CPT];
#pragma unroll
    for (int k = 0; k < NM; ++k)
#pragma unroll
        for (int j = 0; j < CPT; ++j) acc[k][j] = C{0, 0};
    const int u0 = cl_ptr[cl], u1 = cl_ptr[cl + 1];
    const int nchunk = (u1 - u0 + kClsS - 1) / kClsS;
    // staging: entry e = (s, k, l) of chunk c -> tfs[b][s][k][l]
    C tv[EPT];
    auto stage_load = [&](int c) {
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = min((int)threadIdx.x + NTH * i, NE - 1);
            const int s = e / (NM * 64), k = (e / 64) % NM, l = e % 64;
            const int u = u0 + c * kClsS + s;
            const int vid = cl_vec[(size_t)min(u, u1 - 1) * 8 + k];
            const C v = TFw[(size_t)vid * K + min(kt * 64 + l, K - 1)];
            tv[i] = u < u1 ? v : C{0, 0};   // the last chunk's padding applies exact zeros
        }
    };
    auto stage_store = [&](int b) {
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = (int)threadIdx.x + NTH * i;
            if (NE % NTH == 0 || e < NE) {
                const int s = e / (NM * 64), k = (e / 64) % NM, l = e % 64;
                tfs[b][s][k][l] = tv[i];
            }
        }
    };
    C sa[CPT], sb[CPT];
    auto fetch = [&](C (&sv)[CPT], int u) {
        const C* S = Sm + ((size_t)cl_src[min(u, u1 - 1)] * NC * 3 + c0) * K + qc;
#pragma unroll
        for (int j = 0; j < CPT; ++j) sv[j] = S[(size_t)j * K];
    };
    auto apply = [&](const C (&sv)[CPT], int b, int s) {
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            const C t = tfs[b][s][k][lane];
#pragma unroll
            for (int j = 0; j < CPT; ++j) cmac(acc[k][j], t, sv[j]);
        }
    };
    if (nchunk > 0) {
        stage_load(0);
        stage_store(0);
        fetch(sa, u0);
        __syncthreads();
        for (int c = 0; c < nchunk; ++c) {
            const int b = c & 1, ub = u0 + c * kClsS;
            stage_load(c + 1 < nchunk ? c + 1 : c);   // the last chunk reloads itself (never stored)
            // kClsS is even: the chunk's first source is in set A
#pragma unroll
            for (int s = 0; s < kClsS; s += 2) {
                fetch(sb, ub + s + 1);
                __builtin_amdgcn_sched_barrier(0);
                apply(sa, b, s);
                fetch(sa, ub + s + 2);
                __builtin_amdgcn_sched_barrier(0);
                apply(sb, b, s + 1);
            }
            if (c + 1 < nchunk) stage_store(b ^ 1);
            __syncthreads();
        }
    }
    if (Bmp != nullptr) {
        const C* P = Bmp + ((size_t)parent[child[0]] * NC * 3 + c0) * Kp;   // siblings share the parent
        const C* sh[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k) sh[k] = shdn + (size_t)oct[child[k]] * Kp;
        for (int j = t_ptr[qc]; j < t_ptr[qc + 1]; ++j) {
            const int kp = t_kp[j];
            const float w = t_w[j];
            C pv[CPT];
#pragma unroll
            for (int jj = 0; jj < CPT; ++jj) pv[jj] = P[(size_t)jj * Kp + kp];
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                C e = sh[k][kp];
                e.x *= w;
                e.y *= w;
#pragma unroll
                for (int jj = 0; jj < CPT; ++jj) cfma(acc[k][jj], e, pv[jj]);
            }
        }
    }
    if (q >= K) return;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        C* o = Bm + ((size_t)child[k] * NC * 3 + c0) * K + q;
#pragma unroll
        for (int j = 0; j < CPT; ++j) o[(size_t)j * K] = acc[k][j];
    }
}

template <int NC, int G>
__global__ __launch_bounds__(64 * G) void k_m2l_cls(int ncl, int K, int ktiles, int wc,
                                                    const int* __restrict__ cl_child, const int* __restrict__ cl_ptr,
                                                    const int* __restrict__ cl_src, const int* __restrict__ cl_vec,
                                                    const float2* __restrict__ TFw, const float2* __restrict__ Sm,
                                                    int Kp, const int* __restrict__ parent,
                                                    const int* __restrict__ oct, const int* __restrict__ t_ptr,
                                                    const int* __restrict__ t_kp, const float* __restrict__ t_w,
                                                    const float2* __restrict__ shdn, const float2* __restrict__ Bmp,
                                                    float2* __restrict__ Bm) {
    __shared__ float2 tfs[2][kClsS][8][64];
    constexpr int CPT = NC * 3 / G;
    int cl, kt;
    window_order(xcd_remap(blockIdx.x, ncl * ktiles), ncl, ktiles, wc, cl, kt);
    const int c0 = (threadIdx.x >> 6) * CPT;
    // members are packed first; the count is uniform over the workgroup
    int nm = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) nm += cl_child[cl * 8 + k] >= 0 ? 1 : 0;
#define M2LS(NM)                                                                                                      \
    m2l_cls_body<NC, G, NM>(cl, K, kt, c0, tfs, cl_child, cl_ptr, cl_src, cl_vec, TFw, Sm, Kp, parent, oct, t_ptr,   \
                            t_kp, t_w, shdn, Bmp, Bm)
    switch (nm) {
        case 8: M2LS(8); break;
        case 7: M2LS(7); break;
        case 6: M2LS(6); break;
        case 5: M2LS(5); break;
        case 4: M2LS(4); break;
        case 3: M2LS(3); break;
        case 2: M2LS(2); break;
        default: M2LS(1); break;
    }
#undef M2LS
}

template <typename R>
void launch_m2l_cls(hipStream_t st, int ncur, int ncl, int K, const int* cl_child, const int* cl_ptr,
                    const int* cl_src, const int* cl_vec, const cplx_t<R>* TFw, const cplx_t<R>* Sm, int Kp,
                    const int* parent, const int* octant, const int* t_ptr, const int* t_kp, const R* t_w,
                    const cplx_t<R>* shift_down, const cplx_t<R>* Bm_parent, cplx_t<R>* Bm) {
    if constexpr (sizeof(R) == 8) {
        throw std::invalid_argument("LDS-staged cluster translation: FP32 far field only");
    } else {
        if (ncl == 0) return;
        const int kt = (K + 63) / 64;
        constexpr int wc = 128;
#define M2LS_L(NC)                                                                                                    \
    k_m2l_cls<NC, NC / 2><<<ncl * kt, 32 * NC, 0, st>>>(ncl, K, kt, wc, cl_child, cl_ptr, cl_src, cl_vec, TFw, Sm,   \
                                                         Kp, parent, octant, t_ptr, t_kp, t_w, shift_down,            \
                                                         Bm_parent, Bm)
        if (ncur == 4) M2LS_L(4);
        else if (ncur == 8) M2LS_L(8);
        else if (ncur == 16) M2LS_L(16);
        else throw std::invalid_argument("LDS-staged cluster translation: 4, 8 or 16 workspace currents per pass");
#undef M2LS_L
    }
}

template <typename R, int NC, int G>
__global__ __launch_bounds__(64 * G) void k_m2l_cl(int ncl, int K, int ktiles, int wc, const int* __restrict__ cl_child,
                                                   const int* __restrict__ cl_ptr, const int* __restrict__ cl_src,
                                                   const int* __restrict__ cl_vec, const cplx_t<R>* __restrict__ TFw,
                                                   const cplx_t<R>* __restrict__ Sm, int Kp,
                                                   const int* __restrict__ parent, const int* __restrict__ oct,
                                                   const int* __restrict__ t_ptr, const int* __restrict__ t_kp,
                                                   const R* __restrict__ t_w, const cplx_t<R>* __restrict__ shdn,
                                                   const cplx_t<R>* __restrict__ Bmp, cplx_t<R>* __restrict__ Bm) {
    constexpr int CPT = NC * 3 / G;
    int cl, kt;
    window_order(xcd_remap(blockIdx.x, ncl * ktiles), ncl, ktiles, wc, cl, kt);
    const int c0 = (threadIdx.x >> 6) * CPT;
    const int q = kt * 64 + (threadIdx.x & 63);
    if (q >= K) return;
    // members are packed first; the count is uniform over the workgroup
    const int4 ch = *reinterpret_cast<const int4*>(cl_child + (size_t)cl * 4);
    const int nm = (ch.y >= 0) + (ch.z >= 0) + (ch.w >= 0) + 1;
#define M2LB(NM)                                                                                                      \
    m2l_cl_body<R, NC, G, NM>(cl, K, q, c0, cl_child, cl_ptr, cl_src, cl_vec, TFw, Sm, Kp, parent, oct, t_ptr, t_kp, \
                              t_w, shdn, Bmp, Bm)
    if (nm == 4) M2LB(4);
    else if (nm == 3) M2LB(3);
    else if (nm == 2) M2LB(2);
    else M2LB(1);
#undef M2LB
}

template <typename R>
void launch_m2l_cl(hipStream_t st, int variant, int ncur, int ncl, int K, const int* cl_child, const int* cl_ptr,
                   const int* cl_src, const int* cl_vec, const cplx_t<R>* TFw, const cplx_t<R>* Sm, int Kp,
                   const int* parent, const int* octant, const int* t_ptr, const int* t_kp, const R* t_w,
                   const cplx_t<R>* shift_down, const cplx_t<R>* Bm_parent, cplx_t<R>* Bm) {
    if (ncl == 0) return;
    const int kt = (K + 63) / 64;
    // windows of 128 clusters: the fetch from memory drops from about 7.6 to
    // 4.5 times the level's workspaces (the kernel's time does not change)
    constexpr int wc = 128;
    // columns per thread: 3 in FP64, 6 in FP32 (3 when there are only 3)
    constexpr bool f64 = sizeof(R) == 8;
#define M2LC_G(NC, G)                                                                                                 \
    k_m2l_cl<R, NC, G><<<ncl * kt, 64 * G, 0, st>>>(ncl, K, kt, wc, cl_child, cl_ptr, cl_src, cl_vec, TFw, Sm, Kp, \
                                                    parent, octant, t_ptr, t_kp, t_w, shift_down, Bm_parent, Bm)
#define M2LC(NC)                                                                                                      \
    {                                                                                                                 \
        if constexpr (!f64 && NC >= 4) {                                                                              \
            if (variant == 1) {                                                                                       \
                M2LC_G(NC, NC / 4);                                                                                   \
            } else {                                                                                                  \
                M2LC_G(NC, NC / 2);                                                                                   \
            }                                                                                                         \
        } else {                                                                                                      \
            constexpr int G = f64 ? NC : (NC == 1 ? 1 : NC / 2);                                                      \
            M2LC_G(NC, G);                                                                                            \
        }                                                                                                             \
    }
    if (ncur == 1) M2LC(1)
    else if (ncur == 2) M2LC(2)
    else if (ncur == 4) M2LC(4)
    else if (ncur == 8) M2LC(8)
    else if constexpr (sizeof(R) == 4) M2LC(16)
    else throw std::invalid_argument("FP64 far field: at most 8 workspace currents per pass");
#undef M2LC
#undef M2LC_G
}

// ---------------------------------------------------------------- disaggregation
// Code/MLFMM.cpp:211-227: w_m += (k eta/4pi)(k/4pi) sum_k Vfmj_m . Bm_leaf.
// Each wavefront owns ROWS consecutive RWG rows (internal order) and streams
// their patterns together, lanes over k, then one wavefront reduction per row
// and right-hand side.  Rows run over the whole RWG range, so small and large
// leaves load the GPU alike.  With RB right-hand sides each pattern load
// serves RB workspaces (current rr of the leaf).
template <typename R, int ROWS, int RB>
__global__ __launch_bounds__(256) void k_disagg_pec(int N, int K, int ldr, int r0, const int* __restrict__ leaf_of,
                                                    const int* __restrict__ perm, const cplx_t<R>* __restrict__ Vf,
                                                    const cplx_t<R>* __restrict__ Bm, double2 c,
                                                    double2* __restrict__ y) {
    using C = cplx_t<R>;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i00 = (blockIdx.x * 4 + wave) * ROWS;
    if (i00 >= N) return;
    const C* V[ROWS];
    const C* B[ROWS];
    C s[ROWS][RB];
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
        const int i = min(i00 + u, N - 1);   // rows past N repeat row N-1 and are not written
        V[u] = Vf + (size_t)i * 3 * K;
        B[u] = Bm + (size_t)leaf_of[i] * RB * 3 * K;
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) s[u][rr] = C{0, 0};
    }
    for (int q = lane; q < K; q += 64) {
#pragma unroll
        for (int u = 0; u < ROWS; ++u) {
            const C v0 = V[u][q], v1 = V[u][K + q], v2 = V[u][2 * K + q];
#pragma unroll
            for (int rr = 0; rr < RB; ++rr) {
                const C* b = B[u] + (size_t)rr * 3 * K;
                cfma(s[u][rr], v0, b[q]);
                cfma(s[u][rr], v1, b[K + q]);
                cfma(s[u][rr], v2, b[2 * K + q]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) s[u][rr] = wave_sum(s[u][rr]);
    // lane u * RB + rr writes row u, right-hand side rr
    if (lane < ROWS * RB) {
        const int u = lane / RB, rr = lane - u * RB;
        C v = s[0][0];
#pragma unroll
        for (int uu = 0; uu < ROWS; ++uu)
#pragma unroll
            for (int q = 0; q < RB; ++q)
                if (uu == u && q == rr) v = s[uu][q];
        if (i00 + u < N) {
            double2* yp = y + (size_t)perm[i00 + u] * ldr + r0 + rr;
            double2 t = *yp;
            cfma(t, c, widen(v));
            *yp = t;
        }
    }
}

template <typename R>
void launch_disagg_pec_b(hipStream_t st, int rb, int N, int K, int ldr, int r0, const int* leaf_of, const int* perm,
                         const cplx_t<R>* Vf, const cplx_t<R>* Bm, double2 c, double2* y) {
#define DIS(ROWS, RB)                                                                                      \
    k_disagg_pec<R, ROWS, RB><<<(N + 4 * ROWS - 1) / (4 * ROWS), 256, 0, st>>>(N, K, ldr, r0, leaf_of, perm, Vf, \
                                                                               Bm, c, y)
    if (rb == 1) DIS(4, 1);
    else if (rb == 2) DIS(2, 2);
    else if (rb == 4) DIS(1, 4);
    else DIS(1, 8);
#undef DIS
}

template <typename R>
void launch_disagg_pec(hipStream_t st, int N, int K, const int* leaf_of, const int* perm, const cplx_t<R>* Vf,
                       const cplx_t<R>* Bm, double2 c, double2* y) {
    launch_disagg_pec_b<R>(st, 1, N, K, 1, 0, leaf_of, perm, Vf, Bm, c, y);
}

// Code/MLFMM.cpp:285-351.  The K integral sum (-khat x V).B with khat = -kp is
// evaluated as sum V.(kp x B).  Sums of RB right-hand sides: s[rr] = {LJ, KJ, LM, KM}.
template <typename R, int RB>
__device__ __forceinline__ void far_sums(int K, const cplx_t<R>* __restrict__ V, const cplx_t<R>* __restrict__ B,
                                         const R* __restrict__ kp, int lane, cplx_t<R> (&s)[RB][4]) {
    using C = cplx_t<R>;
    for (int q = lane; q < K; q += 64) {
        const R kx = kp[q], ky = kp[K + q], kz = kp[2 * K + q];
        const C v0 = V[q], v1 = V[K + q], v2 = V[2 * K + q];
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) {
            const C* BJ = B + (size_t)rr * 6 * K;
            const C* BM = BJ + (size_t)3 * K;
            const C j0 = BJ[q], j1 = BJ[K + q], j2 = BJ[2 * K + q];
            const C m0 = BM[q], m1 = BM[K + q], m2 = BM[2 * K + q];
            cfma(s[rr][0], v0, j0);
            cfma(s[rr][0], v1, j1);
            cfma(s[rr][0], v2, j2);
            cfma(s[rr][2], v0, m0);
            cfma(s[rr][2], v1, m1);
            cfma(s[rr][2], v2, m2);
            // kp x B
            const C cj0 = csub(cscale(j2, ky), cscale(j1, kz));
            const C cj1 = csub(cscale(j0, kz), cscale(j2, kx));
            const C cj2 = csub(cscale(j1, kx), cscale(j0, ky));
            const C cm0 = csub(cscale(m2, ky), cscale(m1, kz));
            const C cm1 = csub(cscale(m0, kz), cscale(m2, kx));
            const C cm2 = csub(cscale(m1, kx), cscale(m0, ky));
            cfma(s[rr][1], v0, cj0);
            cfma(s[rr][1], v1, cj1);
            cfma(s[rr][1], v2, cj2);
            cfma(s[rr][3], v0, cm0);
            cfma(s[rr][3], v1, cm1);
            cfma(s[rr][3], v2, cm2);
        }
    }
}

// One wavefront per RWG row over the whole RWG range (four rows per block).
template <typename R, int RB>
__global__ __launch_bounds__(256) void k_disagg_pmchwt(int ldr, int r0, const int* __restrict__ leaf_of,
                                                       const int* __restrict__ perm, PmchwtFarArgs<R> a,
                                                       double2* __restrict__ y) {
    using C = cplx_t<R>;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wave;
    if (i >= a.N) return;
    const int leaf = leaf_of[i];
    C s1[RB][4], s2[RB][4];
#pragma unroll
    for (int rr = 0; rr < RB; ++rr)
#pragma unroll
        for (int f = 0; f < 4; ++f) s1[rr][f] = s2[rr][f] = C{0, 0};
    far_sums<R, RB>(a.K1, a.Vf1 + (size_t)i * 3 * a.K1, a.B1 + (size_t)leaf * RB * 6 * a.K1, a.kp1, lane, s1);
    far_sums<R, RB>(a.K2, a.Vf2 + (size_t)i * 3 * a.K2, a.B2 + (size_t)leaf * RB * 6 * a.K2, a.kp2, lane, s2);
#pragma unroll
    for (int rr = 0; rr < RB; ++rr)
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            s1[rr][f] = wave_sum(s1[rr][f]);
            s2[rr][f] = wave_sum(s2[rr][f]);
        }
    if (lane < RB) {
        C t1[4], t2[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) t1[f] = s1[0][f], t2[f] = s2[0][f];
#pragma unroll
        for (int rr = 1; rr < RB; ++rr)
            if (lane == rr)
#pragma unroll
                for (int f = 0; f < 4; ++f) t1[f] = s1[rr][f], t2[f] = s2[rr][f];
        const double2 L1J = widen(t1[0]), K1J = widen(t1[1]), L1M = widen(t1[2]), K1M = widen(t1[3]);
        const double2 L2J = widen(t2[0]), K2J = widen(t2[1]), L2M = widen(t2[2]), K2M = widen(t2[3]);
        const int r = perm[i];
        double2* pj = y + (size_t)r * ldr + r0 + lane;
        double2* pm = y + ((size_t)a.N + r) * ldr + r0 + lane;
        double2 wj = *pj, wm = *pm;
        // Code/MLFMM.cpp:339-349
        wj = cadd(wj, cmul(a.c1, L1J));
        wj = csub(wj, cmul(a.c1, K1M));
        wm = cadd(wm, cmul(a.c1, K1J));
        wm = cadd(wm, cmul(a.c1, L1M));
        wj = cadd(wj, cmul(cmul(a.c2, a.etai), L2J));
        wj = csub(wj, cmul(a.c2, K2M));
        wm = cadd(wm, cmul(a.c2, K2J));
        wm = cadd(wm, cmul(cmul(a.c2, L2M), a.inv_etai));
        *pj = wj;
        *pm = wm;
    }
}

template <typename R>
void launch_disagg_pmchwt_b(hipStream_t st, int rb, int ldr, int r0, const int* leaf_of, const int* perm,
                            const PmchwtFarArgs<R>& a, double2* y) {
    const int blocks = (a.N + 3) / 4;
    if (rb == 1) k_disagg_pmchwt<R, 1><<<blocks, 256, 0, st>>>(ldr, r0, leaf_of, perm, a, y);
    else if (rb == 2) k_disagg_pmchwt<R, 2><<<blocks, 256, 0, st>>>(ldr, r0, leaf_of, perm, a, y);
    else if (rb == 4) k_disagg_pmchwt<R, 4><<<blocks, 256, 0, st>>>(ldr, r0, leaf_of, perm, a, y);
    else if constexpr (sizeof(R) == 4) k_disagg_pmchwt<R, 8><<<blocks, 256, 0, st>>>(ldr, r0, leaf_of, perm, a, y);
    else throw std::invalid_argument("FP64 disaggregation: at most 4 right-hand sides per pass");
}

template <typename R>
void launch_disagg_pmchwt(hipStream_t st, const int* leaf_of, const int* perm, const PmchwtFarArgs<R>& a,
                          double2* y) {
    launch_disagg_pmchwt_b<R>(st, 1, 1, 0, leaf_of, perm, a, y);
}

__global__ void k_narrow(size_t n, const double2* __restrict__ in, float2* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        narrow(in[i], out[i]);
}

void launch_narrow(hipStream_t st, size_t n, const double2* in, float2* out) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    k_narrow<<<(unsigned)blocks, 256, 0, st>>>(n, in, out);
}

#define MLFMA_FAR_INSTANTIATE(R)                                                                                    \
    template void launch_aggregate<R>(hipStream_t, int, int, int, int, const int*, const int*, const int*,         \
                                      const cplx_t<R>*, const double2*, cplx_t<R>*);                                      \
    template void launch_m2m<R>(hipStream_t, int, int, int, int, const int*, const int*, const int*, const int*,   \
                                const R*, const cplx_t<R>*, const cplx_t<R>*, cplx_t<R>*);                                     \
    template void launch_m2l_l2l<R>(hipStream_t, int, int, int, const int*, const int*, const int*, const cplx_t<R>*,  \
                                    const cplx_t<R>*, int, const int*, const int*, const int*, const int*, const R*,   \
                                    const cplx_t<R>*, const cplx_t<R>*, cplx_t<R>*);                                           \
    template void launch_disagg_pec<R>(hipStream_t, int, int, const int*, const int*, const cplx_t<R>*, const cplx_t<R>*,  \
                                       double2, double2*);                                                        \
    template void launch_disagg_pmchwt<R>(hipStream_t, const int*, const int*, const PmchwtFarArgs<R>&, double2*); \
    template void launch_aggregate_b<R>(hipStream_t, int, int, int, int, int, int, int, const int*, const int*,     \
                                        const int*, const cplx_t<R>*, const double2*, cplx_t<R>*);                \
    template void launch_disagg_pec_b<R>(hipStream_t, int, int, int, int, int, const int*, const int*,             \
                                         const cplx_t<R>*, const cplx_t<R>*, double2, double2*);                  \
    template void launch_disagg_pmchwt_b<R>(hipStream_t, int, int, int, const int*, const int*,                    \
                                            const PmchwtFarArgs<R>&, double2*);                                   \
    template bool launch_m2m_sep<R>(hipStream_t, int, int, int, int, const int*, const int*, const int*, const int*, \
                                    const R*, const int*, const R*, const cplx_t<R>*, const cplx_t<R>*,           \
                                    cplx_t<R>*);                                                                  \
    template void launch_m2l_cls<R>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,      \
                                    const cplx_t<R>*, const cplx_t<R>*, int, const int*, const int*, const int*,  \
                                    const int*, const R*, const cplx_t<R>*, const cplx_t<R>*, cplx_t<R>*);        \
    template void launch_m2l_cl<R>(hipStream_t, int, int, int, int, const int*, const int*, const int*, const int*,  \
                                   const cplx_t<R>*, const cplx_t<R>*, int, const int*, const int*, const int*,   \
                                   const int*, const R*, const cplx_t<R>*, const cplx_t<R>*, cplx_t<R>*);       \
    template void launch_l2l_sep<R>(hipStream_t, int, int, int, int, int, const int*, const int*, const int*,        \
                                    const int*, const R*, const int*, const int*, const R*, const cplx_t<R>*,     \
                                    const cplx_t<R>*, cplx_t<R>*);
MLFMA_FAR_INSTANTIATE(double)
MLFMA_FAR_INSTANTIATE(float)
#undef MLFMA_FAR_INSTANTIATE

// ---------------------------------------------------------------- near field
// Code/MLFMM.cpp:228-231 (PEC) and :354-366 (PMCHWT).  All near blocks of one
// target leaf form a single column-major nt x C_t matrix per plane (PMCHWT:
// z11, zK, z22 with z12 = -zK, z21 = zK, Code/Zmartix.h:611-614), so one
// workgroup streams one contiguous range of HBM.  For nt < BS the BS threads
// form G = BS / nt column groups of nt rows and group g takes the columns
// c = g (mod G): thread tid reads entry c * nt + tid, so the workgroup reads
// consecutive addresses; U columns per thread are in flight at once.  The
// group partials are summed in LDS in a fixed order.  x is gathered into LDS
// CH columns at a time.  Entries are stored in S (double or float) and
// accumulated in FP64.
template <typename S, int P, int BS, int U, int CH>
__global__ __launch_bounds__(BS) void k_near(NearDev nd, const cplx_t<S>* __restrict__ Z,
                                             const double2* __restrict__ x, double2* __restrict__ y) {
    constexpr int NX = (P == 3) ? 2 : 1;
    __shared__ double2 xs[NX][CH];
    __shared__ double2 red[NX][BS];
    const int t = nd.order[blockIdx.x];
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    if (nt == 0) return;
    const int C = nd.tgt_cols[t];
    const int* cref = nd.colref + nd.col_ptr[t];
    const int plane = C * nt;   // < 2^31: one target's block row
    const cplx_t<S>* Zt = Z + nd.tgt_off[t];
    const int span = nt < BS ? nt : BS;
    const int G = BS / span;
    const int g = threadIdx.x / span, rl = threadIdx.x - g * span;
    for (int r0 = 0; r0 < nt; r0 += span) {
        const int row = r0 + rl;
        const bool act = g < G && row < nt;
        double2 aJ = {0, 0}, aM = {0, 0};
        for (int c0 = 0; c0 < C; c0 += CH) {
            const int cl = min(CH, C - c0);
            __syncthreads();
            for (int c = threadIdx.x; c < cl; c += BS) {
                const int r = cref[c0 + c];
                xs[0][c] = x[r];
                if (P == 3) xs[NX - 1][c] = x[nd.N + r];
            }
            __syncthreads();
            if (act) {
                const cplx_t<S>* z = Zt + c0 * nt + row;
                const int step = G * nt;
                int c = g;
                if (P == 1) {
                    for (; c + (U - 1) * G < cl; c += U * G) {
                        double2 zz[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) zz[u] = widen(z[c * nt + u * step]);
#pragma unroll
                        for (int u = 0; u < U; ++u) cfma(aJ, zz[u], xs[0][c + u * G]);
                    }
                    for (; c < cl; c += G) cfma(aJ, widen(z[c * nt]), xs[0][c]);
                } else {
                    for (; c + (U - 1) * G < cl; c += U * G) {
                        double2 p[U], k[U], m[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int o = c * nt + u * step;
                            p[u] = widen(z[o]);
                            k[u] = widen(z[o + plane]);
                            m[u] = widen(z[o + 2 * plane]);
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const double2 xJ = xs[0][c + u * G], xM = xs[NX - 1][c + u * G];
                            cfma(aJ, p[u], xJ);
                            cfms(aJ, k[u], xM);
                            cfma(aM, k[u], xJ);
                            cfma(aM, m[u], xM);
                        }
                    }
                    for (; c < cl; c += G) {
                        const int o = c * nt;
                        const double2 p0 = widen(z[o]), k0 = widen(z[o + plane]), m0 = widen(z[o + 2 * plane]);
                        const double2 xJ = xs[0][c], xM = xs[NX - 1][c];
                        cfma(aJ, p0, xJ);
                        cfms(aJ, k0, xM);
                        cfma(aM, k0, xJ);
                        cfma(aM, m0, xM);
                    }
                }
            }
        }
        if (G > 1) {   // nt < BS: a single row pass, fold the column groups
            red[0][threadIdx.x] = aJ;
            if (P == 3) red[NX - 1][threadIdx.x] = aM;
            __syncthreads();
            if (threadIdx.x < nt) {
                double2 sJ = red[0][threadIdx.x], sM = red[NX - 1][threadIdx.x];
                for (int h = 1; h < G; ++h) {
                    sJ = cadd(sJ, red[0][h * nt + threadIdx.x]);
                    if (P == 3) sM = cadd(sM, red[NX - 1][h * nt + threadIdx.x]);
                }
                const int r = nd.perm[i0 + threadIdx.x];
                y[r] = sJ;
                if (P == 3) y[nd.N + r] = sM;
            }
        } else if (act) {   // G == 1: threads past nt hold no row
            const int r = nd.perm[i0 + row];
            y[r] = aJ;
            if (P == 3) y[nd.N + r] = aM;
        }
    }
}

template <typename S>
void launch_near(hipStream_t st, const NearDev& nd, const cplx_t<S>* Z, const double2* x, double2* y) {
    if (nd.nleaf == 0) return;
    if (nd.planes == 1) k_near<S, 1, 512, 8, 1024><<<nd.nleaf, 512, 0, st>>>(nd, Z, x, y);
    else k_near<S, 3, 512, 4, 512><<<nd.nleaf, 512, 0, st>>>(nd, Z, x, y);
}
template void launch_near<double>(hipStream_t, const NearDev&, const double2*, const double2*, double2*);
template void launch_near<float>(hipStream_t, const NearDev&, const float2*, const double2*, double2*);

// Batched near product: RB right-hand sides per pass over the entries (vectors
// RHS-fastest, element (u, r) at u * ldr + r).  Same column-group scheme as
// k_near with 256 threads; the x chunk is staged as [column][rhs] and each
// thread keeps RB (PMCHWT: 2 RB) FP64 accumulators.
template <typename S, int P, int RB, int U, int CH>
__global__ __launch_bounds__(256) void k_near_b(NearDev nd, int ldr, int r0, const cplx_t<S>* __restrict__ Z,
                                                const double2* __restrict__ x, double2* __restrict__ y) {
    constexpr int NX = (P == 3) ? 2 : 1;
    constexpr int BS = 256;
    __shared__ double2 xs[NX][CH][RB];
    __shared__ double2 red[NX][BS];
    const int t = nd.order[blockIdx.x];
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    if (nt == 0) return;
    const int C = nd.tgt_cols[t];
    const int* cref = nd.colref + nd.col_ptr[t];
    const int plane = C * nt;
    const cplx_t<S>* Zt = Z + nd.tgt_off[t];
    const int span = nt < BS ? nt : BS;
    const int G = BS / span;
    const int g = threadIdx.x / span, rl = threadIdx.x - g * span;
    for (int rs = 0; rs < nt; rs += span) {
        const int row = rs + rl;
        const bool act = g < G && row < nt;
        double2 aJ[RB], aM[RB];
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) aJ[rr] = aM[rr] = make_double2(0.0, 0.0);
        for (int c0 = 0; c0 < C; c0 += CH) {
            const int cl = min(CH, C - c0);
            __syncthreads();
            for (int e = threadIdx.x; e < cl * RB; e += BS) {
                const int c = e / RB, rr = e - c * RB;
                const int r = cref[c0 + c];
                xs[0][c][rr] = x[(size_t)r * ldr + r0 + rr];
                if (P == 3) xs[NX - 1][c][rr] = x[((size_t)nd.N + r) * ldr + r0 + rr];
            }
            __syncthreads();
            if (act) {
                const cplx_t<S>* z = Zt + c0 * nt + row;
                const int step = G * nt;
                int c = g;
                for (; c + (U - 1) * G < cl; c += U * G) {
                    double2 p[U], k[U], m[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int o = c * nt + u * step;
                        p[u] = widen(z[o]);
                        if (P == 3) {
                            k[u] = widen(z[o + plane]);
                            m[u] = widen(z[o + 2 * plane]);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int rr = 0; rr < RB; ++rr) {
                            const double2 xJ = xs[0][c + u * G][rr];
                            cfma(aJ[rr], p[u], xJ);
                            if (P == 3) {
                                const double2 xM = xs[NX - 1][c + u * G][rr];
                                cfms(aJ[rr], k[u], xM);
                                cfma(aM[rr], k[u], xJ);
                                cfma(aM[rr], m[u], xM);
                            }
                        }
                }
                for (; c < cl; c += G) {
                    const int o = c * nt;
                    const double2 p0 = widen(z[o]);
                    double2 k0 = p0, m0 = p0;
                    if (P == 3) {
                        k0 = widen(z[o + plane]);
                        m0 = widen(z[o + 2 * plane]);
                    }
#pragma unroll
                    for (int rr = 0; rr < RB; ++rr) {
                        const double2 xJ = xs[0][c][rr];
                        cfma(aJ[rr], p0, xJ);
                        if (P == 3) {
                            const double2 xM = xs[NX - 1][c][rr];
                            cfms(aJ[rr], k0, xM);
                            cfma(aM[rr], k0, xJ);
                            cfma(aM[rr], m0, xM);
                        }
                    }
                }
            }
        }
        // fold the column groups one right-hand side at a time
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) {
            if (G > 1) {
                __syncthreads();
                red[0][threadIdx.x] = aJ[rr];
                if (P == 3) red[NX - 1][threadIdx.x] = aM[rr];
                __syncthreads();
                if (threadIdx.x < nt) {
                    double2 sJ = red[0][threadIdx.x], sM = red[NX - 1][threadIdx.x];
                    for (int h = 1; h < G; ++h) {
                        sJ = cadd(sJ, red[0][h * nt + threadIdx.x]);
                        if (P == 3) sM = cadd(sM, red[NX - 1][h * nt + threadIdx.x]);
                    }
                    const size_t r = nd.perm[i0 + threadIdx.x];
                    y[r * ldr + r0 + rr] = sJ;
                    if (P == 3) y[((size_t)nd.N + r) * ldr + r0 + rr] = sM;
                }
            } else if (act) {
                const size_t r = nd.perm[i0 + row];
                y[r * ldr + r0 + rr] = aJ[rr];
                if (P == 3) y[((size_t)nd.N + r) * ldr + r0 + rr] = aM[rr];
            }
        }
    }
}

template <typename S>
void launch_near_b(hipStream_t st, int rb, int ldr, int r0, const NearDev& nd, const cplx_t<S>* Z, const double2* x,
                   double2* y) {
    if (nd.nleaf == 0) return;
    if (rb == 1) {
        if (ldr == 1 && r0 == 0) launch_near<S>(st, nd, Z, x, y);
        else if (nd.planes == 1) k_near_b<S, 1, 1, 8, 512><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
        else k_near_b<S, 3, 1, 4, 256><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
    } else if (nd.planes == 1) {
        if (rb == 2) k_near_b<S, 1, 2, 8, 512><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
        else if (rb == 4) k_near_b<S, 1, 4, 4, 256><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
        else k_near_b<S, 1, 8, 4, 256><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
    } else {
        if (rb == 2) k_near_b<S, 3, 2, 2, 256><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
        else k_near_b<S, 3, 4, 2, 128><<<nd.nleaf, 256, 0, st>>>(nd, ldr, r0, Z, x, y);
    }
}
template void launch_near_b<double>(hipStream_t, int, int, int, const NearDev&, const double2*, const double2*,
                                    double2*);
template void launch_near_b<float>(hipStream_t, int, int, int, const NearDev&, const float2*, const double2*,
                                   double2*);

}  // namespace mlfma
