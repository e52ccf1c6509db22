// Batched near-field product on the matrix cores: for every target leaf t,
// Y_t (nt x R) = Z_t (nt x C_t) X_t (C_t x R), either with the FP32 copy of
// the near field (near_precision = MLFMA_FP32; v_mfma_f32_16x16x4_f32, FP32
// accumulation) or with the FP64 entries (v_mfma_f64_16x16x4_f64, FP64
// accumulation; the batched FP64 products, e.g. the true residuals of a
// batched solve).
// Code/MLFMM.cpp:228-231 (PEC) and :354-366 (PMCHWT: Y_J = Z11 X_J - ZK X_M,
// Y_M = ZK X_J + Z22 X_M, Code/Zmartix.h:611-614) for many right-hand sides.
//
// v_mfma_{f32,f64}_16x16x4 compute D(16x16) += A(16x4) B(4x16); lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16], and D[4 (l / 16) + v][l % 16]
// (FP32) or D[l / 16 + 4 v][l % 16] (FP64), v < 4.
// A complex product Z X over 8 right-hand sides is two real products with
// the 16 output columns read as [Re Y (8) | Im Y (8)]:
//   D += Re Z . [Re X | Im X] + Im Z . [-Im X | Re X].
// A workgroup owns one work item (a target leaf's range of up to 4 16-row
// tiles; items are sorted by decreasing work so large leaves do not set the
// duration) and CT column tiles of 8 right-hand sides (up to 64 in FP32, 32 in
// FP64).  Wavefront w owns row tile w and every column tile, so each
// near-field entry is loaded once, by one wavefront, and the right-hand sides
// come from LDS.  The near-field values of the next k-step are loaded while
// the current one's matrix-core products run.  X is staged in LDS 16 columns at
// a time in FP32 (from an FP32 copy, the next chunk in flight during the
// current one's products) and up to 32 in FP64 (fewer for wide batches, the
// stage stays within 32 KB).
#include "cplx.hpp"
#include "kernels.hpp"

namespace mlfma {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int kRTC = 4;   // row tiles per work item

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc = floatx4;
    static constexpr int kCC = 64;   // staged columns per chunk
    __device__ static acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static void stage(const double2& x, float2& v) { narrow(x, v); }
    __device__ static int row(int lk, int v) { return 4 * lk + v; }   // D row held in element v of lane group lk
};
template <>
struct Mfma<double> {
    using acc = doublex4;
    static constexpr int kCC = 32;
    __device__ static acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static void stage(const double2& x, double2& v) { v = x; }
    __device__ static int row(int lk, int v) { return lk + 4 * v; }   // the FP64 product interleaves its rows
};
}  // namespace

// xf[u * R + r] = FP32 copy of x[u * ldr + r0 + r]; u < nu, r < R
__global__ void k_x_f32(size_t total, int R, int ldr, int r0, const double2* __restrict__ x, float2* __restrict__ xf) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const size_t u = e / R;
    const double2 v = x[u * ldr + r0 + (int)(e - u * R)];
    xf[e] = make_float2((float)v.x, (float)v.y);
}

template <typename T, int P, int CT>
__global__ __launch_bounds__(256) void k_near_mfma(NearDev nd, int ldr, int r0, int R,
                                                   const cplx_t<T>* __restrict__ Z, const double2* __restrict__ x,
                                                   const float2* __restrict__ xf, double2* __restrict__ y) {
    using C2 = cplx_t<T>;
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NX = (P == 3) ? 2 : 1;   // staged currents
    constexpr int NO = (P == 3) ? 2 : 1;   // outputs
    constexpr int RT = 8 * CT;             // right-hand sides per workgroup
    // FP32: X comes from its FP32 copy xf (R per unknown), 16 columns at a
    // time, the next chunk loaded into registers while the current one's
    // products run (source indices two chunks ahead).  FP64: staged from x,
    // chunks halved until the stage fits 32 KB of LDS.
    constexpr bool PF = sizeof(T) == 4;
    constexpr int kCC = PF                                                   ? 16
                        : NX * RT * (int)sizeof(C2) * M::kCC <= 32768       ? M::kCC
                        : NX * RT * (int)sizeof(C2) * M::kCC / 2 <= 32768 ? M::kCC / 2
                                                                          : M::kCC / 4;
    __shared__ C2 xs[NX][kCC][RT];
    const int2 item = nd.items[blockIdx.x];
    const int t = item.x, rp = item.y;
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    const int nrp = min(kRTC, ((nt + 15) >> 4) - rp);
    const int rb0 = blockIdx.y * RT;
    const int rv = min(RT, R - rb0);
    const int C = nd.tgt_cols[t];
    const int* cref = nd.colref + nd.col_ptr[t];
    const int plane = C * nt;
    const C2* Zt = Z + nd.tgt_off[t];
    // the wavefront index as a scalar, so tests on it are uniform branches
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;   // 0: real-part column, 1: imaginary-part column
    const bool wact = wave < nrp;   // this wavefront's row tile exists
    // this lane's row of the A operand, clamped into the leaf: rows past nt
    // give output rows that are not stored, columns past C meet zero-staged X,
    // so the loads need no bounds branches (which would make the compiler wait
    // for the prefetched values before the current products)
    const int arow = min((rp + (wact ? wave : 0)) * 16 + li, nt - 1);
    // near-field values of global column gc for this lane's row
    auto load = [&](int gc, C2& z1, C2& zk, C2& z2) {
        const int off = min(gc, C - 1) * nt + arow;
        z1 = Zt[off];
        if (P == 3) {
            zk = Zt[plane + off];
            z2 = Zt[2 * plane + off];
        }
    };
    acc_t acc[CT][NO];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int o = 0; o < NO; ++o) acc[c][o] = acc_t{0, 0, 0, 0};
    // FP32 staging units of this thread: element e = threadIdx.x + 256 u of
    // the chunk is (current e / (kCC RT), column, right-hand side).  Loads go
    // to clamped (valid) addresses without branches; pmask marks the values
    // that are stored (the rest are staged as zero).
    constexpr int NE = NX * kCC * RT, NU = PF ? (NE + 255) / 256 : 1;
    int cr[NU];        // source unknown of each unit's column in the chunk after the staged one
    float2 pv[NU];     // next chunk's X
    unsigned pmask = 0;
    auto unit = [&](int u, int& xj, int& c, int& r) {
        const int e = threadIdx.x + 256 * u, ec = min(e, NE - 1);
        xj = ec / (kCC * RT);
        const int rem = ec - xj * (kCC * RT);
        c = rem / RT;
        r = rem - c * RT;
        return NE % 256 == 0 || e < NE;
    };
    auto load_cref = [&](int cc) {
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            int xj, c, r;
            unit(u, xj, c, r);
            cr[u] = cref[min(cc + c, C - 1)];
        }
    };
    auto load_x = [&](int cc) {
        pmask = 0;
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            int xj, c, r;
            const bool ok = unit(u, xj, c, r);
            pv[u] = xf[((size_t)cr[u] + (xj ? (size_t)nd.N : 0)) * R + rb0 + min(r, rv - 1)];
            pmask |= (unsigned)(ok & (cc + c < C) & (r < rv)) << u;
        }
    };
    if constexpr (PF) {
        load_cref(0);
        load_x(0);
        load_cref(kCC);
    }
    // k-steps alternate between two register sets (no copies, so the next
    // step's values stay in flight during the current step's products)
    auto step = [&](int kk, const C2& z1, const C2& zk, const C2& z2) {
        if (!wact) return;
        const int cl = kk + lk;   // this lane's column in the chunk (< kCC); zero-staged past kend
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int rc = ct * 8 + (li & 7);
            const C2 xj = xs[0][cl][rc];
            const T bJ1 = part ? xj.y : xj.x, bJ2 = part ? xj.x : -xj.y;
            acc_t a = acc[ct][0];
            a = M::op(z1.x, bJ1, a);
            a = M::op(z1.y, bJ2, a);
            if (P == 3) {
                const C2 xm = xs[NX - 1][cl][rc];
                const T bM1 = part ? xm.y : xm.x, bM2 = part ? xm.x : -xm.y;
                a = M::op(zk.x, -bM1, a);   // - ZK X_M
                a = M::op(zk.y, -bM2, a);
                acc_t b = acc[ct][NO - 1];
                b = M::op(zk.x, bJ1, b);
                b = M::op(zk.y, bJ2, b);
                b = M::op(z2.x, bM1, b);
                b = M::op(z2.y, bM2, b);
                acc[ct][NO - 1] = b;
            }
            acc[ct][0] = a;
        }
    };
    C2 a1, ak, a2, b1, bk, b2;
    load(lk, a1, ak, a2);
    for (int c0 = 0; c0 < C; c0 += kCC) {
        __syncthreads();
        if constexpr (PF) {
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                int xj, c, r;
                if (unit(u, xj, c, r)) {
                    const bool v = (pmask >> u) & 1u;
                    xs[xj][c][r].x = v ? pv[u].x : 0.f;
                    xs[xj][c][r].y = v ? pv[u].y : 0.f;
                }
            }
            __syncthreads();
            load_x(c0 + kCC);
            load_cref(c0 + 2 * kCC);
        } else {
            for (int e = threadIdx.x; e < NX * kCC * RT; e += 256) {
                const int xj = e / (kCC * RT), rem = e - xj * (kCC * RT);
                const int c = rem / RT, r = rem - c * RT;
                C2 v = C2{0, 0};
                if (c0 + c < C && r < rv) {
                    const size_t u = (size_t)cref[c0 + c] + (xj ? (size_t)nd.N : 0);
                    M::stage(x[u * ldr + r0 + rb0 + r], v);
                }
                xs[xj][c][r].x = v.x;
                xs[xj][c][r].y = v.y;
            }
            __syncthreads();
        }
        // Every load is issued on every pass (past the end they read clamped,
        // finite values that meet zero-staged X), so the compiler's wait
        // counts stay exact: each wait covers the data it needs, not the loads
        // issued after it.
#pragma unroll
        for (int kk = 0; kk < kCC; kk += 8) {
            load(c0 + kk + 4 + lk, b1, bk, b2);
            step(kk, a1, ak, a2);
            load(c0 + kk + 8 + lk, a1, ak, a2);
            step(kk + 4, b1, bk, b2);
        }
    }
    if (!wact) return;
    // D[row(lk, v)][li]: lanes li < 8 hold Re of right-hand side ct * 8 + li, lane li + 8 its Im
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T re = acc[ct][o][v];
                const T im = __shfl_xor(re, 8, 64);
                const int row = (rp + wave) * 16 + M::row(lk, v);
                const int rc = ct * 8 + li;
                if (part == 0 && rc < rv && row < nt) {
                    const size_t u = (size_t)nd.perm[i0 + row] + (o ? (size_t)nd.N : 0);
                    y[u * ldr + r0 + rb0 + rc] = make_double2((double)re, (double)im);
                }
            }
}

template <typename T>
static void near_mfma_t(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const cplx_t<T>* Z,
                        const double2* x, float2* xf, double2* y) {
    if (nd.nitem == 0 || R < 1) return;
    constexpr bool f64 = sizeof(T) == 8;   // FP64 accumulators: at most 32 right-hand sides per workgroup
    if (!f64) {
        const size_t total = (size_t)nd.N * (nd.planes == 3 ? 2 : 1) * R;   // unknowns x right-hand sides
        k_x_f32<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, R, ldr, r0, x, xf);
    }
    const int ct = (R > 32 && !f64) ? 8 : (R > 16 ? 4 : (R > 8 ? 2 : 1));
    dim3 grid(nd.nitem, (R + 8 * ct - 1) / (8 * ct));
#define NM(P, CT) k_near_mfma<T, P, CT><<<grid, 256, 0, st>>>(nd, ldr, r0, R, Z, x, xf, y)
#define NMP(P)                                                                                                        \
    if (ct == 8 && !f64) NM(P, 8);                                                                                    \
    else if (ct == 4) NM(P, 4);                                                                                       \
    else if (ct == 2) NM(P, 2);                                                                                       \
    else NM(P, 1);
    if (nd.planes == 1) {
        NMP(1)
    } else {
        NMP(3)
    }
#undef NMP
#undef NM
}

void launch_near_mfma(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const float2* Z, const double2* x,
                      float2* xf, double2* y) {
    near_mfma_t<float>(st, R, ldr, r0, nd, Z, x, xf, y);
}

void launch_near_mfma64(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const double2* Z,
                        const double2* x, double2* y) {
    near_mfma_t<double>(st, R, ldr, r0, nd, Z, x, nullptr, y);
}

}  // namespace mlfma
