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
// the current one's matrix-core products run.  X is staged in LDS up to 64 (FP32) or 32 (FP64) columns at a
// time, fewer for wide batches (the stage stays within 32 KB).
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

template <typename T, int P, int CT>
__global__ __launch_bounds__(256) void k_near_mfma(NearDev nd, int ldr, int r0, int R,
                                                   const cplx_t<T>* __restrict__ Z, const double2* __restrict__ x,
                                                   double2* __restrict__ y) {
    using C2 = cplx_t<T>;
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NX = (P == 3) ? 2 : 1;   // staged currents
    constexpr int NO = (P == 3) ? 2 : 1;   // outputs
    constexpr int RT = 8 * CT;             // right-hand sides per workgroup
    // staged columns per chunk, halved until the stage fits 32 KB of LDS
    constexpr int kCC = NX * RT * (int)sizeof(C2) * M::kCC <= 32768       ? M::kCC
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
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;   // 0: real-part column, 1: imaginary-part column
    const bool wact = wave < nrp;              // this wavefront's row tile exists
    const int arow = (rp + wave) * 16 + li;    // this lane's row of the A operand
    const bool rok = wact && arow < nt;
    // near-field values of global column gc for this lane's row
    auto load = [&](int gc, C2& z1, C2& zk, C2& z2) {
        z1 = zk = z2 = C2{0, 0};
        if (rok && gc < C) {
            const int off = gc * nt + arow;
            z1 = Zt[off];
            if (P == 3) {
                zk = Zt[plane + off];
                z2 = Zt[2 * plane + off];
            }
        }
    };
    acc_t acc[CT][NO];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int o = 0; o < NO; ++o) acc[c][o] = acc_t{0, 0, 0, 0};
    C2 n1, nk, n2;
    load(lk, n1, nk, n2);
    for (int c0 = 0; c0 < C; c0 += kCC) {
        __syncthreads();
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
        const int kend = min(kCC, C - c0);
        for (int kk = 0; kk < kend; kk += 4) {
            const C2 z1 = n1, zk = nk, z2 = n2;
            load(c0 + kk + 4 + lk, n1, nk, n2);   // next k-step
            if (!wact) continue;
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
                        const double2* x, double2* y) {
    if (nd.nitem == 0 || R < 1) return;
    constexpr bool f64 = sizeof(T) == 8;   // FP64 accumulators: at most 32 right-hand sides per workgroup
    const int ct = (R > 32 && !f64) ? 8 : (R > 16 ? 4 : (R > 8 ? 2 : 1));
    dim3 grid(nd.nitem, (R + 8 * ct - 1) / (8 * ct));
#define NM(P, CT) k_near_mfma<T, P, CT><<<grid, 256, 0, st>>>(nd, ldr, r0, R, Z, x, y)
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
                      double2* y) {
    near_mfma_t<float>(st, R, ldr, r0, nd, Z, x, y);
}

void launch_near_mfma64(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const double2* Z,
                        const double2* x, double2* y) {
    near_mfma_t<double>(st, R, ldr, r0, nd, Z, x, y);
}

}  // namespace mlfma
