// Batched near-field product on the matrix cores: for every target leaf t,
// Y_t (nt x R) = Z_t (nt x C_t) X_t (C_t x R) with the FP32 copy of the near
// field (near_precision = MLFMA_FP32), FP32 accumulation.
// Code/MLFMM.cpp:228-231 (PEC) and :354-366 (PMCHWT: Y_J = Z11 X_J - ZK X_M,
// Y_M = ZK X_J + Z22 X_M, Code/Zmartix.h:611-614) for many right-hand sides.
//
// v_mfma_f32_16x16x4_f32 computes D(16x16) += A(16x4) B(4x16); lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16] and D[4 (l / 16) + v][l % 16], v < 4.
// A complex product Z X over 8 right-hand sides is two real products with
// the 16 output columns read as [Re Y (8) | Im Y (8)]:
//   D += Re Z . [Re X | Im X] + Im Z . [-Im X | Re X].
// A workgroup (4 wavefronts) owns one target leaf and 16 right-hand sides
// (two 8-wide column tiles); rows go in 16-row tiles.  With 1 or 2 row tiles
// the wavefronts split the columns and fold their partial tiles in LDS in a
// fixed order.  X is staged in LDS 64 columns at a time, in FP32.
#include "cplx.hpp"
#include "kernels.hpp"

namespace mlfma {

typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int kRT = 16;   // right-hand sides per workgroup
constexpr int kCC = 64;   // staged columns per chunk
}  // namespace

template <int P>
__global__ __launch_bounds__(256) void k_near_mfma(NearDev nd, int ldr, int r0, int R, const float2* __restrict__ Z,
                                                   const double2* __restrict__ x, double2* __restrict__ y) {
    constexpr int NX = (P == 3) ? 2 : 1;   // staged currents
    constexpr int NO = (P == 3) ? 2 : 1;   // outputs
    __shared__ float2 xs[NX][kCC][kRT];
    __shared__ floatx4 red[4][2][NO][64];
    const int t = nd.order[blockIdx.x];
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    if (nt == 0) return;
    const int rb0 = blockIdx.y * kRT;
    const int rv = min(kRT, R - rb0);
    const int C = nd.tgt_cols[t];
    const int* cref = nd.colref + nd.col_ptr[t];
    const int plane = C * nt;
    const float2* Zt = Z + nd.tgt_off[t];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;   // 0: real-part column, 1: imaginary-part column
    const int nrt = (nt + 15) >> 4;
    for (int rp = 0; rp < nrt; rp += 16) {
        const int nrp = min(16, nrt - rp);
        const int CG = nrp == 1 ? 4 : (nrp == 2 ? 2 : 1);   // column groups per row tile
        const int cg = CG > 1 ? wave % CG : 0;
        floatx4 acc[4][2][NO];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int o = 0; o < NO; ++o) acc[s][ct][o] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int c0 = 0; c0 < C; c0 += kCC) {
            __syncthreads();
            for (int e = threadIdx.x; e < NX * kCC * kRT; e += 256) {
                const int xj = e / (kCC * kRT), rem = e - xj * (kCC * kRT);
                const int c = rem / kRT, r = rem - c * kRT;
                float2 v = make_float2(0.f, 0.f);
                if (c0 + c < C && r < rv) {
                    const size_t u = (size_t)cref[c0 + c] + (xj ? (size_t)nd.N : 0);
                    narrow(x[u * ldr + r0 + rb0 + r], v);
                }
                xs[xj][c][r] = v;
            }
            __syncthreads();
            const int kend = min(kCC, C - c0);
            for (int kk = cg * 4; kk < kend; kk += 4 * CG) {
                const int cl = kk + lk;           // this lane's column in the chunk
                const bool cok = cl < kend;
                float bJ1[2], bJ2[2], bM1[2], bM2[2], bN1[2], bN2[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const float2 xj = xs[0][cl < kCC ? cl : 0][ct * 8 + (li & 7)];
                    bJ1[ct] = part ? xj.y : xj.x;
                    bJ2[ct] = part ? xj.x : -xj.y;
                    if (P == 3) {
                        const float2 xm = xs[NX - 1][cl < kCC ? cl : 0][ct * 8 + (li & 7)];
                        bM1[ct] = part ? xm.y : xm.x;     // +X_M
                        bM2[ct] = part ? xm.x : -xm.y;
                        bN1[ct] = -bM1[ct];               // -X_M
                        bN2[ct] = -bM2[ct];
                    }
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int tile = CG > 1 ? wave / CG : wave + 4 * s;
                    if ((CG > 1 && s > 0) || tile >= nrp) continue;
                    const int row = (rp + tile) * 16 + li;
                    const bool ok = cok && row < nt;
                    const int off = (c0 + cl) * nt + row;
                    const float2 z1 = ok ? Zt[off] : make_float2(0.f, 0.f);
                    float2 zk = make_float2(0.f, 0.f), z2 = make_float2(0.f, 0.f);
                    if (P == 3 && ok) {
                        zk = Zt[plane + off];
                        z2 = Zt[2 * plane + off];
                    }
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        floatx4 a = acc[s][ct][0];
                        a = __builtin_amdgcn_mfma_f32_16x16x4f32(z1.x, bJ1[ct], a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_16x16x4f32(z1.y, bJ2[ct], a, 0, 0, 0);
                        if (P == 3) {
                            a = __builtin_amdgcn_mfma_f32_16x16x4f32(zk.x, bN1[ct], a, 0, 0, 0);
                            a = __builtin_amdgcn_mfma_f32_16x16x4f32(zk.y, bN2[ct], a, 0, 0, 0);
                            floatx4 b = acc[s][ct][NO - 1];
                            b = __builtin_amdgcn_mfma_f32_16x16x4f32(zk.x, bJ1[ct], b, 0, 0, 0);
                            b = __builtin_amdgcn_mfma_f32_16x16x4f32(zk.y, bJ2[ct], b, 0, 0, 0);
                            b = __builtin_amdgcn_mfma_f32_16x16x4f32(z2.x, bM1[ct], b, 0, 0, 0);
                            b = __builtin_amdgcn_mfma_f32_16x16x4f32(z2.y, bM2[ct], b, 0, 0, 0);
                            acc[s][ct][NO - 1] = b;
                        }
                        acc[s][ct][0] = a;
                    }
                }
            }
        }
        // fold the column groups of a row tile (fixed order)
        if (CG > 1) {
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int o = 0; o < NO; ++o) red[wave][ct][o][lane] = acc[0][ct][o];
            __syncthreads();
            if (cg == 0)
                for (int g = 1; g < CG; ++g)
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                        for (int o = 0; o < NO; ++o) acc[0][ct][o] += red[wave + g][ct][o][lane];
        }
        // D[4 lk + v][li]: lanes li < 8 hold Re of right-hand side ct * 8 + li, lane li + 8 its Im
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int tile = CG > 1 ? wave / CG : wave + 4 * s;
            const bool owner = (CG == 1 || (cg == 0 && s == 0)) && tile < nrp;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int o = 0; o < NO; ++o)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float re = acc[s][ct][o][v];
                        const float im = __shfl_xor(re, 8, 64);
                        const int r = ct * 8 + li;
                        const int row = (rp + tile) * 16 + 4 * lk + v;
                        if (owner && part == 0 && r < rv && row < nt) {
                            const size_t u = (size_t)nd.perm[i0 + row] + (o ? (size_t)nd.N : 0);
                            y[u * ldr + r0 + rb0 + r] = make_double2(re, im);
                        }
                    }
        }
    }
}

void launch_near_mfma(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const float2* Z, const double2* x,
                      double2* y) {
    if (nd.nleaf == 0 || R < 1) return;
    dim3 grid(nd.nleaf, (R + kRT - 1) / kRT);
    if (nd.planes == 1) k_near_mfma<1><<<grid, 256, 0, st>>>(nd, ldr, r0, R, Z, x, y);
    else k_near_mfma<3><<<grid, 256, 0, st>>>(nd, ldr, r0, R, Z, x, y);
}

}  // namespace mlfma
