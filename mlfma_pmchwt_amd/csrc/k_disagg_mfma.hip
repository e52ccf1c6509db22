// Batched FP32 disaggregation on the matrix cores (far_precision = MLFMA_FP32).
// Code/MLFMM.cpp:211-227 (PEC) and :285-351 (PMCHWT) for rb right-hand sides.
//
// Every output is linear in the finest radiation workspaces, so the constants
// and the k-hat cross products are folded first, per leaf and right-hand side,
// into effective receive-side matrices (k_beff):
//   PEC:     E = c B
//   PMCHWT:  E_J1 = c1 (B_J1 - kp x B_M1),   E_M1 = c1 (kp x B_J1 + B_M1)
//            E_J2 = c2 (etai B_J2 - kp x B_M2), E_M2 = c2 (kp x B_J2 + B_M2 / etai)
// (kp x B reproduces the reference's (-khat x V).B with khat = -kp).  Then, per
// leaf, Y (nt x rb) += sum over wavenumbers V_w (nt x 3K_w) E_w (3K_w x rb), a
// complex GEMM done as two real v_mfma_f32_16x16x4_f32 products on [Re | Im]
// column tiles (same lane layout as k_near_mfma).  Rows go in 16-row tiles,
// one workgroup per leaf (largest first); FP32 accumulation.
#include <algorithm>

#include "cplx.hpp"
#include "kernels.hpp"

namespace mlfma {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// E[leaf][rr][o][3][K] from Bm[leaf][rr * ncur + cur][3][K]; one thread per (leaf, rr, k), flat grid.
__global__ void k_beff(int nleaf, int rb, int K, int pmchwt, const float2* __restrict__ Bm,
                       const float* __restrict__ kp, double2 ca, double2 cb, double2 cc, double2 cd,
                       float2* __restrict__ E) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nleaf * rb * K) return;
    const int k = (int)(e % K);
    const int rr = (int)((e / K) % rb), leaf = (int)(e / ((size_t)K * rb));
    const int ncur = pmchwt ? 2 : 1, no = pmchwt ? 2 : 1;
    const float2* BJ = Bm + ((size_t)leaf * rb * ncur + (size_t)rr * ncur) * 3 * K;
    float2* EJ = E + ((size_t)leaf * rb + rr) * no * 3 * K;
    auto mulc = [](double2 c, float2 v) {   // complex FP64 constant times FP32 value, rounded to FP32
        return make_float2((float)(c.x * v.x - c.y * v.y), (float)(c.x * v.y + c.y * v.x));
    };
    if (!pmchwt) {
        for (int d = 0; d < 3; ++d) EJ[(size_t)d * K + k] = mulc(ca, BJ[(size_t)d * K + k]);
        return;
    }
    const float2* BM = BJ + 3 * (size_t)K;
    float2* EM = EJ + 3 * (size_t)K;
    const float kx = kp[k], ky = kp[K + k], kz = kp[2 * K + k];
    float2 j[3], m[3];
    for (int d = 0; d < 3; ++d) j[d] = BJ[(size_t)d * K + k], m[d] = BM[(size_t)d * K + k];
    // kp x B
    const float2 cj[3] = {csub(cscale(j[2], ky), cscale(j[1], kz)), csub(cscale(j[0], kz), cscale(j[2], kx)),
                          csub(cscale(j[1], kx), cscale(j[0], ky))};
    const float2 cm[3] = {csub(cscale(m[2], ky), cscale(m[1], kz)), csub(cscale(m[0], kz), cscale(m[2], kx)),
                          csub(cscale(m[1], kx), cscale(m[0], ky))};
    // ca: coefficient of B_J in E_J, cb: of kp x B_M in E_J (negated), cc: of kp x B_J in E_M, cd: of B_M in E_M
    for (int d = 0; d < 3; ++d) {
        const float2 ej = csub(mulc(ca, j[d]), mulc(cb, cm[d]));
        const float2 em = cadd(mulc(cc, cj[d]), mulc(cd, m[d]));
        EJ[(size_t)d * K + k] = ej;
        EM[(size_t)d * K + k] = em;
    }
}

void launch_beff(hipStream_t st, int nleaf, int rb, int K, bool pmchwt, const float2* Bm, const float* kp,
                 double2 ca, double2 cb, double2 cc, double2 cd, float2* E) {
    if (nleaf == 0) return;
    const size_t total = (size_t)nleaf * rb * K;
    k_beff<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(nleaf, rb, K, pmchwt ? 1 : 0, Bm, kp, ca, cb, cc, cd, E);
}

// Y[perm[i] (+N for output 1)][r0 + rr] += sum_w V_w[i] . E_w[leaf][rr][o], rr < rb <= 8
template <int NO>
__global__ __launch_bounds__(256) void k_disagg_mfma(int N, int ldr, int r0, int rb, const int* __restrict__ order,
                                                     const int* __restrict__ leaf_ptr, const int* __restrict__ perm,
                                                     int nw, int K1, const float2* __restrict__ V1,
                                                     const float2* __restrict__ E1, int K2,
                                                     const float2* __restrict__ V2, const float2* __restrict__ E2,
                                                     double2* __restrict__ y) {
    const int leaf = order[blockIdx.x];
    const int i0 = leaf_ptr[leaf], nt = leaf_ptr[leaf + 1] - i0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3, rcol = li & 7;
    const bool rok = rcol < rb;
    const int nrt = (nt + 15) >> 4;
    for (int rp = 0; rp < nrt; rp += 16) {
        const int nrp = min(16, nrt - rp);
        floatx4 acc[4][NO];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int o = 0; o < NO; ++o) acc[s][o] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int w = 0; w < nw; ++w) {
            const int K = w ? K2 : K1;
            const int Q = 3 * K;
            const float2* V = (w ? V2 : V1) + (size_t)i0 * Q;
            const float2* E = (w ? E2 : E1) + (size_t)leaf * rb * NO * Q;
            // UK k-steps per iteration: all loads of the group are issued before its MFMAs
            constexpr int UK = 4;
            for (int q00 = 0; q00 < Q; q00 += 4 * UK) {
                float2 e[UK][NO];
#pragma unroll
                for (int u = 0; u < UK; ++u) {
                    const int q = q00 + 4 * u + lk;
#pragma unroll
                    for (int o = 0; o < NO; ++o)
                        e[u][o] = (q < Q && rok) ? E[((size_t)rcol * NO + o) * Q + q] : make_float2(0.f, 0.f);
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int tile = wave + 4 * s;
                    if (tile >= nrp) continue;
                    const int row = (rp + tile) * 16 + li;
                    float2 v[UK];
#pragma unroll
                    for (int u = 0; u < UK; ++u) {
                        const int q = q00 + 4 * u + lk;
                        v[u] = (q < Q && row < nt) ? V[(size_t)row * Q + q] : make_float2(0.f, 0.f);
                    }
#pragma unroll
                    for (int u = 0; u < UK; ++u)
#pragma unroll
                        for (int o = 0; o < NO; ++o) {
                            const float b1 = part ? e[u][o].y : e[u][o].x;    // Re column: Re E, Im column: Im E
                            const float b2 = part ? e[u][o].x : -e[u][o].y;   // Re column: -Im E, Im column: Re E
                            floatx4 a = acc[s][o];
                            a = __builtin_amdgcn_mfma_f32_16x16x4f32(v[u].x, b1, a, 0, 0, 0);
                            a = __builtin_amdgcn_mfma_f32_16x16x4f32(v[u].y, b2, a, 0, 0, 0);
                            acc[s][o] = a;
                        }
                }
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int tile = wave + 4 * s;
#pragma unroll
            for (int o = 0; o < NO; ++o)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float re = acc[s][o][v];
                    const float im = __shfl_xor(re, 8, 64);
                    const int row = (rp + tile) * 16 + 4 * lk + v;
                    if (tile < nrp && part == 0 && rok && row < nt) {
                        double2* yp = y + ((size_t)perm[i0 + row] + (o ? (size_t)N : 0)) * ldr + r0 + rcol;
                        double2 t = *yp;
                        t.x += re;
                        t.y += im;
                        *yp = t;
                    }
                }
        }
    }
}

void launch_disagg_mfma(hipStream_t st, int nleaf, int N, int ldr, int r0, int rb, const int* order,
                        const int* leaf_ptr, const int* perm, bool pmchwt, int K1, const float2* V1,
                        const float2* E1, int K2, const float2* V2, const float2* E2, double2* y) {
    if (nleaf == 0) return;
    if (!pmchwt)
        k_disagg_mfma<1><<<nleaf, 256, 0, st>>>(N, ldr, r0, rb, order, leaf_ptr, perm, 1, K1, V1, E1, 0, nullptr,
                                                nullptr, y);
    else
        k_disagg_mfma<2><<<nleaf, 256, 0, st>>>(N, ldr, r0, rb, order, leaf_ptr, perm, 2, K1, V1, E1, K2, V2, E2, y);
}

}  // namespace mlfma
