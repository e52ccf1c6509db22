// Batched disaggregation on the matrix cores, FP32 far field
// (v_mfma_f32_16x16x4_f32) or FP64 (v_mfma_f64_16x16x4_f64, e.g. the true
// residuals of a batched solve).  Code/MLFMM.cpp:211-227 (PEC) and :285-351
// (PMCHWT) for rb right-hand sides.
//
// Every output is linear in the finest radiation workspaces, so the constants
// and the k-hat cross products are folded first, per leaf and right-hand side,
// into effective receive-side matrices (k_beff):
//   PEC:     E = c B
//   PMCHWT:  E_J1 = c1 (B_J1 - kp x B_M1),   E_M1 = c1 (kp x B_J1 + B_M1)
//            E_J2 = c2 (etai B_J2 - kp x B_M2), E_M2 = c2 (kp x B_J2 + B_M2 / etai)
// (kp x B reproduces the reference's (-khat x V).B with khat = -kp).  Then, per
// leaf, Y (nt x rb) += sum over wavenumbers V_w (nt x 3K_w) E_w (3K_w x rb), a
// complex GEMM done as two real 16x16x4 matrix-core products on [Re | Im]
// column tiles (same lane layouts as k_near_mfma; the FP64 product interleaves
// its result rows).  Rows go in 16-row tiles, one workgroup per leaf (largest
// first); accumulation in the far-field precision.
#include <algorithm>

#include "cplx.hpp"
#include "kernels.hpp"

namespace mlfma {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {
template <typename T>
struct Mx;
template <>
struct Mx<float> {
    using acc = floatx4;
    __device__ static acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * lk + v; }
    // complex FP64 constant times a value, rounded to FP32
    __device__ static float2 mulc(double2 c, float2 v) {
        return make_float2((float)(c.x * v.x - c.y * v.y), (float)(c.x * v.y + c.y * v.x));
    }
};
template <>
struct Mx<double> {
    using acc = doublex4;
    __device__ static acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return lk + 4 * v; }
    __device__ static double2 mulc(double2 c, double2 v) { return cmul(c, v); }
};
}  // namespace

// E[leaf][rr][o][3][K] from Bm[leaf][rr * ncur + cur][3][K]; one thread per (leaf, rr, k), flat grid.
template <typename T>
__global__ void k_beff(int nleaf, int rb, int K, int pmchwt, const cplx_t<T>* __restrict__ Bm,
                       const T* __restrict__ kp, double2 ca, double2 cb, double2 cc, double2 cd,
                       cplx_t<T>* __restrict__ E) {
    using C2 = cplx_t<T>;
    using M = Mx<T>;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nleaf * rb * K) return;
    const int k = (int)(e % K);
    const int rr = (int)((e / K) % rb), leaf = (int)(e / ((size_t)K * rb));
    const int ncur = pmchwt ? 2 : 1, no = pmchwt ? 2 : 1;
    const C2* BJ = Bm + ((size_t)leaf * rb * ncur + (size_t)rr * ncur) * 3 * K;
    C2* EJ = E + ((size_t)leaf * rb + rr) * no * 3 * K;
    if (!pmchwt) {
        for (int d = 0; d < 3; ++d) EJ[(size_t)d * K + k] = M::mulc(ca, BJ[(size_t)d * K + k]);
        return;
    }
    const C2* BM = BJ + 3 * (size_t)K;
    C2* EM = EJ + 3 * (size_t)K;
    const T kx = kp[k], ky = kp[K + k], kz = kp[2 * K + k];
    C2 j[3], m[3];
    for (int d = 0; d < 3; ++d) j[d] = BJ[(size_t)d * K + k], m[d] = BM[(size_t)d * K + k];
    // kp x B
    const C2 cj[3] = {csub(cscale(j[2], ky), cscale(j[1], kz)), csub(cscale(j[0], kz), cscale(j[2], kx)),
                      csub(cscale(j[1], kx), cscale(j[0], ky))};
    const C2 cm[3] = {csub(cscale(m[2], ky), cscale(m[1], kz)), csub(cscale(m[0], kz), cscale(m[2], kx)),
                      csub(cscale(m[1], kx), cscale(m[0], ky))};
    // ca: coefficient of B_J in E_J, cb: of kp x B_M in E_J (negated), cc: of kp x B_J in E_M, cd: of B_M in E_M
    for (int d = 0; d < 3; ++d) {
        const C2 ej = csub(M::mulc(ca, j[d]), M::mulc(cb, cm[d]));
        const C2 em = cadd(M::mulc(cc, cj[d]), M::mulc(cd, m[d]));
        EJ[(size_t)d * K + k] = ej;
        EM[(size_t)d * K + k] = em;
    }
}

template <typename T>
void launch_beff(hipStream_t st, int nleaf, int rb, int K, bool pmchwt, const cplx_t<T>* Bm, const T* kp,
                 double2 ca, double2 cb, double2 cc, double2 cd, cplx_t<T>* E) {
    if (nleaf == 0) return;
    const size_t total = (size_t)nleaf * rb * K;
    k_beff<T><<<(unsigned)((total + 255) / 256), 256, 0, st>>>(nleaf, rb, K, pmchwt ? 1 : 0, Bm, kp, ca, cb, cc, cd,
                                                              E);
}

// Y[perm[i] (+N for output 1)][r0 + rr] += sum_w V_w[i] . E_w[leaf][rr][o], rr < rb <= 8
template <typename T, int NO>
__global__ __launch_bounds__(256) void k_disagg_mfma(int N, int ldr, int r0, int rb, const int* __restrict__ order,
                                                     const int* __restrict__ leaf_ptr, const int* __restrict__ perm,
                                                     int nw, int K1, const cplx_t<T>* __restrict__ V1,
                                                     const cplx_t<T>* __restrict__ E1, int K2,
                                                     const cplx_t<T>* __restrict__ V2,
                                                     const cplx_t<T>* __restrict__ E2, double2* __restrict__ y) {
    using C2 = cplx_t<T>;
    using M = Mx<T>;
    using acc_t = typename M::acc;
    const int leaf = order[blockIdx.x];
    const int i0 = leaf_ptr[leaf], nt = leaf_ptr[leaf + 1] - i0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3, rcol = li & 7;
    const bool rok = rcol < rb;
    const int nrt = (nt + 15) >> 4;
    for (int rp = 0; rp < nrt; rp += 16) {
        const int nrp = min(16, nrt - rp);
        acc_t acc[4][NO];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int o = 0; o < NO; ++o) acc[s][o] = acc_t{0, 0, 0, 0};
        for (int w = 0; w < nw; ++w) {
            const int K = w ? K2 : K1;
            const int Q = 3 * K;
            const C2* V = (w ? V2 : V1) + (size_t)i0 * Q;
            const C2* E = (w ? E2 : E1) + (size_t)leaf * rb * NO * Q;
            // UK k-steps per iteration: all loads of the group are issued before its products
            constexpr int UK = 4;
            for (int q00 = 0; q00 < Q; q00 += 4 * UK) {
                C2 e[UK][NO];
#pragma unroll
                for (int u = 0; u < UK; ++u) {
                    const int q = q00 + 4 * u + lk;
#pragma unroll
                    for (int o = 0; o < NO; ++o)
                        e[u][o] = (q < Q && rok) ? E[((size_t)rcol * NO + o) * Q + q] : C2{0, 0};
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int tile = wave + 4 * s;
                    if (tile >= nrp) continue;
                    const int row = (rp + tile) * 16 + li;
                    C2 v[UK];
#pragma unroll
                    for (int u = 0; u < UK; ++u) {
                        const int q = q00 + 4 * u + lk;
                        v[u] = (q < Q && row < nt) ? V[(size_t)row * Q + q] : C2{0, 0};
                    }
#pragma unroll
                    for (int u = 0; u < UK; ++u)
#pragma unroll
                        for (int o = 0; o < NO; ++o) {
                            const T b1 = part ? e[u][o].y : e[u][o].x;    // Re column: Re E, Im column: Im E
                            const T b2 = part ? e[u][o].x : -e[u][o].y;   // Re column: -Im E, Im column: Re E
                            acc_t a = acc[s][o];
                            a = M::op(v[u].x, b1, a);
                            a = M::op(v[u].y, b2, a);
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
                    const T re = acc[s][o][v];
                    const T im = __shfl_xor(re, 8, 64);
                    const int row = (rp + tile) * 16 + M::row(lk, v);
                    if (tile < nrp && part == 0 && rok && row < nt) {
                        double2* yp = y + ((size_t)perm[i0 + row] + (o ? (size_t)N : 0)) * ldr + r0 + rcol;
                        double2 t = *yp;
                        t.x += (double)re;
                        t.y += (double)im;
                        *yp = t;
                    }
                }
        }
    }
}

template <typename T>
void launch_disagg_mfma(hipStream_t st, int nleaf, int N, int ldr, int r0, int rb, const int* order,
                        const int* leaf_ptr, const int* perm, bool pmchwt, int K1, const cplx_t<T>* V1,
                        const cplx_t<T>* E1, int K2, const cplx_t<T>* V2, const cplx_t<T>* E2, double2* y) {
    if (nleaf == 0) return;
    if (!pmchwt)
        k_disagg_mfma<T, 1><<<nleaf, 256, 0, st>>>(N, ldr, r0, rb, order, leaf_ptr, perm, 1, K1, V1, E1, 0, nullptr,
                                                   nullptr, y);
    else
        k_disagg_mfma<T, 2><<<nleaf, 256, 0, st>>>(N, ldr, r0, rb, order, leaf_ptr, perm, 2, K1, V1, E1, K2, V2, E2,
                                                   y);
}

// ---------------------------------------------------------------- all right-hand sides at once (FP32)
// The far-field passes write their effective matrices into one block
// Eall[w][leaf][o][q][R] (q = d K + k as in E, right-hand sides fastest,
// pass columns r0 .. r0 + rb - 1), and one product per matvec then reads
// each pattern row once for all R right-hand sides instead of once per pass.
// SPLIT (the split-pair disaggregation): each element is stored as its BF16
// pairs, .x = hi Re | hi Im << 16, .y = lo Re | lo Im << 16 (8 bytes, as the
// FP32 value), so the product stages them without converting.
__device__ __forceinline__ unsigned bf16r(float f) {   // round to nearest even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
template <bool SPLIT>
__device__ __forceinline__ void put_e(float2* E, size_t i, float2 v) {
    if constexpr (SPLIT) {
        const unsigned hr = bf16r(v.x), hi = bf16r(v.y);
        const unsigned lr = bf16r(v.x - __uint_as_float(hr << 16)), li = bf16r(v.y - __uint_as_float(hi << 16));
        reinterpret_cast<uint2*>(E)[i] = make_uint2(hr | (hi << 16), lr | (li << 16));
    } else {
        E[i] = v;
    }
}
template <bool SPLIT>
__global__ void k_beff_all(int nleaf, int rb, int K, int pmchwt, const float2* __restrict__ Bm,
                           const float* __restrict__ kp, double2 ca, double2 cb, double2 cc, double2 cd, int R,
                           int r0, float2* __restrict__ Eall) {
    using M = Mx<float>;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nleaf * rb * K) return;
    const int rr = (int)(e % rb);   // right-hand side fastest: the writes below are contiguous in it
    const int k = (int)((e / rb) % K), leaf = (int)(e / ((size_t)K * rb));
    const int ncur = pmchwt ? 2 : 1, no = pmchwt ? 2 : 1;
    const size_t Q = 3 * (size_t)K;
    const float2* BJ = Bm + ((size_t)leaf * rb * ncur + (size_t)rr * ncur) * Q;
    float2* EJ = Eall + (size_t)leaf * no * Q * R + r0 + rr;   // element (o, q) at EJ[(o Q + q) R]
    if (!pmchwt) {
        for (int d = 0; d < 3; ++d) put_e<SPLIT>(EJ, ((size_t)d * K + k) * R, M::mulc(ca, BJ[(size_t)d * K + k]));
        return;
    }
    const float2* BM = BJ + Q;
    float2* EM = EJ + Q * R;
    const float kx = kp[k], ky = kp[K + k], kz = kp[2 * K + k];
    float2 j[3], m[3];
    for (int d = 0; d < 3; ++d) j[d] = BJ[(size_t)d * K + k], m[d] = BM[(size_t)d * K + k];
    const float2 cj[3] = {csub(cscale(j[2], ky), cscale(j[1], kz)), csub(cscale(j[0], kz), cscale(j[2], kx)),
                          csub(cscale(j[1], kx), cscale(j[0], ky))};
    const float2 cm[3] = {csub(cscale(m[2], ky), cscale(m[1], kz)), csub(cscale(m[0], kz), cscale(m[2], kx)),
                          csub(cscale(m[1], kx), cscale(m[0], ky))};
    for (int d = 0; d < 3; ++d) {
        put_e<SPLIT>(EJ, ((size_t)d * K + k) * R, csub(M::mulc(ca, j[d]), M::mulc(cb, cm[d])));
        put_e<SPLIT>(EM, ((size_t)d * K + k) * R, cadd(M::mulc(cc, cj[d]), M::mulc(cd, m[d])));
    }
}

void launch_beff_all(hipStream_t st, int nleaf, int rb, int K, bool pmchwt, const float2* Bm, const float* kp,
                     double2 ca, double2 cb, double2 cc, double2 cd, int R, int r0, bool split, float2* Eall) {
    if (nleaf == 0) return;
    const size_t total = (size_t)nleaf * rb * K;
    const unsigned grid = (unsigned)((total + 255) / 256);
    if (split)
        k_beff_all<true><<<grid, 256, 0, st>>>(nleaf, rb, K, pmchwt ? 1 : 0, Bm, kp, ca, cb, cc, cd, R, r0, Eall);
    else
        k_beff_all<false><<<grid, 256, 0, st>>>(nleaf, rb, K, pmchwt ? 1 : 0, Bm, kp, ca, cb, cc, cd, R, r0, Eall);
}

// Y[perm[i] (+N for output 1)][r] += sum_w V_w[i] . Eall_w[leaf][o][.][r] for all R.
// The work items of the matrix-core near product (a leaf's range of up to 4
// 16-row tiles, largest first); wavefront w owns row tile w and CT column
// tiles of 8 right-hand sides, the pattern values of the next k-step are
// loaded while the current one's products run (alternating register sets,
// clamped branch-free loads), and Eall is staged in LDS kQ rows at a time,
// the next chunk loaded into registers during the current one's products.
template <int NO, int CT>
__global__ __launch_bounds__(256) void k_disagg_all(NearDev nd, int R, int nw, int K1, const float2* __restrict__ V1,
                                                    const float2* __restrict__ E1, int K2,
                                                    const float2* __restrict__ V2,
                                                    const float2* __restrict__ E2, double2* __restrict__ y) {
    using M = Mx<float>;
    constexpr int RT = 8 * CT;
    constexpr int kQ = 16;
    constexpr int NE = NO * kQ * RT, NU = (NE + 255) / 256;
    __shared__ float2 es[NO][kQ][RT];
    const int2 item = nd.items[blockIdx.x];
    const int t = item.x, rp = item.y;
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    const int nrp = min(4, ((nt + 15) >> 4) - rp);
    const int rb0 = blockIdx.y * RT;
    const int rv = min(RT, R - rb0);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;
    const bool wact = wave < nrp;
    const int arow = min((rp + (wact ? wave : 0)) * 16 + li, nt - 1);
    floatx4 acc[CT][NO];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int o = 0; o < NO; ++o) acc[c][o] = floatx4{0.f, 0.f, 0.f, 0.f};
    float2 pe[NU];
    unsigned pmask = 0;
    for (int w = 0; w < nw; ++w) {
        const int Q = 3 * (w ? K2 : K1);
        const float2* Vr = (w ? V2 : V1) + ((size_t)i0 + arow) * Q;   // this lane's pattern row
        const float2* E = (w ? E2 : E1) + (size_t)t * NO * Q * R;
        auto load_e = [&](int q0) {
            pmask = 0;
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = threadIdx.x + 256 * u, ec = min(e, NE - 1);
                const int o = ec / (kQ * RT), rem = ec - o * (kQ * RT);
                const int qq = rem / RT, r = rem - qq * RT;
                pe[u] = E[((size_t)o * Q + min(q0 + qq, Q - 1)) * R + rb0 + min(r, rv - 1)];
                pmask |= (unsigned)((NE % 256 == 0 || e < NE) & (q0 + qq < Q) & (r < rv)) << u;
            }
        };
        auto step = [&](int kk, const float2& v) {
            if (!wact) return;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int o = 0; o < NO; ++o) {
                    const float2 ev = es[o][kk + lk][ct * 8 + (li & 7)];
                    const float b1 = part ? ev.y : ev.x, b2 = part ? ev.x : -ev.y;
                    floatx4 a = acc[ct][o];
                    a = M::op(v.x, b1, a);
                    a = M::op(v.y, b2, a);
                    acc[ct][o] = a;
                }
        };
        load_e(0);
        float2 va = Vr[min(lk, Q - 1)], vb;
        for (int q0 = 0; q0 < Q; q0 += kQ) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = threadIdx.x + 256 * u;
                if (NE % 256 == 0 || e < NE) {
                    const int o = e / (kQ * RT), rem = e - o * (kQ * RT);
                    const int qq = rem / RT, r = rem - qq * RT;
                    es[o][qq][r] = ((pmask >> u) & 1u) ? pe[u] : make_float2(0.f, 0.f);
                }
            }
            __syncthreads();
            load_e(q0 + kQ);
#pragma unroll
            for (int kk = 0; kk < kQ; kk += 8) {
                vb = Vr[min(q0 + kk + 4 + lk, Q - 1)];
                step(kk, va);
                va = Vr[min(q0 + kk + 8 + lk, Q - 1)];
                step(kk + 4, vb);
            }
        }
    }
    if (!wact) return;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float re = acc[ct][o][v];
                const float im = __shfl_xor(re, 8, 64);
                const int row = (rp + wave) * 16 + M::row(lk, v);
                const int rc = ct * 8 + li;
                if (part == 0 && rc < rv && row < nt) {
                    double2* yp = y + ((size_t)nd.perm[i0 + row] + (o ? (size_t)nd.N : 0)) * R + rb0 + rc;
                    double2 tt = *yp;
                    tt.x += (double)re;
                    tt.y += (double)im;
                    *yp = tt;
                }
            }
}

void launch_disagg_all(hipStream_t st, int R, const NearDev& nd, bool pmchwt, int K1, const float2* V1,
                       const float2* E1, int K2, const float2* V2, const float2* E2, double2* y) {
    if (nd.nitem == 0 || R < 1) return;
    const int ct = R > 32 ? 8 : (R > 16 ? 4 : (R > 8 ? 2 : 1));
    dim3 grid(nd.nitem, (R + 8 * ct - 1) / (8 * ct));
#define DA(NO, CT) k_disagg_all<NO, CT><<<grid, 256, 0, st>>>(nd, R, NO, K1, V1, E1, K2, V2, E2, y)
#define DAN(NO)                                                                                                       \
    if (ct == 8) DA(NO, 8);                                                                                           \
    else if (ct == 4) DA(NO, 4);                                                                                      \
    else if (ct == 2) DA(NO, 2);                                                                                      \
    else DA(NO, 1);
    if (pmchwt) {
        DAN(2)
    } else {
        DAN(1)
    }
#undef DAN
#undef DA
}

// ---------------------------------------------------------------- split-pair BF16 disaggregation
// The same all-RHS product with split-pair BF16 operands (far_precision =
// MLFMA_FP32 with near_precision = MLFMA_BF16): each FP32 pattern value f is
// hi = bf16(f) plus lo = bf16(f - hi), stored as four planes [Re hi, Im hi,
// Re lo, Im lo][N][Qp] (Qp = 3K rounded up to 32, zero-padded), and Eall is
// split the same way when staged; the product is Vh Eh + Vh El + Vl Eh on
// v_mfma_f32_16x16x32_bf16 (lane l holds A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16], i < 8, and D[4 (l / 16) + v][l % 16]), about
// 2^-17 of |V||E| from the FP32 product, FP32 accumulation.
namespace {
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
__device__ __forceinline__ unsigned bf16b(float f) {   // round to nearest even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16f(unsigned b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ uint4 neg8b(uint4 v) {
    v.x ^= 0x80008000u;
    v.y ^= 0x80008000u;
    v.z ^= 0x80008000u;
    v.w ^= 0x80008000u;
    return v;
}
__device__ __forceinline__ floatx4 mfb(uint4 a, uint4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

// Split pairs of 8 consecutive staged values (bit i of mask: value i is
// staged, else zero), packed for 16-byte LDS stores: the BF16 hi parts of
// Re and of Im, then the lo parts (lo = BF16 of what hi leaves of the FP32
// value).
template <typename V>
__device__ __forceinline__ void split8(const V (&v)[8], unsigned mask, uint4 (&o)[4]) {
    unsigned w[4][4] = {};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const bool on = (mask >> i) & 1u;
        const float fr = on ? (float)v[i].x : 0.f, fi = on ? (float)v[i].y : 0.f;
        const unsigned hr = bf16b(fr), hi = bf16b(fi);
        const int sh = 16 * (i & 1);
        w[0][i >> 1] |= hr << sh;
        w[1][i >> 1] |= hi << sh;
        w[2][i >> 1] |= bf16b(fr - bf16f(hr)) << sh;
        w[3][i >> 1] |= bf16b(fi - bf16f(hi)) << sh;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) o[p] = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
}
// The same from values already split: .x = hi Re | hi Im << 16, .y = lo Re |
// lo Im << 16 (as the near product stages X), only regrouped.
__device__ __forceinline__ void split8(const uint2 (&v)[8], unsigned mask, uint4 (&o)[4]) {
    unsigned w[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint2 a = ((mask >> (2 * k)) & 1u) ? v[2 * k] : make_uint2(0u, 0u);
        const uint2 b = ((mask >> (2 * k + 1)) & 1u) ? v[2 * k + 1] : make_uint2(0u, 0u);
        w[0][k] = (a.x & 0xffffu) | (b.x << 16);
        w[1][k] = (a.x >> 16) | (b.x & 0xffff0000u);
        w[2][k] = (a.y & 0xffffu) | (b.y << 16);
        w[3][k] = (a.y >> 16) | (b.y & 0xffff0000u);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) o[p] = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
}
}  // namespace

// planes[p][i][q] from V[i][q] (FP32 complex), q < Qp (zero past Q)
__global__ void k_split_rows(int N, int Q, int Qp, const float2* __restrict__ V, unsigned short* __restrict__ planes) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * Qp) return;
    const size_t i = e / Qp;
    const int q = (int)(e - i * Qp);
    const float2 v = q < Q ? V[i * Q + q] : make_float2(0.f, 0.f);
    const unsigned hr = bf16b(v.x), hi = bf16b(v.y);
    const size_t plane = (size_t)N * Qp;
    planes[e] = (unsigned short)hr;
    planes[plane + e] = (unsigned short)hi;
    planes[2 * plane + e] = (unsigned short)bf16b(v.x - bf16f(hr));
    planes[3 * plane + e] = (unsigned short)bf16b(v.y - bf16f(hi));
}

void launch_split_rows(hipStream_t st, int N, int Q, int Qp, const float2* V, unsigned short* planes) {
    const size_t total = (size_t)N * Qp;
    if (total == 0) return;
    k_split_rows<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(N, Q, Qp, V, planes);
}

template <int NO, int CT>
__global__ __launch_bounds__(256) void k_disagg_bf3(NearDev nd, int R, int nw, int K1,
                                                    const unsigned short* __restrict__ P1,
                                                    const uint2* __restrict__ E1, int K2,
                                                    const unsigned short* __restrict__ P2,
                                                    const uint2* __restrict__ E2, double2* __restrict__ y) {
    constexpr int RT = 8 * CT;
    constexpr int kQ = 32, kP = kQ + 8;   // staged q per chunk, padded LDS row (80 B)
    // staging units: a thread loads 8 consecutive q of one (output, rhs) and
    // stores their split pairs with four 16-byte writes, which 8 consecutive
    // lanes (consecutive rhs, rows 80 B apart) spread over all banks
    constexpr int NUS = NO * (kQ / 8) * RT;
    static_assert(NUS <= 256, "one staging unit per thread");
    // [hi, lo][output][Re, -Im, Im][rhs][q]: a lane of the real-part columns
    // (part 0) reads planes 0 and 1 as its two B operands, of the imaginary
    // part (part 1) planes 2 and 0, so the products need no sign flips
    __shared__ __align__(16) unsigned short es[2][NO][3][RT][kP];
    // chunks of 4 items (mostly the row ranges of one leaf, which read the
    // same effective matrices) with their ny right-hand-side groups (which
    // read the same pattern planes) run together on one XCD
    const int ny = (R + RT - 1) / RT;
    int it, gy;
    xcd_chunk(blockIdx.x, 4, ny, it, gy);
    if (it >= nd.nitem) return;
    const int2 item = nd.items[it];
    const int t = item.x, rp = item.y;
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    const int nrp = min(4, ((nt + 15) >> 4) - rp);
    const int rb0 = gy * RT;
    const int rv = min(RT, R - rb0);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;
    const bool wact = wave < nrp;
    const int arow = min((rp + (wact ? wave : 0)) * 16 + li, nt - 1);
    floatx4 acc[CT][NO];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int o = 0; o < NO; ++o) acc[c][o] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bool sact = NUS == 256 || (int)threadIdx.x < NUS;
    const int su = min((int)threadIdx.x, NUS - 1);   // spare threads repeat the last unit (not stored)
    const int sr = su % RT, sg = (su / RT) % (kQ / 8), so = su / (RT * (kQ / 8));
    uint2 pe[8];   // BF16 pairs as the passes stored them
    unsigned pmask = 0;
    for (int w = 0; w < nw; ++w) {
        const int Q = 3 * (w ? K2 : K1), Qp = (Q + 31) & ~31;
        const unsigned short* Pl = w ? P2 : P1;
        const size_t plane = (size_t)nd.N * Qp;
        const unsigned short* arow_p = Pl + ((size_t)i0 + arow) * Qp + 8 * lk;   // this lane's 8 q of plane 0
        const uint2* E = (w ? E2 : E1) + (size_t)t * NO * Q * R;
        auto load_a = [&](int q0, uint4* a) {
            const int qc = min(q0, Qp - kQ);
#pragma unroll
            for (int p = 0; p < 4; ++p) a[p] = *reinterpret_cast<const uint4*>(arow_p + p * plane + qc);
        };
        auto load_e = [&](int q0) {
            pmask = 0;
            const uint2* Er = E + (size_t)so * Q * R + rb0 + min(sr, rv - 1);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int q = q0 + 8 * sg + i;
                pe[i] = Er[(size_t)min(q, Q - 1) * R];
                pmask |= (unsigned)((q < Q) & (sr < rv)) << i;
            }
        };
        // one 32-q chunk: stage Eall's (split when stored), then the next
        // chunk's loads, then the products from the current patterns (cur);
        // chunks alternate between two register sets, so nothing is copied
        // and the next patterns stay in flight during the products
        auto chunk = [&](int q0, const uint4* cur, uint4* nxt) {
            __syncthreads();
            if (sact) {
                uint4 sp[4];
                split8(pe, pmask, sp);
                *reinterpret_cast<uint4*>(&es[0][so][0][sr][8 * sg]) = sp[0];
                *reinterpret_cast<uint4*>(&es[0][so][1][sr][8 * sg]) = neg8b(sp[1]);
                *reinterpret_cast<uint4*>(&es[0][so][2][sr][8 * sg]) = sp[1];
                *reinterpret_cast<uint4*>(&es[1][so][0][sr][8 * sg]) = sp[2];
                *reinterpret_cast<uint4*>(&es[1][so][1][sr][8 * sg]) = neg8b(sp[3]);
                *reinterpret_cast<uint4*>(&es[1][so][2][sr][8 * sg]) = sp[3];
            }
            __syncthreads();
            load_e(q0 + kQ);
            load_a(q0 + kQ, nxt);
            if (!wact) return;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int rc = ct * 8 + (li & 7);
#pragma unroll
                for (int o = 0; o < NO; ++o) {
                    floatx4 d = acc[ct][o];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {   // E part: hi, then lo (lo meets only the hi patterns)
                        const uint4 b1 = *reinterpret_cast<const uint4*>(&es[h][o][2 * part][rc][8 * lk]);
                        const uint4 b2 = *reinterpret_cast<const uint4*>(&es[h][o][1 - part][rc][8 * lk]);
                        d = mfb(cur[0], b1, d);
                        d = mfb(cur[1], b2, d);
                        if (h == 0) {
                            d = mfb(cur[2], b1, d);
                            d = mfb(cur[3], b2, d);
                        }
                    }
                    acc[ct][o] = d;
                }
            }
        };
        uint4 aa[4], ab[4];
        load_e(0);
        load_a(0, aa);
        // chunks in pairs (a chunk past Qp reads clamped patterns and meets zero-staged Eall)
        for (int q0 = 0; q0 < Qp; q0 += 2 * kQ) {
            chunk(q0, aa, ab);
            chunk(q0 + kQ, ab, aa);
        }
    }
    if (!wact) return;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float re = acc[ct][o][v];
                const float im = __shfl_xor(re, 8, 64);
                const int row = (rp + wave) * 16 + 4 * lk + v;
                const int rc = ct * 8 + li;
                if (part == 0 && rc < rv && row < nt) {
                    double2* yp = y + ((size_t)nd.perm[i0 + row] + (o ? (size_t)nd.N : 0)) * R + rb0 + rc;
                    double2 tt = *yp;
                    tt.x += (double)re;
                    tt.y += (double)im;
                    *yp = tt;
                }
            }
}

void launch_disagg_bf3(hipStream_t st, int R, const NearDev& nd, bool pmchwt, int K1, const unsigned short* P1,
                       const uint2* E1, int K2, const unsigned short* P2, const uint2* E2, double2* y) {
    if (nd.nitem == 0 || R < 1) return;
    const int ct = R > 16 ? 4 : (R > 8 ? 2 : 1);
    const int ny = (R + 8 * ct - 1) / (8 * ct);
    const unsigned grid = xcd_chunk_grid(nd.nitem, 4, ny);
#define DB(NO, CT) k_disagg_bf3<NO, CT><<<grid, 256, 0, st>>>(nd, R, NO, K1, P1, E1, K2, P2, E2, y)
#define DBN(NO)                                                                                                       \
    if (ct == 4) DB(NO, 4);                                                                                           \
    else if (ct == 2) DB(NO, 2);                                                                                      \
    else DB(NO, 1);
    if (pmchwt) {
        DBN(2)
    } else {
        DBN(1)
    }
#undef DBN
#undef DB
}

// ---------------------------------------------------------------- aggregation, all right-hand sides (FP32)
// Code/ProcessMLFMM.h:33-56 (mexp) for all R right-hand sides in one product:
// per leaf, S (3K x ncur R) = V^T (3K x nt) X (nt x ncur R) on the matrix
// cores, so each pattern row is read once per matvec.  Column c is current
// c % ncur of right-hand side c / ncur.  The result goes straight into the
// leaf workspaces of every far-field pass: pass (r0, rb) (the split of R into
// passes of the batched matvec, rb the largest power of two <= cap that
// fits) at Sall + nleaf ncur r0 Q, element (leaf, rr ncur + cur, q) as in a
// single pass's workspace.  A workgroup owns one leaf and 4 16-row q tiles
// (one per wavefront) and CT column tiles; X is staged in LDS kJ rows at a
// time, the next chunk in flight during the current one's products.
template <int NCUR, int CT>
__global__ __launch_bounds__(256) void k_aggregate_all(int nleaf, int K, int N, int R, int cap,
                                                       const int* __restrict__ leaf_ptr,
                                                       const int* __restrict__ perm, const float2* __restrict__ Vs,
                                                       const double2* __restrict__ x, float2* __restrict__ Sall) {
    using M = Mx<float>;
    constexpr int CC = 8 * CT;   // columns per workgroup
    constexpr int kJ = 16;
    constexpr int NE = kJ * CC, NU = (NE + 255) / 256;
    __shared__ float2 xs[kJ][CC];
    const int Q = 3 * K, QG = (Q + 63) / 64;
    // the sample groups of one leaf (which read the same X rows) with their
    // column groups (which read the same pattern rows) run together on one XCD
    const int ncol = NCUR * R, ncg = (ncol + CC - 1) / CC;
    int lq, cg;
    xcd_chunk(blockIdx.x, QG, ncg, lq, cg);
    if (lq >= nleaf * QG) return;
    const int leaf = lq / QG, qg = lq - leaf * QG;
    const int i0 = leaf_ptr[leaf], nt = leaf_ptr[leaf + 1] - i0;
    const int cb0 = cg * CC;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;
    const int q = (qg * 4 + wave) * 16 + li;   // this lane's row of the A operand (V^T)
    const int qc = min(q, Q - 1);
    floatx4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (nt > 0) {
        float2 px[NU];
        unsigned pmask = 0;
        auto load_x = [&](int j0) {
            pmask = 0;
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = threadIdx.x + 256 * u, ec = min(e, NE - 1);
                const int jj = ec / CC, cc = ec - jj * CC;
                const int c = min(cb0 + cc, ncol - 1);
                const int r = c / NCUR, cur = c - r * NCUR;
                const double2 v = x[((size_t)cur * N + perm[i0 + min(j0 + jj, nt - 1)]) * R + r];
                px[u] = make_float2((float)v.x, (float)v.y);
                pmask |= (unsigned)((NE % 256 == 0 || e < NE) & (j0 + jj < nt) & (cb0 + cc < ncol)) << u;
            }
        };
        const float2* Vq = Vs + (size_t)i0 * Q + qc;   // element (j, q) at Vq[j Q]
        auto vload = [&](int j) { return Vq[(size_t)min(j, nt - 1) * Q]; };
        auto step = [&](int kk, const float2& v) {
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const float2 xv = xs[kk + lk][ct * 8 + (li & 7)];
                const float b1 = part ? xv.y : xv.x, b2 = part ? xv.x : -xv.y;
                floatx4 a = acc[ct];
                a = M::op(v.x, b1, a);
                a = M::op(v.y, b2, a);
                acc[ct] = a;
            }
        };
        load_x(0);
        float2 va = vload(lk), vb;
        for (int j0 = 0; j0 < nt; j0 += kJ) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = threadIdx.x + 256 * u;
                if (NE % 256 == 0 || e < NE) {
                    const int jj = e / CC, cc = e - jj * CC;
                    xs[jj][cc] = ((pmask >> u) & 1u) ? px[u] : make_float2(0.f, 0.f);
                }
            }
            __syncthreads();
            load_x(j0 + kJ);
#pragma unroll
            for (int kk = 0; kk < kJ; kk += 8) {
                vb = vload(j0 + kk + 4 + lk);
                step(kk, va);
                va = vload(j0 + kk + 8 + lk);
                step(kk + 4, vb);
            }
        }
    }
    // D[4 lk + v][li]: lanes li < 8 hold Re of column ct * 8 + li, lane li + 8 its Im
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int c = cb0 + ct * 8 + li;
        const int r = c / NCUR, cur = c - r * NCUR;
        int p0 = 0, rb = 1;   // the pass holding right-hand side r
        for (;;) {
            rb = 1;
            while (rb * 2 <= cap && p0 + rb * 2 <= R) rb *= 2;
            if (r < p0 + rb || p0 + rb >= R) break;
            p0 += rb;
        }
        float2* S = Sall + (size_t)nleaf * NCUR * p0 * Q + ((size_t)leaf * rb + (r - p0)) * NCUR * Q +
                    (size_t)cur * Q;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float re = acc[ct][v];
            const float im = __shfl_xor(re, 8, 64);
            const int qq = (qg * 4 + wave) * 16 + M::row(lk, v);
            if (part == 0 && c < ncol && qq < Q) S[qq] = make_float2(re, im);
        }
    }
}

void launch_aggregate_all(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, const int* leaf_ptr,
                          const int* perm, const float2* Vs, const double2* x, float2* Sall) {
    if (nleaf == 0 || R < 1) return;
    const int ncol = ncur * R;
    const int ct = ncol > 32 ? 8 : (ncol > 16 ? 4 : (ncol > 8 ? 2 : 1));
    const int QG = (3 * K + 63) / 64, ncg = (ncol + 8 * ct - 1) / (8 * ct);
    const unsigned grid = xcd_chunk_grid(nleaf * QG, QG, ncg);
#define AA(NC, CT) k_aggregate_all<NC, CT><<<grid, 256, 0, st>>>(nleaf, K, N, R, cap, leaf_ptr, perm, Vs, x, Sall)
#define AAN(NC)                                                                                                       \
    if (ct == 8) AA(NC, 8);                                                                                           \
    else if (ct == 4) AA(NC, 4);                                                                                      \
    else if (ct == 2) AA(NC, 2);                                                                                      \
    else AA(NC, 1);
    if (ncur == 2) {
        AAN(2)
    } else {
        AAN(1)
    }
#undef AAN
#undef AA
}

// ---------------------------------------------------------------- aggregation, split-pair BF16
// The same all-RHS product S (3K x ncur R) = V^T (3K x nt) X (nt x ncur R) per
// leaf on the BF16 matrix cores from split pairs (near_precision = MLFMA_BF16,
// FP32 far field): V = Vh + Vl from transposed planes split once at set-up,
// X = Xh + Xl split while it is staged, S = Vh Xh + Vh Xl + Vl Xh (about 2^-17
// of |V||X| from the FP32 product), FP32 accumulation.  The planes are
// [hi Re, hi Im, lo Re, lo Im][Qp][Np]: row q of every pattern, the RWGs of a
// leaf from a 32-aligned column joff[leaf] on, zero past the leaf and past Q, so
// a lane's A fragment (8 consecutive RWGs of one sample) is one 16-byte load.
// A workgroup owns one leaf, 64 samples (a 16-row tile per wavefront) and CT
// column tiles of 8 complex columns, columns current-major (c = cur R + r, so
// a staged column group reads consecutive right-hand sides); X is staged 32
// RWGs at a time, the next block's X and patterns loaded during the current
// block's products.
__global__ void k_split_cols(int Q, int Qp, int Np, const int* __restrict__ jmap, const float2* __restrict__ V,
                             unsigned short* __restrict__ planes) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Qp * Np) return;
    const int q = (int)(e / Np), jp = (int)(e - (size_t)q * Np);
    const int i = jmap[jp];
    const float2 v = (i >= 0 && q < Q) ? V[(size_t)i * Q + q] : make_float2(0.f, 0.f);
    const unsigned hr = bf16b(v.x), hi = bf16b(v.y);
    const size_t plane = (size_t)Qp * Np;
    planes[e] = (unsigned short)hr;
    planes[plane + e] = (unsigned short)hi;
    planes[2 * plane + e] = (unsigned short)bf16b(v.x - bf16f(hr));
    planes[3 * plane + e] = (unsigned short)bf16b(v.y - bf16f(hi));
}

void launch_split_cols(hipStream_t st, int Q, int Qp, int Np, const int* jmap, const float2* V,
                       unsigned short* planes) {
    const size_t total = (size_t)Qp * Np;
    if (total == 0) return;
    k_split_cols<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(Q, Qp, Np, jmap, V, planes);
}

template <int NCUR, int CT, typename TX>
__global__ __launch_bounds__(256) void k_aggregate_bf3(int nleaf, int K, int N, int R, int cap, int Qp, int Np,
                                                       const int* __restrict__ leaf_ptr,
                                                       const int* __restrict__ joff, const int* __restrict__ perm,
                                                       const unsigned short* __restrict__ planes,
                                                       const TX* __restrict__ x, float2* __restrict__ Sall) {
    constexpr int CC = 8 * CT;            // complex columns per workgroup
    constexpr int kJ = 32, kP = kJ + 8;   // staged RWGs per block, padded LDS row (80 B)
    // staging units: a thread loads 8 consecutive RWGs of one column and
    // stores their split pairs with four 16-byte writes, which 8 consecutive
    // lanes (consecutive columns, rows 80 B apart) spread over all banks
    constexpr int NUS = (kJ / 8) * CC;
    static_assert(NUS <= 256, "one staging unit per thread");
    // [hi, lo][Re, -Im, Im][column][RWG]: a lane of the real-part columns
    // (part 0) reads planes 0 and 1 as its two B operands, of the imaginary
    // part (part 1) planes 2 and 0, so the products need no sign flips
    __shared__ __align__(16) unsigned short xs[2][3][CC][kP];
    const int Q = 3 * K, QG = Qp / 64;
    const int ncol = NCUR * R, ncg = (ncol + CC - 1) / CC;
    int lq, cg;
    xcd_chunk(blockIdx.x, QG, ncg, lq, cg);
    if (lq >= nleaf * QG) return;
    const int leaf = lq / QG, qg = lq - leaf * QG;
    const int i0 = leaf_ptr[leaf], nt = leaf_ptr[leaf + 1] - i0;
    const int nkb = (nt + kJ - 1) / kJ;
    const int cb0 = cg * CC;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;
    const size_t plane = (size_t)Qp * Np;
    const unsigned short* arow = planes + (size_t)((qg * 4 + wave) * 16 + li) * Np + joff[leaf] + 8 * lk;
    floatx4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (nt > 0) {
        // patterns of block kb, clamped to the last block (a block past nkb
        // meets zero-staged X), without branches
        auto load_a = [&](int kb, uint4* a) {
            const int kc = min(kb, nkb - 1) * kJ;
#pragma unroll
            for (int p = 0; p < 4; ++p) a[p] = *reinterpret_cast<const uint4*>(arow + p * plane + kc);
        };
        const bool sact = NUS == 256 || (int)threadIdx.x < NUS;
        const int su = min((int)threadIdx.x, NUS - 1);   // spare threads repeat the last unit (not stored)
        const int sc = su % CC, sg = su / CC;
        const int scol = min(cb0 + sc, ncol - 1);
        const int scur = scol / R;
        const TX* xc = x + (size_t)scur * N * R + (scol - scur * R);
        const bool scok = cb0 + sc < ncol;
        TX px[8];
        int pidx[8];   // unknowns of the block whose X is loaded next (read one block earlier)
        unsigned pmask = 0;
        auto load_perm = [&](int kb) {
#pragma unroll
            for (int i = 0; i < 8; ++i) pidx[i] = perm[i0 + min(kb * kJ + 8 * sg + i, nt - 1)];
        };
        auto load_x = [&](int kb) {
            pmask = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int jg = kb * kJ + 8 * sg + i;
                px[i] = xc[(size_t)pidx[i] * R];
                pmask |= (unsigned)((jg < nt) & scok) << i;
            }
        };
        auto store_x = [&] {
            if (!sact) return;
            uint4 sp[4];
            split8(px, pmask, sp);
            *reinterpret_cast<uint4*>(&xs[0][0][sc][8 * sg]) = sp[0];
            *reinterpret_cast<uint4*>(&xs[0][1][sc][8 * sg]) = neg8b(sp[1]);
            *reinterpret_cast<uint4*>(&xs[0][2][sc][8 * sg]) = sp[1];
            *reinterpret_cast<uint4*>(&xs[1][0][sc][8 * sg]) = sp[2];
            *reinterpret_cast<uint4*>(&xs[1][1][sc][8 * sg]) = neg8b(sp[3]);
            *reinterpret_cast<uint4*>(&xs[1][2][sc][8 * sg]) = sp[3];
        };
        // one block: stage its X, then the next block's loads, then the
        // products from the current patterns (cur); blocks alternate between
        // two register sets, so nothing is copied
        auto block = [&](int kb, const uint4* cur, uint4* nxt) {
            __syncthreads();
            store_x();
            __syncthreads();
            load_x(kb + 1);
            load_perm(kb + 2);
            load_a(kb + 1, nxt);
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int rc = ct * 8 + (li & 7);
                floatx4 d = acc[ct];
#pragma unroll
                for (int h = 0; h < 2; ++h) {   // X part: hi, then lo (lo meets only the hi patterns)
                    const uint4 b1 = *reinterpret_cast<const uint4*>(&xs[h][2 * part][rc][8 * lk]);   // Re X | Im X
                    const uint4 b2 = *reinterpret_cast<const uint4*>(&xs[h][1 - part][rc][8 * lk]);   // -Im X | Re X
                    d = mfb(cur[0], b1, d);
                    d = mfb(cur[1], b2, d);
                    if (h == 0) {
                        d = mfb(cur[2], b1, d);
                        d = mfb(cur[3], b2, d);
                    }
                }
                acc[ct] = d;
            }
        };
        uint4 aa[4], ab[4];
        load_perm(0);
        load_x(0);
        load_perm(1);
        load_a(0, aa);
        for (int kb = 0; kb < nkb; kb += 2) {
            block(kb, aa, ab);
            block(kb + 1, ab, aa);
        }
    }
    // D[4 lk + v][li]: lanes li < 8 hold Re of column ct * 8 + li, lane li + 8 its Im
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int c = cb0 + ct * 8 + li;
        const int cur = c / R, r = c - cur * R;
        int p0 = 0, rb = 1;   // the pass holding right-hand side r
        for (;;) {
            rb = 1;
            while (rb * 2 <= cap && p0 + rb * 2 <= R) rb *= 2;
            if (r < p0 + rb || p0 + rb >= R) break;
            p0 += rb;
        }
        float2* S = Sall + (size_t)nleaf * NCUR * p0 * Q + ((size_t)leaf * rb + (r - p0)) * NCUR * Q +
                    (size_t)cur * Q;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float re = acc[ct][v];
            const float im = __shfl_xor(re, 8, 64);
            const int qq = (qg * 4 + wave) * 16 + 4 * lk + v;
            if (part == 0 && c < ncol && qq < Q) S[qq] = make_float2(re, im);
        }
    }
}

template <typename TX>
static void aggregate_bf3_t(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, int Qp, int Np,
                            const int* leaf_ptr, const int* joff, const int* perm, const unsigned short* planes,
                            const TX* x, float2* Sall) {
    if (nleaf == 0 || R < 1) return;
    const int ncol = ncur * R;
    const int ct = ncol > 32 ? 8 : (ncol > 16 ? 4 : (ncol > 8 ? 2 : 1));
    const int QG = Qp / 64, ncg = (ncol + 8 * ct - 1) / (8 * ct);
    const unsigned grid = xcd_chunk_grid(nleaf * QG, QG, ncg);
#define AB(NC, CT)                                                                                                    \
    k_aggregate_bf3<NC, CT, TX><<<grid, 256, 0, st>>>(nleaf, K, N, R, cap, Qp, Np, leaf_ptr, joff, perm, planes, x, \
                                                      Sall)
#define ABN(NC)                                                                                                       \
    if (ct == 8) AB(NC, 8);                                                                                           \
    else if (ct == 4) AB(NC, 4);                                                                                      \
    else if (ct == 2) AB(NC, 2);                                                                                      \
    else AB(NC, 1);
    if (ncur == 1) {
        ABN(1)
    } else {
        ABN(2)
    }
#undef ABN
#undef AB
}

void launch_aggregate_bf3(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, int Qp, int Np,
                          const int* leaf_ptr, const int* joff, const int* perm, const unsigned short* planes,
                          const double2* x, float2* Sall) {
    aggregate_bf3_t(st, ncur, nleaf, K, N, R, cap, Qp, Np, leaf_ptr, joff, perm, planes, x, Sall);
}

void launch_aggregate_bf3(hipStream_t st, int ncur, int nleaf, int K, int N, int R, int cap, int Qp, int Np,
                          const int* leaf_ptr, const int* joff, const int* perm, const unsigned short* planes,
                          const uint2* xsp, float2* Sall) {
    aggregate_bf3_t(st, ncur, nleaf, K, N, R, cap, Qp, Np, leaf_ptr, joff, perm, planes, xsp, Sall);
}

#define MLFMA_DISAGG_MFMA_INSTANTIATE(T)                                                                            \
    template void launch_beff<T>(hipStream_t, int, int, int, bool, const cplx_t<T>*, const T*, double2, double2,   \
                                 double2, double2, cplx_t<T>*);                                                     \
    template void launch_disagg_mfma<T>(hipStream_t, int, int, int, int, int, const int*, const int*, const int*,  \
                                        bool, int, const cplx_t<T>*, const cplx_t<T>*, int, const cplx_t<T>*,      \
                                        const cplx_t<T>*, double2*);
MLFMA_DISAGG_MFMA_INSTANTIATE(float)
MLFMA_DISAGG_MFMA_INSTANTIATE(double)
#undef MLFMA_DISAGG_MFMA_INSTANTIATE

}  // namespace mlfma
