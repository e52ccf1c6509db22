// Near-field product of the polynomial preconditioner in BF16 on the matrix
// cores (precond_precision = MLFMA_BF16): for every target leaf t,
// Y_t (nt x R) = Z_t (nt x C_t) X_t (C_t x R), entries and X rounded to BF16,
// FP32 accumulation.  Same products as Code/MLFMM.cpp:228-231 (PEC) and
// :354-366 (PMCHWT: Y_J = Z11 X_J - ZK X_M, Y_M = ZK X_J + Z22 X_M); the
// reference has no preconditioner, and only the preconditioner uses this copy
// (the solver's operator and residuals keep their own precision).
//
// v_mfma_f32_16x16x32_bf16: D(16x16) += A(16x32) B(32x16); lane l holds
// A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16], i < 8, and
// D[4 (l / 16) + v][l % 16], v < 4.  The complex product uses the same
// [Re Y (8) | Im Y (8)] column split as the FP32 kernel:
//   D += Re Z . [Re X | Im X] + Im Z . [-Im X | Re X].
//
// Storage: per target, per 32-column block kb, 16-row tile rt and plane p, a
// 2 KB record holding the A fragments of the 64 lanes for Re Z, then for Im Z
// (8 BF16 = 16 B per lane), so a wavefront reads each record with one
// coalesced 16-byte load per lane.  Rows past nt and columns past C_t are zero.
#include <type_traits>
#include <vector>

#include "cplx.hpp"
#include "kernels.hpp"

namespace mlfma {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {
constexpr int kRTC = 4;        // row tiles per work item (the items of NearDev)
constexpr int kKC = 64;        // staged columns per chunk (two 32-column blocks)
constexpr int kKP = kKC + 8;   // padded LDS row (144 B, keeps 16-byte alignment)

// round to nearest even (finite inputs)
__device__ __forceinline__ unsigned bf16_bits(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ uint4 neg8(uint4 v) {   // flips the sign of 8 packed BF16 values
    v.x ^= 0x80008000u;
    v.y ^= 0x80008000u;
    v.z ^= 0x80008000u;
    v.w ^= 0x80008000u;
    return v;
}

__device__ __forceinline__ floatx4 mfma(uint4 a, uint4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}
}  // namespace

__device__ __forceinline__ float bf16_float(unsigned b) { return __uint_as_float(b << 16); }

// One workgroup per target leaf: writes its records from the FP64 entries.
// S (split pairs): each record also holds the BF16 rounding of what the first
// BF16 value leaves of the FP32 entry, hi = bf16(f), lo = bf16(f - hi), so
// hi + lo carries about 16 significant bits; records are then 4 KB:
// [hi Re][hi Im][lo Re][lo Im], 64 lanes x 16 B each.
template <bool S>
__global__ void k_pack_bf16(NearDev nd, const double2* __restrict__ Z, const long long* __restrict__ bf_off,
                            uint4* __restrict__ out) {
    constexpr int W = S ? 256 : 128;   // uint4 per record
    const int t = blockIdx.x;
    const int nt = nd.leaf_ptr[t + 1] - nd.leaf_ptr[t];
    if (nt == 0) return;
    const int C = nd.tgt_cols[t];
    const int nrt = (nt + 15) >> 4, nkb = (C + 31) >> 5;
    const size_t plane = (size_t)C * nt;
    const double2* Zt = Z + nd.tgt_off[t];
    uint4* o = out + bf_off[t];
    const int units = nkb * nrt * nd.planes * 64;
    for (int u = threadIdx.x; u < units; u += blockDim.x) {
        const int lane = u & 63, rec = u >> 6;   // rec = (kb * nrt + rt) * planes + p
        const int p = rec % nd.planes, rt = (rec / nd.planes) % nrt, kb = rec / (nd.planes * nrt);
        const int row = rt * 16 + (lane & 15), cb = kb * 32 + 8 * (lane >> 4);
        unsigned re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0}, rl[4] = {0, 0, 0, 0}, il[4] = {0, 0, 0, 0};
        for (int i = 0; i < 8; ++i) {
            const int col = cb + i;
            double2 z = make_double2(0.0, 0.0);
            if (row < nt && col < C) z = Zt[p * plane + (size_t)col * nt + row];
            const float fr = (float)z.x, fi = (float)z.y;
            const unsigned hr = bf16_bits(fr), hi = bf16_bits(fi);
            re[i >> 1] |= hr << (16 * (i & 1));
            im[i >> 1] |= hi << (16 * (i & 1));
            if (S) {
                rl[i >> 1] |= bf16_bits(fr - bf16_float(hr)) << (16 * (i & 1));
                il[i >> 1] |= bf16_bits(fi - bf16_float(hi)) << (16 * (i & 1));
            }
        }
        o[(size_t)rec * W + lane] = make_uint4(re[0], re[1], re[2], re[3]);
        o[(size_t)rec * W + 64 + lane] = make_uint4(im[0], im[1], im[2], im[3]);
        if (S) {
            o[(size_t)rec * W + 128 + lane] = make_uint4(rl[0], rl[1], rl[2], rl[3]);
            o[(size_t)rec * W + 192 + lane] = make_uint4(il[0], il[1], il[2], il[3]);
        }
    }
}

// xb[u * R + r] = (BF16 Re, BF16 Im) of x[u * ldr + r0 + r]; u < nu, r < R
__global__ void k_x_bf16(size_t total, int R, int ldr, int r0, const double2* __restrict__ x, unsigned* __restrict__ xb) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const size_t u = e / R;
    const int r = (int)(e - u * R);
    const double2 v = x[u * ldr + r0 + r];
    xb[e] = bf16_bits((float)v.x) | (bf16_bits((float)v.y) << 16);
}

// split pairs of x (FP64 or FP32 input): .x = (hi Re, hi Im), .y = (lo Re, lo
// Im) of the FP32 value
template <typename TX>
__global__ void k_x_bf16x2(size_t total, int R, int ldr, int r0, const TX* __restrict__ x, uint2* __restrict__ xb) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const size_t u = e / R;
    const int r = (int)(e - u * R);
    const TX v = x[u * ldr + r0 + r];
    const float fr = (float)v.x, fi = (float)v.y;
    const unsigned hr = bf16_bits(fr), hi = bf16_bits(fi);
    xb[e] = make_uint2(hr | (hi << 16), bf16_bits(fr - bf16_float(hr)) | (bf16_bits(fi - bf16_float(hi)) << 16));
}

// Polynomial preconditioner in FP32 vectors around the single-BF16 products
// (t_0 = D^-1 x, y = a_0 t_0, then per product t_k = t_{k-1} - D^-1 Z t_{k-1},
// y += a_k t_k): the first step, which also rounds t_0 to BF16 for the first
// product.  x, t, y: nu x R, right-hand side fastest.
__global__ void k_poly_init_bf16(size_t total, int R, const float2* __restrict__ dinv, const float2* __restrict__ x,
                                 float2 a0, float2* __restrict__ t, float2* __restrict__ y, unsigned* __restrict__ xb) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const float2 v = cmul(dinv[e / R], x[e]);
    t[e] = v;
    y[e] = cmul(a0, v);
    xb[e] = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
}

// A workgroup owns one work item (a target leaf's range of up to 4 16-row
// tiles) and CT <= 4 column tiles of 8 right-hand sides (at 64 right-hand
// sides, two workgroups of 32 with 4 waves per SIMD measured faster than one
// of 64 with 3; the two share the item's records through L2).  Wavefront w owns row tile w of the item and every column
// tile, so each record is read from memory once per workgroup, by one
// wavefront, for up to 32 right-hand sides; the next
// 32-column block's records are loaded while the current block's products
// run.  X is staged in LDS 64 columns at a time as BF16, [plane][rhs][column]
// rows of 144 B (planes Re, -Im, Im, and -Re for X_M): a thread gathers 8
// consecutive columns of one right-hand side (each load coalesced over the
// wavefront's right-hand sides) and writes them with 16-byte stores, which 8
// consecutive lanes spread over all 32 banks.
//
// S (split pairs, the operator's batched products with near_precision =
// MLFMA_BF16): Z = Zh + Zl and X = Xh + Xl in BF16, and each product is
// Zh Xh + Zh Xl + Zl Xh (the dropped Zl Xl and the representation error are
// about 2^-17 of |Z||X|), three matrix-core products per record.
//
// F (single BF16, the polynomial preconditioner's products in FP32 vectors):
// instead of storing Z X, the epilogue takes the polynomial's step for its
// rows, t -= D^-1 (Z X), y += a t, and rounds the new t to BF16 into the
// next product's input (pe.xbn, a buffer other than the one read here, since
// other workgroups still gather from it); every (row, right-hand side) has one
// owner, so t and y are updated without races.
template <int P, int CT, bool S, bool F>
__global__ __launch_bounds__(256) void k_near_bf16(NearDev nd, int R, int ldr, int r0, const uint4* __restrict__ Zb,
                                                   const long long* __restrict__ bf_off,
                                                   const void* __restrict__ xbv, double2* __restrict__ y,
                                                   PolyEpi pe) {
    constexpr int NX = (P == 3) ? 2 : 1;   // staged currents
    constexpr int NO = (P == 3) ? 2 : 1;   // outputs
    constexpr int RT = 8 * CT;             // right-hand sides per workgroup
    constexpr int NS = S ? 2 : 1;          // staged X parts (hi, lo)
    constexpr int W = S ? 256 : 128;       // uint4 per record
    using xb_t = typename std::conditional<S, uint2, unsigned>::type;
    const xb_t* __restrict__ xb = static_cast<const xb_t*>(xbv);
    // [part][current][plane][rhs][column], planes Re, -Im, Im and (X_M only)
    // -Re: every B operand of the complex products, with the signs of
    // Y_J = Z11 X_J - ZK X_M, is one LDS read with a lane-dependent plane (the
    // real-part columns, part 0, and the imaginary-part columns, part 1, take
    // different planes), so the products need no sign flips
    constexpr int NPL = P == 3 ? 4 : 3;
    __shared__ __align__(16) unsigned short xs[NS][NX][NPL][RT][kKP];
    // Chunks of 4 items (mostly the row ranges of one leaf, which gather the
    // same X columns) with their ny right-hand-side groups (which read the
    // same records) run together on one XCD, so the later reads hit its L2.
    const int ny = (R + RT - 1) / RT;
    int it, gy;
    xcd_chunk(blockIdx.x, 4, ny, it, gy);
    if (it >= nd.nitem) return;
    const int2 item = nd.items[it];
    const int t = item.x, rp = item.y;
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    const int nrt = (nt + 15) >> 4;
    const int nrp = min(kRTC, nrt - rp);
    const int rb0 = gy * RT;
    const int rv = min(RT, R - rb0);
    const int C = nd.tgt_cols[t];
    const int nkb = (C + 31) >> 5;
    const int* cref = nd.colref + nd.col_ptr[t];
    // the wavefront index as a scalar, so tests on it are uniform branches
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int part = li >> 3;          // 0: real-part column, 1: imaginary-part column
    const bool wact = wave < nrp;      // this wavefront's row tile exists
    const uint4* zrec = Zb + bf_off[t] + (size_t)(rp + (wact ? wave : 0)) * P * W + lane;
    const size_t kstride = (size_t)nrt * P * W;   // records of one 32-column block
    // records of block kb (clamped to the last block: a block past nkb is
    // loaded but never used), without bounds branches, which would make the
    // compiler wait for the prefetched records before the current products
    // zr/zi[p] hold the hi parts of plane p, zr/zi[P + p] the lo parts (S)
    auto load = [&](int kb, uint4* zr, uint4* zi) {
        const uint4* rec = zrec + (size_t)min(kb, nkb - 1) * kstride;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            zr[p] = rec[p * W];
            zi[p] = rec[p * W + 64];
            if (S) {
                zr[P + p] = rec[p * W + 128];
                zi[P + p] = rec[p * W + 192];
            }
        }
    };
    floatx4 acc[CT][NO];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int o = 0; o < NO; ++o) acc[c][o] = floatx4{0.f, 0.f, 0.f, 0.f};
    // X staging, one chunk ahead: a thread's units (current xj, 8-column group
    // cg, right-hand side r) load the next chunk's X into registers while the
    // current chunk's products run, and the source indices two chunks ahead.
    constexpr int NE = NX * 8 * RT, NU = (NE + 255) / 256;
    constexpr bool kShare = NU == 2 && NE == 512;   // both units have the same (cg, r)
    constexpr int NCR = kShare ? 1 : NU;
    int ucg[NU], ur[NU], uxj[NU];
    bool uok[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int e = threadIdx.x + 256 * u, ec = min(e, NE - 1);   // spare threads repeat the last unit
        uok[u] = NE % 256 == 0 || e < NE;
        uxj[u] = ec / (8 * RT);
        const int rem = ec - uxj[u] * (8 * RT);
        ucg[u] = rem / RT;
        ur[u] = rem - ucg[u] * RT;
    }
    // Loads go to clamped (valid) addresses without branches; bit 8 u + i of
    // xmask marks the values that are stored (the rest are staged as zero).
    int cr[NCR][8];        // source unknowns of the columns of the chunk after the staged one
    xb_t xv[NU][8];        // next chunk's X (BF16 Re | Im << 16; S: .x hi, .y lo)
    unsigned xmask = 0;
    auto load_cref = [&](int cc) {
#pragma unroll
        for (int u = 0; u < NCR; ++u)
#pragma unroll
            for (int i = 0; i < 8; ++i) cr[u][i] = cref[min(cc + ucg[u] * 8 + i, C - 1)];
    };
    auto load_x = [&](int cc) {
        xmask = 0;
#pragma unroll
        for (int u = 0; u < NU; ++u)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int src = cr[kShare ? 0 : u][i];
                xv[u][i] = xb[((size_t)src + (uxj[u] ? (size_t)nd.N : 0)) * R + rb0 + min(ur[u], rv - 1)];
                xmask |= (unsigned)(uok[u] & (cc + ucg[u] * 8 + i < C) & (ur[u] < rv)) << (8 * u + i);
            }
    };
    load_cref(0);
    load_x(0);
    load_cref(kKC);
    // products of one 32-column block; zr/zi as filled by load
    auto step = [&](int kb2, const uint4* zr, const uint4* zi) {
        if (!wact) return;
        const int kof = kb2 * 32 + 8 * lk;
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int rc = ct * 8 + (li & 7);
            floatx4 a = acc[ct][0], b = acc[ct][NO - 1];
            auto ld = [&](int xp, int xj, int pl) { return *reinterpret_cast<const uint4*>(&xs[xp][xj][pl][rc][kof]); };
            // (Z part, X part) pairs: (hi, hi), and with S also (hi, lo), (lo, hi)
#pragma unroll
            for (int q = 0; q < (S ? 3 : 1); ++q) {
                const int zp = q == 2 ? P : 0, xp = q == 1 ? 1 : 0;
                const uint4 bJ1 = ld(xp, 0, 2 * part);   // Re X | Im X
                const uint4 bJ2 = ld(xp, 0, 1 - part);   // -Im X | Re X
                a = mfma(zr[zp], bJ1, a);
                a = mfma(zi[zp], bJ2, a);
                if (P == 3) {
                    const uint4 bM1 = ld(xp, NX - 1, 2 * part), bM2 = ld(xp, NX - 1, 1 - part);
                    a = mfma(zr[zp + 1], ld(xp, NX - 1, 3 - 2 * part), a);   // - ZK X_M: -Re | -Im
                    a = mfma(zi[zp + 1], ld(xp, NX - 1, 2 + part), a);       //           Im | -Re
                    b = mfma(zr[zp + 1], bJ1, b);
                    b = mfma(zi[zp + 1], bJ2, b);
                    b = mfma(zr[zp + 2], bM1, b);
                    b = mfma(zi[zp + 2], bM2, b);
                }
            }
            acc[ct][0] = a;
            if (P == 3) acc[ct][NO - 1] = b;
        }
    };
    uint4 ar[NS * P], ai[NS * P], br[NS * P], bi[NS * P];
    load(0, ar, ai);
    for (int c0 = 0; c0 < C; c0 += kKC) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            if (!uok[u]) continue;
#pragma unroll
            for (int h = 0; h < NS; ++h) {
                unsigned re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    unsigned v;
                    if constexpr (S) v = h ? xv[u][i].y : xv[u][i].x;
                    else v = xv[u][i];
                    v = ((xmask >> (8 * u + i)) & 1u) ? v : 0u;
                    re[i >> 1] |= (v & 0xffffu) << (16 * (i & 1));
                    im[i >> 1] |= (v >> 16) << (16 * (i & 1));
                }
                const uint4 vr = make_uint4(re[0], re[1], re[2], re[3]), vi = make_uint4(im[0], im[1], im[2], im[3]);
                *reinterpret_cast<uint4*>(&xs[h][uxj[u]][0][ur[u]][ucg[u] * 8]) = vr;
                *reinterpret_cast<uint4*>(&xs[h][uxj[u]][1][ur[u]][ucg[u] * 8]) = neg8(vi);
                *reinterpret_cast<uint4*>(&xs[h][uxj[u]][2][ur[u]][ucg[u] * 8]) = vi;
                if (P == 3 && uxj[u] == NX - 1)
                    *reinterpret_cast<uint4*>(&xs[h][NX - 1][NPL - 1][ur[u]][ucg[u] * 8]) = neg8(vr);
            }
        }
        __syncthreads();
        // Every load below is issued on every pass (past the end they read
        // clamped, finite records and indices that meet zero-staged X), so
        // the compiler's wait counts stay exact: each wait covers the data it
        // needs, not the loads issued after it.
        load_x(c0 + kKC);
        load_cref(c0 + 2 * kKC);
        // even block in set A, odd block in set B (no copies, so the next
        // block's records stay in flight during the current block's products)
        const int kb = c0 >> 5;
        load(kb + 1, br, bi);
        step(0, ar, ai);
        load(kb + 2, ar, ai);
        step(1, br, bi);
    }
    if (!wact) return;
    // D[4 lk + v][li]: lanes li < 8 hold Re of right-hand side ct * 8 + li, lane li + 8 its Im
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
                    const size_t u = (size_t)nd.perm[i0 + row] + (o ? (size_t)nd.N : 0);
                    const size_t e = u * ldr + r0 + rb0 + rc;
                    if constexpr (F) {
                        float2 v = pe.t[e];
                        cfms(v, pe.dinv[u], make_float2(re, im));
                        pe.t[e] = v;
                        float2 w = pe.y[e];
                        cfma(w, pe.a, v);
                        pe.y[e] = w;
                        if (pe.xbn) pe.xbn[e] = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
                    } else {
                        y[e] = make_double2(re, im);
                    }
                }
            }
}

long long bf16_near_layout(const std::vector<int>& leaf_ptr, const std::vector<int>& tgt_cols, int planes,
                           bool split, std::vector<long long>& bf_off) {
    const int nleaf = (int)tgt_cols.size();
    bf_off.assign(nleaf + 1, 0);
    for (int t = 0; t < nleaf; ++t) {
        const int nt = leaf_ptr[t + 1] - leaf_ptr[t];
        const long long recs = nt == 0 ? 0 : (long long)((tgt_cols[t] + 31) / 32) * ((nt + 15) / 16) * planes;
        bf_off[t + 1] = bf_off[t] + recs * (split ? 256 : 128);
    }
    return bf_off[nleaf];
}

void launch_pack_bf16(hipStream_t st, const NearDev& nd, const double2* Z, const long long* bf_off, bool split,
                      uint4* out) {
    if (nd.nleaf == 0) return;
    if (split) k_pack_bf16<true><<<nd.nleaf, 256, 0, st>>>(nd, Z, bf_off, out);
    else k_pack_bf16<false><<<nd.nleaf, 256, 0, st>>>(nd, Z, bf_off, out);
}

// the products of xb (already staged as BF16: single values, or split pairs
// with S); F: the polynomial step of pe instead of y = Z X
template <bool S, bool F>
static void near_bf16_products(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                               const long long* bf_off, const void* xb, double2* y, const PolyEpi& pe) {
    const int ct = R > 16 ? 4 : (R > 8 ? 2 : 1);
    const int ny = (R + 8 * ct - 1) / (8 * ct);
    const unsigned grid = xcd_chunk_grid(nd.nitem, 4, ny);
#define NB(P, CT) k_near_bf16<P, CT, S, F><<<grid, 256, 0, st>>>(nd, R, ldr, r0, Zb, bf_off, xb, y, pe)
#define NBP(P)                                                                                                        \
    if (ct == 4) NB(P, 4);                                                                                            \
    else if (ct == 2) NB(P, 2);                                                                                       \
    else NB(P, 1);
    if (nd.planes == 1) {
        NBP(1)
    } else {
        NBP(3)
    }
#undef NBP
#undef NB
}

static size_t near_unknowns(const NearDev& nd) { return (size_t)nd.N * (nd.planes == 3 ? 2 : 1); }

void launch_near_bf16(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                      const long long* bf_off, const double2* x, unsigned* xb, double2* y) {
    if (nd.nitem == 0 || R < 1) return;
    const size_t total = near_unknowns(nd) * R;
    k_x_bf16<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, R, ldr, r0, x, xb);
    near_bf16_products<false, false>(st, R, ldr, r0, nd, Zb, bf_off, xb, y, PolyEpi{});
}

template <typename TX>
static void near_bf16x3_t(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                          const long long* bf_off, const TX* x, uint2* xb, double2* y) {
    if (nd.nitem == 0 || R < 1) return;
    const size_t total = near_unknowns(nd) * R;
    k_x_bf16x2<TX><<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, R, ldr, r0, x, xb);
    near_bf16_products<true, false>(st, R, ldr, r0, nd, Zb, bf_off, xb, y, PolyEpi{});
}

void launch_near_bf16x3(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                        const long long* bf_off, const double2* x, uint2* xb, double2* y) {
    near_bf16x3_t(st, R, ldr, r0, nd, Zb, bf_off, x, xb, y);
}

void launch_near_bf16x3(hipStream_t st, int R, int ldr, int r0, const NearDev& nd, const uint4* Zb,
                        const long long* bf_off, const float2* x, uint2* xb, double2* y) {
    near_bf16x3_t(st, R, ldr, r0, nd, Zb, bf_off, x, xb, y);
}

void launch_poly_init_bf16(hipStream_t st, int R, const NearDev& nd, const float2* dinv, const float2* x, float2 a0,
                           float2* t, float2* y, unsigned* xb) {
    const size_t total = near_unknowns(nd) * R;
    if (total == 0) return;
    k_poly_init_bf16<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, R, dinv, x, a0, t, y, xb);
}

void launch_near_bf16_poly(hipStream_t st, int R, const NearDev& nd, const uint4* Zb, const long long* bf_off,
                           const unsigned* xb, const PolyEpi& pe) {
    if (nd.nitem == 0 || R < 1) return;
    near_bf16_products<false, true>(st, R, R, 0, nd, Zb, bf_off, xb, nullptr, pe);
}

}  // namespace mlfma
