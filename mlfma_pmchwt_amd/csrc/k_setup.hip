// FP64 setup kernels: radiation / receiving patterns, near-field Galerkin
// blocks, plane-wave excitation and far-field (RCS) sums.
#include "cplx.hpp"
#include "geo.hpp"
#include "setup.hpp"

namespace mlfma {

namespace {
constexpr double kPiD = 3.14159265358979323846;
constexpr double kEps0D = 2.2204e-16;   // Code/PhyConstant.h:13

struct R3 {
    double x, y, z;
};
__device__ __forceinline__ R3 mk(const double* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ R3 sub3(R3 a, R3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ R3 add3(R3 a, R3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ R3 scl3(R3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot3(R3 a, R3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ R3 crs3(R3 a, R3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double nrm3(R3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ R3 gpt(const GaussRule& g, int q, const double* t) {
    return {g.l1[q] * t[0] + g.l2[q] * t[3] + g.l3[q] * t[6], g.l1[q] * t[1] + g.l2[q] * t[4] + g.l3[q] * t[7],
            g.l1[q] * t[2] + g.l2[q] * t[5] + g.l3[q] * t[8]};
}
// exp(-j k R) for complex k = kr + j ki
__device__ __forceinline__ double2 expmjk(double2 k, double R) {
    double s, c;
    sincos(k.x * R, &s, &c);
    const double m = (k.y != 0.0) ? exp(k.y * R) : 1.0;
    return make_double2(m * c, -m * s);
}

// Analytic integrals over a flat source triangle of 1/R and 1/R^3 (vector and
// scalar parts) for one observation point on that triangle (the identical-face
// terms); Code/Operator.h:72-305 and :585-674.  The point's height D above the
// triangle's plane is exactly 0: the reference computes D = n.(r - v) in floating
// point, and that rounding noise (~1e-16 |r|) picks between the two branches of the
// 1/R^3 scalar integral, one of which divides an O(1) angle difference by the noise;
// D = 0 gives the reference's value wherever its noise stays below 10 eps0.
struct SingFrame {
    R3 rho[3], l[3], u[3], n;
};
struct SingPt {
    R3 Iv, Iv3, rho;
    double Is, Is3;
};

__device__ void sing_frame(const double* t, const double* nn, SingFrame& f) {
    f.n = mk(nn);
    for (int v = 0; v < 3; ++v) {
        R3 p = mk(t + 3 * v);
        f.rho[v] = sub3(p, scl3(f.n, dot3(f.n, p)));
    }
    for (int e = 0; e < 3; ++e) {
        R3 d = sub3(f.rho[(e + 1) % 3], f.rho[e]);
        double L = nrm3(d);
        f.l[e] = {d.x / L, d.y / L, d.z / L};
        f.u[e] = crs3(f.l[e], f.n);
    }
}

__device__ void sing_point(R3 r, const double* t, const SingFrame& f, bool want3, SingPt& o) {
    const double th = 50.0 * kEps0D;
    const R3 rho = sub3(r, scl3(f.n, dot3(f.n, r)));
    o.rho = rho;
    o.Iv = {0, 0, 0};
    o.Iv3 = {0, 0, 0};
    o.Is = 0;
    o.Is3 = 0;
    for (int e = 0; e < 3; ++e) {
        // edge e runs from vertex e ("minus" end) to vertex e+1 ("plus" end)
        const R3 dp = sub3(f.rho[(e + 1) % 3], rho), dm = sub3(f.rho[e], rho);
        const double Lp = dot3(dp, f.l[e]), Lm = dot3(dm, f.l[e]);
        const double P0 = fabs(dot3(dp, f.u[e]));
        const double Pp = nrm3(dp), Pm = nrm3(dm);
        R3 P0v = {0, 0, 0};
        if (!(P0 < th)) {
            const R3 pm = sub3(dp, scl3(f.l[e], Lp));
            P0v = {pm.x / P0, pm.y / P0, pm.z / P0};
        }
        const double D = 0.0;
        const double R0 = sqrt(P0 * P0 + D * D), Rp = sqrt(Pp * Pp + D * D), Rm = sqrt(Pm * Pm + D * D);
        const double LOGp = (Rp + Lp < th) ? 0.0 : log(Rp + Lp);
        const double LOGm = (Rm + Lm < th) ? 0.0 : log(Rm + Lm);
        const double c1 = 0.5 * (R0 * R0 * (LOGp - LOGm) + Rp * Lp - Rm * Lm);
        o.Iv = add3(o.Iv, scl3(f.u[e], c1));
        const double ad = fabs(D);
        const double denp = R0 * R0 + ad * Rp, denm = R0 * R0 + ad * Rm;
        const double atp = (denp < th) ? 0.0 : atan((P0 * Lp) / denp);
        const double atm = (denm < th) ? 0.0 : atan((P0 * Lm) / denm);
        const double pu = dot3(P0v, f.u[e]);
        o.Is = o.Is + pu * (P0 * (LOGp - LOGm) - ad * (atp - atm));
        if (want3) {
            o.Iv3 = sub3(o.Iv3, scl3(f.u[e], LOGp - LOGm));
            const double t10 = 10.0 * kEps0D;
            if (ad >= t10) {
                const double ap = atan((P0 * Lp) / (R0 * R0 + ad * Rp));
                const double am = atan((P0 * Lm) / (R0 * R0 + ad * Rm));
                o.Is3 = o.Is3 + pu * ((1.0 / ad) * (ap - am));
            } else if (P0 >= t10) {
                o.Is3 = o.Is3 - pu * (Lp / (P0 * Rp) - Lm / (P0 * Rm));
            } else {
                o.Is3 = 0.0;   // the reference resets the accumulator (Code/Operator.h:671-673)
            }
        }
    }
}

// Galerkin entries between field RWG F and source RWG S.
//   L(k):  Code/Operator.h:709-881     (EFIE = eta * L, Code/EFIE.cpp:6-194)
//   K(k):  Code/Operator.h:883-1046
//   M(k):  MFIE, Code/MFIE.cpp:6-212
template <int NK, bool WL, bool WK, bool WM>
__device__ void pair_ops(const RwgGeo& F, const RwgGeo& S, const GaussRule& g, const double2* ks, double2* Lo,
                         double2* Ko, double2* Mo) {
    for (int q = 0; q < NK; ++q) Lo[q] = Ko[q] = make_double2(0, 0);
    *Mo = make_double2(0, 0);
    for (int i = 0; i < 2; ++i) {
        const double* tF = F.tri[i];
        const R3 vF = mk(F.vf[i]);
        const R3 nF = mk(F.nrm[i]);
        const double AF = F.area[i], recAf = 1.0 / (2.0 * AF);
        for (int j = 0; j < 2; ++j) {
            const double* tS = S.tri[j];
            const R3 vS = mk(S.vf[j]);
            const double AS = S.area[j], recAs = 1.0 / (2.0 * AS);
            const bool same = (F.face[i] == S.face[j]);
            double2 aL[NK], aK[NK], aM = make_double2(0, 0);
            for (int q = 0; q < NK; ++q) aL[q] = aK[q] = make_double2(0, 0);
            for (int a = 0; a < g.n; ++a) {
                const R3 rf = gpt(g, a, tF);
                const R3 pf = sub3(rf, vF);
                const double wF = g.w[a];
                for (int b = 0; b < g.n; ++b) {
                    const R3 rs = gpt(g, b, tS);
                    const R3 ps = sub3(rs, vS);
                    const R3 Rv = sub3(rf, rs);
                    const double R = nrm3(Rv);
                    const double w = wF * g.w[b];
                    const double dL = WL ? dot3(pf, ps) : 0.0;
                    const double dK = (WK || (WM && !same)) ? dot3(pf, crs3(Rv, ps)) : 0.0;
                    double dM = 0.0;
                    if (WM) dM = same ? dot3(crs3(sub3(rf, vS), ps), crs3(pf, nF)) : dot3(pf, crs3(nF, crs3(Rv, ps)));
#pragma unroll
                    for (int q = 0; q < NK; ++q) {
                        const double2 k = ks[q];
                        const double2 jk = make_double2(-k.y, k.x);
                        double2 G, G3;
                        if (!same) {
                            const double2 e = expmjk(k, R);
                            G = cscale(e, 1.0 / R);
                            // e / R^3 * (1 + j k R)
                            G3 = cscale(cmul(e, make_double2(1.0 - k.y * R, k.x * R)), 1.0 / (R * R * R));
                        } else if (R < kEps0D) {
                            G = make_double2(k.y, -k.x);   // -j k
                            G3 = make_double2(0, 0);
                        } else {
                            const double2 e = expmjk(k, R);
                            G = make_double2(e.x / R - 1.0 / R, e.y / R);
                            G3 = cscale(cmul(e, make_double2(1.0 - k.y * R, k.x * R)), 1.0 / (R * R * R));
                            const double2 k2 = cmul(k, k);
                            G3.x -= 1.0 / (R * R * R) + 0.5 * k2.x / R;
                            G3.y -= 0.5 * k2.y / R;
                        }
                        if (WL) {
                            // jk w (G pf.ps - 4 G / k^2)
                            const double2 c4 = cdiv(make_double2(4.0, 0.0), cmul(k, k));
                            const double2 t = csub(cscale(G, dL), cmul(G, c4));
                            cfma(aL[q], cscale(jk, w), t);
                        }
                        if (WK) cfma(aK[q], make_double2(w, 0.0), cscale(G3, dK));
                        if (WM && q == 0) cfma(aM, make_double2(w, 0.0), cscale(G3, dM));
                    }
                }
            }
            if (same) {
                SingFrame fr;
                sing_frame(tS, S.nrm[j], fr);
                const R3 nS = fr.n;
                const R3 rho0N = sub3(vS, scl3(nS, dot3(nS, vS)));
                for (int a = 0; a < g.n; ++a) {
                    const R3 rf = gpt(g, a, tF);
                    const double wF = g.w[a];
                    SingPt sp;
                    sing_point(rf, tS, fr, WK || WM, sp);
                    const R3 RWGf = scl3(sub3(rf, vF), recAf);
                    const R3 pv1 = sub3(sp.rho, rho0N);
                    const R3 Ri = sub3(rf, vS);
                    if (WM) {   // identity term, Code/MFIE.cpp:175-192
                        const R3 RWGs = scl3(Ri, recAs);
                        aM.x += 4.0 * kPiD * wF * AF * dot3(RWGf, RWGs);
                    }
#pragma unroll
                    for (int q = 0; q < NK; ++q) {
                        const double2 k = ks[q];
                        const double2 jk = make_double2(-k.y, k.x);
                        const double2 k2 = cmul(k, k);
                        if (WL) {
                            const R3 Ikon = scl3(add3(sp.Iv, scl3(pv1, sp.Is)), recAs);
                            const double2 t =
                                csub(make_double2(dot3(RWGf, Ikon), 0.0),
                                     cscale(cdiv(make_double2(1.0, 0.0), k2), sp.Is / (AF * AS)));
                            cfma(aL[q], cscale(jk, wF * 2.0 * AF), t);
                        }
                        if (WK || (WM && q == 0)) {
                            // Ikon = (k^2/2)(Iv + pv1 Is)/(2As) + (Iv3 + pv1 Is3)/(2As)
                            const double2 h = cscale(k2, 0.5);
                            const R3 a1 = add3(sp.Iv, scl3(pv1, sp.Is));
                            const R3 a3 = add3(sp.Iv3, scl3(pv1, sp.Is3));
                            const double2 I0 = make_double2(h.x * a1.x * recAs + a3.x * recAs, h.y * a1.x * recAs);
                            const double2 I1 = make_double2(h.x * a1.y * recAs + a3.y * recAs, h.y * a1.y * recAs);
                            const double2 I2 = make_double2(h.x * a1.z * recAs + a3.z * recAs, h.y * a1.z * recAs);
                            // Ri x Ikon
                            const double2 X0 = csub(cscale(I2, Ri.y), cscale(I1, Ri.z));
                            const double2 X1 = csub(cscale(I0, Ri.z), cscale(I2, Ri.x));
                            const double2 X2 = csub(cscale(I1, Ri.x), cscale(I0, Ri.y));
                            if (WK) {
                                double2 d = cadd(cadd(cscale(X0, RWGf.x), cscale(X1, RWGf.y)), cscale(X2, RWGf.z));
                                aK[q] = cadd(aK[q], cscale(d, wF * 2.0 * AF));
                            }
                            if (WM && q == 0) {
                                const R3 c2 = crs3(RWGf, nF);
                                double2 d = cadd(cadd(cscale(X0, c2.x), cscale(X1, c2.y)), cscale(X2, c2.z));
                                aM = cadd(aM, cscale(d, wF * 2.0 * AF));
                            }
                        }
                    }
                }
            }
            const double sg = ((i == 0) ? -1.0 : 1.0) * ((j == 0) ? -1.0 : 1.0);
            const double sc = (F.len * S.len) / (4.0 * kPiD) * sg;
            for (int q = 0; q < NK; ++q) {
                if (WL) Lo[q] = cadd(Lo[q], cscale(aL[q], sc));
                if (WK) Ko[q] = cadd(Ko[q], cscale(aK[q], sc));
            }
            if (WM) *Mo = cadd(*Mo, cscale(aM, sc));
        }
    }
}

}  // namespace

// ---------------------------------------------------------------- near field
// One workgroup per (target leaf, source leaf) block, written into its
// columns of the target's block row (layout of NearDev, geo.hpp).
// Code/Zmartix.h:188-278 (EFIE), :372-465 (CFIE), :541-643 (PMCHWT).
template <int MODE>
__global__ __launch_bounds__(256) void k_near_fill(NearDev nd, const RwgGeo* __restrict__ geo, GaussRule g,
                                                   NearConsts c, double2* __restrict__ Z) {
    const int p = blockIdx.x;
    const int t = nd.pair_tgt[p], s = nd.pair_src[p];
    const int i0 = nd.leaf_ptr[t], nt = nd.leaf_ptr[t + 1] - i0;
    const int j0 = nd.leaf_ptr[s], ns = nd.leaf_ptr[s + 1] - j0;
    const size_t plane = (size_t)nd.tgt_cols[t] * nt;
    double2* out = Z + nd.tgt_off[t] + (size_t)nd.pair_col[p] * nt;
    for (int e = threadIdx.x; e < nt * ns; e += blockDim.x) {
        const int row = e % nt, col = e / nt;
        const RwgGeo& F = geo[i0 + row];
        const RwgGeo& S = geo[j0 + col];
        if (MODE == 2) {
            double2 L[2], K[2], M;
            const double2 ks[2] = {c.k1, c.k2};
            pair_ops<2, true, true, false>(F, S, g, ks, L, K, &M);
            const double2 z11 = cadd(L[0], cmul(c.etai, L[1]));
            const double2 zk = cadd(K[0], K[1]);
            const double2 z22 = cadd(L[0], cmul(c.inv_etai, L[1]));
            out[(size_t)col * nt + row] = z11;
            out[plane + (size_t)col * nt + row] = zk;
            out[2 * plane + (size_t)col * nt + row] = z22;
        } else if (MODE == 0) {
            double2 L, K, M;
            pair_ops<1, true, false, false>(F, S, g, &c.k1, &L, &K, &M);
            out[(size_t)col * nt + row] = cmul(c.eta1, L);
        } else {
            double2 L, K, M;
            pair_ops<1, true, false, true>(F, S, g, &c.k1, &L, &K, &M);
            // alpha E + (1 - alpha) eta M, E = eta L  (Code/Zmartix.h:432)
            const double2 E = cmul(c.eta1, L);
            out[(size_t)col * nt + row] = cadd(cscale(E, c.alpha), cscale(cmul(c.eta1, M), 1.0 - c.alpha));
        }
    }
}

void launch_near_fill(hipStream_t st, int mode, const NearDev& nd, const RwgGeo* geo, const GaussRule& g,
                      const NearConsts& c, double2* Z) {
    if (nd.npairs == 0) return;
    if (mode == 2) k_near_fill<2><<<nd.npairs, 256, 0, st>>>(nd, geo, g, c, Z);
    else if (mode == 1) k_near_fill<1><<<nd.npairs, 256, 0, st>>>(nd, geo, g, c, Z);
    else k_near_fill<0><<<nd.npairs, 256, 0, st>>>(nd, geo, g, c, Z);
}

// Single-entry evaluation (tests): out[0..3] = L(k1), L(k2), K(k1), K(k2); out[4] = MFIE(k1)
__global__ void k_pair_probe(const RwgGeo* geo, int f, int s, GaussRule g, NearConsts c, double2* out) {
    double2 L[2], K[2], M;
    const double2 ks[2] = {c.k1, c.k2};
    pair_ops<2, true, true, false>(geo[f], geo[s], g, ks, L, K, &M);
    out[0] = L[0], out[1] = L[1], out[2] = K[0], out[3] = K[1];
    double2 L1, K1;
    pair_ops<1, false, false, true>(geo[f], geo[s], g, &c.k1, &L1, &K1, &M);
    out[4] = M;
}

void launch_pair_probe(hipStream_t st, const RwgGeo* geo, int f, int s, const GaussRule& g, const NearConsts& c,
                       double2* out) {
    k_pair_probe<<<1, 1, 0, st>>>(geo, f, s, g, c, out);
}

// ---------------------------------------------------------------- patterns
// Code/ProcessFMM.h:17-180: per RWG (internal order) and k-space sample,
//   Vs = sum_q w l e^{+jk khat.(r-c)} (r - v_free)                (KernelVsmi_Efie)
//   Vf = sum_q w l e^{-jk khat.(r-c)} (-khat x (khat x (r - v)))   (KernelVfmj_Efie)
//   or alpha * that + (1-alpha) khat x (n x (r - v))               (KernelVfmj_Cfie)
// positive face minus negative face.
template <bool CFIE>
__global__ __launch_bounds__(256) void k_patterns(int K, const RwgGeo* __restrict__ geo, const int* __restrict__ leaf_of,
                                                  const double* __restrict__ centre, const double* __restrict__ kp,
                                                  double2 k, GaussRule g, double alpha, double2* __restrict__ Vs,
                                                  double2* __restrict__ Vf) {
    const int i = blockIdx.x;
    const int q = blockIdx.y * blockDim.x + threadIdx.x;
    if (q >= K) return;
    const RwgGeo& G = geo[i];
    const int lf = leaf_of[i];
    const R3 c = {centre[3 * lf], centre[3 * lf + 1], centre[3 * lf + 2]};
    const R3 kh = {kp[q], kp[K + q], kp[2 * K + q]};
    double2 s0 = {0, 0}, s1 = {0, 0}, s2 = {0, 0}, f0 = {0, 0}, f1 = {0, 0}, f2 = {0, 0};
    for (int a = 0; a < g.n; ++a) {
        const double w = g.w[a] * G.len;
        for (int side = 1; side >= 0; --side) {
            const double sg = side == 1 ? 1.0 : -1.0;
            const R3 r = gpt(g, a, G.tri[side]);
            const R3 pf = sub3(r, mk(G.vf[side]));
            const double ph = dot3(sub3(r, c), kh);
            double sn, cs;
            sincos(k.x * ph, &sn, &cs);
            const double mf = (k.y != 0.0) ? exp(k.y * ph) : 1.0;    // |e^{-jk ph}|
            const double ms = (k.y != 0.0) ? exp(-k.y * ph) : 1.0;   // |e^{+jk ph}|
            const double2 es = make_double2(sg * w * ms * cs, sg * w * ms * sn);
            const double2 ef = make_double2(sg * w * mf * cs, -sg * w * mf * sn);
            s0 = cadd(s0, cscale(es, pf.x));
            s1 = cadd(s1, cscale(es, pf.y));
            s2 = cadd(s2, cscale(es, pf.z));
            const R3 t1 = crs3(kh, pf);
            const R3 t2 = crs3(kh, t1);
            R3 v = {-t2.x, -t2.y, -t2.z};
            if (CFIE) {
                const R3 m = crs3(kh, crs3(mk(G.nrm[side]), pf));
                v = add3(scl3(v, alpha), scl3(m, 1.0 - alpha));
            }
            f0 = cadd(f0, cscale(ef, v.x));
            f1 = cadd(f1, cscale(ef, v.y));
            f2 = cadd(f2, cscale(ef, v.z));
        }
    }
    double2* o = Vs + (size_t)i * 3 * K + q;
    o[0] = s0, o[K] = s1, o[2 * K] = s2;
    o = Vf + (size_t)i * 3 * K + q;
    o[0] = f0, o[K] = f1, o[2 * K] = f2;
}

void launch_patterns(hipStream_t st, bool cfie, int N, int K, const RwgGeo* geo, const int* leaf_of,
                     const double* centre, const double* kp, double2 k, const GaussRule& g, double alpha, double2* Vs,
                     double2* Vf) {
    dim3 grid(N, (K + 255) / 256);   // gridDim.y is limited to 65535
    if (cfie) k_patterns<true><<<grid, 256, 0, st>>>(K, geo, leaf_of, centre, kp, k, g, alpha, Vs, Vf);
    else k_patterns<false><<<grid, 256, 0, st>>>(K, geo, leaf_of, centre, kp, k, g, alpha, Vs, Vf);
}

// ---------------------------------------------------------------- excitation
// Code/Vm.h:10-203: E-testing <f, E e^{-jk kinc.r}>, H-testing with (n x H).rho
// (CFIE) or H.rho (PMCHWT); one thread per RWG, written at the reference index.
__global__ void k_excitation(int N, int mode, const RwgGeo* __restrict__ geo, const int* __restrict__ perm,
                             GaussRule g, ExcitationArgs a, double2* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const RwgGeo& G = geo[i];
    double2 vE = {0, 0}, vH = {0, 0};
    const R3 kin = {a.kinc[0], a.kinc[1], a.kinc[2]};
    const R3 E = {a.einc[0], a.einc[1], a.einc[2]};
    const R3 H = {a.hinc[0], a.hinc[1], a.hinc[2]};
    for (int q = 0; q < g.n; ++q) {
        for (int side = 1; side >= 0; --side) {
            const R3 r = gpt(g, q, G.tri[side]);
            // rho = r - v+ on the positive face, v- - r on the negative face
            const R3 rho = side == 1 ? sub3(r, mk(G.vf[1])) : sub3(mk(G.vf[0]), r);
            const double2 e = expmjk(a.k, dot3(kin, r));
            const double wl = G.len * g.w[q];
            vE = cadd(vE, cscale(e, wl * dot3(E, rho)));
            if (mode == 1) vH = cadd(vH, cscale(e, wl * dot3(crs3(mk(G.nrm[side]), H), rho)));
            if (mode == 2) vH = cadd(vH, cscale(e, wl * dot3(H, rho)));
        }
    }
    const int r = perm[i];
    if (mode == 0) v[r] = vE;
    else if (mode == 1) v[r] = cadd(cscale(vE, a.alpha), cscale(vH, 1.0 - a.alpha));
    else {
        v[r] = vE;
        v[N + r] = vH;
    }
}

void launch_excitation(hipStream_t st, int N, int mode, const RwgGeo* geo, const int* perm, const GaussRule& g,
                       const ExcitationArgs& a, double2* v) {
    k_excitation<<<(N + 255) / 256, 256, 0, st>>>(N, mode, geo, perm, g, a, v);
}

// ---------------------------------------------------------------- far field
// Code/RCS.h:54-192: per observation direction, block partial sums of
//   PEC:    jk eta sum_n I_n F_n,  F_n = l sum_q w (I - rr).rho e^{+jk r.p}
//   PMCHWT: jk1 sum_n (F_n I_J - (r x rho part) I_M)
__global__ __launch_bounds__(256) void k_farfield(int N, int pmchwt, const RwgGeo* __restrict__ geo,
                                                  const int* __restrict__ perm, GaussRule g, double2 k,
                                                  const double* __restrict__ dirs, const double2* __restrict__ I,
                                                  size_t istride, double2* __restrict__ partial) {
    __shared__ double2 red[3][4];
    const int ang = blockIdx.y;
    I += (size_t)ang * istride;   // monostatic: every angle has its own current vector
    const R3 rd = {dirs[3 * ang], dirs[3 * ang + 1], dirs[3 * ang + 2]};
    double2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const RwgGeo& G = geo[i];
        double2 FJ[3] = {{0, 0}, {0, 0}, {0, 0}}, FM[3] = {{0, 0}, {0, 0}, {0, 0}};
        for (int q = 0; q < g.n; ++q) {
            for (int side = 1; side >= 0; --side) {
                const R3 r = gpt(g, q, G.tri[side]);
                const R3 rho = side == 1 ? sub3(r, mk(G.vf[1])) : sub3(mk(G.vf[0]), r);
                // e^{+jk r.p}
                double2 e = expmjk(make_double2(-k.x, -k.y), dot3(r, rd));
                const double wl = G.len * g.w[q];
                const R3 t = sub3(rho, scl3(rd, dot3(rd, rho)));
                FJ[0] = cadd(FJ[0], cscale(e, wl * t.x));
                FJ[1] = cadd(FJ[1], cscale(e, wl * t.y));
                FJ[2] = cadd(FJ[2], cscale(e, wl * t.z));
                if (pmchwt) {
                    const R3 cr = crs3(rd, rho);
                    FM[0] = cadd(FM[0], cscale(e, wl * cr.x));
                    FM[1] = cadd(FM[1], cscale(e, wl * cr.y));
                    FM[2] = cadd(FM[2], cscale(e, wl * cr.z));
                }
            }
        }
        const int r = perm[i];
        const double2 IJ = I[r];
        for (int c = 0; c < 3; ++c) acc[c] = cmul(FJ[c], IJ);
        if (pmchwt) {
            const double2 IM = I[N + r];
            for (int c = 0; c < 3; ++c) acc[c] = csub(acc[c], cmul(FM[c], IM));
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int c = 0; c < 3; ++c) {
        double2 s = wave_sum(acc[c]);
        if (lane == 0) red[c][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        double2 s = cadd(cadd(red[c][0], red[c][1]), cadd(red[c][2], red[c][3]));
        partial[((size_t)ang * gridDim.x + blockIdx.x) * 3 + c] = s;
    }
}

void launch_farfield(hipStream_t st, int N, int pmchwt, const RwgGeo* geo, const int* perm, const GaussRule& g,
                     double2 k, int nang, const double* dirs, const double2* I, size_t istride, double2* partial,
                     int nblocks) {
    dim3 grid(nblocks, nang);
    k_farfield<<<grid, 256, 0, st>>>(N, pmchwt, geo, perm, g, k, dirs, I, istride, partial);
}

}  // namespace mlfma
