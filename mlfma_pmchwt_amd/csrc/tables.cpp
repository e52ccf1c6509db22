// Angular grids, Lagrange interpolation tables, translation operators and
// phase-shift tables of the MLFMA levels (host FP64, OpenMP).
#include <omp.h>

#include <algorithm>
#include <cmath>

#include "mlfma.h"
#include "model.hpp"

namespace mlfma {

// Code/EMSource.cpp:6-32
Medium make_medium(double freq, cplx epsr, cplx mur, bool dielectric) {
    Medium m;
    const double eta0 = 119.9169832 * kPi;
    m.freq = freq;
    m.lambda = kC0 / freq;
    m.k0 = 2.0 * kPi / m.lambda;
    m.k1 = cplx(m.k0, 0.0);
    m.eta1 = m.eta2 = cplx(eta0, 0.0);
    m.etai = cplx(1.0, 0.0);
    m.k2 = cplx(0.0, 0.0);
    if (dielectric) {
        m.k2 = cplx(m.k0, 0.0) * std::sqrt(epsr * mur);
        m.etai = std::sqrt(mur / epsr);
        m.eta2 = cplx(eta0, 0.0) * std::sqrt(mur / epsr);
    }
    return m;
}

// Dunavant (1985) symmetric rules, weights halved; expansion order of the
// symmetry orbits as in Code/GaussPoints.cpp:16-24.  The 1- and 4-point rules
// use the 12-digit centroid 0.333333333333 of the reference.
bool dunavant(int n, std::vector<double>& l1, std::vector<double>& l2, std::vector<double>& l3,
              std::vector<double>& w) {
    struct O {
        int k;
        double a, b, c, w;
    };
    static const O r1[] = {{1, 0.333333333333, 0, 0, 0.5}};
    static const O r3[] = {{3, 0.666666666666667, 0.166666666666667, 0, 0.1666666666666665}};
    static const O r4[] = {{1, 0.333333333333, 0, 0, -0.28125}, {3, 0.6, 0.2, 0, 0.2604166666666665}};
    static const O r6[] = {{3, 0.108103018168070, 0.445948490915965, 0, 0.111690794839006},
                           {3, 0.816847572980459, 0.091576213509771, 0, 0.054975871827661}};
    static const O r7[] = {{1, 0.333333333333333, 0, 0, 0.1125},
                           {3, 0.059715871789770, 0.470142064105115, 0, 0.066197076394253},
                           {3, 0.797426985353087, 0.101286507323456, 0, 0.0629695902724135}};
    static const O r12[] = {{3, 0.501426509658179, 0.249286745170910, 0, 0.0583931378631895},
                            {3, 0.873821971016996, 0.063089014491502, 0, 0.0254224531851035},
                            {6, 0.053145049844817, 0.310352451033785, 0.636502499121399, 0.041425537809187}};
    static const O r13[] = {{1, 0.333333333333333, 0, 0, -0.074785044733841},
                            {3, 0.479308067841920, 0.260345966079040, 0, 0.087807628716604},
                            {3, 0.869739794195568, 0.065130102902216, 0, 0.026673617804419},
                            {6, 0.048690315425316, 0.312865496004874, 0.638444188569810, 0.0385568804451285}};
    static const O r16[] = {{1, 0.333333333333333, 0, 0, 0.0721578038388935},
                            {3, 0.081414823414554, 0.459292588292723, 0, 0.0475458171336425},
                            {3, 0.658861384496480, 0.170569307751760, 0, 0.051608685267359},
                            {3, 0.898905543365938, 0.050547228317031, 0, 0.016229248811599},
                            {6, 0.008394777409958, 0.263112829634638, 0.728492392955404, 0.0136151570872175}};
    const O* r;
    int m;
    switch (n) {
        case 1: r = r1, m = 1; break;
        case 3: r = r3, m = 1; break;
        case 4: r = r4, m = 2; break;
        case 6: r = r6, m = 2; break;
        case 7: r = r7, m = 3; break;
        case 12: r = r12, m = 3; break;
        case 13: r = r13, m = 4; break;
        case 16: r = r16, m = 5; break;
        default: return false;
    }
    l1.clear(), l2.clear(), l3.clear(), w.clear();
    auto put = [&](double a, double b, double c, double ww) {
        l1.push_back(a), l2.push_back(b), l3.push_back(c), w.push_back(ww);
    };
    for (int i = 0; i < m; ++i) {
        const O& o = r[i];
        if (o.k == 1) {
            put(o.a, o.a, o.a, o.w);
        } else if (o.k == 3) {
            put(o.a, o.b, o.b, o.w), put(o.b, o.a, o.b, o.w), put(o.b, o.b, o.a, o.w);
        } else {
            put(o.a, o.b, o.c, o.w), put(o.a, o.c, o.b, o.w), put(o.b, o.a, o.c, o.w);
            put(o.b, o.c, o.a, o.w), put(o.c, o.a, o.b, o.w), put(o.c, o.b, o.a, o.w);
        }
    }
    return true;
}

// P_n and P'_n by the three-term recurrences.
static void legendre_pd(int n, double x, double& p, double& d) {
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (n == 0) {
        p = 1.0, d = 0.0;
        return;
    }
    for (int l = 1; l < n; ++l) {
        double p2 = ((2.0 * l + 1.0) * x * p1 - l * p0) / (l + 1.0);
        double d2 = d0 + (2.0 * l + 1.0) * p1;
        p0 = p1, p1 = p2, d0 = d1, d1 = d2;
    }
    p = p1, d = d1;
}

// Code/OverloadAlgo.h:256-288: Newton from cos(pi (i + 0.75)/(n + 0.5)), stop when
// |dx| < 1e-10 (at most 100 steps); nodes ascending, weights 2/((1-x^2) P'^2).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double r = std::cos(kPi * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p, d;
            legendre_pd(n, r, p, d);
            double dx = p / d;
            r -= dx;
            if (std::fabs(dx) < 1e-10) break;
        }
        double p, d;
        legendre_pd(n, r, p, d);
        x[i] = -r;
        x[n - 1 - i] = r;
        w[i] = w[n - 1 - i] = 2.0 / ((1 - r * r) * d * d);
    }
}

// h2_l(z) = j_l(z) - j y_l(z) for complex z: h2_0 = j e^{-jz}/z, h2_1 = e^{-jz}(j/z^2 - 1/z),
// then h_{l+1} = (2l+1)/z h_l - h_{l-1} upward (stable: the Hankel functions are the dominant
// solution of the recurrence once l passes |z|).
void spherical_h2(int L, cplx z, std::vector<cplx>& h) {
    h.assign(L + 1, cplx(0, 0));
    const cplx j(0.0, 1.0);
    const cplx e = std::exp(-j * z);
    h[0] = j * e / z;
    if (L > 0) h[1] = e * (j / (z * z) - 1.0 / z);
    for (int l = 1; l < L; ++l) h[l + 1] = (2.0 * l + 1.0) / z * h[l] - h[l - 1];
}

// j_l by Miller's backward recurrence (normalised on j0 or j1), y_l forward.
void spherical_jy(int L, double x, std::vector<double>& j, std::vector<double>& y) {
    j.assign(L + 1, 0.0);
    y.assign(L + 1, 0.0);
    const double s = std::sin(x), c = std::cos(x);
    y[0] = -c / x;
    if (L > 0) y[1] = -c / (x * x) - s / x;
    for (int l = 1; l < L; ++l) y[l + 1] = (2.0 * l + 1.0) / x * y[l] - y[l - 1];
    int m = std::max(L, (int)x);
    int top = m + 30 + (int)std::sqrt(40.0 * (m + 1));
    std::vector<double> f(top + 2, 0.0);
    f[top] = 1e-300;
    for (int l = top; l >= 1; --l) {
        f[l - 1] = (2.0 * l + 1.0) / x * f[l] - f[l + 1];
        if (std::fabs(f[l - 1]) > 1e250)
            for (int q = l - 1; q <= top + 1; ++q) f[q] *= 1e-250;
    }
    const double j0 = s / x, j1 = s / (x * x) - c / x;
    const double sc = (std::fabs(j0) >= std::fabs(j1)) ? j0 / f[0] : j1 / f[1];
    for (int l = 0; l <= L; ++l) j[l] = f[l] * sc;
}

// Code/AlgoMLFMM.h:141-189: nearest sample by bisection, neighbours with a
// periodic index wrap (the coordinate itself is not unwrapped).
static void stencil(double t, const std::vector<double>& a, int out[3]) {
    const int n = (int)a.size();
    const bool asc = (a[1] - a[0]) > 0;
    int lo = 0, hi = n - 1, mid = 0;
    while (lo <= hi) {
        mid = (lo + hi) / 2;
        if (std::fabs(a[mid] - t) < 1e-12) break;
        if ((a[mid] > t) == asc) hi = mid - 1;
        else lo = mid + 1;
    }
    int c = mid;
    if (mid > 0 && std::fabs(a[mid - 1] - t) < std::fabs(a[c] - t)) c = mid - 1;
    if (mid < n - 1 && std::fabs(a[mid + 1] - t) < std::fabs(a[c] - t)) c = mid + 1;
    if (asc) {
        out[0] = (c == 0) ? n - 1 : c - 1;
        out[2] = (c == n - 1) ? 0 : c + 1;
    } else {
        out[0] = (c == 0) ? 1 : (c == n - 1 ? 0 : c + 1);
        out[2] = (c == 0) ? n - 1 : c - 1;
    }
    out[1] = c;
}

static double lagrange(double x0, double x1, double x2, double x3) {
    return ((x0 - x2) * (x0 - x3)) / ((x1 - x2) * (x1 - x3));
}

void build_wave_tables(const Octree& t, cplx k, double kabs, int tf_mode, WaveTables& W) {
    const int L = t.max_level;
    W.k = k;
    W.grid.assign(L + 1, Grid());
    W.interp.assign(L + 1, InterpTable());
    W.tf.assign(L + 1, {});
    W.qw.assign(L + 1, {});
    W.shift_up.assign(L + 1, {});
    W.shift_down.assign(L + 1, {});

    // Code/AlgoMLFMM.h:34-101
    for (int l = 2; l <= L; ++l) {
        Grid& g = W.grid[l];
        double kd = kabs * std::sqrt(3.0) * t.lv[l].len;
        g.L = (int)std::ceil(kd + 6.0 * std::cbrt(kd));
        std::vector<double> x;
        gauss_legendre(g.L, x, g.wgl);
        g.theta.resize(g.L);
        for (int j = 0; j < g.L; ++j) g.theta[j] = std::acos(x[g.L - 1 - j]);
        g.dphi = kPi / g.L;
        g.phi.resize(2 * g.L);
        for (int j = 0; j < 2 * g.L; ++j) g.phi[j] = j * g.dphi + 0.5 * g.dphi;
        g.kp.clear();
        for (double th : g.theta)
            for (double ph : g.phi) g.kp.push_back({std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)});
        W.qw[l].resize(g.K());
        for (int q = 0; q < g.K(); ++q) W.qw[l][q] = g.wgl[q / (2 * g.L)] * g.dphi;
    }

    // Code/AlgoMLFMM.h:298-368 (child level c -> parent level c-1) and the
    // transposed table used by the gather-form anterpolation.
    for (int c = 3; c <= L; ++c) {
        const Grid& P = W.grid[c - 1];
        const Grid& S = W.grid[c];
        InterpTable& it = W.interp[c];
        it.Kp = P.K();
        it.Kc = S.K();
        it.tap.assign(9 * (size_t)it.Kp, 0);
        it.w.assign(9 * (size_t)it.Kp, 0.0);
        const int pfn = (int)P.phi.size(), psn = (int)S.phi.size();
        it.Lp = (int)P.theta.size();
        it.Lc = (int)S.theta.size();
        it.th_tap.assign(3 * (size_t)it.Lp, 0);
        it.th_w.assign(3 * (size_t)it.Lp, 0.0);
        it.ph_tap.assign(3 * (size_t)pfn, 0);
        it.ph_w.assign(3 * (size_t)pfn, 0.0);
        for (int q = 0; q < it.Kp; ++q) {
            int ti[3], pi[3];
            double t0 = P.theta[q / pfn], p0 = P.phi[q % pfn];
            stencil(t0, S.theta, ti);
            stencil(p0, S.phi, pi);
            double tv[3] = {S.theta[ti[0]], S.theta[ti[1]], S.theta[ti[2]]};
            double pv[3] = {S.phi[pi[0]], S.phi[pi[1]], S.phi[pi[2]]};
            double lt[3] = {lagrange(t0, tv[0], tv[1], tv[2]), lagrange(t0, tv[1], tv[0], tv[2]),
                            lagrange(t0, tv[2], tv[0], tv[1])};
            double lp[3] = {lagrange(p0, pv[0], pv[1], pv[2]), lagrange(p0, pv[1], pv[0], pv[2]),
                            lagrange(p0, pv[2], pv[0], pv[1])};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    it.tap[(size_t)(3 * a + b) * it.Kp + q] = ti[a] * psn + pi[b];
                    it.w[(size_t)(3 * a + b) * it.Kp + q] = lt[a] * lp[b];
                }
            const int i = q / pfn, j = q % pfn;
            for (int a = 0; a < 3; ++a) {
                it.th_tap[3 * (size_t)i + a] = ti[a];
                it.th_w[3 * (size_t)i + a] = lt[a];
                it.ph_tap[3 * (size_t)j + a] = pi[a];
                it.ph_w[3 * (size_t)j + a] = lp[a];
            }
        }
        // transposed stencils by direction (CSR, ascending parent row / column)
        auto transpose = [](int np, int nc, const std::vector<int>& tap, const std::vector<double>& w,
                            std::vector<int>& ptr, std::vector<int>& idx, std::vector<double>& tw) {
            ptr.assign(nc + 1, 0);
            for (int i = 0; i < np; ++i)
                for (int a = 0; a < 3; ++a) ptr[tap[3 * (size_t)i + a] + 1]++;
            for (int c = 0; c < nc; ++c) ptr[c + 1] += ptr[c];
            idx.assign(ptr[nc], 0);
            tw.assign(ptr[nc], 0.0);
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            for (int i = 0; i < np; ++i)
                for (int a = 0; a < 3; ++a) {
                    const int c = tap[3 * (size_t)i + a];
                    idx[fill[c]] = i;
                    tw[fill[c]] = w[3 * (size_t)i + a];
                    fill[c]++;
                }
        };
        transpose(it.Lp, it.Lc, it.th_tap, it.th_w, it.at_ptr, it.at_idx, it.at_w);
        transpose(pfn, psn, it.ph_tap, it.ph_w, it.ap_ptr, it.ap_idx, it.ap_w);
        it.t_ptr.assign(it.Kc + 1, 0);
        for (int q = 0; q < it.Kp; ++q)
            for (int s = 0; s < 9; ++s) it.t_ptr[it.tap[(size_t)s * it.Kp + q] + 1]++;
        for (int i = 0; i < it.Kc; ++i) it.t_ptr[i + 1] += it.t_ptr[i];
        it.t_kp.assign(it.t_ptr[it.Kc], 0);
        it.t_w.assign(it.t_ptr[it.Kc], 0.0);
        std::vector<int> fill(it.t_ptr.begin(), it.t_ptr.end() - 1);
        for (int q = 0; q < it.Kp; ++q)
            for (int s = 0; s < 9; ++s) {
                int kc = it.tap[(size_t)s * it.Kp + q];
                it.t_kp[fill[kc]] = q;
                it.t_w[fill[kc]] = it.w[(size_t)s * it.Kp + q];
                fill[kc]++;
            }

        // phase shifts between a parent centre F and its child in octant o:
        // child - F = (+-h, +-h, +-h), h = child box / 2, sign + when the octant
        // bit is set (Code/OCTree.cpp:328-333).
        const double h = t.lv[c].len * 0.5;
        W.shift_up[c].resize(8 * (size_t)it.Kp);
        W.shift_down[c].resize(8 * (size_t)it.Kp);
        for (int o = 0; o < 8; ++o) {
            double d[3] = {(o & 1) ? h : -h, (o & 2) ? h : -h, (o & 4) ? h : -h};
            for (int q = 0; q < it.Kp; ++q) {
                const Vec3& kp = P.kp[q];
                // m2m: exp(-jk khat.(F - c))  (Code/ProcessMLFMM.h:78-92)
                cplx ku = k * ((-d[0]) * kp.x + (-d[1]) * kp.y + (-d[2]) * kp.z);
                W.shift_up[c][(size_t)o * it.Kp + q] = std::exp(-cplx(0, 1) * ku);
                // l2l: exp(-jk khat.(c - F))  (Code/ProcessMLFMM.h:173-183)
                cplx kd = k * (kp.x * d[0] + kp.y * d[1] + kp.z * d[2]);
                W.shift_down[c][(size_t)o * it.Kp + q] = std::exp(-cplx(0, 1) * kd);
            }
        }
    }

    // Code/AlgoMLFMM.h:192-295: sum_{l<=L'} (-j)^l (2l+1) h2_l(|kR|) P_l(Rhat.khat),
    // L' = min(L, ceil|kR|) (real Hankel argument even when k is complex).  tf_mode bit 1
    // keeps L' = L; bit 2 evaluates h2_l at the complex kR (accuracy mode for lossy media).
    for (int l = 2; l <= L; ++l) {
        const Grid& g = W.grid[l];
        const auto& vecs = t.lv[l].drvec;
        const int nv = (int)vecs.size(), K = g.K();
        W.tf[l].assign((size_t)nv * K, cplx(0, 0));
#pragma omp parallel for schedule(dynamic)
        for (int v = 0; v < nv; ++v) {
            double R = std::sqrt(vecs[v].x * vecs[v].x + vecs[v].y * vecs[v].y + vecs[v].z * vecs[v].z);
            double rx = vecs[v].x / R, ry = vecs[v].y / R, rz = vecs[v].z / R;
            cplx kR = k * R;
            int Lv = (tf_mode & MLFMA_TF_FULL_ORDER) ? g.L : std::min(g.L, (int)std::ceil(std::abs(kR)));
            std::vector<cplx> h2(Lv + 1);
            if (tf_mode & MLFMA_TF_COMPLEX_ARG) {
                spherical_h2(Lv, kR, h2);
            } else {
                std::vector<double> jb, yb;
                spherical_jy(Lv, std::abs(kR), jb, yb);
                for (int n = 0; n <= Lv; ++n) h2[n] = cplx(jb[n], -yb[n]);
            }
            std::vector<cplx> coef(Lv + 1);
            const cplx mj[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
            for (int n = 0; n <= Lv; ++n) coef[n] = mj[n & 3] * h2[n] * (2.0 * n + 1.0);
            for (int q = 0; q < K; ++q) {
                double x = rx * g.kp[q].x + ry * g.kp[q].y + rz * g.kp[q].z;
                double p0 = 1.0, p1 = x;
                cplx s = coef[0];
                if (Lv >= 1) s += coef[1] * p1;
                for (int n = 1; n < Lv; ++n) {
                    double p2 = ((2.0 * n + 1.0) * x * p1 - n * p0) / (n + 1.0);
                    s += coef[n + 1] * p2;
                    p0 = p1, p1 = p2;
                }
                W.tf[l][(size_t)v * K + q] = s;
            }
        }
    }
}

}  // namespace mlfma
