// ORACLE -- test infrastructure only (see orc.hpp).
// Quadrature rules, special functions and the plane-wave source.
#include <cmath>
#include <stdexcept>

#include "orc.hpp"

namespace orc {

// Code/GaussPoints.cpp:11-301 -- Dunavant (1985) symmetric triangle rules with the
// weights halved (sum = 0.5).  Stored here as symmetry orbits and expanded in the
// reference's point order: (a,b,b),(b,a,b),(b,b,a) for 3-orbits and
// (a,b,c),(a,c,b),(b,a,c),(b,c,a),(c,a,b),(c,b,a) for 6-orbits.  The 1- and
// 4-point centroid coordinate is 0.333333333333 (12 digits) as in the reference.
namespace {
struct Orbit {
    int kind;          // 1 centroid, 3 or 6
    double a, b, c, w;
};
void push(Gauss& g, double l1, double l2, double l3, double w) {
    g.l1.push_back(l1);
    g.l2.push_back(l2);
    g.l3.push_back(l3);
    g.w.push_back(w);
}
}  // namespace

Gauss gauss_points(int n) {
    std::vector<Orbit> o;
    const double c12 = 0.333333333333, c15 = 0.333333333333333;
    switch (n) {
        case 1: o = {{1, c12, 0, 0, 0.5}}; break;
        case 3: o = {{3, 0.666666666666667, 0.166666666666667, 0, 0.1666666666666665}}; break;
        case 4: o = {{1, c12, 0, 0, -0.281250000000000}, {3, 0.6, 0.2, 0, 0.2604166666666665}}; break;
        case 6:
            o = {{3, 0.108103018168070, 0.445948490915965, 0, 0.111690794839006},
                 {3, 0.816847572980459, 0.091576213509771, 0, 0.054975871827661}};
            break;
        case 7:
            o = {{1, c15, 0, 0, 0.112500000000000},
                 {3, 0.059715871789770, 0.470142064105115, 0, 0.066197076394253},
                 {3, 0.797426985353087, 0.101286507323456, 0, 0.0629695902724135}};
            break;
        case 12:
            o = {{3, 0.501426509658179, 0.249286745170910, 0, 0.0583931378631895},
                 {3, 0.873821971016996, 0.063089014491502, 0, 0.0254224531851035},
                 {6, 0.053145049844817, 0.310352451033785, 0.636502499121399, 0.041425537809187}};
            break;
        case 13:
            o = {{1, c15, 0, 0, -0.074785044733841},
                 {3, 0.479308067841920, 0.260345966079040, 0, 0.087807628716604},
                 {3, 0.869739794195568, 0.065130102902216, 0, 0.026673617804419},
                 {6, 0.048690315425316, 0.312865496004874, 0.638444188569810, 0.0385568804451285}};
            break;
        case 16:
            o = {{1, c15, 0, 0, 0.0721578038388935},
                 {3, 0.081414823414554, 0.459292588292723, 0, 0.0475458171336425},
                 {3, 0.658861384496480, 0.170569307751760, 0, 0.051608685267359},
                 {3, 0.898905543365938, 0.050547228317031, 0, 0.016229248811599},
                 {6, 0.008394777409958, 0.263112829634638, 0.728492392955404, 0.0136151570872175}};
            break;
        default: throw std::invalid_argument("Unsupported Gauss point count");
    }
    Gauss g;
    g.n = n;
    for (const Orbit& r : o) {
        if (r.kind == 1) {
            push(g, r.a, r.a, r.a, r.w);
        } else if (r.kind == 3) {
            push(g, r.a, r.b, r.b, r.w);
            push(g, r.b, r.a, r.b, r.w);
            push(g, r.b, r.b, r.a, r.w);
        } else {
            push(g, r.a, r.b, r.c, r.w);
            push(g, r.a, r.c, r.b, r.w);
            push(g, r.b, r.a, r.c, r.w);
            push(g, r.b, r.c, r.a, r.w);
            push(g, r.c, r.a, r.b, r.w);
            push(g, r.c, r.b, r.a, r.w);
        }
    }
    return g;
}

// Three-term recurrence (Bonnet); same recurrence as Boost.Math legendre_p for
// integer order (Boost 1.88 legendre_next).
double legendre_p(int l, double x) {
    if (l == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int n = 1; n < l; ++n) {
        double p2 = ((2.0 * n + 1.0) * x * p1 - n * p0) / (n + 1.0);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// P'_{l+1} = P'_{l-1} + (2l+1) P_l
double legendre_p_prime(int l, double x) {
    if (l == 0) return 0.0;
    double pm1 = 1.0, p = x;     // P_{n-1}, P_n
    double dm1 = 0.0, d = 1.0;   // P'_{n-1}, P'_n
    for (int n = 1; n < l; ++n) {
        double pn1 = ((2.0 * n + 1.0) * x * p - n * pm1) / (n + 1.0);
        double dn1 = dm1 + (2.0 * n + 1.0) * p;
        pm1 = p;
        p = pn1;
        dm1 = d;
        d = dn1;
    }
    return d;
}

// Code/OverloadAlgo.h:256-288 -- Newton on P_n from cos(pi (i+0.75)/(n+0.5)),
// tolerance 1e-10, 100 iterations; x ascending, symmetric weights.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double r = std::cos(kPi * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double dx = legendre_p(n, r) / legendre_p_prime(n, r);
            r -= dx;
            if (std::fabs(dx) < 1e-10) break;
        }
        x[i] = -r;
        x[n - i - 1] = r;
        double dp = legendre_p_prime(n, r);
        double wt = 2.0 / ((1 - r * r) * dp * dp);
        w[i] = wt;
        w[n - i - 1] = wt;
    }
}

// Spherical Bessel j_l and y_l of a real positive argument, l = 0..L.
// y_l: upward recurrence (stable).  j_l: Miller's downward recurrence started
// well above max(L, x), normalised to j_0 or j_1 whichever is larger.
// (The reference calls Boost.Math sph_bessel / sph_neumann, Code/Bessel.h:20-47.)
void sph_jy(int L, double x, std::vector<double>& j, std::vector<double>& y) {
    j.assign(L + 1, 0.0);
    y.assign(L + 1, 0.0);
    double s = std::sin(x), c = std::cos(x);
    y[0] = -c / x;
    if (L >= 1) y[1] = -c / (x * x) - s / x;
    for (int l = 1; l < L; ++l) y[l + 1] = (2.0 * l + 1.0) / x * y[l] - y[l - 1];

    int top = std::max(L, (int)x) + 30 + (int)std::sqrt(40.0 * (std::max(L, (int)x) + 1));
    std::vector<double> f(top + 2, 0.0);
    f[top + 1] = 0.0;
    f[top] = 1e-300;
    for (int l = top; l >= 1; --l) {
        f[l - 1] = (2.0 * l + 1.0) / x * f[l] - f[l + 1];
        if (std::fabs(f[l - 1]) > 1e250) {   // rescale to avoid overflow
            for (int q = l - 1; q <= top + 1; ++q) f[q] *= 1e-250;
        }
    }
    double j0 = s / x, j1 = s / (x * x) - c / x;
    double scale = (std::fabs(j0) >= std::fabs(j1)) ? j0 / f[0] : j1 / f[1];
    for (int l = 0; l <= L; ++l) j[l] = f[l] * scale;
}

// Spherical Hankel functions of the second kind h2_l(z), l = 0..L, for a complex
// argument z (accuracy mode of the translation operator; the reference evaluates
// its series at the real |kR|, Code/AlgoMLFMM.h:262-265).  Closed forms
// h2_0 = j e^{-jz}/z, h2_1 = e^{-jz}(j/z^2 - 1/z) and the upward recurrence
// h_{l+1} = (2l+1)/z h_l - h_{l-1}, which is stable for the Hankel functions
// (they are the dominant solution once l exceeds |z|).
void sph_h2(int L, cd z, std::vector<cd>& h) {
    h.assign(L + 1, cd(0, 0));
    const cd j(0.0, 1.0);
    cd e = std::exp(-j * z);
    h[0] = j * e / z;
    if (L >= 1) h[1] = e * (j / (z * z) - 1.0 / z);
    for (int l = 1; l < L; ++l) h[l + 1] = (2.0 * l + 1.0) / z * h[l] - h[l - 1];
}

// Code/EMSource.cpp:6-32 -- k1 = 2 pi / lambda; k2 = k1 sqrt(eps mu);
// etai = sqrt(mu/eps); eta0 = 119.9169832 pi.
Wave make_wave(double freq, cd epsr, cd mur, bool dielectric) {
    Wave w;
    const double eta0 = 119.9169832 * kPi;
    w.freq = freq;
    w.lambda = kC0 / freq;
    w.k0 = 2.0 * kPi / w.lambda;
    w.k1 = cd(w.k0, 0.0);
    w.eta1 = cd(eta0, 0.0);
    w.eta2 = cd(eta0, 0.0);
    w.etai = cd(1.0, 0.0);
    w.k2 = cd(0.0, 0.0);
    if (dielectric) {
        w.epsr = epsr;
        w.mur = mur;
        w.k2 = cd(w.k0, 0.0) * std::sqrt(epsr * mur);
        w.etai = std::sqrt(mur / epsr);
        w.eta2 = cd(eta0, 0.0) * std::sqrt(mur / epsr);
    }
    return w;
}

// Code/EMSource.cpp:34-72 -- incident direction -r_hat, v = -theta_hat,
// h = phi_hat, E = v (pol 0) or h (pol 90), H = normalise(k x E).
PlaneWave init_pw(double pol_deg, double th_deg, double ph_deg) {
    auto rad = [](double a) {
        double b = a * kPi / 180.0;
        return std::fabs(b) < 1e-6 ? 0.0 : b;
    };
    auto snap = [](double& a) {
        if (std::fabs(a) < 1e-6) a = 0.0;
    };
    PlaneWave pw;
    double pol = rad(pol_deg), th = rad(th_deg), ph = rad(ph_deg);
    pw.kinc[0] = -std::sin(th) * std::cos(ph);
    pw.kinc[1] = -std::sin(th) * std::sin(ph);
    pw.kinc[2] = -std::cos(th);
    pw.vth[0] = -std::cos(th) * std::cos(ph);
    pw.vth[1] = -std::cos(th) * std::sin(ph);
    pw.vth[2] = std::sin(th);
    pw.vph[0] = -std::sin(ph);
    pw.vph[1] = std::cos(ph);
    pw.vph[2] = 0.0;
    for (int i = 0; i < 3; ++i) {
        snap(pw.kinc[i]);
        snap(pw.vth[i]);
        snap(pw.vph[i]);
        pw.einc[i] = pw.hinc[i] = 0.0;
    }
    auto setEH = [&](const double* e) {
        for (int i = 0; i < 3; ++i) pw.einc[i] = e[i];
        const double* a = pw.kinc;
        double h[3] = {a[1] * e[2] - a[2] * e[1], a[2] * e[0] - a[0] * e[2], a[0] * e[1] - a[1] * e[0]};
        double nn = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        if (nn > 1e-15)
            for (int i = 0; i < 3; ++i) h[i] /= nn;
        for (int i = 0; i < 3; ++i) pw.hinc[i] = h[i];
    };
    if (std::fabs(pol - rad(90.0)) < 1e-6) setEH(pw.vph);
    if (std::fabs(pol - 0.0) < 1e-6) setEH(pw.vth);
    return pw;
}

}  // namespace orc
