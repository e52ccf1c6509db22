// ORACLE -- test infrastructure only (see orc.hpp).
// Near-field Galerkin entries between two RWG functions: PMCHWT L and K
// operators (Code/Operator.h:709-1046), PEC EFIE (Code/EFIE.cpp:6-194) and
// MFIE (Code/MFIE.cpp:6-212).  Four triangle pairs, N x N Gauss points; the
// singular part is extracted analytically only when field and source triangle
// are the same face (Code/Operator.h:781,952; Code/EFIE.cpp:94).
#include <cmath>

#include "orc.hpp"

namespace orc {
namespace {

struct Tri {
    V3 v1, v2, v3, vf, n;
    double area;
    int face;
};

// i == 0 -> negative face, i == 1 -> positive face (loop order of the reference)
Tri tri_of(const Mesh& m, const Rwg& r, int i) {
    int f = (i == 0) ? r.fneg : r.fpos;
    int vf = (i == 0) ? r.vneg : r.vpos;
    Tri t;
    t.v1 = m.p[m.fv[f][0]];
    t.v2 = m.p[m.fv[f][1]];
    t.v3 = m.p[m.fv[f][2]];
    t.vf = m.p[vf];
    t.n = m.normal[f];
    t.area = m.area[f];
    t.face = m.fid[f];
    return t;
}

inline V3 gpoint(const Gauss& g, int q, const Tri& t) {
    return {g.l1[q] * t.v1.x + g.l2[q] * t.v2.x + g.l3[q] * t.v3.x,
            g.l1[q] * t.v1.y + g.l2[q] * t.v2.y + g.l3[q] * t.v3.y,
            g.l1[q] * t.v1.z + g.l2[q] * t.v2.z + g.l3[q] * t.v3.z};
}
inline V3 scale(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 add(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline double nrm(const V3& a) { return std::sqrt(dot(a, a)); }

// Analytic integrals over the source triangle of 1/R and 1/R^3 (vector and
// scalar parts) for one field point; Code/Operator.h:72-305 (1/R) and
// :585-674 (1/R^3).
struct SingTerms {
    V3 Ivec;
    double Isca = 0;
    V3 Ivec3;
    double Isca3 = 0;
    V3 rho;
};

struct SingFrame {
    V3 rho1, rho2, rho3, l[3], u[3];
};

SingFrame sing_frame(const Tri& s) {
    SingFrame f;
    const V3& n = s.n;
    f.rho1 = sub(s.v1, scale(n, dot(n, s.v1)));
    f.rho2 = sub(s.v2, scale(n, dot(n, s.v2)));
    f.rho3 = sub(s.v3, scale(n, dot(n, s.v3)));
    V3 e1 = sub(f.rho2, f.rho1), e2 = sub(f.rho3, f.rho2), e3 = sub(f.rho1, f.rho3);
    double n1 = nrm(e1), n2 = nrm(e2), n3 = nrm(e3);
    f.l[0] = {e1.x / n1, e1.y / n1, e1.z / n1};
    f.l[1] = {e2.x / n2, e2.y / n2, e2.z / n2};
    f.l[2] = {e3.x / n3, e3.y / n3, e3.z / n3};
    for (int i = 0; i < 3; ++i) f.u[i] = cross(f.l[i], n);
    return f;
}

// on_face: r is a point of the source triangle itself (the identical-face terms of
// the Galerkin entries).  Its height D above the triangle's plane is then exactly 0;
// the reference computes D = n.(r - v) in floating point (Code/Operator.h:585-674),
// whose rounding noise (|D| ~ 1e-16 |r|) decides between the two branches of the
// 1/R^3 scalar integral, one of which divides an O(1) angle difference by that noise.
// The exact D = 0 gives the reference's value wherever its noise stays below
// 10 eps0, and a noise-free value where it does not.
SingTerms sing_terms(const V3& r, const Tri& s, const SingFrame& fr, bool want3, bool on_face) {
    const double th = 50 * kEps0;
    SingTerms out;
    const V3& n = s.n;
    V3 rho = sub(r, scale(n, dot(n, r)));
    out.rho = rho;
    const V3* rp[3] = {&fr.rho2, &fr.rho3, &fr.rho1};   // "plus" end of edge i
    const V3* rm[3] = {&fr.rho1, &fr.rho2, &fr.rho3};   // "minus" end of edge i
    const V3* sv[3] = {&s.v1, &s.v2, &s.v3};
    double Lpl[3], Lmn[3], P0[3], Ppl[3], Pmn[3], D[3], R0[3], Rpl[3], Rmn[3];
    V3 P0v[3];
    for (int i = 0; i < 3; ++i) {
        V3 dp = sub(*rp[i], rho), dm = sub(*rm[i], rho);
        Lpl[i] = dot(dp, fr.l[i]);
        Lmn[i] = dot(dm, fr.l[i]);
        P0[i] = std::fabs(dot(dp, fr.u[i]));
        Ppl[i] = nrm(dp);
        Pmn[i] = nrm(dm);
        if (P0[i] < th) {
            P0v[i] = {0, 0, 0};
        } else {
            V3 pom = sub(dp, scale(fr.l[i], Lpl[i]));
            P0v[i] = {pom.x / P0[i], pom.y / P0[i], pom.z / P0[i]};
        }
        D[i] = on_face ? 0.0 : dot(n, sub(r, *sv[i]));
    }
    for (int i = 0; i < 3; ++i) {
        R0[i] = std::sqrt(P0[i] * P0[i] + D[i] * D[i]);
        Rpl[i] = std::sqrt(Ppl[i] * Ppl[i] + D[i] * D[i]);
        Rmn[i] = std::sqrt(Pmn[i] * Pmn[i] + D[i] * D[i]);
    }
    double LOGp[3], LOGm[3];
    for (int i = 0; i < 3; ++i) {
        LOGp[i] = (Rpl[i] + Lpl[i] < th) ? 0.0 : std::log(Rpl[i] + Lpl[i]);
        LOGm[i] = (Rmn[i] + Lmn[i] < th) ? 0.0 : std::log(Rmn[i] + Lmn[i]);
    }
    V3 Iv = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        double c1 = 0.5 * (R0[i] * R0[i] * (LOGp[i] - LOGm[i]) + Rpl[i] * Lpl[i] - Rmn[i] * Lmn[i]);
        Iv = add(Iv, scale(fr.u[i], c1));
    }
    double Is = 0;
    for (int i = 0; i < 3; ++i) {
        double ad = std::fabs(D[i]);
        double denp = R0[i] * R0[i] + ad * Rpl[i];
        double denm = R0[i] * R0[i] + ad * Rmn[i];
        double atp = (denp < th) ? 0.0 : std::atan((P0[i] * Lpl[i]) / denp);
        double atm = (denm < th) ? 0.0 : std::atan((P0[i] * Lmn[i]) / denm);
        Is = Is + dot(P0v[i], fr.u[i]) * (P0[i] * (LOGp[i] - LOGm[i]) - ad * (atp - atm));
    }
    out.Ivec = Iv;
    out.Isca = Is;
    if (want3) {
        V3 Iv3 = {0, 0, 0};
        for (int i = 0; i < 3; ++i) Iv3 = sub(Iv3, scale(fr.u[i], LOGp[i] - LOGm[i]));
        double Is3 = 0;
        const double t10 = 10.0 * kEps0;
        for (int i = 0; i < 3; ++i) {
            double ad = std::fabs(D[i]);
            if (ad >= t10) {
                double atp = std::atan((P0[i] * Lpl[i]) / (R0[i] * R0[i] + ad * Rpl[i]));
                double atm = std::atan((P0[i] * Lmn[i]) / (R0[i] * R0[i] + ad * Rmn[i]));
                Is3 = Is3 + dot(P0v[i], fr.u[i]) * ((1.0 / ad) * (atp - atm));
            } else if (P0[i] >= t10) {
                Is3 = Is3 - dot(P0v[i], fr.u[i]) * (Lpl[i] / (P0[i] * Rpl[i]) - Lmn[i] / (P0[i] * Rmn[i]));
            } else {
                Is3 = 0.0;   // the reference resets the accumulator here
            }
        }
        out.Ivec3 = Iv3;
        out.Isca3 = Is3;
    }
    return out;
}

// Ikon of singK / singularityMFIE: (k^2/2)(Ivec + (rho - rho0N) Isca)/(2A_s) + (Ivec3 + (rho - rho0N) Isca3)/(2A_s)
void ikon_k(const SingTerms& st, const V3& rho0N, cd k, double recAs, cd out[3]) {
    V3 pv1 = sub(st.rho, rho0N);
    cd konst1 = k * k * 0.5;
    cd konst2 = konst1 * st.Isca;
    double iv[3] = {st.Ivec.x, st.Ivec.y, st.Ivec.z};
    double p1[3] = {pv1.x, pv1.y, pv1.z};
    double iv3[3] = {st.Ivec3.x, st.Ivec3.y, st.Ivec3.z};
    for (int a = 0; a < 3; ++a) {
        cd ik1 = (iv[a] * konst1 + p1[a] * konst2) * recAs;
        double ik3 = (iv3[a] + p1[a] * st.Isca3) * recAs;
        out[a] = ik3 + ik1;
    }
}

inline cd cdot3(const V3& a, const cd b[3]) { return a.x * b[0] + a.y * b[1] + a.z * b[2]; }

}  // namespace

// Code/Operator.h:709-881
cd op_L(const Mesh& m, const Rwg& f, const Rwg& s, cd k, const Gauss& gp) {
    cd Z(0, 0);
    for (int i = 0; i < 2; ++i) {
        Tri F = tri_of(m, f, i);
        double recAf = 1.0 / (2.0 * F.area);
        for (int j = 0; j < 2; ++j) {
            Tri S = tri_of(m, s, j);
            double recAs = 1.0 / (2.0 * S.area);
            cd alok1(0, 0);
            bool same = (F.face == S.face);
            for (int a = 0; a < gp.n; ++a) {
                V3 rf = gpoint(gp, a, F);
                double wF = gp.w[a];
                V3 RWGf = scale(sub(rf, F.vf), recAf);
                for (int b = 0; b < gp.n; ++b) {
                    V3 rs = gpoint(gp, b, S);
                    double wS = gp.w[b];
                    double R = nrm(sub(rf, rs));
                    cd G;
                    if (!same) {
                        G = std::exp(-kJ * k * R) / R;
                    } else {
                        G = (R < kEps0) ? -kJ * k : (std::exp(-kJ * k * R) / R) - 1.0 / R;
                    }
                    V3 RWGs = scale(sub(rs, S.vf), recAs);
                    alok1 += kJ * k * wF * wS * 4.0 * F.area * S.area *
                             (G * dot(RWGf, RWGs) - (1.0 / (k * k)) * G * (1.0 / (F.area * S.area)));
                }
            }
            if (same) {
                SingFrame fr = sing_frame(S);
                V3 rho0N = sub(S.vf, scale(S.n, dot(S.n, S.vf)));
                cd alok(0, 0);
                for (int a = 0; a < gp.n; ++a) {
                    V3 rf = gpoint(gp, a, F);
                    double wF = gp.w[a];
                    SingTerms st = sing_terms(rf, S, fr, false, true);
                    V3 RWGf = scale(sub(rf, F.vf), recAf);
                    V3 Ikon = scale(add(st.Ivec, scale(sub(st.rho, rho0N), st.Isca)), recAs);
                    alok += kJ * k * wF * 2.0 * F.area *
                            (dot(RWGf, Ikon) - (1.0 / (k * k)) * (double(1.0 / (F.area * S.area))) * st.Isca);
                }
                alok1 += alok;
            }
            double sg = double(((i == 0) ? -1 : 1) * ((j == 0) ? -1 : 1));
            Z += (f.len * s.len) / (4.0 * kPi) * sg * alok1;
        }
    }
    return Z;
}

// Code/EFIE.cpp:6-194 and :196-508
cd op_EFIE(const Mesh& m, const Rwg& f, const Rwg& s, cd k, cd eta, const Gauss& gp) {
    cd Z(0, 0);
    for (int i = 0; i < 2; ++i) {
        Tri F = tri_of(m, f, i);
        double recAf = 1.0 / (2.0 * F.area);
        for (int j = 0; j < 2; ++j) {
            Tri S = tri_of(m, s, j);
            double recAs = 1.0 / (2.0 * S.area);
            cd alok1(0, 0);
            bool same = (F.face == S.face);
            for (int a = 0; a < gp.n; ++a) {
                V3 rf = gpoint(gp, a, F);
                double wF = gp.w[a];
                V3 RWGf = scale(sub(rf, F.vf), recAf);
                for (int b = 0; b < gp.n; ++b) {
                    V3 rs = gpoint(gp, b, S);
                    double wS = gp.w[b];
                    double R = nrm(sub(rf, rs));
                    cd G;
                    if (!same) {
                        G = std::exp(-kJ * k * R) / R;
                    } else {
                        G = (R < kEps0) ? -kJ * k : (std::exp(-kJ * k * R) / R) - 1.0 / R;
                    }
                    V3 RWGs = scale(sub(rs, S.vf), recAs);
                    alok1 += wF * wS * 4.0 * F.area * S.area *
                             (kJ * k * eta * G * dot(RWGf, RWGs) - ((kJ * eta) / k) * G * (1.0 / (F.area * S.area)));
                }
            }
            if (same) {
                SingFrame fr = sing_frame(S);
                V3 rho0N = sub(S.vf, scale(S.n, dot(S.n, S.vf)));
                cd alok(0, 0);
                for (int a = 0; a < gp.n; ++a) {
                    V3 rf = gpoint(gp, a, F);
                    double wF = gp.w[a];
                    SingTerms st = sing_terms(rf, S, fr, false, true);
                    V3 RWGf = scale(sub(rf, F.vf), recAf);
                    V3 Ikon = scale(add(st.Ivec, scale(sub(st.rho, rho0N), st.Isca)), recAs);
                    alok += wF * 2.0 * F.area *
                            (kJ * k * eta * dot(RWGf, Ikon) -
                             ((kJ * eta) / k) * (double(1.0 / F.area)) * (double(1.0 / S.area)) * st.Isca);
                }
                alok1 += alok;
            }
            double sg = double(((i == 0) ? -1 : 1) * ((j == 0) ? -1 : 1));
            Z += (f.len * s.len) / (4.0 * kPi) * sg * alok1;
        }
    }
    return Z;
}

// Code/Operator.h:883-1046 (with singK :325-707)
cd op_K(const Mesh& m, const Rwg& f, const Rwg& s, cd k, const Gauss& gp) {
    cd Z(0, 0);
    for (int i = 0; i < 2; ++i) {
        Tri F = tri_of(m, f, i);
        double recAf = 1.0 / (2.0 * F.area);
        for (int j = 0; j < 2; ++j) {
            Tri S = tri_of(m, s, j);
            double recAs = 1.0 / (2.0 * S.area);
            cd alok1(0, 0);
            bool same = (F.face == S.face);
            for (int a = 0; a < gp.n; ++a) {
                V3 rf = gpoint(gp, a, F);
                double wF = gp.w[a];
                V3 pf = sub(rf, F.vf);
                for (int b = 0; b < gp.n; ++b) {
                    V3 rs = gpoint(gp, b, S);
                    double wS = gp.w[b];
                    V3 Rv = sub(rf, rs);
                    double R = nrm(Rv);
                    cd G;
                    if (!same) {
                        G = (std::exp(-kJ * k * R) / (R * R * R)) * (1.0 + kJ * k * R);
                    } else if (R < kEps0) {
                        G = cd(0, 0);
                    } else {
                        G = (std::exp(-kJ * k * R) / (R * R * R)) * (1.0 + kJ * k * R) - 1.0 / (R * R * R) -
                            0.5 * (k * k) * (1.0 / R);
                    }
                    V3 ps = sub(rs, S.vf);
                    V3 c1 = cross(Rv, ps);
                    alok1 += wF * wS * G * dot(pf, c1);
                }
            }
            if (same) {
                SingFrame fr = sing_frame(S);
                V3 rho0N = sub(S.vf, scale(S.n, dot(S.n, S.vf)));
                cd alok(0, 0);
                for (int a = 0; a < gp.n; ++a) {
                    V3 rf = gpoint(gp, a, F);
                    double wF = gp.w[a];
                    SingTerms st = sing_terms(rf, S, fr, true, true);
                    V3 RWGf = scale(sub(rf, F.vf), recAf);
                    V3 Ri = sub(rf, S.vf);
                    cd Ik[3];
                    ikon_k(st, rho0N, k, recAs, Ik);
                    cd K1[3] = {Ri.y * Ik[2] - Ri.z * Ik[1], Ri.z * Ik[0] - Ri.x * Ik[2], Ri.x * Ik[1] - Ri.y * Ik[0]};
                    alok += wF * 2.0 * F.area * cdot3(RWGf, K1);
                }
                alok1 += alok;
            }
            double sg = ((i == 0) ? -1.0 : 1.0) * ((j == 0) ? -1.0 : 1.0);
            Z += sg * (f.len * s.len) / (4.0 * kPi) * alok1;
        }
    }
    return Z;
}

// Code/MFIE.cpp:6-212 (with singularityMFIE :212-595)
cd op_MFIE(const Mesh& m, const Rwg& f, const Rwg& s, cd k, const Gauss& gp) {
    cd Z(0, 0);
    for (int i = 0; i < 2; ++i) {
        Tri F = tri_of(m, f, i);
        double recAf = 1.0 / (2.0 * F.area);
        for (int j = 0; j < 2; ++j) {
            Tri S = tri_of(m, s, j);
            double recAs = 1.0 / (2.0 * S.area);
            cd alok1(0, 0);
            bool same = (F.face == S.face);
            for (int a = 0; a < gp.n; ++a) {
                V3 rf = gpoint(gp, a, F);
                double wF = gp.w[a];
                V3 pf = sub(rf, F.vf);
                for (int b = 0; b < gp.n; ++b) {
                    V3 rs = gpoint(gp, b, S);
                    double wS = gp.w[b];
                    V3 Rv = sub(rf, rs);
                    double R = nrm(Rv);
                    V3 ps = sub(rs, S.vf);
                    if (!same) {
                        cd gG = (std::exp(-kJ * k * R) / (R * R * R)) * (1.0 + kJ * k * R);
                        V3 c1 = cross(Rv, ps);
                        V3 c2 = cross(F.n, c1);
                        alok1 += wF * wS * (gG * dot(pf, c2));
                    } else {
                        cd G = (R < kEps0) ? cd(0, 0)
                                           : (std::exp(-kJ * k * R) / (R * R * R)) * (1.0 + kJ * k * R) -
                                                 1.0 / (R * R * R) - 0.5 * (k * k) * (1.0 / R);
                        V3 Ri = sub(rf, S.vf);
                        V3 c1 = cross(Ri, ps);
                        V3 c2 = cross(pf, F.n);
                        alok1 += wF * wS * (G * dot(c1, c2));
                    }
                }
            }
            if (same) {
                cd alok(0, 0);
                for (int a = 0; a < gp.n; ++a) {   // identity term
                    V3 rf = gpoint(gp, a, F);
                    double wF = gp.w[a];
                    V3 RWGf = scale(sub(rf, F.vf), recAf);
                    V3 RWGs = scale(sub(rf, S.vf), recAs);
                    alok += 4.0 * kPi * wF * F.area * dot(RWGf, RWGs);
                }
                SingFrame fr = sing_frame(S);
                V3 rho0N = sub(S.vf, scale(S.n, dot(S.n, S.vf)));
                for (int a = 0; a < gp.n; ++a) {
                    V3 rf = gpoint(gp, a, F);
                    double wF = gp.w[a];
                    SingTerms st = sing_terms(rf, S, fr, true, true);
                    V3 RWGf = scale(sub(rf, F.vf), recAf);
                    V3 Ri = sub(rf, S.vf);
                    cd Ik[3];
                    ikon_k(st, rho0N, k, recAs, Ik);
                    cd K1[3] = {Ri.y * Ik[2] - Ri.z * Ik[1], Ri.z * Ik[0] - Ri.x * Ik[2], Ri.x * Ik[1] - Ri.y * Ik[0]};
                    V3 K2 = cross(RWGf, F.n);
                    alok += wF * 2.0 * F.area * cdot3(K2, K1);
                }
                alok1 += alok;
            }
            double sg = ((i == 0) ? -1.0 : 1.0) * ((j == 0) ? -1.0 : 1.0);
            Z += (f.len * s.len) / (4.0 * kPi) * sg * alok1;
        }
    }
    return Z;
}

}  // namespace orc

namespace orc {
// Analytic triangle integrals at one observation point (test access):
// out = Ivec(3), Isca, Ivec3(3), Isca3, rho(3).
void sing_probe(const double* tri9, const double* n3, const double* r3, double* out) {
    Tri S;
    S.v1 = {tri9[0], tri9[1], tri9[2]};
    S.v2 = {tri9[3], tri9[4], tri9[5]};
    S.v3 = {tri9[6], tri9[7], tri9[8]};
    S.n = {n3[0], n3[1], n3[2]};
    S.vf = S.v1;
    S.area = 0;
    S.face = 0;
    SingFrame fr = sing_frame(S);
    SingTerms st = sing_terms({r3[0], r3[1], r3[2]}, S, fr, true, false);
    const double o[11] = {st.Ivec.x, st.Ivec.y, st.Ivec.z, st.Isca, st.Ivec3.x, st.Ivec3.y, st.Ivec3.z, st.Isca3,
                          st.rho.x, st.rho.y, st.rho.z};
    for (int i = 0; i < 11; ++i) out[i] = o[i];
}
}  // namespace orc
