// ORACLE -- test infrastructure only (see orc.hpp).
//
// Checks the oracle against the parts of the reference that compile here
// without any stand-in: Code/ProcessMLFMM.h (stages), Code/GMRES.h,
// Code/CGS.h, Code/GaussPoints.cpp and Code/EMSource.cpp (+ Code/Complex3D.cpp),
// and Code/RWGGenerator.cpp (which needs only -fpermissive: its file-static
// FindPoint differs in linkage from the declaration in Code/NASReader.h).
// Everything else in the reference needs Boost.Math, which this image lacks.
//
// The reference headers are included from /root/reference/Code at build time
// (see oracle/Makefile); nothing of them is copied here.  The binary prints one
// JSON object per sub-command on stdout.
//
//   ref_harness gauss
//   ref_harness emsource
//   ref_harness stages <nas> <ie> <eps_re> <eps_im> <seed>
//   ref_harness solvers <nas> <ie> <eps_re> <eps_im>
//   ref_harness solution <nas> <ie> <eps_re> <eps_im> <out.bin>
//   ref_harness rwg <nas> [out.bin]   (out: int32 n x 6 file-order indices, then float64 n lengths)
#include "PhyConstant.h"
#include "MyStruct.h"
#include "OCTree.h"
#include "ProcessMLFMM.h"
#include "GMRES.h"
#include "CGS.h"
#include "GaussPoints.h"
#include "EMSource.h"
#include "RWGGenerator.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "orc_oracle.hpp"

namespace {

using cdd = std::complex<double>;

int cmd_gauss() {
    const int ns[] = {1, 3, 4, 6, 7, 12, 13, 16};
    long mismatches = 0;
    for (int n : ns) {
        gaussPoints g(n);
        orc::Gauss o = orc::gauss_points(n);
        for (int i = 0; i < n; ++i) {
            if (g.l1[i] != o.l1[i] || g.l2[i] != o.l2[i] || g.l3[i] != o.l3[i] || g.weight[i] != o.w[i]) ++mismatches;
        }
    }
    std::printf("{\"cmd\":\"gauss\",\"mismatches\":%ld}\n", mismatches);
    return 0;
}

int cmd_emsource() {
    long mism = 0;
    double maxd = 0;
    const double freqs[] = {299792458.0, 1.0e9, 3.0e8};
    const cdd eps[] = {{4.0, 0.0}, {2.5, -0.1}, {8.0, 0.0}};
    const cdd mus[] = {{1.0, 0.0}, {1.0, 0.0}, {2.0, -0.3}};
    for (double f : freqs)
        for (int e = 0; e < 3; ++e) {
            EMSource w(f, 0.0);
            w.initDielectric(eps[e], mus[e]);
            orc::Wave o = orc::make_wave(f, eps[e], mus[e], true);
            cdd a[5] = {w.k1(), w.k2(), w.eta1(), w.eta2(), w.etai()};
            cdd b[5] = {o.k1, o.k2, o.eta1, o.eta2, o.etai};
            for (int i = 0; i < 5; ++i) {
                if (a[i] != b[i]) ++mism;
                maxd = std::max(maxd, std::abs(a[i] - b[i]) / std::max(1e-300, std::abs(a[i])));
            }
        }
    const double pols[] = {0.0, 90.0};
    const double ths[] = {0.0, 30.0, 90.0, 137.5, 180.0};
    const double phs[] = {0.0, 45.0, 90.0, 270.0};
    for (double p : pols)
        for (double t : ths)
            for (double ph : phs) {
                EMSource w;
                w.initPW(p, t, ph);
                orc::PlaneWave o = orc::init_pw(p, t, ph);
                for (int i = 0; i < 3; ++i) {
                    double a[5] = {w.vW(i), w.vE(i), w.vH(i), w.vTh(i), w.vPh(i)};
                    double b[5] = {o.kinc[i], o.einc[i], o.hinc[i], o.vth[i], o.vph[i]};
                    for (int q = 0; q < 5; ++q) {
                        if (a[q] != b[q]) ++mism;
                        maxd = std::max(maxd, std::fabs(a[q] - b[q]));
                    }
                }
            }
    std::printf("{\"cmd\":\"emsource\",\"mismatches\":%ld,\"max_diff\":%.3e}\n", mism, maxd);
    return 0;
}

// Reference-typed copy of the oracle's tree and tables.
struct RefView {
    std::vector<RWGBase> rwgs;
    std::vector<std::vector<OCTree::Node>> store;
    std::vector<std::vector<OCTree::Node*>> nodes;
    std::vector<std::vector<kp_Point>> kp1, kp2;
    std::vector<std::vector<InterpData>> ip1, ip2;
    std::vector<std::vector<double>> wgl1, wgl2, phi1, phi2;
    std::vector<double> wphi1, wphi2;
    std::vector<std::vector<Complex3D>> Vs1, Vs2;
    std::vector<std::vector<std::vector<cdd>>> TF1, TF2;
};

Complex3D toC3(const orc::C3& c) { return Complex3D(c.x, c.y, c.z); }

void build_view(const orc::Oracle& o, RefView& v) {
    const auto& t = o.tree;
    v.rwgs.resize(o.N);
    for (int i = 0; i < o.N; ++i) {
        std::memset(&v.rwgs[i], 0, sizeof(RWGBase));
        v.rwgs[i].rwgid = i;
    }
    v.store.resize(t.maxLevel + 1);
    v.nodes.resize(t.maxLevel + 1);
    for (int l = 0; l <= t.maxLevel; ++l) {
        v.store[l].resize(t.lv[l].size());
        for (size_t i = 0; i < t.lv[l].size(); ++i) {
            OCTree::Node& n = v.store[l][i];
            const orc::Node& s = t.lv[l][i];
            n.nodeNumber = (int)i;
            n.level = l;
            n.cubeLength = s.len;
            n.center = nodePoint(s.c.x, s.c.y, s.c.z);
            n.mortonCode = s.code;
            for (int r : s.rwg) n.rwgIndices.push_back(&v.rwgs[r]);
        }
    }
    for (int l = 0; l <= t.maxLevel; ++l) {
        for (size_t i = 0; i < t.lv[l].size(); ++i) {
            OCTree::Node& n = v.store[l][i];
            const orc::Node& s = t.lv[l][i];
            if (l < t.maxLevel)
                for (int c : s.sons) n.realSons.push_back(&v.store[l + 1][c]);
            for (int d : s.dr) n.distantRelatives.push_back(&v.store[l][d]);
            n.dRVecId = s.drv;
            v.nodes[l].push_back(&n);
        }
    }
    auto grids = [&](const std::vector<orc::Grid>& g, std::vector<std::vector<kp_Point>>& kp,
                     std::vector<std::vector<double>>& wgl, std::vector<std::vector<double>>& phi,
                     std::vector<double>& wphi) {
        for (const auto& G : g) {
            std::vector<kp_Point> k;
            for (const auto& p : G.kp) k.emplace_back(p.x, p.y, p.z);
            kp.push_back(k);
            wgl.push_back(G.wgl);
            phi.push_back(G.phi);
            wphi.push_back(G.dphi);
        }
    };
    auto interp = [&](const std::vector<std::vector<orc::Interp>>& ip, const std::vector<orc::Grid>& g,
                      std::vector<std::vector<InterpData>>& out) {
        int span = t.maxLevel - 1;
        for (size_t level = 0; level < ip.size(); ++level) {
            int psn = (int)g[span - 2 - level].phi.size();
            std::vector<InterpData> row;
            for (size_t q = 0; q < ip[level].size(); ++q) {
                const orc::Interp& d = ip[level][q];
                std::array<double, 9> w;
                for (int a = 0; a < 9; ++a) w[a] = d.w[a];
                row.emplace_back((int)q, std::array<int, 3>{d.tIdx[0], d.tIdx[1], d.tIdx[2]},
                                 std::array<int, 3>{d.pIdx[0], d.pIdx[1], d.pIdx[2]}, w, psn);
            }
            out.push_back(row);
        }
    };
    auto pats = [&](const std::vector<std::vector<orc::C3>>& P, std::vector<std::vector<Complex3D>>& out) {
        out.resize(P.size());
        for (size_t i = 0; i < P.size(); ++i)
            for (const auto& c : P[i]) out[i].push_back(toC3(c));
    };
    grids(o.g1, v.kp1, v.wgl1, v.phi1, v.wphi1);
    interp(o.ip1, o.g1, v.ip1);
    pats(o.Vs1, v.Vs1);
    v.TF1 = o.TF1;
    if (o.ie == 2) {
        grids(o.g2, v.kp2, v.wgl2, v.phi2, v.wphi2);
        interp(o.ip2, o.g2, v.ip2);
        pats(o.Vs2, v.Vs2);
        v.TF2 = o.TF2;
    }
}

orc::Oracle* make_oracle(const char* nas, int ie, double er, double ei, int mode) {
    auto* o = new orc::Oracle();
    o->ie = ie;
    o->alpha = 0.2;
    if (orc::read_nas(nas, o->mesh) != 0) {
        delete o;
        return nullptr;
    }
    o->gp = orc::gauss_points(7);
    o->wave = orc::make_wave(299792458.0, cdd(er, ei), cdd(1.0, 0.0), ie == 2);
    orc::build_rwg(o->mesh, o->rwgs);
    orc::build_tree(o->rwgs, o->wave.lambda, 0.3, o->tree);
    o->build_base();
    o->build_interp();
    o->build_tf();
    if (mode & 1) o->build_patterns();
    if (mode & 2) o->build_near();
    o->init_workspace();
    return o;
}

int cmd_stages(const char* nas, int ie, double er, double ei, unsigned seed) {
    orc::Oracle* o = make_oracle(nas, ie, er, ei, 1);
    if (!o) return 2;
    RefView v;
    build_view(*o, v);
    int n = o->unknowns();
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<cdd> x(n);
    for (auto& c : x) c = cdd(nd(rng), nd(rng));
    double worst = 0;
    long compared = 0;
    int passes = (ie == 2) ? 4 : 1;
    for (int p = 0; p < passes; ++p) {
        int wn = (p < 2) ? 1 : 2;
        int cur = p % 2;
        const auto& kp = wn == 1 ? v.kp1 : v.kp2;
        const auto& ip = wn == 1 ? v.ip1 : v.ip2;
        const auto& TF = wn == 1 ? v.TF1 : v.TF2;
        const auto& wgl = wn == 1 ? v.wgl1 : v.wgl2;
        const auto& wphi = wn == 1 ? v.wphi1 : v.wphi2;
        const auto& phi = wn == 1 ? v.phi1 : v.phi2;
        const auto& Vs = wn == 1 ? v.Vs1 : v.Vs2;
        cdd k = wn == 1 ? o->wave.k1 : o->wave.k2;
        int lvl = o->tree.maxLevel - 1;
        std::vector<std::vector<std::vector<Complex3D>>> Sm(lvl), Bm(lvl);
        for (int level = 0; level < lvl; ++level) {
            int nodes = (int)v.nodes[level + 2].size();
            int K = (int)kp[o->levelSpan - 1 - level].size();
            Sm[level].assign(nodes, std::vector<Complex3D>(K));
            Bm[level].assign(nodes, std::vector<Complex3D>(K));
        }
        ProcessMLFMM::clearSmBm(Sm, Bm);
        ProcessMLFMM::mexp(Sm, v.nodes[o->tree.maxLevel], Vs, kp[0], x.data(), cur == 0 ? 0 : o->N,
                           o->tree.maxLevel);
        ProcessMLFMM::m2m(Sm, v.nodes, kp, ip, o->tree.maxLevel, o->levelSpan, k);
        ProcessMLFMM::m2l(Sm, Bm, v.nodes, TF, kp, wgl, wphi, phi, o->tree.maxLevel, o->levelSpan);
        ProcessMLFMM::l2l(Bm, v.nodes, kp, phi, ip, o->tree.maxLevel, o->levelSpan, k);

        orc::Ws& S = wn == 1 ? (cur == 0 ? o->SmJ1 : o->SmM1) : (cur == 0 ? o->SmJ2 : o->SmM2);
        orc::Ws& B = wn == 1 ? (cur == 0 ? o->BmJ1 : o->BmM1) : (cur == 0 ? o->BmJ2 : o->BmM2);
        o->clear(S, B);
        o->pass(S, B, wn, x.data(), cur == 0 ? 0 : o->N);
        for (int which = 0; which < 2; ++which) {
            const auto& R = which == 0 ? Sm : Bm;
            const orc::Ws& W = which == 0 ? S : B;
            for (int level = 0; level < lvl; ++level) {
                double num = 0, den = 0;
                for (size_t nn = 0; nn < R[level].size(); ++nn)
                    for (size_t q = 0; q < R[level][nn].size(); ++q) {
                        const Complex3D& a = R[level][nn][q];
                        const orc::C3& b = W[level][nn][q];
                        num += std::norm(a.x - b.x) + std::norm(a.y - b.y) + std::norm(a.z - b.z);
                        den += std::norm(a.x) + std::norm(a.y) + std::norm(a.z);
                        ++compared;
                    }
                double rel = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
                worst = std::max(worst, rel);
            }
        }
    }
    std::printf("{\"cmd\":\"stages\",\"ie\":%d,\"compared\":%ld,\"max_rel\":%.3e}\n", ie, compared, worst);
    delete o;
    return 0;
}

int cmd_solvers(const char* nas, int ie, double er, double ei) {
    orc::Oracle* o = make_oracle(nas, ie, er, ei, 3);
    if (!o) return 2;
    int n = o->unknowns();
    orc::PlaneWave pw = orc::init_pw(0.0, 0.0, 0.0);
    std::vector<cdd> b;
    if (ie == 0) orc::rhs_E(*o, o->wave.k1, pw.kinc, pw.einc, b);
    else if (ie == 1) orc::rhs_C(*o, o->wave.k1, pw.kinc, pw.einc, pw.hinc, o->alpha, b);
    else orc::rhs_PMCHWT(*o, o->wave.k1, pw.kinc, pw.einc, pw.hinc, b);
    auto ax = [o](int, const cdd* a, cdd* w) { o->apply(a, w); };
    const int mr = 30;
    const double tol = 1e-6;
    std::vector<cdd> xr(n, cdd(0, 0)), xo(n, cdd(0, 0));
    bool cr = MGMRES::mgmres(n, xr.data(), b.data(), 4 * n, mr, 1e-12, tol, ax, false);
    int it = 0;
    double res = 0;
    bool co = orc::gmres_oracle(*o, n, xo.data(), b.data(), 4 * n, mr, 1e-12, tol, &it, &res);
    double num = 0, den = 0;
    for (int i = 0; i < n; ++i) {
        num += std::norm(xr[i] - xo[i]);
        den += std::norm(xr[i]);
    }
    double g_rel = std::sqrt(num / den);
    std::vector<cdd> yr(n), yo(n);
    std::fill(xr.begin(), xr.end(), cdd(0, 0));
    std::fill(xo.begin(), xo.end(), cdd(0, 0));
    bool ccr = MCGS::CGS::solve(n, xr.data(), b.data(), 4 * n, tol, ax, false);
    int it2 = 0;
    double res2 = 0;
    bool cco = orc::cgs_oracle(*o, n, xo.data(), b.data(), 4 * n, tol, &it2, &res2);
    num = den = 0;
    for (int i = 0; i < n; ++i) {
        num += std::norm(xr[i] - xo[i]);
        den += std::norm(xr[i]);
    }
    double c_rel = std::sqrt(num / den);
    std::printf("{\"cmd\":\"solvers\",\"ie\":%d,\"n\":%d,\"gmres_ref_conv\":%d,\"gmres_orc_conv\":%d,"
                "\"gmres_iters\":%d,\"gmres_x_rel\":%.3e,\"cgs_ref_conv\":%d,\"cgs_orc_conv\":%d,"
                "\"cgs_iters\":%d,\"cgs_x_rel\":%.3e}\n",
                ie, n, (int)cr, (int)co, it, g_rel, (int)ccr, (int)cco, it2, c_rel);
    delete o;
    return 0;
}

// Reference GMRES (Code/GMRES.h) solution of the EFIE/CFIE/PMCHWT system built
// by the oracle, written as raw complex128: rhs then x.
int cmd_solution(const char* nas, int ie, double er, double ei, const char* out) {
    orc::Oracle* o = make_oracle(nas, ie, er, ei, 3);
    if (!o) return 2;
    int n = o->unknowns();
    orc::PlaneWave pw = orc::init_pw(0.0, 0.0, 0.0);
    std::vector<cdd> b;
    if (ie == 0) orc::rhs_E(*o, o->wave.k1, pw.kinc, pw.einc, b);
    else if (ie == 1) orc::rhs_C(*o, o->wave.k1, pw.kinc, pw.einc, pw.hinc, o->alpha, b);
    else orc::rhs_PMCHWT(*o, o->wave.k1, pw.kinc, pw.einc, pw.hinc, b);
    auto ax = [o](int, const cdd* a, cdd* w) { o->apply(a, w); };
    std::vector<cdd> x(n, cdd(0, 0));
    bool c = MGMRES::mgmres(n, x.data(), b.data(), 4 * n, 30, 1e-12, 1e-6, ax, false);
    FILE* f = std::fopen(out, "wb");
    if (!f) return 3;
    std::fwrite(b.data(), sizeof(cdd), n, f);
    std::fwrite(x.data(), sizeof(cdd), n, f);
    std::fclose(f);
    std::printf("{\"cmd\":\"solution\",\"n\":%d,\"converged\":%d}\n", n, (int)c);
    delete o;
    return 0;
}

// RWG numbering: the reference's RWGGenerator::GenerateRWG on the mesh the
// oracle read, against the oracle's build_rwg: same count and, per function in
// order, the same common-edge vertex ids, positive / negative face ids, free
// vertex ids and edge length (bit-exact).
int cmd_rwg(const char* nas, const char* out) {
    orc::Mesh m;
    if (orc::read_nas(nas, m) != 0) return 2;
    std::vector<Point> pts(m.p.size());
    for (size_t i = 0; i < pts.size(); ++i) pts[i] = Point{m.pid[i], m.p[i].x, m.p[i].y, m.p[i].z};
    std::vector<FaceElement> faces(m.fv.size());
    for (size_t f = 0; f < faces.size(); ++f) {
        faces[f].faceId = m.fid[f];
        faces[f].vertex1 = &pts[m.fv[f][0]];
        faces[f].vertex2 = &pts[m.fv[f][1]];
        faces[f].vertex3 = &pts[m.fv[f][2]];
    }
    std::vector<RWGBase> ref;
    if (!RWGGenerator::GenerateRWG(faces, pts, ref)) return 3;
    std::vector<orc::Rwg> orw;
    orc::build_rwg(m, orw);
    long mism = 0;
    const size_t n = std::min(ref.size(), orw.size());
    for (size_t i = 0; i < n; ++i) {
        const RWGBase& a = ref[i];
        const orc::Rwg& b = orw[i];
        if (a.commonEdgeVertex1->id != m.pid[b.e0] || a.commonEdgeVertex2->id != m.pid[b.e1] ||
            a.positiveFace->faceId != m.fid[b.fpos] || a.negativeFace->faceId != m.fid[b.fneg] ||
            a.freeVertexPositive->id != m.pid[b.vpos] || a.freeVertexNegative->id != m.pid[b.vneg] ||
            a.edgeLength != b.len || a.rwgid != (int)i)
            ++mism;
    }
    if (out) {   // the reference's numbering as indices: pos face, neg face, edge v1, edge v2, free pos, free neg
        std::FILE* f = std::fopen(out, "wb");
        if (!f) return 4;
        for (const RWGBase& a : ref) {
            const int v[6] = {(int)(a.positiveFace - faces.data()), (int)(a.negativeFace - faces.data()),
                              (int)(a.commonEdgeVertex1 - pts.data()), (int)(a.commonEdgeVertex2 - pts.data()),
                              (int)(a.freeVertexPositive - pts.data()), (int)(a.freeVertexNegative - pts.data())};
            std::fwrite(v, sizeof(int), 6, f);
        }
        for (const RWGBase& a : ref) std::fwrite(&a.edgeLength, sizeof(double), 1, f);
        std::fclose(f);
    }
    std::printf("{\"cmd\":\"rwg\",\"ref_count\":%zu,\"orc_count\":%zu,\"mismatches\":%ld}\n", ref.size(), orw.size(),
                mism);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    std::string c = argv[1];
    if (c == "gauss") return cmd_gauss();
    if (c == "emsource") return cmd_emsource();
    if (c == "stages" && argc >= 7)
        return cmd_stages(argv[2], std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]),
                          (unsigned)std::atoi(argv[6]));
    if (c == "solution" && argc >= 7)
        return cmd_solution(argv[2], std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]), argv[6]);
    if (c == "rwg" && argc >= 3) return cmd_rwg(argv[2], argc >= 4 ? argv[3] : nullptr);
    if (c == "solvers" && argc >= 6) return cmd_solvers(argv[2], std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]));
    return 1;
}
