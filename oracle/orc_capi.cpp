// ORACLE -- test infrastructure only (see orc.hpp).
// Flat C entry points so tests can drive the oracle through ctypes.
#include <omp.h>

#include <cstring>
#include <memory>

#include "orc_oracle.hpp"

using namespace orc;

namespace {
inline const cd* C(const double* p) { return reinterpret_cast<const cd*>(p); }
inline cd* C(double* p) { return reinterpret_cast<cd*>(p); }
Oracle* O(void* h) { return static_cast<Oracle*>(h); }

void finish(Oracle* o, double freq, double er_re, double er_im, double mr_re, double mr_im, double leafCoeff,
            int ngauss, int mode) {
    o->gp = gauss_points(ngauss);
    o->wave = make_wave(freq, cd(er_re, er_im), cd(mr_re, mr_im), o->ie == 2);
    build_rwg(o->mesh, o->rwgs);
    build_tree(o->rwgs, o->wave.lambda, leafCoeff, o->tree);
    o->build_base();
    o->build_interp();
    o->build_tf();
    if (mode & 4) {
        o->fill_synthetic(12345u);
    } else {
        if (mode & 1) o->build_patterns();
        if (mode & 2) o->build_near();
    }
    o->init_workspace();
}
}  // namespace

extern "C" {

// mode bit 1: radiation patterns, bit 2: near field, bit 4: synthetic values
void* orc_create(int np, const int* ids, const double* xyz, int nf, const int* fids, const int* fvid, int ie,
                 double freq, double er_re, double er_im, double mr_re, double mr_im, double alpha,
                 double leafCoeff, int ngauss, int tfFull, int mode) {
    try {
        auto o = std::make_unique<Oracle>();
        o->ie = ie;
        o->alpha = alpha;
        o->tfFull = tfFull;
        mesh_from_arrays(o->mesh, np, ids, xyz, nf, fids, fvid);
        finish(o.get(), freq, er_re, er_im, mr_re, mr_im, leafCoeff, ngauss, mode);
        return o.release();
    } catch (...) {
        return nullptr;
    }
}

void* orc_create_nas(const char* path, int ie, double freq, double er_re, double er_im, double mr_re,
                     double mr_im, double alpha, double leafCoeff, int ngauss, int tfFull, int mode) {
    try {
        auto o = std::make_unique<Oracle>();
        o->ie = ie;
        o->alpha = alpha;
        o->tfFull = tfFull;
        if (read_nas(path, o->mesh) != 0) return nullptr;
        finish(o.get(), freq, er_re, er_im, mr_re, mr_im, leafCoeff, ngauss, mode);
        return o.release();
    } catch (...) {
        return nullptr;
    }
}

void orc_destroy(void* h) { delete O(h); }

// out: N, unknowns, maxLevel, levelSpan, leaves
void orc_dims(void* h, int* out) {
    Oracle* o = O(h);
    out[0] = o->N;
    out[1] = o->unknowns();
    out[2] = o->tree.maxLevel;
    out[3] = o->levelSpan;
    out[4] = (int)o->tree.lv[o->tree.maxLevel].size();
}

void orc_mesh_geometry(void* h, double* area, double* normal) {
    Oracle* o = O(h);
    for (size_t f = 0; f < o->mesh.area.size(); ++f) {
        area[f] = o->mesh.area[f];
        normal[3 * f] = o->mesh.normal[f].x;
        normal[3 * f + 1] = o->mesh.normal[f].y;
        normal[3 * f + 2] = o->mesh.normal[f].z;
    }
}

void orc_rwg(void* h, int* ints, double* len, double* mid) {
    Oracle* o = O(h);
    for (int i = 0; i < o->N; ++i) {
        const Rwg& r = o->rwgs[i];
        int v[6] = {r.fpos, r.fneg, r.e0, r.e1, r.vpos, r.vneg};
        std::memcpy(ints + 6 * i, v, sizeof(v));
        len[i] = r.len;
        mid[3 * i] = r.mid.x;
        mid[3 * i + 1] = r.mid.y;
        mid[3 * i + 2] = r.mid.z;
    }
}

int orc_level_count(void* h, int level) { return (int)O(h)->tree.lv[level].size(); }

void orc_level_nodes(void* h, int level, unsigned long long* codes, double* centres, double* len, int* parent) {
    const auto& L = O(h)->tree.lv[level];
    for (size_t i = 0; i < L.size(); ++i) {
        codes[i] = L[i].code;
        centres[3 * i] = L[i].c.x;
        centres[3 * i + 1] = L[i].c.y;
        centres[3 * i + 2] = L[i].c.z;
        len[i] = L[i].len;
        parent[i] = L[i].parent;
    }
}

static const std::vector<int>& pick(const Node& n, int which) {
    switch (which) {
        case 0: return n.nearn;
        case 1: return n.dr;
        case 2: return n.drv;
        case 3: return n.sons;
        default: return n.rwg;
    }
}

// which: 0 near, 1 distant relatives, 2 their vector ids, 3 sons, 4 rwgs
int orc_level_list_total(void* h, int level, int which) {
    int t = 0;
    for (const Node& n : O(h)->tree.lv[level]) t += (int)pick(n, which).size();
    return t;
}

void orc_level_list(void* h, int level, int which, int* ptr, int* idx) {
    int t = 0;
    const auto& L = O(h)->tree.lv[level];
    for (size_t i = 0; i < L.size(); ++i) {
        ptr[i] = t;
        for (int v : pick(L[i], which)) idx[t++] = v;
    }
    ptr[L.size()] = t;
}

int orc_drvec_count(void* h, int level) { return (int)O(h)->tree.drvec[level - 2].size(); }

void orc_drvec(void* h, int level, double* out) {
    const auto& v = O(h)->tree.drvec[level - 2];
    for (size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}

int orc_grid_L(void* h, int wn, int rev) { return (wn == 1 ? O(h)->g1 : O(h)->g2)[rev].L; }

void orc_grid(void* h, int wn, int rev, double* theta, double* phi, double* wgl, double* dphi, double* kp) {
    const Grid& g = (wn == 1 ? O(h)->g1 : O(h)->g2)[rev];
    std::memcpy(theta, g.theta.data(), g.theta.size() * sizeof(double));
    std::memcpy(phi, g.phi.data(), g.phi.size() * sizeof(double));
    std::memcpy(wgl, g.wgl.data(), g.wgl.size() * sizeof(double));
    *dphi = g.dphi;
    for (size_t i = 0; i < g.kp.size(); ++i) {
        kp[3 * i] = g.kp[i].x;
        kp[3 * i + 1] = g.kp[i].y;
        kp[3 * i + 2] = g.kp[i].z;
    }
}

// level 0 = tree level 2; K = parent-grid size
void orc_interp(void* h, int wn, int level, int* tp, double* w) {
    const auto& ip = (wn == 1 ? O(h)->ip1 : O(h)->ip2)[level];
    for (size_t q = 0; q < ip.size(); ++q)
        for (int t = 0; t < 9; ++t) {
            tp[9 * q + t] = ip[q].tp[t];
            w[9 * q + t] = ip[q].w[t];
        }
}

void orc_tf(void* h, int wn, int level, double* out) {
    const auto& T = (wn == 1 ? O(h)->TF1 : O(h)->TF2)[level];
    size_t o = 0;
    for (const auto& row : T)
        for (const cd& c : row) {
            out[o++] = c.real();
            out[o++] = c.imag();
        }
}

// which: 0 Vsmi, 1 Vfmj; out K x (x,y,z) complex
void orc_pattern(void* h, int wn, int which, int rwg, double* out) {
    Oracle* o = O(h);
    const auto& P = (wn == 1) ? (which == 0 ? o->Vs1 : o->Vf1) : (which == 0 ? o->Vs2 : o->Vf2);
    const auto& row = P[rwg];
    for (size_t q = 0; q < row.size(); ++q) {
        out[6 * q + 0] = row[q].x.real();
        out[6 * q + 1] = row[q].x.imag();
        out[6 * q + 2] = row[q].y.real();
        out[6 * q + 3] = row[q].y.imag();
        out[6 * q + 4] = row[q].z.real();
        out[6 * q + 5] = row[q].z.imag();
    }
}

int orc_near_row_len(void* h, int row) { return (int)O(h)->Zn[row].size(); }

void orc_near_row(void* h, int row, double* vals, int* ids) {
    const auto& z = O(h)->Zn[row];
    const auto& zi = O(h)->Zid[row];
    for (size_t j = 0; j < z.size(); ++j) {
        vals[2 * j] = z[j].real();
        vals[2 * j + 1] = z[j].imag();
        ids[j] = zi[j];
    }
}

void orc_apply(void* h, const double* x, double* y) { O(h)->apply(C(x), C(y)); }

// Run clear + (mexp, m2m, m2l, l2l) for one wavenumber / current on x.
void orc_stage_pass(void* h, int wn, int cur, const double* x) {
    Oracle* o = O(h);
    orc::Ws* S;
    orc::Ws* B;
    if (wn == 1) {
        S = cur == 0 ? &o->SmJ1 : &o->SmM1;
        B = cur == 0 ? &o->BmJ1 : &o->BmM1;
    } else {
        S = cur == 0 ? &o->SmJ2 : &o->SmM2;
        B = cur == 0 ? &o->BmJ2 : &o->BmM2;
    }
    o->clear(*S, *B);
    o->pass(*S, *B, wn, C(x), cur == 0 ? 0 : o->N);
}

// which: 0 Sm, 1 Bm; level 0 = tree level 2; out K x 3 complex
void orc_get_ws(void* h, int wn, int cur, int which, int level, int node, double* out) {
    Oracle* o = O(h);
    const orc::Ws* W;
    if (wn == 1) W = which == 0 ? (cur == 0 ? &o->SmJ1 : &o->SmM1) : (cur == 0 ? &o->BmJ1 : &o->BmM1);
    else W = which == 0 ? (cur == 0 ? &o->SmJ2 : &o->SmM2) : (cur == 0 ? &o->BmJ2 : &o->BmM2);
    const auto& v = (*W)[level][node];
    for (size_t q = 0; q < v.size(); ++q) {
        out[6 * q + 0] = v[q].x.real();
        out[6 * q + 1] = v[q].x.imag();
        out[6 * q + 2] = v[q].y.real();
        out[6 * q + 3] = v[q].y.imag();
        out[6 * q + 4] = v[q].z.real();
        out[6 * q + 5] = v[q].z.imag();
    }
}

int orc_gmres(void* h, const double* rhs, double* x, int itr_max, int mr, double tol_abs, double tol_rel, int* iters,
              double* resid) {
    Oracle* o = O(h);
    return gmres_oracle(*o, o->unknowns(), C(x), C(rhs), itr_max, mr, tol_abs, tol_rel, iters, resid) ? 1 : 0;
}

int orc_cgs(void* h, const double* b, double* x, int max_iter, double tol, int* iters, double* resid) {
    Oracle* o = O(h);
    return cgs_oracle(*o, o->unknowns(), C(x), C(b), max_iter, tol, iters, resid) ? 1 : 0;
}

// Excitation for the handle's equation (Code/MLFMM.cpp:399-452 computeV lambdas).
void orc_rhs(void* h, double pol, double th, double ph, double* out) {
    Oracle* o = O(h);
    PlaneWave pw = init_pw(pol, th, ph);
    std::vector<cd> V;
    if (o->ie == 0) rhs_E(*o, o->wave.k1, pw.kinc, pw.einc, V);
    else if (o->ie == 1) rhs_C(*o, o->wave.k1, pw.kinc, pw.einc, pw.hinc, o->alpha, V);
    else rhs_PMCHWT(*o, o->wave.k1, pw.kinc, pw.einc, pw.hinc, V);
    std::memcpy(out, V.data(), V.size() * sizeof(cd));
}

// RCS (dBsm) at one observation angle: out = {rcs, rcs_v, rcs_h, Fx, Fy, Fz (complex)}
// Code/RCS.h:247-259 / :397-410.
void orc_rcs(void* h, const double* I, double pol, double th, double ph, double* out) {
    Oracle* o = O(h);
    cd F[3];
    if (o->ie == 2) rcs_pmchwt(*o, C(I), C(I) + o->N, o->wave.k1, th, ph, F);
    else rcs_pec(*o, C(I), o->wave.k1, o->wave.eta1, th, ph, F);
    PlaneWave pw = init_pw(pol, th, ph);
    cd Ev = pw.vth[0] * F[0] + pw.vth[1] * F[1] + pw.vth[2] * F[2];
    cd Eh = pw.vph[0] * F[0] + pw.vph[1] * F[1] + pw.vph[2] * F[2];
    out[0] = 10.0 * std::log10((std::norm(Ev) + std::norm(Eh)) / (4.0 * kPi));
    out[1] = 10.0 * std::log10(std::norm(Ev) / (4.0 * kPi));
    out[2] = 10.0 * std::log10(std::norm(Eh) / (4.0 * kPi));
    for (int c = 0; c < 3; ++c) {
        out[3 + 2 * c] = F[c].real();
        out[4 + 2 * c] = F[c].imag();
    }
}

void orc_dense(void* h, double* Z) { O(h)->dense(C(Z)); }

// which: 0 L(k1) 1 L(k2) 2 K(k1) 3 K(k2) 4 EFIE(k1, eta1) 5 MFIE(k1)
void orc_op(void* h, int which, int f, int s, double* out) {
    Oracle* o = O(h);
    const Rwg& F = o->rwgs[f];
    const Rwg& S = o->rwgs[s];
    cd v;
    switch (which) {
        case 0: v = op_L(o->mesh, F, S, o->wave.k1, o->gp); break;
        case 1: v = op_L(o->mesh, F, S, o->wave.k2, o->gp); break;
        case 2: v = op_K(o->mesh, F, S, o->wave.k1, o->gp); break;
        case 3: v = op_K(o->mesh, F, S, o->wave.k2, o->gp); break;
        case 4: v = op_EFIE(o->mesh, F, S, o->wave.k1, o->wave.eta1, o->gp); break;
        default: v = op_MFIE(o->mesh, F, S, o->wave.k1, o->gp); break;
    }
    out[0] = v.real();
    out[1] = v.imag();
}

// out: k1, k2, eta1, eta2, etai (complex)
void orc_wave(void* h, double* out) {
    const Wave& w = O(h)->wave;
    cd v[5] = {w.k1, w.k2, w.eta1, w.eta2, w.etai};
    std::memcpy(out, v, sizeof(v));
}

// --- free numerics
void orc_gauss(int n, double* l1, double* l2, double* l3, double* w) {
    Gauss g = gauss_points(n);
    for (int i = 0; i < n; ++i) {
        l1[i] = g.l1[i];
        l2[i] = g.l2[i];
        l3[i] = g.l3[i];
        w[i] = g.w[i];
    }
}
void orc_gauss_legendre(int n, double* x, double* w) {
    std::vector<double> a, b;
    gauss_legendre(n, a, b);
    std::memcpy(x, a.data(), n * sizeof(double));
    std::memcpy(w, b.data(), n * sizeof(double));
}
void orc_sph_jy(int L, double x, double* j, double* y) {
    std::vector<double> a, b;
    sph_jy(L, x, a, b);
    std::memcpy(j, a.data(), (L + 1) * sizeof(double));
    std::memcpy(y, b.data(), (L + 1) * sizeof(double));
}
void orc_sing(const double* tri9, const double* n3, const double* r3, double* out) { sing_probe(tri9, n3, r3, out); }
int orc_threads() { return omp_get_max_threads(); }
double orc_legendre_p(int l, double x) { return legendre_p(l, x); }
double orc_legendre_p_prime(int l, double x) { return legendre_p_prime(l, x); }
// out: kinc, einc, hinc, vth, vph
void orc_plane_wave(double pol, double th, double ph, double* out) {
    PlaneWave pw = init_pw(pol, th, ph);
    for (int i = 0; i < 3; ++i) {
        out[i] = pw.kinc[i];
        out[3 + i] = pw.einc[i];
        out[6 + i] = pw.hinc[i];
        out[9 + i] = pw.vth[i];
        out[12 + i] = pw.vph[i];
    }
}
}
