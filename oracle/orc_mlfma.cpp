// ORACLE -- test infrastructure only (see orc.hpp).
// MLFMA setup tables, stages and the assembled matvec.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <random>

#include "orc_oracle.hpp"

namespace orc {

// Code/AlgoMLFMM.h:21-138 -- per level L = ceil(kD + 6 cbrt(kD)), D = sqrt(3) box;
// theta = acos(GL nodes) ascending (L points), phi_j = (j+1/2) pi/L (2L points).
static void build_grids(const Tree& t, double kabs, int levelSpan, std::vector<Grid>& g) {
    g.assign(levelSpan, Grid());
    for (int i = 0; i < levelSpan; ++i) {
        int phys = t.maxLevel - i;
        double D = std::sqrt(3.0) * t.lv[phys][0].len;
        int L = (int)std::ceil(kabs * D + 6.0 * std::cbrt(kabs * D));
        Grid& G = g[i];
        G.L = L;
        std::vector<double> x;
        gauss_legendre(L, x, G.wgl);
        G.theta.clear();
        for (int j = L - 1; j >= 0; --j) G.theta.push_back(std::acos(x[j]));
        G.dphi = kPi / L;
        G.phi.clear();
        for (int j = 0; j < 2 * L; ++j) G.phi.push_back(j * G.dphi + 0.5 * G.dphi);
        G.kp.clear();
        for (double th : G.theta)
            for (double ph : G.phi)
                G.kp.push_back({std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)});
    }
}

void Oracle::build_base() {
    N = (int)rwgs.size();
    levelSpan = tree.maxLevel - 1;
    build_grids(tree, std::abs(wave.k1), levelSpan, g1);
    if (ie == 2) build_grids(tree, std::abs(wave.k2), levelSpan, g2);
}

// Code/AlgoMLFMM.h:141-189 -- binary search, nearest of mid-1/mid/mid+1, periodic
// wrap of the neighbour index without unwrapping the coordinate.
static void interp_pos(double target, const std::vector<double>& a, int out[3]) {
    int n = (int)a.size();
    double delta = a[1] - a[0];
    bool asc = delta > 0;
    int lo = 0, hi = n - 1, mid = 0;
    while (lo <= hi) {
        mid = (lo + hi) / 2;
        if (std::fabs(a[mid] - target) < 1e-12) break;
        if (asc) {
            if (a[mid] > target) hi = mid - 1;
            else lo = mid + 1;
        } else {
            if (a[mid] > target) lo = mid + 1;
            else hi = mid - 1;
        }
    }
    int cl = mid;
    if (mid > 0 && std::fabs(a[mid - 1] - target) < std::fabs(a[cl] - target)) cl = mid - 1;
    if (mid < n - 1 && std::fabs(a[mid + 1] - target) < std::fabs(a[cl] - target)) cl = mid + 1;
    int prev, next;
    if (asc) {
        if (cl == 0) { prev = n - 1; next = 1; }
        else if (cl == n - 1) { prev = cl - 1; next = 0; }
        else { prev = cl - 1; next = cl + 1; }
    } else {
        if (cl == 0) { prev = 1; next = n - 1; }
        else if (cl == n - 1) { prev = 0; next = cl - 1; }
        else { prev = cl + 1; next = cl - 1; }
    }
    out[0] = prev;
    out[1] = cl;
    out[2] = next;
}

static double lag(double x0, double x1, double x2, double x3) {
    return ((x0 - x2) * (x0 - x3)) / ((x1 - x2) * (x1 - x3));
}

// Code/AlgoMLFMM.h:298-368 -- 3x3 Lagrange from the child grid to the parent grid.
static void build_interp_tables(const std::vector<Grid>& g, int maxLevel, std::vector<std::vector<Interp>>& ip) {
    int span = maxLevel - 1, levels = maxLevel - 2;
    ip.assign(levels, {});
    for (int level = 0; level < levels; ++level) {
        const Grid& F = g[span - 1 - level];
        const Grid& S = g[span - 2 - level];
        int tfn = (int)F.theta.size(), pfn = (int)F.phi.size(), psn = (int)S.phi.size();
        ip[level].resize((size_t)tfn * pfn);
        for (int t = 0; t < tfn; ++t) {
            for (int p = 0; p < pfn; ++p) {
                Interp d;
                double t0 = F.theta[t];
                interp_pos(t0, S.theta, d.tIdx);
                double tv[3] = {S.theta[d.tIdx[0]], S.theta[d.tIdx[1]], S.theta[d.tIdx[2]]};
                double lT[3] = {lag(t0, tv[0], tv[1], tv[2]), lag(t0, tv[1], tv[0], tv[2]), lag(t0, tv[2], tv[0], tv[1])};
                double p0 = F.phi[p];
                interp_pos(p0, S.phi, d.pIdx);
                double pv[3] = {S.phi[d.pIdx[0]], S.phi[d.pIdx[1]], S.phi[d.pIdx[2]]};
                double lP[3] = {lag(p0, pv[0], pv[1], pv[2]), lag(p0, pv[1], pv[0], pv[2]), lag(p0, pv[2], pv[0], pv[1])};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        d.w[a * 3 + b] = lT[a] * lP[b];
                        d.tp[a * 3 + b] = d.tIdx[a] * psn + d.pIdx[b];
                    }
                ip[level][(size_t)t * pfn + p] = d;
            }
        }
    }
}

void Oracle::build_interp() {
    build_interp_tables(g1, tree.maxLevel, ip1);
    if (ie == 2) build_interp_tables(g2, tree.maxLevel, ip2);
}

// Code/AlgoMLFMM.h:192-295 -- TF = sum_{l<=L'} (-j)^l (2l+1) h2_l(|kR|) P_l(Rhat.khat),
// L' = min(L_level, ceil|kR|), Hankel argument real |kR| even for complex k.
// mode bit 1: L' = L_level; bit 2: h2_l at the complex kR (accuracy mode).
static void build_tf_tables(const Tree& t, const std::vector<Grid>& g, cd k, int mode,
                            std::vector<std::vector<std::vector<cd>>>& TF) {
    int span = t.maxLevel - 1;
    TF.assign(span, {});
    const cd jl[4] = {{1.0, 0.0}, {0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}};
    for (int level = 0; level < span; ++level) {
        const auto& vecs = t.drvec[level];
        const Grid& G = g[span - 1 - level];
        int nv = (int)vecs.size(), K = (int)G.kp.size();
        TF[level].assign(nv, std::vector<cd>(K));
#pragma omp parallel for schedule(dynamic)
        for (int v = 0; v < nv; ++v) {
            double r[3] = {vecs[v].x, vecs[v].y, vecs[v].z};
            double R = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            r[0] /= R;
            r[1] /= R;
            r[2] /= R;
            cd kR = k * R;
            int L = (mode & 1) ? G.L : std::min(G.L, (int)std::ceil(std::abs(kR)));
            std::vector<cd> h(L + 1);
            if (mode & 2) {
                sph_h2(L, kR, h);
            } else {
                std::vector<double> jv, yv;
                sph_jy(L, std::abs(kR), jv, yv);
                for (int l = 0; l <= L; ++l) h[l] = cd(jv[l], -yv[l]);
            }
            std::vector<double> P(L + 1);
            for (int q = 0; q < K; ++q) {
                double x = r[0] * G.kp[q].x + r[1] * G.kp[q].y + r[2] * G.kp[q].z;
                P[0] = 1.0;
                if (L >= 1) P[1] = x;
                for (int l = 1; l < L; ++l) P[l + 1] = ((2.0 * l + 1.0) * x * P[l] - l * P[l - 1]) / (l + 1.0);
                cd sum(0, 0);
                for (int l = 0; l <= L; ++l) sum += jl[l & 3] * h[l] * (2.0 * l + 1.0) * P[l];
                TF[level][v][q] = sum;
            }
        }
    }
}

void Oracle::build_tf() {
    build_tf_tables(tree, g1, wave.k1, tfFull, TF1);
    if (ie == 2) build_tf_tables(tree, g2, wave.k2, tfFull, TF2);
}

// Code/ProcessFMM.h:17-180 -- radiation (Vsmi) and receiving (Vfmj) patterns of
// every RWG about its leaf centre.
namespace {
enum class Kern { Vs, VfE, VfC };
C3 kernel(Kern kind, const V3& r, const V3& c, const V3& vfree, const V3& kp, cd k, double w, const V3& n, double alpha) {
    C3 res;
    if (kind == Kern::Vs) {
        V3 p = sub(c, r);
        V3 pf = sub(r, vfree);
        cd e = std::exp(-kJ * (k * dot(p, kp)));
        res.x = w * e * pf.x;
        res.y = w * e * pf.y;
        res.z = w * e * pf.z;
        return res;
    }
    V3 p = sub(r, c);
    V3 pf = sub(r, vfree);
    cd e = std::exp(-kJ * (k * dot(p, kp)));
    V3 t1 = cross(kp, pf);
    V3 t2 = cross(kp, t1);
    V3 efie = {-t2.x, -t2.y, -t2.z};
    if (kind == Kern::VfE) {
        res.x = w * e * efie.x;
        res.y = w * e * efie.y;
        res.z = w * e * efie.z;
        return res;
    }
    V3 nc = cross(n, pf);
    V3 mfie = cross(kp, nc);
    res.x = w * e * (alpha * efie.x + (1 - alpha) * mfie.x);
    res.y = w * e * (alpha * efie.y + (1 - alpha) * mfie.y);
    res.z = w * e * (alpha * efie.z + (1 - alpha) * mfie.z);
    return res;
}
}  // namespace

static void build_pattern(const Oracle& o, const std::vector<V3>& kps, cd k, Kern kind,
                          std::vector<std::vector<C3>>& out) {
    const Tree& t = o.tree;
    const Mesh& m = o.mesh;
    const auto& leaves = t.lv[t.maxLevel];
    out.assign(o.N, std::vector<C3>(kps.size()));
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < (int)leaves.size(); ++li) {
        const Node& nd = leaves[li];
        for (int id : nd.rwg) {
            const Rwg& r = o.rwgs[id];
            const V3& fP = m.p[r.vpos];
            const V3& fN = m.p[r.vneg];
            const V3& nP = m.normal[r.fpos];
            const V3& nN = m.normal[r.fneg];
            const auto& tp = m.fv[r.fpos];
            const auto& tn = m.fv[r.fneg];
            for (size_t q = 0; q < kps.size(); ++q) {
                C3 acc{{0, 0}, {0, 0}, {0, 0}};
                for (int i = 0; i < o.gp.n; ++i) {
                    double w = o.gp.w[i] * r.len;
                    const double l1 = o.gp.l1[i], l2 = o.gp.l2[i], l3 = o.gp.l3[i];
                    V3 rp = {l1 * m.p[tp[0]].x + l2 * m.p[tp[1]].x + l3 * m.p[tp[2]].x,
                             l1 * m.p[tp[0]].y + l2 * m.p[tp[1]].y + l3 * m.p[tp[2]].y,
                             l1 * m.p[tp[0]].z + l2 * m.p[tp[1]].z + l3 * m.p[tp[2]].z};
                    V3 rn = {l1 * m.p[tn[0]].x + l2 * m.p[tn[1]].x + l3 * m.p[tn[2]].x,
                             l1 * m.p[tn[0]].y + l2 * m.p[tn[1]].y + l3 * m.p[tn[2]].y,
                             l1 * m.p[tn[0]].z + l2 * m.p[tn[1]].z + l3 * m.p[tn[2]].z};
                    C3 a = kernel(kind, rp, nd.c, fP, kps[q], k, w, nP, o.alpha);
                    C3 b = kernel(kind, rn, nd.c, fN, kps[q], k, w, nN, o.alpha);
                    acc.x += a.x - b.x;
                    acc.y += a.y - b.y;
                    acc.z += a.z - b.z;
                }
                out[id][q] = acc;
            }
        }
    }
}

void Oracle::build_patterns() {
    if (ie == 2) {
        build_pattern(*this, g1[0].kp, wave.k1, Kern::Vs, Vs1);
        build_pattern(*this, g1[0].kp, wave.k1, Kern::VfE, Vf1);
        build_pattern(*this, g2[0].kp, wave.k2, Kern::Vs, Vs2);
        build_pattern(*this, g2[0].kp, wave.k2, Kern::VfE, Vf2);
    } else {
        build_pattern(*this, g1[0].kp, wave.k1, Kern::Vs, Vs1);
        build_pattern(*this, g1[0].kp, wave.k1, ie == 1 ? Kern::VfC : Kern::VfE, Vf1);
    }
}

// Code/Zmartix.h:188-643 -- ragged rows, columns = near-neighbour leaves in list
// order, then the own leaf; PMCHWT rows interleave (J, M) source columns.
void Oracle::build_near() {
    const auto& leaves = tree.lv[tree.maxLevel];
    int rows = unknowns();
    Zn.assign(rows, {});
    Zid.assign(rows, {});
    for (const Node& nd : leaves) {
        int cnt = (int)nd.rwg.size();
        for (int nb : nd.nearn) cnt += (int)leaves[nb].rwg.size();
        for (int id : nd.rwg) {
            int c = (ie == 2) ? 2 * cnt : cnt;
            Zn[id].assign(c, cd(0, 0));
            Zid[id].assign(c, 0);
            if (ie == 2) {
                Zn[N + id].assign(c, cd(0, 0));
                Zid[N + id].assign(c, 0);
            }
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < (int)leaves.size(); ++li) {
        const Node& nd = leaves[li];
        for (int fid : nd.rwg) {
            const Rwg& F = rwgs[fid];
            int count = 0;
            for (int kk = 0; kk <= (int)nd.nearn.size(); ++kk) {
                const Node& src = (kk == (int)nd.nearn.size()) ? nd : leaves[nd.nearn[kk]];
                for (int sid : src.rwg) {
                    const Rwg& S = rwgs[sid];
                    if (ie == 0) {
                        Zn[fid][count] = op_EFIE(mesh, F, S, wave.k1, wave.eta1, gp);
                        Zid[fid][count] = sid;
                        count++;
                    } else if (ie == 1) {
                        cd E = op_EFIE(mesh, F, S, wave.k1, wave.eta1, gp);
                        cd M = op_MFIE(mesh, F, S, wave.k1, gp);
                        Zn[fid][count] = alpha * E + (1.0 - alpha) * wave.eta1 * M;
                        Zid[fid][count] = sid;
                        count++;
                    } else {
                        cd L1 = op_L(mesh, F, S, wave.k1, gp);
                        cd L2 = op_L(mesh, F, S, wave.k2, gp);
                        cd K1 = op_K(mesh, F, S, wave.k1, gp);
                        cd K2 = op_K(mesh, F, S, wave.k2, gp);
                        cd z11 = L1 + wave.etai * L2;
                        cd z12 = -(K1 + K2);
                        cd z21 = K1 + K2;
                        cd z22 = L1 + (1.0 / wave.etai) * L2;
                        Zn[fid][count] = z11;
                        Zid[fid][count] = sid;
                        count++;
                        Zn[fid][count] = z12;
                        Zid[fid][count] = N + sid;
                        count++;
                        Zn[N + fid][count - 2] = z21;
                        Zn[N + fid][count - 1] = z22;
                        Zid[N + fid][count - 2] = sid;
                        Zid[N + fid][count - 1] = N + sid;
                    }
                }
            }
        }
    }
}

// Values of the right shape without the setup cost; used only to time the CPU
// matvec on workloads whose FP64 setup would take hours on the host (the
// matvec's running time does not depend on the values).
void Oracle::fill_synthetic(unsigned seed) {
    auto val = [seed](uint64_t a, uint64_t b) {
        uint64_t h = (a * 0x9E3779B97F4A7C15ULL) ^ (b + 0x632BE59BD9B4E019ULL + seed);
        h ^= h >> 31;
        h *= 0xD6E8FEB86659FD93ULL;
        h ^= h >> 32;
        return cd((double)(h & 0xffff) / 32768.0 - 1.0, (double)((h >> 16) & 0xffff) / 32768.0 - 1.0);
    };
    auto fillP = [&](std::vector<std::vector<C3>>& V, int K, uint64_t tag) {
        V.assign(N, {});
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            V[i].resize(K);
            for (int q = 0; q < K; ++q) {
                uint64_t a = ((uint64_t)i << 20) ^ (uint64_t)q ^ (tag << 56);
                V[i][q] = {val(a, 1), val(a, 2), val(a, 3)};
            }
        }
    };
    fillP(Vs1, (int)g1[0].kp.size(), 1);
    fillP(Vf1, (int)g1[0].kp.size(), 2);
    if (ie == 2) {
        fillP(Vs2, (int)g2[0].kp.size(), 3);
        fillP(Vf2, (int)g2[0].kp.size(), 4);
    }
    const auto& leaves = tree.lv[tree.maxLevel];
    int rows = unknowns();
    Zn.assign(rows, {});
    Zid.assign(rows, {});
    const int h2 = (ie == 2) ? 2 : 1;
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < (int)leaves.size(); ++li) {
        const Node& nd = leaves[li];
        std::vector<int> cols;
        for (int nb : nd.nearn)
            for (int s : leaves[nb].rwg) cols.push_back(s);
        for (int s : nd.rwg) cols.push_back(s);
        for (int id : nd.rwg) {
            for (int q = 0; q < h2; ++q) {
                int row = id + q * N;
                auto& z = Zn[row];
                auto& zi = Zid[row];
                z.resize(cols.size() * h2);
                zi.resize(cols.size() * h2);
                size_t e = 0;
                for (int s : cols)
                    for (int h = 0; h < h2; ++h) {
                        z[e] = val(((uint64_t)row << 24) ^ e, 7);
                        zi[e] = s + h * N;
                        ++e;
                    }
            }
        }
    }
}

// Code/MLFMM.cpp:115-161
void Oracle::init_workspace() {
    auto initW = [&](Ws& Sm, Ws& Bm, const std::vector<Grid>& g) {
        int lvl = tree.maxLevel - 1;
        Sm.assign(lvl, {});
        Bm.assign(lvl, {});
        for (int level = 0; level < lvl; ++level) {
            int nodes = (int)tree.lv[level + 2].size();
            int K = (int)g[levelSpan - 1 - level].kp.size();
            Sm[level].assign(nodes, std::vector<C3>(K));
            Bm[level].assign(nodes, std::vector<C3>(K));
        }
    };
    initW(SmJ1, BmJ1, g1);
    if (ie == 2) {
        initW(SmM1, BmM1, g1);
        initW(SmJ2, BmJ2, g2);
        initW(SmM2, BmM2, g2);
    }
}

// Code/ProcessMLFMM.h:14-28
void Oracle::clear(Ws& Sm, Ws& Bm) {
    for (size_t l = 0; l < Sm.size(); ++l)
#pragma omp parallel for
        for (int n = 0; n < (int)Sm[l].size(); ++n) {
            for (auto& c : Sm[l][n]) c = C3{};
            for (auto& c : Bm[l][n]) c = C3{};
        }
}

// Code/ProcessMLFMM.h:33-56
void Oracle::mexp(Ws& Sm, const std::vector<std::vector<C3>>& Vs, const cd* x, int off) {
    const auto& leaves = tree.lv[tree.maxLevel];
    int lvl = tree.maxLevel - 2;
#pragma omp parallel for
    for (int n = 0; n < (int)leaves.size(); ++n) {
        auto& S = Sm[lvl][n];
        int K = (int)S.size();
        for (int id : leaves[n].rwg) {
            const auto& V = Vs[id];
            cd xv = x[off + id];
            for (int q = 0; q < K; ++q) {
                S[q].x += V[q].x * xv;
                S[q].y += V[q].y * xv;
                S[q].z += V[q].z * xv;
            }
        }
    }
}

// Code/ProcessMLFMM.h:61-107
void Oracle::m2m(Ws& Sm, const std::vector<Grid>& g, const std::vector<std::vector<Interp>>& ip, cd k) {
    for (int level = tree.maxLevel - 3; level > -1; --level) {
        const Grid& G = g[levelSpan - 1 - level];
        const auto& nodes = tree.lv[level + 2];
        const auto& sons = tree.lv[level + 3];
        int K = (int)G.kp.size();
#pragma omp parallel for schedule(dynamic)
        for (int n = 0; n < (int)nodes.size(); ++n) {
            const Node& nd = nodes[n];
            for (int s : nd.sons) {
                const Node& so = sons[s];
                double v[3] = {nd.c.x - so.c.x, nd.c.y - so.c.y, nd.c.z - so.c.z};
                const auto& Sc = Sm[level + 1][s];
                auto& Sp = Sm[level][n];
                for (int q = 0; q < K; ++q) {
                    const Interp& d = ip[level][q];
                    const V3& kp = G.kp[q];
                    cd kr = k * (v[0] * kp.x + v[1] * kp.y + v[2] * kp.z);
                    cd e = std::exp(-kJ * kr);
                    C3 r{};
                    for (int t = 0; t < 9; ++t) {
                        const C3& f = Sc[d.tp[t]];
                        r.x += f.x * d.w[t];
                        r.y += f.y * d.w[t];
                        r.z += f.z * d.w[t];
                    }
                    Sp[q].x += e * r.x;
                    Sp[q].y += e * r.y;
                    Sp[q].z += e * r.z;
                }
            }
        }
    }
}

// Code/ProcessMLFMM.h:112-147 -- quadrature weight w_theta * dphi applied here.
void Oracle::m2l(const Ws& Sm, Ws& Bm, const std::vector<std::vector<std::vector<cd>>>& TF, const std::vector<Grid>& g) {
    for (int level = 0; level < tree.maxLevel - 1; ++level) {
        const Grid& G = g[levelSpan - 1 - level];
        int K = (int)G.kp.size();
        int pN = (int)G.phi.size();
        const auto& nodes = tree.lv[level + 2];
#pragma omp parallel for
        for (int n = 0; n < (int)nodes.size(); ++n) {
            const Node& nd = nodes[n];
            auto& B = Bm[level][n];
            for (size_t d = 0; d < nd.dr.size(); ++d) {
                const auto& T = TF[level][nd.drv[d]];
                const auto& S = Sm[level][nd.dr[d]];
                for (int q = 0; q < K; ++q) {
                    int t = q / pN;
                    double w = G.wgl[t] * G.dphi;
                    B[q].x += T[q] * S[q].x * w;
                    B[q].y += T[q] * S[q].y * w;
                    B[q].z += T[q] * S[q].z * w;
                }
            }
        }
    }
}

// Code/ProcessMLFMM.h:152-202 -- phase shift then transpose (scatter) interpolation.
void Oracle::l2l(Ws& Bm, const std::vector<Grid>& g, const std::vector<std::vector<Interp>>& ip, cd k) {
    for (int level = 0; level < tree.maxLevel - 2; ++level) {
        const Grid& GF = g[levelSpan - 1 - level];
        const Grid& GS = g[levelSpan - 2 - level];
        const auto& nodes = tree.lv[level + 2];
        const auto& sons = tree.lv[level + 3];
        int K = (int)GF.kp.size();
        int phiC = (int)GS.phi.size();
#pragma omp parallel for
        for (int f = 0; f < (int)nodes.size(); ++f) {
            const Node& nd = nodes[f];
            for (int s : nd.sons) {
                const Node& so = sons[s];
                double r[3] = {so.c.x - nd.c.x, so.c.y - nd.c.y, so.c.z - nd.c.z};
                auto& Bc = Bm[level + 1][s];
                for (int q = 0; q < K; ++q) {
                    const V3& kp = GF.kp[q];
                    cd kr = k * (kp.x * r[0] + kp.y * r[1] + kp.z * r[2]);
                    cd e = std::exp(-kJ * kr);
                    const Interp& d = ip[level][q];
                    const C3& B2 = Bm[level][f][q];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            auto& dst = Bc[d.tIdx[a] * phiC + d.pIdx[b]];
                            double w = d.w[a * 3 + b];
                            dst.x += e * B2.x * w;
                            dst.y += e * B2.y * w;
                            dst.z += e * B2.z * w;
                        }
                }
            }
        }
    }
}

void Oracle::pass(Ws& Sm, Ws& Bm, int wn, const cd* x, int off) {
    const auto& g = (wn == 1) ? g1 : g2;
    const auto& ip = (wn == 1) ? ip1 : ip2;
    const auto& TF = (wn == 1) ? TF1 : TF2;
    const auto& Vs = (wn == 1) ? Vs1 : Vs2;
    cd k = (wn == 1) ? wave.k1 : wave.k2;
    mexp(Sm, Vs, x, off);
    m2m(Sm, g, ip, k);
    m2l(Sm, Bm, TF, g);
    l2l(Bm, g, ip, k);
}

// Code/MLFMM.cpp:189-369 -- the ax(n, x, w) lambda.
void Oracle::apply(const cd* x, cd* w) {
    int n = unknowns();
    for (int i = 0; i < n; ++i) w[i] = cd(0, 0);
    const auto& leaves = tree.lv[tree.maxLevel];
    int fin = tree.maxLevel - 2;
    if (ie != 2) {
        cd k = wave.k1, eta = wave.eta1;
        clear(SmJ1, BmJ1);
        pass(SmJ1, BmJ1, 1, x, 0);
        int K = (int)g1[0].kp.size();
#pragma omp parallel for schedule(dynamic)
        for (int li = 0; li < (int)leaves.size(); ++li) {
            const auto& B = BmJ1[fin][li];
            for (int id : leaves[li].rwg) {
                const auto& V = Vf1[id];
                cd val(0, 0);
                for (int q = 0; q < K; ++q) val += V[q].x * B[q].x + V[q].y * B[q].y + V[q].z * B[q].z;
                val *= (k * eta / (4.0 * kPi)) * (k / (4.0 * kPi));
                for (size_t j = 0; j < Zn[id].size(); ++j) val += x[Zid[id][j]] * Zn[id][j];
                w[id] = val;
            }
        }
        return;
    }
    cd etai = wave.etai;
    cd c1 = wave.k1 * wave.k1 / (16.0 * kPi * kPi);
    cd c2 = wave.k2 * wave.k2 / (16.0 * kPi * kPi);
    clear(SmJ1, BmJ1);
    clear(SmM1, BmM1);
    clear(SmJ2, BmJ2);
    clear(SmM2, BmM2);
    pass(SmJ1, BmJ1, 1, x, 0);
    pass(SmM1, BmM1, 1, x, N);
    pass(SmJ2, BmJ2, 2, x, 0);
    pass(SmM2, BmM2, 2, x, N);
    auto intL = [](const std::vector<C3>& V, const std::vector<C3>& B) {
        cd v(0, 0);
        for (size_t q = 0; q < V.size(); ++q) v += V[q].x * B[q].x + V[q].y * B[q].y + V[q].z * B[q].z;
        return v;
    };
    auto intK = [](const std::vector<C3>& V, const std::vector<C3>& B, const std::vector<V3>& kp) {
        cd v(0, 0);
        for (size_t q = 0; q < V.size(); ++q) {
            double kh[3] = {-kp[q].x, -kp[q].y, -kp[q].z};
            cd cx = kh[1] * V[q].z - kh[2] * V[q].y;
            cd cy = kh[2] * V[q].x - kh[0] * V[q].z;
            cd cz = kh[0] * V[q].y - kh[1] * V[q].x;
            v += cx * B[q].x + cy * B[q].y + cz * B[q].z;
        }
        return v;
    };
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < (int)leaves.size(); ++li) {
        const auto& BJ1 = BmJ1[fin][li];
        const auto& BM1 = BmM1[fin][li];
        const auto& BJ2 = BmJ2[fin][li];
        const auto& BM2 = BmM2[fin][li];
        for (int id : leaves[li].rwg) {
            const auto& V1 = Vf1[id];
            const auto& V2 = Vf2[id];
            cd L1J = intL(V1, BJ1), K1J = intK(V1, BJ1, g1[0].kp), K1M = intK(V1, BM1, g1[0].kp), L1M = intL(V1, BM1);
            cd L2J = intL(V2, BJ2), K2J = intK(V2, BJ2, g2[0].kp), K2M = intK(V2, BM2, g2[0].kp), L2M = intL(V2, BM2);
            w[id] += c1 * L1J;
            w[id] += -c1 * K1M;
            w[N + id] += c1 * K1J;
            w[N + id] += c1 * L1M;
            w[id] += c2 * etai * L2J;
            w[id] += -c2 * K2M;
            w[N + id] += c2 * K2J;
            w[N + id] += c2 * L2M / etai;
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < (int)leaves.size(); ++li) {
        for (int id : leaves[li].rwg) {
            for (size_t j = 0; j < Zn[id].size(); ++j) w[id] += x[Zid[id][j]] * Zn[id][j];
            for (size_t j = 0; j < Zn[N + id].size(); ++j) w[N + id] += x[Zid[N + id][j]] * Zn[N + id][j];
        }
    }
}

// Full Galerkin matrix from the same near-field entry routines (MoM-style,
// Code/Zmartix.h:35-186 and :467-539); for small test problems only.
void Oracle::dense(cd* Z) {
    int n = unknowns();
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < N; ++m) {
        for (int s = 0; s < N; ++s) {
            const Rwg& F = rwgs[m];
            const Rwg& S = rwgs[s];
            if (ie == 0) {
                Z[(size_t)m * n + s] = op_EFIE(mesh, F, S, wave.k1, wave.eta1, gp);
            } else if (ie == 1) {
                cd E = op_EFIE(mesh, F, S, wave.k1, wave.eta1, gp);
                cd M = op_MFIE(mesh, F, S, wave.k1, gp);
                Z[(size_t)m * n + s] = alpha * E + (1.0 - alpha) * wave.eta1 * M;
            } else {
                cd L1 = op_L(mesh, F, S, wave.k1, gp), L2 = op_L(mesh, F, S, wave.k2, gp);
                cd K1 = op_K(mesh, F, S, wave.k1, gp), K2 = op_K(mesh, F, S, wave.k2, gp);
                Z[(size_t)m * n + s] = L1 + wave.etai * L2;
                Z[(size_t)m * n + N + s] = -(K1 + K2);
                Z[(size_t)(N + m) * n + s] = K1 + K2;
                Z[(size_t)(N + m) * n + N + s] = L1 + (1.0 / wave.etai) * L2;
            }
        }
    }
}

}  // namespace orc
