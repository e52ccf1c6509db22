// ORACLE -- test infrastructure only (see orc.hpp).
// Krylov solvers, excitation vectors and far-field (RCS) evaluation.
#include <algorithm>
#include <cmath>

#include "orc_oracle.hpp"

namespace orc {

namespace {
cd dotc(int n, const cd* a, const cd* b) {
    cd s = 0.0;
    for (int i = 0; i < n; ++i) s += std::conj(a[i]) * b[i];
    return s;
}
// Code/GMRES.h:22-48
void gen_givens(cd a, cd b, cd& c, cd& s) {
    const double eps = 1e-16;
    double an = std::abs(a), bn = std::abs(b);
    if (bn < eps) {
        c = 1.0;
        s = 0.0;
    } else if (an < eps) {
        c = 0.0;
        s = 1.0;
    } else {
        double r = std::sqrt(std::real(a * std::conj(a) + b * std::conj(b)));
        c = a / r;
        s = b / r;
    }
}
void app_givens(cd& x, cd& y, cd c, cd s) {
    cd t = std::conj(c) * x + std::conj(s) * y;
    y = -s * x + c * y;
    x = t;
}
}  // namespace

// Code/GMRES.h:65-217 -- restarted GMRES(mr): modified Gram-Schmidt followed by
// one re-orthogonalisation pass, complex Givens, exact residual per restart.
template <typename Ax>
bool gmres(int n, cd* x, const cd* rhs, int itr_max, int mr, double tol_abs, double tol_rel, Ax ax, int* iters,
           double* resid) {
    const double eps = 1e-16;
    std::vector<cd> r(n), w(n), z(n), v((size_t)n * (mr + 1)), h((size_t)(mr + 1) * mr), cs(mr), sn(mr), s(mr + 1),
        y(mr);
    int total = 0;
    ax(n, x, w.data());
    for (int i = 0; i < n; ++i) r[i] = rhs[i] - w[i];
    double bnorm = std::sqrt(std::real(dotc(n, rhs, rhs)));
    if (bnorm < eps) bnorm = 1.0;
    double rnorm = std::sqrt(std::real(dotc(n, r.data(), r.data())));
    double err = rnorm / bnorm;
    *iters = 0;
    *resid = err;
    if (err < tol_rel || err < tol_abs) return true;
    for (int cyc = 1; cyc <= itr_max; ++cyc) {
        for (int i = 0; i < n; ++i) v[i] = r[i] / rnorm;
        std::fill(s.begin(), s.end(), cd(0, 0));
        s[0] = rnorm;
        std::fill(h.begin(), h.end(), cd(0, 0));
        std::fill(cs.begin(), cs.end(), cd(0, 0));
        std::fill(sn.begin(), sn.end(), cd(0, 0));
        int kused = 0;
        for (int k = 0; k < mr; ++k) {
            if (total >= itr_max) break;
            total++;
            kused = k + 1;
            for (int i = 0; i < n; ++i) z[i] = v[i + (size_t)k * n];
            ax(n, z.data(), w.data());
            for (int j = 0; j <= k; ++j) {
                h[j + k * (mr + 1)] = dotc(n, &v[(size_t)j * n], w.data());
                for (int i = 0; i < n; ++i) w[i] -= h[j + k * (mr + 1)] * v[i + (size_t)j * n];
            }
            for (int j = 0; j <= k; ++j) {
                cd t = dotc(n, &v[(size_t)j * n], w.data());
                h[j + k * (mr + 1)] += t;
                for (int i = 0; i < n; ++i) w[i] -= t * v[i + (size_t)j * n];
            }
            double hn = std::sqrt(std::real(dotc(n, w.data(), w.data())));
            h[k + 1 + k * (mr + 1)] = hn;
            if (hn < eps * rnorm) break;
            for (int i = 0; i < n; ++i) v[i + (size_t)(k + 1) * n] = w[i] / hn;
            for (int j = 0; j < k; ++j) app_givens(h[j + k * (mr + 1)], h[j + 1 + k * (mr + 1)], cs[j], sn[j]);
            gen_givens(h[k + k * (mr + 1)], h[k + 1 + k * (mr + 1)], cs[k], sn[k]);
            app_givens(h[k + k * (mr + 1)], h[k + 1 + k * (mr + 1)], cs[k], sn[k]);
            h[k + 1 + k * (mr + 1)] = cd(0, 0);
            app_givens(s[k], s[k + 1], cs[k], sn[k]);
            err = std::abs(s[k + 1]) / bnorm;
            if (err < tol_rel || err < tol_abs) break;
        }
        int m = std::min(kused, mr);
        for (int i = m - 1; i >= 0; --i) {
            y[i] = s[i];
            for (int j = i + 1; j < m; ++j) y[i] -= h[i + j * (mr + 1)] * y[j];
            if (std::abs(h[i + i * (mr + 1)]) > eps) y[i] /= h[i + i * (mr + 1)];
            else y[i] = cd(0, 0);
        }
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < n; ++i) x[i] += y[j] * v[i + (size_t)j * n];
        ax(n, x, w.data());
        for (int i = 0; i < n; ++i) r[i] = rhs[i] - w[i];
        rnorm = std::sqrt(std::real(dotc(n, r.data(), r.data())));
        err = rnorm / bnorm;
        *iters = total;
        *resid = err;
        if (err < tol_rel || err < tol_abs) return true;
        if (total >= itr_max) return false;
        if (err > 100.0 * tol_rel && cyc > 8 && err > 1000.0 * tol_rel) return false;
    }
    return false;
}

bool gmres_oracle(Oracle& o, int n, cd* x, const cd* rhs, int itr_max, int mr, double tol_abs, double tol_rel,
                  int* iters, double* resid) {
    auto ax = [&o](int, const cd* a, cd* b) { o.apply(a, b); };
    return gmres(n, x, rhs, itr_max, mr, tol_abs, tol_rel, ax, iters, resid);
}

// Code/CGS.h:104-234 -- bilinear (unconjugated) inner products, no preconditioner.
bool cgs_oracle(Oracle& o, int n, cd* x, const cd* b, int max_iter, double tol, int* iters, double* resid) {
    const double eps = 1.0e-16;
    auto nrm = [n](const cd* a) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += std::norm(a[i]);
        return std::sqrt(s);
    };
    auto bil = [n](const cd* a, const cd* c) {
        cd s(0, 0);
        for (int i = 0; i < n; ++i) s += a[i] * c[i];
        return s;
    };
    std::vector<cd> r(n), Rt(n), u(n), p(n), V(n), q(n), uq(n), Q(n);
    o.apply(x, V.data());
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - V[i];
        Rt[i] = r[i];
    }
    double bn = nrm(b);
    if (bn < eps) bn = 1.0;
    double rss = nrm(r.data()) / bn;
    *iters = 0;
    *resid = rss;
    if (rss < tol) return true;
    cd rho_prev(0, 0);
    for (int it = 1; it <= max_iter; ++it) {
        cd rho = bil(r.data(), Rt.data());
        if (std::abs(rho) < eps) return false;
        if (it == 1) {
            u = r;
            p = u;
        } else {
            cd beta = rho / rho_prev;
            for (int i = 0; i < n; ++i) {
                u[i] = r[i] + beta * q[i];
                p[i] = u[i] + beta * (q[i] + beta * p[i]);
            }
        }
        o.apply(p.data(), V.data());
        cd den = bil(Rt.data(), V.data());
        if (std::abs(den) < eps) return false;
        cd al = rho / den;
        for (int i = 0; i < n; ++i) {
            q[i] = u[i] - al * V[i];
            uq[i] = u[i] + q[i];
        }
        for (int i = 0; i < n; ++i) x[i] += al * uq[i];
        o.apply(uq.data(), Q.data());
        for (int i = 0; i < n; ++i) r[i] -= al * Q[i];
        rss = nrm(r.data()) / bn;
        *iters = it;
        *resid = rss;
        if (rss < tol) return true;
        rho_prev = rho;
    }
    return false;
}

namespace {
inline V3 gpt(const Oracle& o, int f, int q) {
    const auto& t = o.mesh.fv[f];
    const V3& a = o.mesh.p[t[0]];
    const V3& b = o.mesh.p[t[1]];
    const V3& c = o.mesh.p[t[2]];
    const double l1 = o.gp.l1[q], l2 = o.gp.l2[q], l3 = o.gp.l3[q];
    return {l1 * a.x + l2 * b.x + l3 * c.x, l1 * a.y + l2 * b.y + l3 * c.y, l1 * a.z + l2 * b.z + l3 * c.z};
}
}  // namespace

// Code/Vm.h:10-71
void rhs_E(const Oracle& o, cd k, const double* kinc, const double* einc, std::vector<cd>& V) {
    V.assign(o.N, cd(0, 0));
    for (int id = 0; id < o.N; ++id) {
        const Rwg& r = o.rwgs[id];
        for (int q = 0; q < o.gp.n; ++q) {
            V3 p = gpt(o, r.fpos, q);
            const V3& fp = o.mesh.p[r.vpos];
            double rho[3] = {p.x - fp.x, p.y - fp.y, p.z - fp.z};
            cd e = std::exp(-kJ * (k * (kinc[0] * p.x + kinc[1] * p.y + kinc[2] * p.z)));
            V[id] = V[id] + r.len * o.gp.w[q] * (einc[0] * rho[0] + einc[1] * rho[1] + einc[2] * rho[2]) * e;
            p = gpt(o, r.fneg, q);
            const V3& fn = o.mesh.p[r.vneg];
            double rn[3] = {fn.x - p.x, fn.y - p.y, fn.z - p.z};
            e = std::exp(-kJ * (k * (kinc[0] * p.x + kinc[1] * p.y + kinc[2] * p.z)));
            V[id] = V[id] + r.len * o.gp.w[q] * (einc[0] * rn[0] + einc[1] * rn[1] + einc[2] * rn[2]) * e;
        }
    }
}

// Code/Vm.h:73-134 -- (n x H).rho, Code/OverloadAlgo.h:300-307
void rhs_M(const Oracle& o, cd k, const double* kinc, const double* hinc, std::vector<cd>& V) {
    V.assign(o.N, cd(0, 0));
    auto cdot = [](const V3& n, const double* h, const double* r) {
        double cx = n.y * h[2] - n.z * h[1], cy = n.z * h[0] - n.x * h[2], cz = n.x * h[1] - n.y * h[0];
        return cx * r[0] + cy * r[1] + cz * r[2];
    };
    for (int id = 0; id < o.N; ++id) {
        const Rwg& r = o.rwgs[id];
        const V3& np = o.mesh.normal[r.fpos];
        const V3& nn = o.mesh.normal[r.fneg];
        for (int q = 0; q < o.gp.n; ++q) {
            V3 p = gpt(o, r.fpos, q);
            const V3& fp = o.mesh.p[r.vpos];
            double rho[3] = {p.x - fp.x, p.y - fp.y, p.z - fp.z};
            cd e = std::exp(-kJ * (k * (kinc[0] * p.x + kinc[1] * p.y + kinc[2] * p.z)));
            V[id] = V[id] + r.len * o.gp.w[q] * cdot(np, hinc, rho) * e;
            p = gpt(o, r.fneg, q);
            const V3& fn = o.mesh.p[r.vneg];
            double rn[3] = {fn.x - p.x, fn.y - p.y, fn.z - p.z};
            e = std::exp(-kJ * (k * (kinc[0] * p.x + kinc[1] * p.y + kinc[2] * p.z)));
            V[id] = V[id] + r.len * o.gp.w[q] * cdot(nn, hinc, rn) * e;
        }
    }
}

// Code/Vm.h:136-148
void rhs_C(const Oracle& o, cd k, const double* kinc, const double* einc, const double* hinc, double a,
           std::vector<cd>& V) {
    std::vector<cd> V1, V2;
    rhs_E(o, k, kinc, einc, V1);
    rhs_M(o, k, kinc, hinc, V2);
    V.resize(o.N);
    for (int i = 0; i < o.N; ++i) V[i] = a * V1[i] + (1.0 - a) * V2[i];
}

// Code/Vm.h:150-203
void rhs_PMCHWT(const Oracle& o, cd k1, const double* kinc, const double* einc, const double* hinc,
                std::vector<cd>& V) {
    int N = o.N;
    V.assign(2 * N, cd(0, 0));
    for (int id = 0; id < N; ++id) {
        const Rwg& r = o.rwgs[id];
        for (int q = 0; q < o.gp.n; ++q) {
            V3 p = gpt(o, r.fpos, q);
            const V3& fp = o.mesh.p[r.vpos];
            double rho[3] = {p.x - fp.x, p.y - fp.y, p.z - fp.z};
            cd e = std::exp(-kJ * (k1 * (kinc[0] * p.x + kinc[1] * p.y + kinc[2] * p.z)));
            V[id] += r.len * o.gp.w[q] * (einc[0] * rho[0] + einc[1] * rho[1] + einc[2] * rho[2]) * e;
            V[N + id] += r.len * o.gp.w[q] * (hinc[0] * rho[0] + hinc[1] * rho[1] + hinc[2] * rho[2]) * e;
            p = gpt(o, r.fneg, q);
            const V3& fn = o.mesh.p[r.vneg];
            double rn[3] = {fn.x - p.x, fn.y - p.y, fn.z - p.z};
            e = std::exp(-kJ * (k1 * (kinc[0] * p.x + kinc[1] * p.y + kinc[2] * p.z)));
            V[id] += r.len * o.gp.w[q] * (einc[0] * rn[0] + einc[1] * rn[1] + einc[2] * rn[2]) * e;
            V[N + id] += r.len * o.gp.w[q] * (hinc[0] * rn[0] + hinc[1] * rn[1] + hinc[2] * rn[2]) * e;
        }
    }
}

// Code/RCS.h:54-114 -- jk eta sum_n I_n l_n sum_q w (I - r r).rho e^{+jk r.p}
void rcs_pec(const Oracle& o, const cd* I, cd k, cd eta, double th, double ph, cd out[3]) {
    double t = th / 180.0 * kPi, p = ph / 180.0 * kPi;
    double r[3] = {std::sin(t) * std::cos(p), std::sin(t) * std::sin(p), std::cos(t)};
    out[0] = out[1] = out[2] = 0.0;
    auto tens = [&](const double* rho, cd* F, cd e, double s) {
        double f0 = (1 - r[0] * r[0]) * rho[0] - r[0] * r[1] * rho[1] - r[0] * r[2] * rho[2];
        double f1 = -r[1] * r[0] * rho[0] + (1 - r[1] * r[1]) * rho[1] - r[1] * r[2] * rho[2];
        double f2 = -r[2] * r[0] * rho[0] - r[2] * r[1] * rho[1] + (1 - r[2] * r[2]) * rho[2];
        F[0] += s * f0 * e;
        F[1] += s * f1 * e;
        F[2] += s * f2 * e;
    };
    for (int id = 0; id < o.N; ++id) {
        const Rwg& rw = o.rwgs[id];
        cd F[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < o.gp.n; ++q) {
            double wf = o.gp.w[q];
            V3 a = gpt(o, rw.fpos, q);
            const V3& fp = o.mesh.p[rw.vpos];
            double rho[3] = {a.x - fp.x, a.y - fp.y, a.z - fp.z};
            tens(rho, F, std::exp(kJ * (k * (a.x * r[0] + a.y * r[1] + a.z * r[2]))), rw.len * wf);
            V3 b = gpt(o, rw.fneg, q);
            const V3& fn = o.mesh.p[rw.vneg];
            double rn[3] = {fn.x - b.x, fn.y - b.y, fn.z - b.z};
            tens(rn, F, std::exp(kJ * (k * (b.x * r[0] + b.y * r[1] + b.z * r[2]))), rw.len * wf);
        }
        cd coeff = kJ * k * eta;
        for (int c = 0; c < 3; ++c) out[c] += coeff * F[c] * I[id];
    }
}

// Code/RCS.h:117-192 -- -(-jk1) F_J I_J + (-jk1) F_M I_M, F_M = r x rho
void rcs_pmchwt(const Oracle& o, const cd* IJ, const cd* IM, cd k1, double th, double ph, cd out[3]) {
    double t = th / 180.0 * kPi, p = ph / 180.0 * kPi;
    double r[3] = {std::sin(t) * std::cos(p), std::sin(t) * std::sin(p), std::cos(t)};
    out[0] = out[1] = out[2] = 0.0;
    for (int id = 0; id < o.N; ++id) {
        const Rwg& rw = o.rwgs[id];
        cd FJ[3] = {0.0, 0.0, 0.0}, FM[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < o.gp.n; ++q) {
            double wf = o.gp.w[q];
            for (int side = 0; side < 2; ++side) {
                V3 a = gpt(o, side == 0 ? rw.fpos : rw.fneg, q);
                const V3& fv = o.mesh.p[side == 0 ? rw.vpos : rw.vneg];
                double rho[3];
                if (side == 0) {
                    rho[0] = a.x - fv.x; rho[1] = a.y - fv.y; rho[2] = a.z - fv.z;
                } else {
                    rho[0] = fv.x - a.x; rho[1] = fv.y - a.y; rho[2] = fv.z - a.z;
                }
                cd e = std::exp(kJ * (k1 * (a.x * r[0] + a.y * r[1] + a.z * r[2])));
                double f0 = (1 - r[0] * r[0]) * rho[0] - r[0] * r[1] * rho[1] - r[0] * r[2] * rho[2];
                double f1 = -r[1] * r[0] * rho[0] + (1 - r[1] * r[1]) * rho[1] - r[1] * r[2] * rho[2];
                double f2 = -r[2] * r[0] * rho[0] - r[2] * r[1] * rho[1] + (1 - r[2] * r[2]) * rho[2];
                FJ[0] += rw.len * wf * f0 * e;
                FJ[1] += rw.len * wf * f1 * e;
                FJ[2] += rw.len * wf * f2 * e;
                double c0 = r[1] * rho[2] - r[2] * rho[1], c1 = r[2] * rho[0] - r[0] * rho[2],
                       c2 = r[0] * rho[1] - r[1] * rho[0];
                FM[0] += rw.len * wf * c0 * e;
                FM[1] += rw.len * wf * c1 * e;
                FM[2] += rw.len * wf * c2 * e;
            }
        }
        cd coeff = -kJ * k1;
        for (int c = 0; c < 3; ++c) out[c] += -coeff * FJ[c] * IJ[id] + coeff * FM[c] * IM[id];
    }
}

}  // namespace orc
