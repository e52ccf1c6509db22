// C++ caller of the drop-in MLFMM wrapper (include/mlfma_dropin.hpp), written the
// way the reference drives its matvec: a host GMRES template with the argument
// order and `ax` callable of MGMRES::mgmres (Code/GMRES.h:50-66) receives
// MLFMM::ax(), so every Arnoldi step is one applyAx call through the C ABI.
//
//   dropin_gmres <mesh.nas> <integral_equation> <eps_re> <eps_im> <dump.txt>
//
// Writes the one-shot y = A * 1 dump to dump.txt, solves the theta = 0 plane-wave
// excitation with the host GMRES(30) to 1e-6 and with MLFMM::matrix_solver (the
// device GMRES), and prints one JSON line with both iteration counts, the true
// residuals and the relative difference of the two solutions.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mlfma_dropin.hpp"

using mlfma_dropin::Complex;

static double norm2(const std::vector<Complex>& v) {
    double s = 0;
    for (const Complex& a : v) s += std::norm(a);
    return std::sqrt(s);
}

// Restarted GMRES(mr), modified Gram-Schmidt Arnoldi, complex Givens rotations on
// the Hessenberg columns; returns true when ||b - A x|| <= max(tol_abs, tol_rel ||b||).
template <typename AxFunc>
bool gmres(int n, Complex* x, const Complex* rhs, int itr_max, int mr, double tol_abs, double tol_rel, AxFunc ax,
           int* iterations) {
    std::vector<Complex> b(rhs, rhs + n), r(n), w(n);
    const double tol = std::max(tol_abs, tol_rel * norm2(b));
    std::vector<std::vector<Complex>> V(mr + 1, std::vector<Complex>(n));
    std::vector<std::vector<Complex>> H(mr + 1, std::vector<Complex>(mr, Complex(0, 0)));
    std::vector<Complex> cs(mr), sn(mr), g(mr + 1);
    int its = 0;
    while (true) {
        ax(n, x, w.data());
        for (int i = 0; i < n; ++i) r[i] = b[i] - w[i];
        const double beta = norm2(r);
        if (beta <= tol || its >= itr_max) {
            *iterations = its;
            return beta <= tol;
        }
        for (int i = 0; i < n; ++i) V[0][i] = r[i] / beta;
        std::fill(g.begin(), g.end(), Complex(0, 0));
        g[0] = beta;
        int k = 0;
        for (; k < mr && its < itr_max; ++k, ++its) {
            ax(n, V[k].data(), w.data());
            for (int j = 0; j <= k; ++j) {
                Complex h(0, 0);
                for (int i = 0; i < n; ++i) h += std::conj(V[j][i]) * w[i];
                H[j][k] = h;
                for (int i = 0; i < n; ++i) w[i] -= h * V[j][i];
            }
            const double hn = norm2(w);
            H[k + 1][k] = hn;
            if (hn > 0)
                for (int i = 0; i < n; ++i) V[k + 1][i] = w[i] / hn;
            for (int j = 0; j < k; ++j) {   // earlier rotations on the new column
                const Complex t = std::conj(cs[j]) * H[j][k] + std::conj(sn[j]) * H[j + 1][k];
                H[j + 1][k] = -sn[j] * H[j][k] + cs[j] * H[j + 1][k];
                H[j][k] = t;
            }
            const double den = std::sqrt(std::norm(H[k][k]) + std::norm(H[k + 1][k]));
            cs[k] = H[k][k] / den;
            sn[k] = H[k + 1][k] / den;
            H[k][k] = den;
            H[k + 1][k] = 0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = std::conj(cs[k]) * g[k];
            if (std::abs(g[k + 1]) <= tol) {
                ++k;
                ++its;
                break;
            }
        }
        std::vector<Complex> y(k);   // back substitution H[0:k,0:k] y = g[0:k]
        for (int i = k - 1; i >= 0; --i) {
            Complex s = g[i];
            for (int j = i + 1; j < k; ++j) s -= H[i][j] * y[j];
            y[i] = s / H[i][i];
        }
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < n; ++i) x[i] += y[j] * V[j][i];
    }
}

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s mesh.nas integral_equation eps_re eps_im dump.txt\n", argv[0]);
        return 2;
    }
    try {
        mlfma_params p;
        mlfma_params_default(&p);
        p.integral_equation = std::atoi(argv[2]);
        p.eps_r_re = std::atof(argv[3]);
        p.eps_r_im = std::atof(argv[4]);
        mlfma_dropin::MLFMM m(argv[1], p);
        const int n = m.size();
        mlfma_dropin::MatVecExport::runAndExport(n, m.ax(), argv[5], p.integral_equation);

        bool threw = false;   // the reference's argument contract
        try {
            m.applyAx(0, nullptr, nullptr);
        } catch (const std::invalid_argument&) {
            threw = true;
        }

        std::vector<Complex> rhs = m.excitation(0.0, 0.0, 0.0);
        std::vector<Complex> xh(n, Complex(0, 0)), xd(n, Complex(0, 0)), w(n);
        int its_h = 0;
        const bool conv_h = gmres(n, xh.data(), rhs.data(), 3000, 30, 0.0, 1e-6, m.ax(), &its_h);
        mlfma_solve_stats st{};
        const bool conv_d = m.matrix_solver(n, xd.data(), rhs.data(), 3000, 30, 0.0, 1e-6, MLFMA_SOLVER_GMRES, &st);
        m.applyAx(n, xh.data(), w.data());
        for (int i = 0; i < n; ++i) w[i] = rhs[i] - w[i];
        const double res_h = norm2(w) / norm2(rhs);
        std::vector<Complex> d(n);
        for (int i = 0; i < n; ++i) d[i] = xh[i] - xd[i];
        std::printf("{\"n\": %d, \"invalid_argument_thrown\": %s, \"host_gmres_converged\": %s, \"host_iterations\": %d, "
                    "\"host_true_residual\": %.6e, \"device_converged\": %s, \"device_iterations\": %d, "
                    "\"device_true_residual\": %.6e, \"solution_rel_diff\": %.6e}\n",
                    n, threw ? "true" : "false", conv_h ? "true" : "false", its_h, res_h, conv_d ? "true" : "false",
                    st.iterations, st.rel_residual, norm2(d) / norm2(xd));
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
