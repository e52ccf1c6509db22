// mlfma_dropin.hpp -- C++ drop-in over the C ABI of libmlfma.so (mlfma.h) for
// code written against the reference's MLFMM class.
//
// * MLFMM::applyAx(int n, const std::complex<double>* x, std::complex<double>* w)
//   has the signature the reference's design proposes for exporting its matvec
//   (README/矩阵向量乘法导出头文件设计.md:392-437): w is overwritten with A x in
//   the reference's unknown order, and n <= 0 or a null pointer throws
//   std::invalid_argument.  It is the `ax` callable MGMRES::mgmres and
//   MCGS::CGS::solve take (Code/GMRES.h:50-66, Code/CGS.h:54-60; the lambda of
//   Code/MLFMM.cpp:189-369): pass MLFMM::ax() where the reference passes its lambda.
// * MLFMM::matrix_solver keeps the argument list of Code/MLFMM.cpp:370-397 and
//   runs the solve on the device (mlfma_solve); it returns false when the
//   iteration limit is reached, as the reference's solvers do.
// * MatVecExport writes the one-shot y = A * 1 dump of the same design
//   (README/矩阵向量乘法导出头文件设计.md:75-115): '#' header lines, then
//   "index x_real x_imag y_real y_imag abs_y" per unknown, 17 significant digits.
//
// Errors of the library other than bad arguments surface as std::runtime_error
// carrying mlfma_last_error().  One MLFMM object owns one handle (one GPU
// operator); it is movable, not copyable, and not thread-safe.
#pragma once

#include <complex>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mlfma.h"

namespace mlfma_dropin {

using Complex = std::complex<double>;
using Vector = std::vector<Complex>;

inline void check(int rc, const mlfma_handle* h, const char* what) {
    if (rc == MLFMA_OK || rc == MLFMA_ERR_NOT_CONVERGED) return;
    const std::string msg = std::string(what) + ": " + mlfma_last_error(h);
    if (rc == MLFMA_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

class MLFMM {
public:
    // Builds the operator of a NAS mesh (Code/NASReader.cpp:60-144) with the given parameters;
    // mlfma_params_default() gives the reference's compiled constants.
    MLFMM(const std::string& nas_path, const mlfma_params& params) {
        check(mlfma_create_nas(&params, nas_path.c_str(), &h_), nullptr, "mlfma_create_nas");
        long long d[6] = {0, 0, 0, 0, 0, 0};
        check(mlfma_dims(h_, d), h_, "mlfma_dims");
        n_ = static_cast<int>(d[1]);
    }
    MLFMM(const MLFMM&) = delete;
    MLFMM& operator=(const MLFMM&) = delete;
    MLFMM(MLFMM&& o) noexcept : h_(o.h_), n_(o.n_) { o.h_ = nullptr; }
    MLFMM& operator=(MLFMM&& o) noexcept {
        if (this != &o) {
            if (h_) mlfma_destroy(h_);
            h_ = o.h_;
            n_ = o.n_;
            o.h_ = nullptr;
        }
        return *this;
    }
    ~MLFMM() {
        if (h_) mlfma_destroy(h_);
    }

    int size() const { return n_; }
    mlfma_handle* handle() const { return h_; }

    // w = A x (w overwritten), the reference's ax(n, x, w).
    void applyAx(int n, const Complex* x, Complex* w) {
        if (n <= 0 || x == nullptr || w == nullptr) throw std::invalid_argument("MLFMM::applyAx invalid input.");
        if (n != n_) throw std::invalid_argument("MLFMM::applyAx: n is not the operator's unknown count.");
        check(mlfma_apply(h_, n, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(w)), h_,
              "mlfma_apply");
    }

    // The callable a MGMRES::mgmres / MCGS::CGS::solve style solver is given.
    auto ax() {
        return [this](int n, const Complex* x, Complex* w) { applyAx(n, x, w); };
    }

    // Plane-wave excitation (Code/Vm.h:10-203), pol/theta/phi in degrees.
    Vector excitation(double pol_deg, double theta_deg, double phi_deg) {
        Vector v(static_cast<size_t>(n_));
        check(mlfma_excitation(h_, pol_deg, theta_deg, phi_deg, reinterpret_cast<double*>(v.data())), h_,
              "mlfma_excitation");
        return v;
    }

    // Code/MLFMM.cpp:370-397: x holds the initial guess and receives the solution.
    // solver: MLFMA_SOLVER_GMRES (restart mr) or MLFMA_SOLVER_CGS.  Returns converged.
    bool matrix_solver(int n, Complex x[], Complex rhs[], int itr_max, int mr, double tol_abs, double tol_rel,
                       int solver = MLFMA_SOLVER_GMRES, mlfma_solve_stats* stats = nullptr) {
        if (n <= 0 || x == nullptr || rhs == nullptr) throw std::invalid_argument("MLFMM::matrix_solver invalid input.");
        mlfma_solve_params sp{solver, itr_max, mr, tol_rel, tol_abs, 0};
        mlfma_solve_stats st{};
        const int rc = mlfma_solve(h_, &sp, n, reinterpret_cast<const double*>(rhs), reinterpret_cast<double*>(x), &st);
        check(rc, h_, "mlfma_solve");
        if (stats) *stats = st;
        return rc == MLFMA_OK;
    }

private:
    mlfma_handle* h_ = nullptr;
    int n_ = 0;
};

// One-shot matvec export of the reference's design (README/矩阵向量乘法导出头文件设计.md:20-180).
namespace MatVecExport {

struct Result {
    Vector x;
    Vector y;
};

inline Vector makeOnesVector(int n) {
    if (n <= 0) throw std::invalid_argument("MatVecExport::makeOnesVector: n must be positive.");
    return Vector(static_cast<size_t>(n), Complex(1.0, 0.0));
}

template <typename Ax>
Vector applyOnce(int n, const Vector& x, Ax&& ax) {
    if (n <= 0) throw std::invalid_argument("MatVecExport::applyOnce: n must be positive.");
    if (static_cast<int>(x.size()) != n) throw std::invalid_argument("MatVecExport::applyOnce: x size mismatch.");
    Vector y(static_cast<size_t>(n), Complex(0.0, 0.0));
    ax(n, x.data(), y.data());
    return y;
}

inline void exportResultTxt(const std::string& path, const Result& r, const std::string& algorithm,
                            int selectAlgorithm, int selectIntegralEqu) {
    if (r.x.size() != r.y.size()) throw std::invalid_argument("MatVecExport::exportResultTxt: x and y size mismatch.");
    std::ofstream out(path);
    if (!out) throw std::runtime_error("MatVecExport::exportResultTxt: cannot open " + path);
    out << std::setprecision(17);
    out << "# MatVec one-shot result\n"
        << "# algorithm = " << algorithm << '\n'
        << "# selectAlgorithm = " << selectAlgorithm << '\n'
        << "# selectIntegralEqu = " << selectIntegralEqu << '\n'
        << "# size = " << r.x.size() << '\n'
        << "# x is fixed to complex<double>(1.0, 0.0)\n"
        << "# columns: index x_real x_imag y_real y_imag abs_y\n";
    for (size_t i = 0; i < r.x.size(); ++i)
        out << i << '\t' << r.x[i].real() << '\t' << r.x[i].imag() << '\t' << r.y[i].real() << '\t' << r.y[i].imag()
            << '\t' << std::abs(r.y[i]) << '\n';
    if (!out) throw std::runtime_error("MatVecExport::exportResultTxt: write failed for " + path);
}

// y = A * 1 through ax, written to path (algorithm "MLFMM"; selectAlgorithm 2 = MLFMA, Code/main.cpp:16-26).
template <typename Ax>
Result runAndExport(int n, Ax&& ax, const std::string& path, int selectIntegralEqu) {
    Result r;
    r.x = makeOnesVector(n);
    r.y = applyOnce(n, r.x, ax);
    exportResultTxt(path, r, "MLFMM", 2, selectIntegralEqu);
    return r;
}

}  // namespace MatVecExport
}  // namespace mlfma_dropin
