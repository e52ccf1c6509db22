// ORACLE -- test infrastructure only (see orc.hpp).
// The assembled reference-shaped MLFMA operator (Code/MLFMM.h:11-82).
#pragma once
#include "orc.hpp"

namespace orc {

using Ws = std::vector<std::vector<std::vector<C3>>>;   // [level][node][k], level 0 = tree level 2

struct Oracle {
    int ie = 2;              // 0 EFIE, 1 CFIE, 2 PMCHWT (Code/RCSExportConfig.h:17)
    double alpha = 0.2;      // CFIE weight (Code/PhyConstant.h:15)
    // Translation-operator mode, a bit set.  0 = the reference's series (parity):
    // length min(L, ceil|kR|), Hankel functions at the real |kR|.  Bit 1: length L.
    // Bit 2: Hankel functions at the complex kR (accuracy mode for lossy media).
    int tfFull = 0;
    Mesh mesh;
    std::vector<Rwg> rwgs;
    Tree tree;
    Gauss gp;
    Wave wave;
    int N = 0, levelSpan = 0;
    std::vector<Grid> g1, g2;                         // [rev], rev 0 = finest level
    std::vector<std::vector<Interp>> ip1, ip2;        // [level], level 0 = tree level 2
    std::vector<std::vector<std::vector<cd>>> TF1, TF2;   // [level][vec][k]
    std::vector<std::vector<C3>> Vs1, Vf1, Vs2, Vf2;  // [rwg][k]
    std::vector<std::vector<cd>> Zn;                  // near-field rows
    std::vector<std::vector<int>> Zid;
    Ws SmJ1, BmJ1, SmM1, BmM1, SmJ2, BmJ2, SmM2, BmM2;

    void build_base();                 // grids (Code/AlgoMLFMM.h:21-138)
    void build_interp();               // Code/AlgoMLFMM.h:298-368
    void build_tf();                   // Code/AlgoMLFMM.h:192-295
    void build_patterns();             // Code/ProcessFMM.h:183-233
    void build_near();                 // Code/Zmartix.h:188-643
    void init_workspace();             // Code/MLFMM.cpp:115-161
    void fill_synthetic(unsigned seed);   // placeholder pattern/near values of the true shape
    int unknowns() const { return ie == 2 ? 2 * N : N; }

    // stages, Code/ProcessMLFMM.h
    void clear(Ws& Sm, Ws& Bm);
    void mexp(Ws& Sm, const std::vector<std::vector<C3>>& Vs, const cd* x, int off);
    void m2m(Ws& Sm, const std::vector<Grid>& g, const std::vector<std::vector<Interp>>& ip, cd k);
    void m2l(const Ws& Sm, Ws& Bm, const std::vector<std::vector<std::vector<cd>>>& TF, const std::vector<Grid>& g);
    void l2l(Ws& Bm, const std::vector<Grid>& g, const std::vector<std::vector<Interp>>& ip, cd k);
    void pass(Ws& Sm, Ws& Bm, int wn, const cd* x, int off);   // mexp, m2m, m2l, l2l

    void apply(const cd* x, cd* w);    // Code/MLFMM.cpp:189-369
    void dense(cd* Z);                 // all-pairs Galerkin matrix (test aid, O(N^2))
};

// Code/GMRES.h:65-217 and Code/CGS.h:104-234
template <typename Ax>
bool gmres(int n, cd* x, const cd* rhs, int itr_max, int mr, double tol_abs, double tol_rel, Ax ax, int* iters,
           double* resid);
bool gmres_oracle(Oracle& o, int n, cd* x, const cd* rhs, int itr_max, int mr, double tol_abs, double tol_rel,
                  int* iters, double* resid);
bool cgs_oracle(Oracle& o, int n, cd* x, const cd* b, int max_iter, double tol, int* iters, double* resid);

// Code/Vm.h:10-203
void rhs_E(const Oracle& o, cd k, const double* kinc, const double* einc, std::vector<cd>& V);
void rhs_M(const Oracle& o, cd k, const double* kinc, const double* hinc, std::vector<cd>& V);
void rhs_C(const Oracle& o, cd k, const double* kinc, const double* einc, const double* hinc, double a,
           std::vector<cd>& V);
void rhs_PMCHWT(const Oracle& o, cd k1, const double* kinc, const double* einc, const double* hinc,
                std::vector<cd>& V);
// Code/RCS.h:54-192 -- far-field vector (x,y,z) of the solved currents
void rcs_pec(const Oracle& o, const cd* I, cd k, cd eta, double th, double ph, cd out[3]);
void rcs_pmchwt(const Oracle& o, const cd* IJ, const cd* IM, cd k1, double th, double ph, cd out[3]);

}  // namespace orc
