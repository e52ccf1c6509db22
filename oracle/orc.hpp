// ORACLE -- test infrastructure only.
//
// CPU FP64 restatement of the reference MLFMA / PMCHWT pipeline
// (CCCCl0uD/MLFMA_PMCHWT, Code/*). It exists to check the HIP product path
// and to provide the timed CPU baseline.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it; the product library never
// links or calls it.
//
// Every function cites the reference file:line whose behaviour it restates.
// Node order: the reference orders octree nodes by std::unordered_map
// iteration (Code/OCTree.cpp:298,368), which is implementation-defined; the
// oracle orders every level by Morton code instead.  This changes only the
// floating-point summation order of the matvec, never which terms are summed.
#pragma once
#include <array>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

using cd = std::complex<double>;

// Code/PhyConstant.h:7-15
constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = 299792458.0;
constexpr double kEps0 = 2.2204e-16;
const cd kJ(0.0, 1.0);

struct V3 {
    double x = 0, y = 0, z = 0;
};
struct C3 {
    cd x, y, z;
};

inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------- mesh / RWG
struct Mesh {
    std::vector<int> pid;                   // point ids as written in the NAS file
    std::vector<V3> p;                      // point coordinates
    std::vector<int> fid;                   // face ids
    std::vector<std::array<int, 3>> fv;     // face vertex indices into p
    std::vector<double> area;
    std::vector<V3> normal, centre;
};

struct Rwg {
    int fpos = -1, fneg = -1;   // face indices (first / second face in file order)
    int e0 = -1, e1 = -1;       // common edge vertex indices (smaller id, larger id)
    int vpos = -1, vneg = -1;   // free vertex indices
    double len = 0;
    V3 mid;
};

// ---------------------------------------------------------------- octree
struct Node {
    uint64_t code = 0;
    int level = 0;
    double len = 0;
    V3 c;
    int parent = -1;
    std::vector<int> sons, rwg, nearn, dr, drv;
};

struct Tree {
    int maxLevel = 0;
    double leafLen = 0;
    V3 bmin;
    std::vector<std::vector<Node>> lv;       // lv[level][i], Morton-sorted
    std::vector<std::vector<V3>> drvec;      // drvec[level-2][v]
};

// ---------------------------------------------------------------- numerics
struct Gauss {
    int n = 0;
    std::vector<double> l1, l2, l3, w;
};

struct Wave {
    double freq = 0, lambda = 0, k0 = 0;
    cd k1, k2, eta1, eta2, etai, epsr{1, 0}, mur{1, 0};
};

struct PlaneWave {
    double kinc[3], einc[3], hinc[3], vth[3], vph[3];
};

struct Interp {
    int tIdx[3], pIdx[3], tp[9];
    double w[9];
};

struct Grid {      // angular grid of one tree level (one wavenumber)
    int L = 0;
    std::vector<double> theta, phi, wgl;   // wgl in Gauss-Legendre node order
    double dphi = 0;
    std::vector<V3> kp;                    // (theta, phi) row-major
};

// ---------------------------------------------------------------- API used by orc_capi.cpp
void mesh_from_arrays(Mesh& m, int np, const int* ids, const double* xyz, int nf, const int* fids, const int* fvid);
int read_nas(const std::string& path, Mesh& m);
void build_rwg(const Mesh& m, std::vector<Rwg>& rwgs);
void build_tree(const std::vector<Rwg>& rwgs, double wavelength, double leafCoeff, Tree& t);

Gauss gauss_points(int n);
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
double legendre_p(int l, double x);
double legendre_p_prime(int l, double x);
void sph_jy(int L, double x, std::vector<double>& j, std::vector<double>& y);
void sph_h2(int L, cd z, std::vector<cd>& h);

Wave make_wave(double freq, cd epsr, cd mur, bool dielectric);
PlaneWave init_pw(double pol_deg, double th_deg, double ph_deg);

// near-field Galerkin entries (Code/Operator.h, Code/EFIE.cpp, Code/MFIE.cpp)
struct TriRef {
    V3 v1, v2, v3, vfree, n;
    double area;
    int id;
};
cd op_L(const Mesh& m, const Rwg& f, const Rwg& s, cd k, const Gauss& gp);
cd op_K(const Mesh& m, const Rwg& f, const Rwg& s, cd k, const Gauss& gp);
cd op_EFIE(const Mesh& m, const Rwg& f, const Rwg& s, cd k, cd eta, const Gauss& gp);
cd op_MFIE(const Mesh& m, const Rwg& f, const Rwg& s, cd k, const Gauss& gp);
void sing_probe(const double* tri9, const double* n3, const double* r3, double* out);

}  // namespace orc
