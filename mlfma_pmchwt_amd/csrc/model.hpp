// Host-side model of one scattering problem: mesh, RWG basis, octree and the
// MLFMA tables, in flat arrays ready to upload to the device.
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace mlfma {

using cplx = std::complex<double>;

constexpr double kPi = 3.14159265358979323846;   // Code/PhyConstant.h:9
constexpr double kC0 = 299792458.0;              // Code/PhyConstant.h:11
constexpr double kEps0 = 2.2204e-16;             // Code/PhyConstant.h:13

struct Vec3 {
    double x = 0, y = 0, z = 0;
};

// ------------------------------------------------------------------ mesh
struct Mesh {
    std::vector<int> point_id;
    std::vector<Vec3> pts;
    std::vector<int> face_id;
    std::vector<int> tri;          // 3 point indices per face (file order)
    std::vector<double> area;
    std::vector<Vec3> normal;      // snapped to 0 below 1e-6 (Code/NASReader.cpp:52-56)
};

// RWG functions in reference numbering (sorted edge key, Code/RWGGenerator.cpp:39-113).
struct Basis {
    int n = 0;
    std::vector<int> fpos, fneg;   // face indices
    std::vector<int> e0, e1;       // common-edge point indices (smaller id first)
    std::vector<int> vpos, vneg;   // free vertices
    std::vector<double> len;
    std::vector<Vec3> mid;
};

// ------------------------------------------------------------------ octree
struct Level {
    int n = 0;
    double len = 0;                    // box edge
    std::vector<uint64_t> code;        // Morton code, ascending
    std::vector<Vec3> centre;
    std::vector<int> parent;           // index into level-1
    std::vector<int> son_ptr, son_idx;
    std::vector<int> near_ptr, near_idx;   // 26-neighbour list, (dx,dy,dz) order
    std::vector<int> dr_ptr, dr_idx, dr_vec;   // distant relatives and vector ids
    std::vector<Vec3> drvec;           // unique translation vectors
};

struct Octree {
    int max_level = 0;
    double leaf_len = 0;
    Vec3 bmin;
    std::vector<Level> lv;             // lv[0..max_level]
    // leaves: RWG functions in leaf-major (internal) order
    std::vector<int> perm;             // internal -> reference RWG id
    std::vector<int> leaf_ptr;         // internal RWG range of each leaf
};

// ------------------------------------------------------------------ angular tables
struct Grid {
    int L = 0;
    std::vector<double> theta, phi, wgl;   // wgl in Gauss-Legendre node order
    double dphi = 0;
    std::vector<Vec3> kp;                  // (theta, phi) row-major
    int K() const { return (int)kp.size(); }
};

// Per tree level l >= 3: interpolation child grid (l) -> parent grid (l-1).
struct InterpTable {
    int Kp = 0, Kc = 0;
    std::vector<int> tap;        // [9][Kp] child sample index (plane per tap)
    std::vector<double> w;       // [9][Kp]
    // the same stencil by direction: tap (3 a + b) is child sample
    // th_tap[i][a] * 2 Lc + ph_tap[j][b] with weight th_w[i][a] * ph_w[j][b]
    // for parent sample i * 2 Lp + j
    int Lp = 0, Lc = 0;
    std::vector<int> th_tap, ph_tap;    // [Lp][3], [2 Lp][3]
    std::vector<double> th_w, ph_w;     // [Lp][3], [2 Lp][3]
    // transposed (gather form of the anterpolation): for each child sample the
    // parent samples that use it, ascending
    std::vector<int> t_ptr;      // Kc + 1
    std::vector<int> t_kp;
    std::vector<double> t_w;
    // the same by direction: for each child theta row the (parent row, weight)
    // pairs with th_tap[row][a] = child row, ascending parent row, and for each
    // child phi column the (parent column, weight) pairs with ph_tap (the
    // weight of a parent sample in a child sample is the product of the two)
    std::vector<int> at_ptr, at_idx;    // Lc + 1, entries
    std::vector<double> at_w;
    std::vector<int> ap_ptr, ap_idx;    // 2 Lc + 1, entries
    std::vector<double> ap_w;
};

struct WaveTables {
    cplx k;
    std::vector<Grid> grid;                  // indexed by tree level (2..max_level)
    std::vector<InterpTable> interp;         // indexed by child tree level (3..max_level)
    std::vector<std::vector<cplx>> tf;       // [level][vec*K + q], quadrature weight NOT applied
    std::vector<std::vector<double>> qw;     // [level][q] = w_theta * dphi
    std::vector<std::vector<cplx>> shift_up;     // [child level][oct*Kp + q] exp(-jk khat.(F - c))
    std::vector<std::vector<cplx>> shift_down;   // [child level][oct*Kp + q] exp(-jk khat.(c - F))
};

struct Medium {
    double freq = 0, lambda = 0, k0 = 0;
    cplx k1, k2, eta1, eta2, etai;
};

// ------------------------------------------------------------------ builders (host)
int read_nas(const std::string& path, std::vector<int>& ids, std::vector<double>& xyz, std::vector<int>& fids,
             std::vector<int>& fvid, std::string& err);
bool build_mesh(int np, const int* ids, const double* xyz, int nf, const int* fids, const int* fvid, Mesh& m,
                std::string& err);
void build_basis(const Mesh& m, Basis& b);
void build_octree(const Basis& b, double wavelength, double leaf_coeff, Octree& t);

Medium make_medium(double freq, cplx epsr, cplx mur, bool dielectric);
void build_wave_tables(const Octree& t, cplx k, double kabs, int tf_mode, WaveTables& w);

// Dunavant rule (weights halved), Code/GaussPoints.cpp:29-284
bool dunavant(int n, std::vector<double>& l1, std::vector<double>& l2, std::vector<double>& l3,
              std::vector<double>& w);
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
void spherical_jy(int L, double x, std::vector<double>& j, std::vector<double>& y);
void spherical_h2(int L, cplx z, std::vector<cplx>& h);

}  // namespace mlfma
