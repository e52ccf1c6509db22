// Mesh, RWG numbering and the Morton octree, built with sorts over flat arrays.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>
#include <numeric>
#include <sstream>
#include <unordered_map>

#include "model.hpp"

namespace mlfma {

// Code/NASReader.cpp:60-144 -- "GRID* id x y", continuation "* id z",
// "CTRIA3 fid pid v1 v2 v3"; tokens are whitespace separated.
int read_nas(const std::string& path, std::vector<int>& ids, std::vector<double>& xyz, std::vector<int>& fids,
             std::vector<int>& fvid, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return -1;
    }
    std::string line;
    bool pending = false;
    int pid = 0;
    double px = 0, py = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.rfind("GRID*", 0) == 0) {
            std::istringstream is(line.substr(5));
            if (!(is >> pid >> px >> py)) {
                err = "invalid GRID* line: " + line;
                return -1;
            }
            pending = true;
        } else if (pending && !line.empty() && line[0] == '*') {
            std::istringstream is(line.substr(1));
            int id2;
            double z;
            if (!(is >> id2 >> z) || id2 != pid) {
                err = "invalid GRID* continuation: " + line;
                return -1;
            }
            ids.push_back(pid);
            xyz.insert(xyz.end(), {px, py, z});
            pending = false;
        } else if (line.rfind("CTRIA3", 0) == 0) {
            std::istringstream is(line.substr(6));
            int f, prop, a, b, c;
            if (!(is >> f >> prop >> a >> b >> c)) {
                err = "invalid CTRIA3 line: " + line;
                return -1;
            }
            fids.push_back(f);
            fvid.insert(fvid.end(), {a, b, c});
        }
    }
    if (pending) {
        err = "unfinished GRID* entry at end of file";
        return -1;
    }
    return 0;
}

bool build_mesh(int np, const int* ids, const double* xyz, int nf, const int* fids, const int* fvid, Mesh& m,
                std::string& err) {
    m.point_id.assign(ids, ids + np);
    m.pts.resize(np);
    std::unordered_map<int, int> at;
    at.reserve(np * 2);
    for (int i = 0; i < np; ++i) {
        m.pts[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        at[ids[i]] = i;
    }
    m.face_id.assign(fids, fids + nf);
    m.tri.resize(3 * (size_t)nf);
    m.area.resize(nf);
    m.normal.resize(nf);
    for (int f = 0; f < nf; ++f) {
        for (int k = 0; k < 3; ++k) {
            auto it = at.find(fvid[3 * f + k]);
            if (it == at.end()) {
                err = "face " + std::to_string(fids[f]) + " references a missing point";
                return false;
            }
            m.tri[3 * f + k] = it->second;
        }
        // Code/NASReader.cpp:25-58
        const Vec3& a = m.pts[m.tri[3 * f]];
        const Vec3& b = m.pts[m.tri[3 * f + 1]];
        const Vec3& c = m.pts[m.tri[3 * f + 2]];
        double abx = b.x - a.x, aby = b.y - a.y, abz = b.z - a.z;
        double acx = c.x - a.x, acy = c.y - a.y, acz = c.z - a.z;
        Vec3 cr = {aby * acz - abz * acy, abz * acx - abx * acz, abx * acy - aby * acx};
        double len = std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
        m.area[f] = len / 2.0;
        Vec3 n = cr;
        if (len > 0) n = {n.x / len, n.y / len, n.z / len};
        if (std::fabs(n.x) < 1e-6) n.x = 0;
        if (std::fabs(n.y) < 1e-6) n.y = 0;
        if (std::fabs(n.z) < 1e-6) n.z = 0;
        m.normal[f] = n;
        if (!(m.area[f] > 0)) {
            err = "degenerate face " + std::to_string(fids[f]);
            return false;
        }
    }
    return true;
}

// Code/RWGGenerator.cpp:39-113: every edge shared by exactly two faces becomes a
// basis function; functions are numbered in ascending (min id, max id) order;
// the positive face is the one that comes first in the file.
void build_basis(const Mesh& m, Basis& b) {
    const int nf = (int)m.face_id.size();
    struct E {
        int lo, hi, face;
    };
    std::vector<E> e(3 * (size_t)nf);
    for (int f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) {
            int a = m.point_id[m.tri[3 * f + k]], c = m.point_id[m.tri[3 * f + (k + 1) % 3]];
            e[3 * f + k] = {std::min(a, c), std::max(a, c), f};
        }
    std::stable_sort(e.begin(), e.end(), [](const E& x, const E& y) {
        return x.lo < y.lo || (x.lo == y.lo && x.hi < y.hi);
    });
    std::unordered_map<int, int> at;
    for (int i = 0; i < (int)m.point_id.size(); ++i) at[m.point_id[i]] = i;
    b = Basis();
    for (size_t i = 0; i < e.size();) {
        size_t j = i;
        while (j < e.size() && e[j].lo == e[i].lo && e[j].hi == e[i].hi) ++j;
        if (j - i == 2) {
            int fp = e[i].face, fn = e[i + 1].face;
            int p0 = at[e[i].lo], p1 = at[e[i].hi];
            auto freev = [&](int f) {
                for (int k = 0; k < 3; ++k) {
                    int id = m.point_id[m.tri[3 * f + k]];
                    if (id != e[i].lo && id != e[i].hi) return m.tri[3 * f + k];
                }
                return -1;
            };
            b.fpos.push_back(fp);
            b.fneg.push_back(fn);
            b.e0.push_back(p0);
            b.e1.push_back(p1);
            b.vpos.push_back(freev(fp));
            b.vneg.push_back(freev(fn));
            const Vec3& a = m.pts[p0];
            const Vec3& c = m.pts[p1];
            b.len.push_back(std::sqrt(std::pow(a.x - c.x, 2) + std::pow(a.y - c.y, 2) + std::pow(a.z - c.z, 2)));
            b.mid.push_back({(a.x + c.x) / 2.0, (a.y + c.y) / 2.0, (a.z + c.z) / 2.0});
        }
        i = j;
    }
    b.n = (int)b.len.size();
}

namespace {
uint64_t spread(uint32_t x) {
    uint64_t v = x & 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}
uint32_t gather(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffffULL;
    return (uint32_t)x;
}
// Morton interleave, x in bit 0 (Code/OCTree.h:110-124)
uint64_t morton(int x, int y, int z) { return spread(x) | (spread(y) << 1) | (spread(z) << 2); }
int find(const std::vector<uint64_t>& codes, uint64_t c) {
    auto it = std::lower_bound(codes.begin(), codes.end(), c);
    return (it != codes.end() && *it == c) ? (int)(it - codes.begin()) : -1;
}
}  // namespace

void build_octree(const Basis& b, double wavelength, double leaf_coeff, Octree& t) {
    // Code/OCTree.cpp:161-242 -- cube around the RWG edge midpoints, enlarged 0.1%,
    // depth ceil(log2(root / (coeff*lambda))) clamped to [2, 15].
    Vec3 lo = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
               std::numeric_limits<double>::max()};
    Vec3 hi = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
               std::numeric_limits<double>::lowest()};
    for (const Vec3& p : b.mid) {
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    double root = std::max(hi.x - lo.x, std::max(hi.y - lo.y, hi.z - lo.z)) * 1.001;
    double target = leaf_coeff * wavelength;
    double ratio = std::max(1.0, root / target);
    int L = std::max(2, std::min((int)std::ceil(std::log2(ratio)), 15));
    root = target * (1 << L);
    Vec3 c = {0.5 * (lo.x + hi.x), 0.5 * (lo.y + hi.y), 0.5 * (lo.z + hi.z)};
    t.bmin = {c.x - 0.5 * root, c.y - 0.5 * root, c.z - 0.5 * root};
    t.leaf_len = root / (1 << L);
    t.max_level = L;
    t.lv.assign(L + 1, Level());

    // leaves (Code/OCTree.cpp:244-307): sort functions by leaf code, keep id order inside
    const int n = b.n, maxIdx = (1 << L) - 1;
    std::vector<uint64_t> code(n);
    std::vector<int> ixyz(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        int ix = (int)((b.mid[i].x - t.bmin.x) / t.leaf_len);
        int iy = (int)((b.mid[i].y - t.bmin.y) / t.leaf_len);
        int iz = (int)((b.mid[i].z - t.bmin.z) / t.leaf_len);
        ix = std::max(0, std::min(ix, maxIdx));
        iy = std::max(0, std::min(iy, maxIdx));
        iz = std::max(0, std::min(iz, maxIdx));
        ixyz[3 * i] = ix;
        ixyz[3 * i + 1] = iy;
        ixyz[3 * i + 2] = iz;
        code[i] = morton(ix, iy, iz);
    }
    t.perm.resize(n);
    std::iota(t.perm.begin(), t.perm.end(), 0);
    std::stable_sort(t.perm.begin(), t.perm.end(), [&](int a, int c2) { return code[a] < code[c2]; });
    Level& leaf = t.lv[L];
    leaf.len = t.leaf_len;
    t.leaf_ptr.clear();
    for (int i = 0; i < n; ++i) {
        int r = t.perm[i];
        if (i == 0 || code[r] != code[t.perm[i - 1]]) {
            t.leaf_ptr.push_back(i);
            leaf.code.push_back(code[r]);
            leaf.centre.push_back({t.bmin.x + (ixyz[3 * r] + 0.5) * t.leaf_len,
                                   t.bmin.y + (ixyz[3 * r + 1] + 0.5) * t.leaf_len,
                                   t.bmin.z + (ixyz[3 * r + 2] + 0.5) * t.leaf_len});
        }
    }
    t.leaf_ptr.push_back(n);
    leaf.n = (int)leaf.code.size();

    // parents (Code/OCTree.cpp:309-375): code >> 3, centre from the first child
    for (int l = L; l > 0; --l) {
        Level& ch = t.lv[l];
        Level& pa = t.lv[l - 1];
        pa.len = ch.len * 2.0;
        ch.parent.assign(ch.n, -1);
        pa.son_ptr.clear();
        pa.son_idx.clear();
        for (int i = 0; i < ch.n; ++i) {
            uint64_t pc = ch.code[i] >> 3;
            if (pa.code.empty() || pa.code.back() != pc) {
                int ci = (int)(ch.code[i] & 7);
                double off = ch.len * 0.5;
                Vec3 pcn = ch.centre[i];
                pcn.x += (ci & 1) ? -off : off;
                pcn.y += (ci & 2) ? -off : off;
                pcn.z += (ci & 4) ? -off : off;
                pa.code.push_back(pc);
                pa.centre.push_back(pcn);
                pa.son_ptr.push_back((int)pa.son_idx.size());
            }
            ch.parent[i] = (int)pa.code.size() - 1;
            pa.son_idx.push_back(i);
        }
        pa.son_ptr.push_back((int)pa.son_idx.size());
        pa.n = (int)pa.code.size();
    }
    t.lv[0].parent.assign(t.lv[0].n, -1);
    t.lv[L].son_ptr.assign(t.lv[L].n + 1, 0);

    // near neighbours (Code/OCTree.cpp:385-426)
    for (int l = 0; l <= L; ++l) {
        Level& v = t.lv[l];
        int g = 1 << l;
        v.near_ptr.assign(1, 0);
        v.near_idx.clear();
        for (int i = 0; i < v.n; ++i) {
            int x = gather(v.code[i]), y = gather(v.code[i] >> 1), z = gather(v.code[i] >> 2);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        if (!dx && !dy && !dz) continue;
                        int a = x + dx, bb = y + dy, cc = z + dz;
                        if (a < 0 || bb < 0 || cc < 0 || a >= g || bb >= g || cc >= g) continue;
                        int j = find(v.code, morton(a, bb, cc));
                        if (j >= 0) v.near_idx.push_back(j);
                    }
            v.near_ptr.push_back((int)v.near_idx.size());
        }
    }

    // distant relatives (Code/OCTree.cpp:428-498) and unique translation vectors
    // in first-met order (:500-551)
    for (int l = 0; l <= L; ++l) {
        Level& v = t.lv[l];
        v.dr_ptr.assign(1, 0);
        v.dr_idx.clear();
        v.dr_vec.clear();
        v.drvec.clear();
        if (l < 2) {
            v.dr_ptr.assign(v.n + 1, 0);
            continue;
        }
        const Level& pl = t.lv[l - 1];
        int g = 1 << (l - 1);
        std::unordered_map<uint64_t, int> vid;
        for (int i = 0; i < v.n; ++i) {
            uint64_t pc = pl.code[v.parent[i]];
            int px = gather(pc), py = gather(pc >> 1), pz = gather(pc >> 2);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        int a = px + dx, bb = py + dy, cc = pz + dz;
                        if (a < 0 || bb < 0 || cc < 0 || a >= g || bb >= g || cc >= g) continue;
                        for (int cx = 0; cx < 2; ++cx)
                            for (int cy = 0; cy < 2; ++cy)
                                for (int cz = 0; cz < 2; ++cz) {
                                    int j = find(v.code, morton(2 * a + cx, 2 * bb + cy, 2 * cc + cz));
                                    if (j < 0 || j == i) continue;
                                    const Vec3& p = v.centre[i];
                                    const Vec3& q = v.centre[j];
                                    double d = std::max({std::fabs(p.x - q.x), std::fabs(p.y - q.y),
                                                         std::fabs(p.z - q.z)});
                                    if (d <= 1.74 * v.len) continue;
                                    int kx = (int)std::round((p.x - q.x) / v.len);
                                    int ky = (int)std::round((p.y - q.y) / v.len);
                                    int kz = (int)std::round((p.z - q.z) / v.len);
                                    uint64_t key = ((uint64_t)(kx + 1024) << 22) | ((uint64_t)(ky + 1024) << 11) |
                                                   (uint64_t)(kz + 1024);
                                    auto it = vid.find(key);
                                    int vi;
                                    if (it == vid.end()) {
                                        vi = (int)v.drvec.size();
                                        vid.emplace(key, vi);
                                        v.drvec.push_back({kx * v.len, ky * v.len, kz * v.len});
                                    } else {
                                        vi = it->second;
                                    }
                                    v.dr_idx.push_back(j);
                                    v.dr_vec.push_back(vi);
                                }
                    }
            v.dr_ptr.push_back((int)v.dr_idx.size());
        }
    }
}

}  // namespace mlfma
