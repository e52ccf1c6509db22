// ORACLE -- test infrastructure only (see orc.hpp).
// Mesh, RWG and Morton octree restated from the reference.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <unordered_map>

#include "orc.hpp"

namespace orc {

// Code/NASReader.cpp:25-58 -- area, centroid, normal snapped to 0 below 1e-6.
static void face_geometry(Mesh& m) {
    size_t nf = m.fv.size();
    m.area.resize(nf);
    m.normal.resize(nf);
    m.centre.resize(nf);
    for (size_t f = 0; f < nf; ++f) {
        const V3& a = m.p[m.fv[f][0]];
        const V3& b = m.p[m.fv[f][1]];
        const V3& c = m.p[m.fv[f][2]];
        V3 ab = sub(b, a), ac = sub(c, a);
        V3 cr = {ab.y * ac.z - ab.z * ac.y, ab.z * ac.x - ab.x * ac.z, ab.x * ac.y - ab.y * ac.x};
        double len = std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
        m.area[f] = len / 2.0;
        m.centre[f] = {(a.x + b.x + c.x) / 3.0, (a.y + b.y + c.y) / 3.0, (a.z + b.z + c.z) / 3.0};
        V3 n = cr;
        if (len > 0) n = {n.x / len, n.y / len, n.z / len};
        if (std::fabs(n.x) < 1e-6) n.x = 0;
        if (std::fabs(n.y) < 1e-6) n.y = 0;
        if (std::fabs(n.z) < 1e-6) n.z = 0;
        m.normal[f] = n;
    }
}

void mesh_from_arrays(Mesh& m, int np, const int* ids, const double* xyz, int nf, const int* fids,
                      const int* fvid) {
    m = Mesh();
    m.pid.assign(ids, ids + np);
    m.p.resize(np);
    std::unordered_map<int, int> idx;
    for (int i = 0; i < np; ++i) {
        m.p[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        idx[ids[i]] = i;   // later duplicates win, like the reference's linear scan
    }
    m.fid.assign(fids, fids + nf);
    m.fv.resize(nf);
    for (int f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) m.fv[f][k] = idx.at(fvid[3 * f + k]);
    face_geometry(m);
}

// Code/NASReader.cpp:60-144 -- "GRID* id x y" + "* id z" and "CTRIA3 fid pid v1 v2 v3".
int read_nas(const std::string& path, Mesh& m) {
    std::ifstream in(path);
    if (!in) return -1;
    std::vector<int> ids, fids, fvid;
    std::vector<double> xyz;
    std::string line;
    bool cont = false;
    int curId = 0;
    double cx = 0, cy = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.compare(0, 5, "GRID*") == 0) {
            std::istringstream is(line.substr(5));
            if (!(is >> curId >> cx >> cy)) return -2;
            cont = true;
        } else if (cont && !line.empty() && line[0] == '*') {
            std::istringstream is(line.substr(1));
            int id2;
            double z;
            if (!(is >> id2 >> z)) return -3;
            if (id2 != curId) return -4;
            ids.push_back(curId);
            xyz.push_back(cx);
            xyz.push_back(cy);
            xyz.push_back(z);
            cont = false;
        } else if (line.compare(0, 6, "CTRIA3") == 0) {
            std::istringstream is(line.substr(6));
            int f, pidv, a, b, c;
            if (!(is >> f >> pidv >> a >> b >> c)) return -5;
            fids.push_back(f);
            fvid.push_back(a);
            fvid.push_back(b);
            fvid.push_back(c);
        }
    }
    if (cont) return -6;
    mesh_from_arrays(m, (int)ids.size(), ids.data(), xyz.data(), (int)fids.size(), fids.data(),
                     fvid.data());
    return 0;
}

// Code/RWGGenerator.cpp:39-113 -- edges keyed by sorted vertex ids (std::map order),
// interior edges only, positive face = first face in file order.
void build_rwg(const Mesh& m, std::vector<Rwg>& rwgs) {
    std::map<std::pair<int, int>, std::vector<int>> edges;
    for (int f = 0; f < (int)m.fv.size(); ++f) {
        for (int e = 0; e < 3; ++e) {
            int a = m.pid[m.fv[f][e]], b = m.pid[m.fv[f][(e + 1) % 3]];
            edges[{std::min(a, b), std::max(a, b)}].push_back(f);
        }
    }
    std::unordered_map<int, int> idx;
    for (int i = 0; i < (int)m.pid.size(); ++i) idx[m.pid[i]] = i;
    rwgs.clear();
    for (auto& kv : edges) {
        if (kv.second.size() != 2) continue;
        Rwg r;
        r.fpos = kv.second[0];
        r.fneg = kv.second[1];
        r.e0 = idx[kv.first.first];
        r.e1 = idx[kv.first.second];
        auto freev = [&](int f) {
            for (int k = 0; k < 3; ++k) {
                int id = m.pid[m.fv[f][k]];
                if (id != kv.first.first && id != kv.first.second) return m.fv[f][k];
            }
            return -1;
        };
        r.vpos = freev(r.fpos);
        r.vneg = freev(r.fneg);
        const V3& p1 = m.p[r.e0];
        const V3& p2 = m.p[r.e1];
        r.len = std::sqrt(std::pow(p1.x - p2.x, 2) + std::pow(p1.y - p2.y, 2) + std::pow(p1.z - p2.z, 2));
        r.mid = {(p1.x + p2.x) / 2.0, (p1.y + p2.y) / 2.0, (p1.z + p2.z) / 2.0};
        rwgs.push_back(r);
    }
}

// Code/OCTree.h:110-141
static uint64_t spread3(uint32_t x) {
    uint64_t v = x;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}
static uint64_t morton(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
static uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffffULL;
    return (uint32_t)x;
}

void build_tree(const std::vector<Rwg>& rwgs, double wavelength, double leafCoeff, Tree& t) {
    // Code/OCTree.cpp:161-242 -- cubic bounding box of edge midpoints.
    V3 lo = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
             std::numeric_limits<double>::max()};
    V3 hi = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
             std::numeric_limits<double>::lowest()};
    for (const Rwg& r : rwgs) {
        lo.x = std::min(lo.x, r.mid.x);
        lo.y = std::min(lo.y, r.mid.y);
        lo.z = std::min(lo.z, r.mid.z);
        hi.x = std::max(hi.x, r.mid.x);
        hi.y = std::max(hi.y, r.mid.y);
        hi.z = std::max(hi.z, r.mid.z);
    }
    double rootLen = std::max(hi.x - lo.x, std::max(hi.y - lo.y, hi.z - lo.z)) * 1.001;
    double leafTarget = leafCoeff * wavelength;
    double ratio = rootLen / leafTarget;
    if (ratio < 1.0) ratio = 1.0;
    int maxLevel = (int)std::ceil(std::log2(ratio));
    maxLevel = std::max(2, std::min(maxLevel, 15));
    rootLen = leafTarget * (1 << maxLevel);
    V3 c = {0.5 * (lo.x + hi.x), 0.5 * (lo.y + hi.y), 0.5 * (lo.z + hi.z)};
    t.bmin = {c.x - 0.5 * rootLen, c.y - 0.5 * rootLen, c.z - 0.5 * rootLen};
    t.leafLen = rootLen / (1 << maxLevel);
    t.maxLevel = maxLevel;
    t.lv.assign(maxLevel + 1, {});

    // Code/OCTree.cpp:244-307 -- leaves keyed by Morton code; RWGs in id order.
    std::map<uint64_t, Node> leaves;
    int maxIdx = (1 << maxLevel) - 1;
    for (int id = 0; id < (int)rwgs.size(); ++id) {
        const V3& p = rwgs[id].mid;
        int ix = (int)((p.x - t.bmin.x) / t.leafLen);
        int iy = (int)((p.y - t.bmin.y) / t.leafLen);
        int iz = (int)((p.z - t.bmin.z) / t.leafLen);
        ix = std::max(0, std::min(ix, maxIdx));
        iy = std::max(0, std::min(iy, maxIdx));
        iz = std::max(0, std::min(iz, maxIdx));
        uint64_t code = morton(ix, iy, iz);
        auto it = leaves.find(code);
        if (it == leaves.end()) {
            Node n;
            n.code = code;
            n.level = maxLevel;
            n.len = t.leafLen;
            n.c = {t.bmin.x + (ix + 0.5) * t.leafLen, t.bmin.y + (iy + 0.5) * t.leafLen,
                   t.bmin.z + (iz + 0.5) * t.leafLen};
            it = leaves.emplace(code, n).first;
        }
        it->second.rwg.push_back(id);
    }
    for (auto& kv : leaves) t.lv[maxLevel].push_back(kv.second);

    // Code/OCTree.cpp:309-375 -- parents from code>>3, centre recovered from the
    // first child met (children visited in Morton order here).
    for (int level = maxLevel; level > 0; --level) {
        std::map<uint64_t, int> pidx;
        std::vector<Node> parents;
        for (int ci = 0; ci < (int)t.lv[level].size(); ++ci) {
            Node& ch = t.lv[level][ci];
            uint64_t pc = ch.code >> 3;
            int cidx = (int)(ch.code & 7);
            auto it = pidx.find(pc);
            if (it == pidx.end()) {
                Node p;
                p.code = pc;
                p.level = level - 1;
                p.len = ch.len * 2.0;
                double off = ch.len * 0.5;
                p.c = ch.c;
                p.c.x += (cidx & 1) ? -off : off;
                p.c.y += (cidx & 2) ? -off : off;
                p.c.z += (cidx & 4) ? -off : off;
                it = pidx.emplace(pc, (int)parents.size()).first;
                parents.push_back(p);
            }
            Node& p = parents[it->second];
            p.sons.push_back(ci);
            p.rwg.insert(p.rwg.end(), ch.rwg.begin(), ch.rwg.end());
            ch.parent = it->second;
        }
        // keep parents Morton-sorted; parent index was assigned in first-meet order,
        // which is Morton order because children are Morton-sorted
        t.lv[level - 1] = parents;
    }

    // Code/OCTree.cpp:385-426 -- 26-neighbour lists, (dx,dy,dz) loop order.
    for (int level = 0; level <= maxLevel; ++level) {
        std::unordered_map<uint64_t, int> map;
        for (int i = 0; i < (int)t.lv[level].size(); ++i) map[t.lv[level][i].code] = i;
        int g = 1 << level;
        for (Node& n : t.lv[level]) {
            int ix = compact3(n.code), iy = compact3(n.code >> 1), iz = compact3(n.code >> 2);
            n.nearn.clear();
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        if (!dx && !dy && !dz) continue;
                        int nx = ix + dx, ny = iy + dy, nz = iz + dz;
                        if (nx < 0 || ny < 0 || nz < 0 || nx >= g || ny >= g || nz >= g) continue;
                        auto it = map.find(morton(nx, ny, nz));
                        if (it != map.end()) n.nearn.push_back(it->second);
                    }
        }
    }

    // Code/OCTree.cpp:428-498 -- distant relatives: children of the parent's
    // neighbourhood that fail isNear (Chebyshev distance <= 1.74 * box).
    for (int level = 2; level <= maxLevel; ++level) {
        std::unordered_map<uint64_t, int> map;
        for (int i = 0; i < (int)t.lv[level].size(); ++i) map[t.lv[level][i].code] = i;
        int g = 1 << (level - 1);
        for (int ni = 0; ni < (int)t.lv[level].size(); ++ni) {
            Node& n = t.lv[level][ni];
            const Node& par = t.lv[level - 1][n.parent];
            int px = compact3(par.code), py = compact3(par.code >> 1), pz = compact3(par.code >> 2);
            n.dr.clear();
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        int nx = px + dx, ny = py + dy, nz = pz + dz;
                        if (nx < 0 || ny < 0 || nz < 0 || nx >= g || ny >= g || nz >= g) continue;
                        for (int cx = 0; cx < 2; ++cx)
                            for (int cy = 0; cy < 2; ++cy)
                                for (int cz = 0; cz < 2; ++cz) {
                                    auto it = map.find(morton(nx * 2 + cx, ny * 2 + cy, nz * 2 + cz));
                                    if (it == map.end() || it->second == ni) continue;
                                    const Node& o = t.lv[level][it->second];
                                    double d = std::max({std::fabs(n.c.x - o.c.x), std::fabs(n.c.y - o.c.y),
                                                         std::fabs(n.c.z - o.c.z)});
                                    if (!(d <= 1.74 * n.len)) n.dr.push_back(it->second);
                                }
                    }
        }
    }

    // Code/OCTree.cpp:500-551 -- unique integer translation vectors per level.
    t.drvec.assign(maxLevel - 1, {});
    for (int level = 2; level <= maxLevel; ++level) {
        std::map<std::array<int, 3>, int> vm;
        auto& vecs = t.drvec[level - 2];
        for (Node& n : t.lv[level]) {
            n.drv.clear();
            for (int o : n.dr) {
                const Node& d = t.lv[level][o];
                std::array<int, 3> key = {(int)std::round((n.c.x - d.c.x) / n.len),
                                          (int)std::round((n.c.y - d.c.y) / n.len),
                                          (int)std::round((n.c.z - d.c.z) / n.len)};
                auto it = vm.find(key);
                int vi;
                if (it == vm.end()) {
                    vi = (int)vecs.size();
                    vm[key] = vi;
                    vecs.push_back({key[0] * n.len, key[1] * n.len, key[2] * n.len});
                } else {
                    vi = it->second;
                }
                n.drv.push_back(vi);
            }
        }
    }
}

}  // namespace orc
