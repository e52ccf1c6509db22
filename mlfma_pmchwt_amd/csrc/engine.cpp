// Engine: setup orchestration, the matvec pipeline and device Krylov solvers.
#include "engine.hpp"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <type_traits>

#include "kernels.hpp"
#include "precond.hpp"
#include "setup.hpp"

namespace mlfma {

namespace {
double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline double2 d2(cplx c) { return make_double2(c.real(), c.imag()); }
inline cplx cx(double2 d) { return cplx(d.x, d.y); }
constexpr int kRedBlocks = 256;   // fixed partition of every reduction
}  // namespace

// Code/EMSource.cpp:34-72: incidence -r_hat, v = -theta_hat, h = phi_hat,
// E = v (pol 0) or h (pol 90), H = normalise(k x E).
void plane_wave(double pol_deg, double th_deg, double ph_deg, double* k, double* e, double* h, double* vth,
                double* vph) {
    auto rad = [](double a) {
        double b = a * kPi / 180.0;
        return std::fabs(b) < 1e-6 ? 0.0 : b;
    };
    auto snap = [](double a) { return std::fabs(a) < 1e-6 ? 0.0 : a; };
    const double pol = rad(pol_deg), th = rad(th_deg), ph = rad(ph_deg);
    k[0] = snap(-std::sin(th) * std::cos(ph));
    k[1] = snap(-std::sin(th) * std::sin(ph));
    k[2] = snap(-std::cos(th));
    vth[0] = snap(-std::cos(th) * std::cos(ph));
    vth[1] = snap(-std::cos(th) * std::sin(ph));
    vth[2] = snap(std::sin(th));
    vph[0] = snap(-std::sin(ph));
    vph[1] = snap(std::cos(ph));
    vph[2] = 0.0;
    for (int i = 0; i < 3; ++i) e[i] = h[i] = 0.0;
    const double* src = nullptr;
    if (std::fabs(pol - rad(90.0)) < 1e-6) src = vph;
    if (std::fabs(pol) < 1e-6) src = vth;
    if (src) {
        for (int i = 0; i < 3; ++i) e[i] = src[i];
        double c[3] = {k[1] * e[2] - k[2] * e[1], k[2] * e[0] - k[0] * e[2], k[0] * e[1] - k[1] * e[0]};
        double nn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int i = 0; i < 3; ++i) h[i] = nn > 1e-15 ? c[i] / nn : c[i];
    }
}

Engine::Engine(const mlfma_params& p) : P(p) {
    ie = p.integral_equation;
    apply_prec = p.far_precision;
    nw = (ie == MLFMA_PMCHWT) ? 2 : 1;
    ncur = nw;
    MLFMA_HIP(hipSetDevice(p.device));
    MLFMA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
}

Engine::~Engine() {
    if (st) hipStreamSynchronize(st);
    if (h_get_) hipHostFree(h_get_);
    if (h_put_) hipHostFree(h_put_);
    for (auto& kv : graphs_) hipGraphExecDestroy(kv.second);
    for (hipEvent_t e : ev_)
        if (e) hipEventDestroy(e);
    for (auto& a : allocs_) hipFree(a.first);
    if (st) hipStreamDestroy(st);
}

template <typename T>
T* Engine::dalloc(size_t count) {
    if (capturing_) throw Error(MLFMA_ERR_STATE, "device allocation inside a recorded matvec");
    if (count == 0) count = 1;
    void* p = nullptr;
    MLFMA_HIP(hipMalloc(&p, count * sizeof(T)));
    allocs_[p] = count * sizeof(T);
    dev_bytes += count * sizeof(T);
    return static_cast<T*>(p);
}

template <typename T>
T* Engine::upload(const std::vector<T>& v) {
    T* d = dalloc<T>(v.size());
    if (!v.empty()) MLFMA_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

void Engine::build(int np, const int* ids, const double* xyz, int nf, const int* fids, const int* fvid) {
    const double t0 = wall();
    rcs_scratch_init(P.device);   // this device's copy of the allocation-free RCS scratch
    std::string err;
    if (!build_mesh(np, ids, xyz, nf, fids, fvid, mesh, err)) throw Error(MLFMA_ERR_MESH, err);
    build_basis(mesh, basis);
    N = basis.n;
    if (N == 0) throw Error(MLFMA_ERR_MESH, "mesh has no interior edges");
    nunk = (ie == MLFMA_PMCHWT) ? 2 * N : N;
    med = make_medium(P.frequency_hz, cplx(P.eps_r_re, P.eps_r_im), cplx(P.mu_r_re, P.mu_r_im), ie == MLFMA_PMCHWT);
    build_octree(basis, med.lambda, P.leaf_coeff, tree);
    const int L = tree.max_level;
    nleaf = tree.lv[L].n;
    inv_perm.assign(N, 0);
    leaf_of.assign(N, 0);
    for (int i = 0; i < N; ++i) inv_perm[tree.perm[i]] = i;
    max_leaf = 0;
    for (int t = 0; t < nleaf; ++t) {
        max_leaf = std::max(max_leaf, tree.leaf_ptr[t + 1] - tree.leaf_ptr[t]);
        for (int i = tree.leaf_ptr[t]; i < tree.leaf_ptr[t + 1]; ++i) leaf_of[i] = t;
    }
    build_wave_tables(tree, med.k1, std::abs(med.k1), P.tf_mode, W[0]);
    if (nw == 2) build_wave_tables(tree, med.k2, std::abs(med.k2), P.tf_mode, W[1]);
    const double t1 = wall();

    std::vector<double> l1, l2, l3, w;
    if (!dunavant(P.gauss_points, l1, l2, l3, w)) throw Error(MLFMA_ERR_ARG, "unsupported Gauss point count");
    gauss_.n = P.gauss_points;
    for (int q = 0; q < gauss_.n; ++q) gauss_.l1[q] = l1[q], gauss_.l2[q] = l2[q], gauss_.l3[q] = l3[q], gauss_.w[q] = w[q];

    std::vector<RwgGeo> geo(N);
    for (int i = 0; i < N; ++i) {
        const int r = tree.perm[i];
        RwgGeo& g = geo[i];
        const int faces[2] = {basis.fneg[r], basis.fpos[r]};
        const int frees[2] = {basis.vneg[r], basis.vpos[r]};
        for (int s = 0; s < 2; ++s) {
            const int f = faces[s];
            for (int v = 0; v < 3; ++v) {
                const Vec3& p = mesh.pts[mesh.tri[3 * f + v]];
                g.tri[s][3 * v] = p.x, g.tri[s][3 * v + 1] = p.y, g.tri[s][3 * v + 2] = p.z;
            }
            const Vec3& fv = mesh.pts[frees[s]];
            g.vf[s][0] = fv.x, g.vf[s][1] = fv.y, g.vf[s][2] = fv.z;
            g.nrm[s][0] = mesh.normal[f].x, g.nrm[s][1] = mesh.normal[f].y, g.nrm[s][2] = mesh.normal[f].z;
            g.area[s] = mesh.area[f];
            g.face[s] = mesh.face_id[f];
        }
        g.len = basis.len[r];
    }
    d_geo_ = upload(geo);
    d_perm_ = upload(tree.perm);
    d_leaf_ptr_ = upload(tree.leaf_ptr);
    d_leaf_of_ = upload(leaf_of);
    {
        std::vector<int> order(nleaf);
        for (int t = 0; t < nleaf; ++t) order[t] = t;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return tree.leaf_ptr[a + 1] - tree.leaf_ptr[a] > tree.leaf_ptr[b + 1] - tree.leaf_ptr[b];
        });
        d_leaf_order_ = upload(order);
    }
    std::vector<double> lc(3 * (size_t)nleaf);
    for (int t = 0; t < nleaf; ++t) {
        lc[3 * t] = tree.lv[L].centre[t].x;
        lc[3 * t + 1] = tree.lv[L].centre[t].y;
        lc[3 * t + 2] = tree.lv[L].centre[t].z;
    }
    d_leaf_c_ = upload(lc);
    build_device_tree();
    for (int wv = 0; wv < nw; ++wv) build_wave(wv);
    const double t2 = wall();
    build_near();
    MLFMA_HIP(hipStreamSynchronize(st));
    MLFMA_HIP(hipGetLastError());
    const double t3 = wall();
    if (P.verbose)
        std::fprintf(stderr,
                     "[mlfma] N=%d unknowns=%d levels=%d leaves=%d max_leaf=%d | host tables %.3fs, patterns "
                     "%.3fs, near field %.3fs (%lld entries), device %.2f GB\n",
                     N, nunk, L, nleaf, max_leaf, t1 - t0, t2 - t1, t3 - t2, near_entries, dev_bytes / 1e9);
}

void Engine::build_device_tree() {
    const int L = tree.max_level;
    lv_.assign(L + 1, LevelDev());
    for (int l = 2; l <= L; ++l) {
        const Level& v = tree.lv[l];
        LevelDev& d = lv_[l];
        d.n = v.n;
        d.dr_ptr = upload(v.dr_ptr);
        d.dr_idx = upload(v.dr_idx);
        d.dr_vec = upload(v.dr_vec);
        if (l < L) {
            d.son_ptr = upload(v.son_ptr);
            d.son_idx = upload(v.son_idx);
        }
        d.parent = upload(v.parent);
        std::vector<int> oct(v.n);
        for (int i = 0; i < v.n; ++i) oct[i] = (int)(v.code[i] & 7);
        d.oct = upload(oct);
        // sibling clusters: boxes are Morton-sorted, so the children of one
        // parent are consecutive; slots clusters of up to `slots` children
        auto clusters = [&](int slots, std::vector<int>& cl_child, std::vector<int>& cl_ptr, std::vector<int>& cl_src,
                            std::vector<int>& cl_vec, long long* pairs) {
            cl_ptr.assign(1, 0);
            for (int b0 = 0; b0 < v.n;) {
                int b1 = b0;
                while (b1 < v.n && v.parent[b1] == v.parent[b0] && b1 - b0 < slots) ++b1;
                std::vector<int> srcs;
                for (int b = b0; b < b1; ++b)
                    for (int j = v.dr_ptr[b]; j < v.dr_ptr[b + 1]; ++j) srcs.push_back(v.dr_idx[j]);
                std::sort(srcs.begin(), srcs.end());
                srcs.erase(std::unique(srcs.begin(), srcs.end()), srcs.end());
                const size_t base = cl_src.size();
                cl_src.insert(cl_src.end(), srcs.begin(), srcs.end());
                // a member without a source reads the zero row after the last
                // translation vector; absent members keep -1
                cl_vec.resize(cl_src.size() * slots, -1);
                for (size_t u = base; u < cl_src.size(); ++u)
                    for (int b = b0; b < b1; ++b) cl_vec[u * slots + (b - b0)] = (int)v.drvec.size();
                for (int b = b0; b < b1; ++b)
                    for (int j = v.dr_ptr[b]; j < v.dr_ptr[b + 1]; ++j) {
                        const size_t u =
                            base + (std::lower_bound(srcs.begin(), srcs.end(), v.dr_idx[j]) - srcs.begin());
                        cl_vec[u * slots + (b - b0)] = v.dr_vec[j];
                    }
                for (int k = 0; k < slots; ++k) cl_child.push_back(b0 + k < b1 ? b0 + k : -1);
                cl_ptr.push_back((int)cl_src.size());
                if (pairs) *pairs += v.dr_ptr[b1] - v.dr_ptr[b0];
                b0 = b1;
            }
        };
        std::vector<int> cl_child, cl_ptr, cl_src, cl_vec;
        clusters(4, cl_child, cl_ptr, cl_src, cl_vec, &d.cl_pairs);
        {
            std::vector<int> c8, p8, s8, v8;
            clusters(8, c8, p8, s8, v8, nullptr);
            d.ncl8 = (int)p8.size() - 1;
            d.cl8_child = upload(c8);
            d.cl8_ptr = upload(p8);
            d.cl8_src = upload(s8);
            d.cl8_vec = upload(v8);
        }
        d.ncl = (int)cl_ptr.size() - 1;
        d.cl_entries = (long long)cl_src.size();
        d.cl_child = upload(cl_child);
        d.cl_ptr = upload(cl_ptr);
        d.cl_src = upload(cl_src);
        d.cl_vec = upload(cl_vec);
    }
}

void Engine::build_wave(int w) {
    const WaveTables& T = W[w];
    WaveDev& D = wd_[w];
    const int L = tree.max_level;
    D.K.assign(L + 1, 0);
    D.tap.assign(L + 1, nullptr);
    D.t_ptr.assign(L + 1, nullptr);
    D.t_kp.assign(L + 1, nullptr);
    D.th_tap.assign(L + 1, nullptr);
    D.ph_tap.assign(L + 1, nullptr);
    D.at_ptr.assign(L + 1, nullptr);
    D.at_idx.assign(L + 1, nullptr);
    D.ap_ptr.assign(L + 1, nullptr);
    D.ap_idx.assign(L + 1, nullptr);
    D.Lp.assign(L + 1, 0);
    D.Lc.assign(L + 1, 0);
    for (int l = 2; l <= L; ++l) D.K[l] = T.grid[l].K();
    for (int c = 3; c <= L; ++c) {
        const InterpTable& it = T.interp[c];
        D.tap[c] = upload(it.tap);
        D.th_tap[c] = upload(it.th_tap);
        D.ph_tap[c] = upload(it.ph_tap);
        D.at_ptr[c] = upload(it.at_ptr);
        D.at_idx[c] = upload(it.at_idx);
        D.ap_ptr[c] = upload(it.ap_ptr);
        D.ap_idx[c] = upload(it.ap_idx);
        D.Lp[c] = it.Lp;
        D.Lc[c] = it.Lc;
        D.t_ptr[c] = upload(it.t_ptr);
        D.t_kp[c] = upload(it.t_kp);
    }
    build_far<double>(w, T);
    const int no = ie == MLFMA_PMCHWT ? 2 : 1;
    if (rb_max<double>() > 1) D.E64 = dalloc<double2>((size_t)nleaf * rb_max<double>() * no * 3 * D.K[L]);
    if (P.far_precision == MLFMA_FP32) {
        build_far<float>(w, T);
        if (rb_max<float>() > 1) D.E = dalloc<float2>((size_t)nleaf * rb_max<float>() * no * 3 * D.K[L]);
        if (P.near_precision == MLFMA_BF16) {   // split-pair planes of the receive patterns
            const int Q = 3 * D.K[L], Qp = (Q + 31) & ~31;
            d_Vb_[w] = dalloc<unsigned short>((size_t)4 * N * Qp);
            launch_split_rows(st, N, Q, Qp, far<float>(w).Vf, d_Vb_[w]);
            // and transposed planes of the radiation patterns (all-RHS aggregation):
            // the RWGs of a leaf from a 32-aligned column on
            if (!d_jmap_) {
                std::vector<int> joff(nleaf + 1, 0), jmap;
                for (int t = 0; t < nleaf; ++t) {
                    const int nt = tree.leaf_ptr[t + 1] - tree.leaf_ptr[t];
                    joff[t + 1] = joff[t] + (nt + 31) / 32 * 32;
                }
                jmap.assign(joff[nleaf], -1);
                for (int t = 0; t < nleaf; ++t)
                    for (int i = tree.leaf_ptr[t]; i < tree.leaf_ptr[t + 1]; ++i) jmap[joff[t] + i - tree.leaf_ptr[t]] = i;
                Np_ = joff[nleaf];
                d_joff_ = upload(joff);
                d_jmap_ = upload(jmap);
            }
            Qa_[w] = (Q + 63) / 64 * 64;
            d_Va_[w] = dalloc<unsigned short>((size_t)4 * Qa_[w] * Np_);
            launch_split_cols(st, Q, Qa_[w], Np_, d_jmap_, far<float>(w).Vs, d_Va_[w]);
            MLFMA_HIP(hipGetLastError());
        }
    }
}

template <>
float2* Engine::beff_buf<float>(int w) {
    return wd_[w].E;
}
template <>
double2* Engine::beff_buf<double>(int w) {
    return wd_[w].E64;
}

template <typename R>
FarDev<R>& Engine::far(int w) {
    if constexpr (std::is_same<R, double>::value) return wd_[w].f64;
    else return wd_[w].f32;
}

namespace {
template <typename R>
cplx_t<R> to_cx(cplx c) {
    return cplx_t<R>{(R)c.real(), (R)c.imag()};
}
template <typename R>
std::vector<R> to_real(const std::vector<double>& v) {
    return std::vector<R>(v.begin(), v.end());
}
}  // namespace

// Patterns are computed in FP64 on the device and narrowed for FP32; tables
// are computed in FP64 on the host (translation operators, interpolation
// weights, phase shifts) and rounded once.
template <typename R>
void Engine::build_far(int w, const WaveTables& T) {
    WaveDev& D = wd_[w];
    FarDev<R>& F = far<R>(w);
    const int L = tree.max_level;
    F.Sm.assign(L + 1, nullptr);
    F.Bm.assign(L + 1, nullptr);
    F.TFw.assign(L + 1, nullptr);
    F.tw.assign(L + 1, nullptr);
    F.th_w.assign(L + 1, nullptr);
    F.ph_w.assign(L + 1, nullptr);
    F.at_w.assign(L + 1, nullptr);
    F.ap_w.assign(L + 1, nullptr);
    F.t_w.assign(L + 1, nullptr);
    F.sh_up.assign(L + 1, nullptr);
    F.sh_dn.assign(L + 1, nullptr);
    const int K0 = D.K[L];
    std::vector<double> kp(3 * (size_t)K0);
    for (int q = 0; q < K0; ++q) {
        kp[q] = T.grid[L].kp[q].x;
        kp[K0 + q] = T.grid[L].kp[q].y;
        kp[2 * K0 + q] = T.grid[L].kp[q].z;
    }
    F.kp0 = upload(to_real<R>(kp));
    F.Vs = dalloc<cplx_t<R>>((size_t)N * 3 * K0);
    F.Vf = dalloc<cplx_t<R>>((size_t)N * 3 * K0);
    if constexpr (std::is_same<R, double>::value) {
        launch_patterns(st, ie == MLFMA_CFIE, N, K0, d_geo_, d_leaf_of_, d_leaf_c_, F.kp0, d2(T.k), gauss_,
                        P.cfie_alpha, F.Vs, F.Vf);
    } else {
        const FarDev<double>& G = wd_[w].f64;
        launch_narrow(st, (size_t)N * 3 * K0, G.Vs, F.Vs);
        launch_narrow(st, (size_t)N * 3 * K0, G.Vf, F.Vf);
    }
    MLFMA_HIP(hipGetLastError());
    for (int l = 2; l <= L; ++l) {
        const size_t ws = (size_t)tree.lv[l].n * ncur * rb_max<R>() * 3 * D.K[l];
        F.Sm[l] = dalloc<cplx_t<R>>(ws);
        F.Bm[l] = dalloc<cplx_t<R>>(ws);
        // quadrature weight w_theta * dphi folded into the translation operator
        // (the reference applies it inside m2l, Code/ProcessMLFMM.h:138-142)
        // plus one all-zero row after the last vector (read by the cluster
        // kernel for members without the source)
        std::vector<cplx_t<R>> tfw(T.tf[l].size() + D.K[l], cplx_t<R>{0, 0});
        const int K = D.K[l];
        for (size_t e = 0; e < T.tf[l].size(); ++e) tfw[e] = to_cx<R>(T.tf[l][e] * T.qw[l][e % K]);
        F.TFw[l] = upload(tfw);
    }
    for (int c = 3; c <= L; ++c) {
        const InterpTable& it = T.interp[c];
        F.tw[c] = upload(to_real<R>(it.w));
        F.th_w[c] = upload(to_real<R>(it.th_w));
        F.ph_w[c] = upload(to_real<R>(it.ph_w));
        F.at_w[c] = upload(to_real<R>(it.at_w));
        F.ap_w[c] = upload(to_real<R>(it.ap_w));
        F.t_w[c] = upload(to_real<R>(it.t_w));
        std::vector<cplx_t<R>> su(T.shift_up[c].size()), sd(T.shift_down[c].size());
        for (size_t e = 0; e < su.size(); ++e) su[e] = to_cx<R>(T.shift_up[c][e]);
        for (size_t e = 0; e < sd.size(); ++e) sd[e] = to_cx<R>(T.shift_down[c][e]);
        F.sh_up[c] = upload(su);
        F.sh_dn[c] = upload(sd);
    }
}

// Near field: for every target leaf its near neighbours in list order, then
// itself (the column order of Code/Zmartix.h:595-629).  All blocks of one target
// are stored as one column-major nt x C_t matrix per plane (NearDev, geo.hpp).
void Engine::build_near() {
    const Level& lf = tree.lv[tree.max_level];
    const int planes = (ie == MLFMA_PMCHWT) ? 3 : 1;
    const std::vector<int>& lp = tree.leaf_ptr;
    pair_ptr.assign(1, 0);
    pair_tgt.clear();
    pair_src.clear();
    pair_col.clear();
    tgt_off.assign(nleaf + 1, 0);
    tgt_cols.assign(nleaf, 0);
    std::vector<int> col_ptr(nleaf + 1, 0), colref;
    for (int t = 0; t < nleaf; ++t) {
        const int nt = lp[t + 1] - lp[t];
        int cols = 0;
        auto add = [&](int s) {
            pair_tgt.push_back(t);
            pair_src.push_back(s);
            pair_col.push_back(cols);
            for (int j = lp[s]; j < lp[s + 1]; ++j) colref.push_back(tree.perm[j]);
            cols += lp[s + 1] - lp[s];
        };
        for (int j = lf.near_ptr[t]; j < lf.near_ptr[t + 1]; ++j) add(lf.near_idx[j]);
        add(t);
        pair_ptr.push_back((int)pair_src.size());
        tgt_cols[t] = cols;
        col_ptr[t + 1] = (int)colref.size();
        if ((long long)planes * cols * nt >= (1LL << 31))
            throw Error(MLFMA_ERR_ARG, "near-field block row of one leaf exceeds 2^31 entries (leaf too large)");
        tgt_off[t + 1] = tgt_off[t] + (long long)planes * cols * nt;
    }
    near_entries = tgt_off[nleaf];
    // launch order: largest block rows first, so the last wave of workgroups is short
    std::vector<int> order(nleaf);
    for (int t = 0; t < nleaf; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long long)tgt_cols[a] * (lp[a + 1] - lp[a]) > (long long)tgt_cols[b] * (lp[b + 1] - lp[b]);
    });
    // matrix-core items: row ranges of at most 64 rows, so that large leaves do
    // not set the kernel's duration
    std::vector<int2> items;
    for (int t = 0; t < nleaf; ++t)
        for (int rt = 0; rt * 16 < lp[t + 1] - lp[t]; rt += 4) items.push_back(make_int2(t, rt));
    std::stable_sort(items.begin(), items.end(), [&](int2 a, int2 b) {
        auto work = [&](int2 i) {
            return (long long)tgt_cols[i.x] * std::min(64, lp[i.x + 1] - lp[i.x] - 16 * i.y);
        };
        return work(a) > work(b);
    });
    nd_.nitem = (int)items.size();
    nd_.items = upload(items);
    nd_.planes = planes;
    nd_.nleaf = nleaf;
    nd_.npairs = (int)pair_src.size();
    nd_.N = N;
    nd_.leaf_ptr = d_leaf_ptr_;
    nd_.perm = d_perm_;
    nd_.order = upload(order);
    nd_.tgt_off = upload(tgt_off);
    nd_.tgt_cols = upload(tgt_cols);
    nd_.col_ptr = upload(col_ptr);
    nd_.colref = upload(colref);
    nd_.pair_tgt = upload(pair_tgt);
    nd_.pair_src = upload(pair_src);
    nd_.pair_col = upload(pair_col);
    d_Z_ = dalloc<double2>((size_t)near_entries);
    NearConsts c;
    c.k1 = d2(med.k1);
    c.k2 = d2(med.k2);
    c.eta1 = d2(med.eta1);
    c.etai = d2(med.etai);
    c.inv_etai = d2(1.0 / med.etai);
    c.alpha = P.cfie_alpha;
    launch_near_fill(st, ie, nd_, d_geo_, gauss_, c, d_Z_);
    MLFMA_HIP(hipGetLastError());
    if (P.near_precision != MLFMA_FP64) {   // FP32 storage (also the single-vector products of MLFMA_BF16)
        d_Z32_ = dalloc<float2>((size_t)near_entries);
        launch_narrow(st, (size_t)near_entries, d_Z_, d_Z32_);
        MLFMA_HIP(hipGetLastError());
    }
    if (P.near_precision == MLFMA_BF16) {   // split-pair records of the batched matrix-core products
        std::vector<long long> off;
        const long long recs = bf16_near_layout(tree.leaf_ptr, tgt_cols, planes, true, off);
        d_bsoff_ = upload(off);
        d_Zbs_ = dalloc<uint4>((size_t)recs);
        launch_pack_bf16(st, nd_, d_Z_, d_bsoff_, true, d_Zbs_);
        MLFMA_HIP(hipGetLastError());
    }
    if (P.precond_precision == MLFMA_BF16) {
        std::vector<long long> bf_off;
        const long long recs = bf16_near_layout(tree.leaf_ptr, tgt_cols, planes, false, bf_off);
        d_bfoff_ = upload(bf_off);
        d_Zbf_ = dalloc<uint4>((size_t)recs);
        launch_pack_bf16(st, nd_, d_Z_, d_bfoff_, false, d_Zbf_);
        MLFMA_HIP(hipGetLastError());
    }
    d_dinv_ = dalloc<double2>(nunk);
    launch_near_diag_inv(st, nd_, d_Z_, d_dinv_);
    if (d_Zbf_) {   // the FP32-vector polynomial's diagonal
        d_dinvf_ = dalloc<float2>(nunk);
        launch_narrow(st, nunk, d_dinv_, d_dinvf_);
    }
    MLFMA_HIP(hipGetLastError());
}

// Copy of one leaf-pair block, planes x (ns x nt column-major).
std::vector<double2> Engine::near_block(int p) {
    const int t = pair_tgt[p], s = pair_src[p];
    const int nt = tree.leaf_ptr[t + 1] - tree.leaf_ptr[t], ns = tree.leaf_ptr[s + 1] - tree.leaf_ptr[s];
    const int planes = (ie == MLFMA_PMCHWT) ? 3 : 1;
    const size_t blk = (size_t)nt * ns, plane = (size_t)tgt_cols[t] * nt;
    std::vector<double2> out(planes * blk);
    for (int pl = 0; pl < planes; ++pl)
        MLFMA_HIP(hipMemcpy(out.data() + pl * blk, d_Z_ + tgt_off[t] + pl * plane + (size_t)pair_col[p] * nt,
                            blk * sizeof(double2), hipMemcpyDeviceToHost));
    return out;
}

// Right-hand sides per pass of the batched matvec: nc = rb * ncur workspace
// currents, at most 8 in FP64 and 16 in FP32 (register budget of the
// per-box translation and vector disaggregation kernels), rb at most 8 (one
// matrix-core column tile of the disaggregation) and at most max_rhs.
template <typename R>
int Engine::rb_max() const {
    const int cap = std::min(8, (std::is_same<R, double>::value ? 8 : 16) / ncur);
    int rb = 1;
    while (rb * 2 <= cap && rb * 2 <= P.max_rhs) rb *= 2;
    return rb;
}

// mexp, m2m (fine -> coarse), then m2l with the gather-form l2l (coarse -> fine)
// for right-hand sides r0 .. r0 + rb - 1 of RHS-fastest vectors with stride ldr.
template <typename R>
void Engine::far_pass(int w, const double2* x, int ldr, int r0, int rb) {
    WaveDev& D = wd_[w];
    FarDev<R>& F = far<R>(w);
    const int L = tree.max_level;
    const int nc = ncur * rb;
    mark(6 * w + 0);
    cplx_t<R>* leaf_ws = F.Sm[L];
    if constexpr (std::is_same<R, float>::value) {
        if (eall_R_ > 0)   // aggregated for all right-hand sides already: this pass's leaf workspaces
            F.Sm[L] = d_Sall_[w] + (size_t)nleaf * ncur * r0 * 3 * D.K[L];
    }
    if (F.Sm[L] == leaf_ws)
        launch_aggregate_b<R>(st, ncur, rb, nleaf, D.K[L], N, ldr, r0, d_leaf_order_, d_leaf_ptr_, d_perm_, F.Vs, x,
                              F.Sm[L]);
    mark(6 * w + 1);
    mark(6 * w + 2);
    for (int l = L - 1; l >= 2; --l) interpolate<R>(w, l, nc);
    mark(6 * w + 3);
    mark(6 * w + 4);
    for (int l = 2; l <= L; ++l) translate<R>(w, l, nc);
    F.Sm[L] = leaf_ws;
    mark(6 * w + 5);
}

// m2m of parent level l from level l + 1: the separable form where it
// applies, else the 9-tap gather (m2m_separable_ off: always the 9-tap form).
template <typename R>
void Engine::interpolate(int w, int l, int nc) {
    WaveDev& D = wd_[w];
    FarDev<R>& F = far<R>(w);
    const int c = l + 1;
    if (m2m_separable_ && launch_m2m_sep<R>(st, nc, lv_[l].n, D.Lp[c], D.Lc[c], lv_[l].son_ptr, lv_[l].son_idx,
                                            lv_[c].oct, D.th_tap[c], F.th_w[c], D.ph_tap[c], F.ph_w[c], F.sh_up[c],
                                            F.Sm[c], F.Sm[l]))
        return;
    launch_m2m<R>(st, nc, lv_[l].n, D.K[l], D.K[c], lv_[l].son_ptr, lv_[l].son_idx, lv_[c].oct, D.tap[c], F.tw[c],
                  F.sh_up[c], F.Sm[c], F.Sm[l]);
}

// m2l + l2l of one level: sibling-cluster kernel (sources staged once per
// cluster) or the per-box gather kernel.
template <typename R>
void Engine::translate(int w, int l, int nc) {
    WaveDev& D = wd_[w];
    FarDev<R>& F = far<R>(w);
    const LevelDev& v = lv_[l];
    // the anterpolation from the parents either fused into the translation
    // (gather over the 9-tap stencil's transpose) or as a separable pass after it
    const int rw = (l > 2 && l2l_separable_) ? l2l_sep_rows(nc, D.Lp[l], D.Lc[l], sizeof(cplx_t<R>)) : 0;
    const bool up = l > 2 && rw == 0;
    if (m2l_cluster_ && m2l_variant_ == 2 && std::is_same<R, float>::value && (nc == 4 || nc == 8 || nc == 16))
        launch_m2l_cls<R>(st, nc, v.ncl8, D.K[l], v.cl8_child, v.cl8_ptr, v.cl8_src, v.cl8_vec, F.TFw[l], F.Sm[l],
                          up ? D.K[l - 1] : 0, v.parent, v.oct, D.t_ptr[l], D.t_kp[l], F.t_w[l], F.sh_dn[l],
                          up ? F.Bm[l - 1] : nullptr, F.Bm[l]);
    else if (m2l_cluster_)
        launch_m2l_cl<R>(st, m2l_variant_, nc, v.ncl, D.K[l], v.cl_child, v.cl_ptr, v.cl_src, v.cl_vec, F.TFw[l], F.Sm[l],
                         up ? D.K[l - 1] : 0, v.parent, v.oct, D.t_ptr[l], D.t_kp[l], F.t_w[l], F.sh_dn[l],
                         up ? F.Bm[l - 1] : nullptr, F.Bm[l]);
    else
        launch_m2l_l2l<R>(st, nc, v.n, D.K[l], v.dr_ptr, v.dr_idx, v.dr_vec, F.TFw[l], F.Sm[l], up ? D.K[l - 1] : 0,
                          v.parent, v.oct, D.t_ptr[l], D.t_kp[l], F.t_w[l], F.sh_dn[l], up ? F.Bm[l - 1] : nullptr,
                          F.Bm[l]);
    if (rw > 0)
        launch_l2l_sep<R>(st, nc, v.n, D.Lp[l], D.Lc[l], rw, v.parent, v.oct, D.at_ptr[l], D.at_idx[l], F.at_w[l],
                          D.ap_ptr[l], D.ap_idx[l], F.ap_w[l], F.sh_dn[l], F.Bm[l - 1], F.Bm[l]);
}

// y += far field (finest Bm of every wavenumber), same right-hand-side range
template <typename R>
void Engine::disaggregate(double2* y, int ldr, int r0, int rb) {
    const int L = tree.max_level;
    if (std::is_same<R, float>::value && eall_R_ > 0) {   // deferred: this pass's columns of Eall only
        const bool pm = ie == MLFMA_PMCHWT;
        const bool split = d_Vb_[0] != nullptr;   // the split-pair product reads Eall as BF16 pairs
        const FarDev<float>& F1 = far<float>(0);
        if (!pm) {
            const cplx c = (med.k1 * med.eta1 / (4.0 * kPi)) * (med.k1 / (4.0 * kPi));   // Code/MLFMM.cpp:226
            const double2 z = make_double2(0, 0);
            launch_beff_all(st, nleaf, rb, wd_[0].K[L], false, F1.Bm[L], F1.kp0, d2(c), z, z, z, eall_R_, r0, split,
                            d_Eall_[0]);
        } else {
            const FarDev<float>& F2 = far<float>(1);
            const cplx c1 = med.k1 * med.k1 / (16.0 * kPi * kPi), c2 = med.k2 * med.k2 / (16.0 * kPi * kPi);
            launch_beff_all(st, nleaf, rb, wd_[0].K[L], true, F1.Bm[L], F1.kp0, d2(c1), d2(c1), d2(c1), d2(c1),
                            eall_R_, r0, split, d_Eall_[0]);
            launch_beff_all(st, nleaf, rb, wd_[1].K[L], true, F2.Bm[L], F2.kp0, d2(c2 * med.etai), d2(c2), d2(c2),
                            d2(c2 / med.etai), eall_R_, r0, split, d_Eall_[1]);
        }
        return;
    }
    if (rb >= 2 && disagg_mfma_ && beff_buf<R>(0)) {
        const bool pm = ie == MLFMA_PMCHWT;
        const FarDev<R>& F1 = far<R>(0);
        if (!pm) {
            const cplx c = (med.k1 * med.eta1 / (4.0 * kPi)) * (med.k1 / (4.0 * kPi));   // Code/MLFMM.cpp:226
            const double2 z = make_double2(0, 0);
            launch_beff<R>(st, nleaf, rb, wd_[0].K[L], false, F1.Bm[L], F1.kp0, d2(c), z, z, z, beff_buf<R>(0));
            launch_disagg_mfma<R>(st, nleaf, N, ldr, r0, rb, d_leaf_order_, d_leaf_ptr_, d_perm_, false, wd_[0].K[L],
                                  F1.Vf, beff_buf<R>(0), 0, nullptr, nullptr, y);
        } else {
            const FarDev<R>& F2 = far<R>(1);
            const cplx c1 = med.k1 * med.k1 / (16.0 * kPi * kPi), c2 = med.k2 * med.k2 / (16.0 * kPi * kPi);
            launch_beff<R>(st, nleaf, rb, wd_[0].K[L], true, F1.Bm[L], F1.kp0, d2(c1), d2(c1), d2(c1), d2(c1),
                           beff_buf<R>(0));
            launch_beff<R>(st, nleaf, rb, wd_[1].K[L], true, F2.Bm[L], F2.kp0, d2(c2 * med.etai), d2(c2), d2(c2),
                           d2(c2 / med.etai), beff_buf<R>(1));
            launch_disagg_mfma<R>(st, nleaf, N, ldr, r0, rb, d_leaf_order_, d_leaf_ptr_, d_perm_, true, wd_[0].K[L],
                                  F1.Vf, beff_buf<R>(0), wd_[1].K[L], F2.Vf, beff_buf<R>(1), y);
        }
        return;
    }
    if (ie != MLFMA_PMCHWT) {
        // (k eta / 4 pi)(k / 4 pi), Code/MLFMM.cpp:226
        const cplx c = (med.k1 * med.eta1 / (4.0 * kPi)) * (med.k1 / (4.0 * kPi));
        const FarDev<R>& F = far<R>(0);
        launch_disagg_pec_b<R>(st, rb, N, wd_[0].K[L], ldr, r0, d_leaf_of_, d_perm_, F.Vf, F.Bm[L], d2(c), y);
    } else {
        const FarDev<R>& F1 = far<R>(0);
        const FarDev<R>& F2 = far<R>(1);
        PmchwtFarArgs<R> a;
        a.K1 = wd_[0].K[L];
        a.K2 = wd_[1].K[L];
        a.N = N;
        a.Vf1 = F1.Vf;
        a.Vf2 = F2.Vf;
        a.B1 = F1.Bm[L];
        a.B2 = F2.Bm[L];
        a.kp1 = F1.kp0;
        a.kp2 = F2.kp0;
        a.c1 = d2(med.k1 * med.k1 / (16.0 * kPi * kPi));   // Code/MLFMM.cpp:252-253
        a.c2 = d2(med.k2 * med.k2 / (16.0 * kPi * kPi));
        a.etai = d2(med.etai);
        a.inv_etai = d2(1.0 / med.etai);
        launch_disagg_pmchwt_b<R>(st, rb, ldr, r0, d_leaf_of_, d_perm_, a, y);
    }
}

void Engine::mark(int slot) {
    if (prof_on_) MLFMA_HIP(hipEventRecord(ev_[slot], st));
}

void Engine::set_profiling(bool on) {
    if (on && !ev_[0])
        for (hipEvent_t& e : ev_) MLFMA_HIP(hipEventCreate(&e));
    prof_on_ = on;
    for (int g = 0; g < 5; ++g) prof_ms[g] = 0, prof_launch[g] = 0;
    prof_apply = 0;
}

void Engine::collect_profile(int count) {
    MLFMA_HIP(hipEventSynchronize(ev_[15]));
    auto span = [&](int a) {
        float ms = 0;
        MLFMA_HIP(hipEventElapsedTime(&ms, ev_[a], ev_[a + 1]));
        return (double)ms;
    };
    const int L = tree.max_level;
    for (int w = 0; w < nw; ++w) {
        prof_ms[0] += span(6 * w);
        prof_ms[1] += span(6 * w + 2);
        prof_ms[2] += span(6 * w + 4);
        prof_launch[0] += eall_R_ > 0 ? 0 : 1;   // aggregated for all right-hand sides beforehand
        prof_launch[1] += L - 2;
        prof_launch[2] += L - 1;
    }
    prof_ms[3] += span(12);
    prof_ms[4] += span(14);
    prof_launch[3] += 1;
    prof_launch[4] += 1;
    prof_apply += count;
}

// One pass of the matvec for right-hand sides r0 .. r0 + rb - 1 of RHS-fastest
// vectors with stride ldr (single vector: ldr = 1, r0 = 0, rb = 1).
void Engine::apply_pass(const double2* x, double2* y, int ldr, int r0, int rb, int prec, bool with_near) {
    const bool f32 = prec == MLFMA_FP32 && P.far_precision == MLFMA_FP32;
    for (int w = 0; w < nw; ++w) {
        if (f32) far_pass<float>(w, x, ldr, r0, rb);
        else far_pass<double>(w, x, ldr, r0, rb);
    }
    mark(12);
    if (with_near) {
        if (prec == MLFMA_FP32 && d_Z32_) launch_near_b<float>(st, rb, ldr, r0, nd_, d_Z32_, x, y);
        else launch_near_b<double>(st, rb, ldr, r0, nd_, d_Z_, x, y);
    }
    mark(13);
    mark(14);
    if (f32) disaggregate<float>(y, ldr, r0, rb);
    else disaggregate<double>(y, ldr, r0, rb);
    mark(15);
}

void Engine::apply_one(const double2* x, double2* y, int prec) { apply_pass(x, y, 1, 0, 1, prec, true); }

// Y = A X for R right-hand sides, RHS-fastest (element (u, r) at u * R + r).
// With the FP32 near field the near product of all R runs first on the
// matrix cores; the far field follows in passes of rb right-hand sides.
// Whether apply_block can read an FP32 input directly: on that path only the
// matrix-core near product of split pairs reads X (the all-RHS split-pair
// aggregation reads the pairs it leaves).
bool Engine::fp32_input_ok(int R, int prec) const {
    if (!(prec == MLFMA_FP32 && P.far_precision == MLFMA_FP32 && R >= 2 && near_mfma_ && disagg_mfma_ && d_Zbs_))
        return false;
    for (int w = 0; w < nw; ++w)
        if (!d_Va_[w]) return false;
    return true;
}

// Y = A X from an FP32 X (the FP32 Krylov bases): read directly where every
// reader of X takes FP32, else widened into scratch (nunk x R FP64) first.
void Engine::apply_block_f(const float2* X, double2* Y, int R, int prec, double2* scratch) {
    if (fp32_input_ok(R, prec)) {
        apply_block(nullptr, Y, R, prec, X);
    } else {
        launch_widen(st, (size_t)nunk * R, X, scratch);
        apply_block(scratch, Y, R, prec);
    }
}

void Engine::apply_block(const double2* X, double2* Y, int R, int prec, const float2* Xf) {
    if (Xf && !fp32_input_ok(R, prec)) throw Error(MLFMA_ERR_STATE, "FP32 matvec input on a path that reads FP64");
    const bool f32 = prec == MLFMA_FP32 && P.far_precision == MLFMA_FP32;
    int cap = f32 ? rb_max<float>() : rb_max<double>();
    if (rb_tuned_[f32 ? 1 : 0] > 0) cap = std::min(cap, rb_tuned_[f32 ? 1 : 0]);
    const bool mfma_near = R >= 2 && near_mfma_;
    // the split-pair near product leaves X's BF16 pairs in d_xf_ (all R, RHS-fastest)
    const bool xsplit = mfma_near && prec == MLFMA_FP32 && d_Zbs_;
    if (mfma_near) {
        if (prof_on_) MLFMA_HIP(hipEventRecord(ev_[16], st));
        if (prec == MLFMA_FP32 && d_Zbs_) {
            ensure_xf(R);
            if (Xf) launch_near_bf16x3(st, R, R, 0, nd_, d_Zbs_, d_bsoff_, Xf, reinterpret_cast<uint2*>(d_xf_), Y);
            else launch_near_bf16x3(st, R, R, 0, nd_, d_Zbs_, d_bsoff_, X, reinterpret_cast<uint2*>(d_xf_), Y);
        } else if (prec == MLFMA_FP32 && d_Z32_) {
            ensure_xf(R);
            launch_near_mfma(st, R, R, 0, nd_, d_Z32_, X, d_xf_, Y);
        } else {
            launch_near_mfma64(st, R, R, 0, nd_, d_Z_, X, Y);
        }
        if (prof_on_) {
            MLFMA_HIP(hipEventRecord(ev_[17], st));
            MLFMA_HIP(hipEventSynchronize(ev_[17]));
            float ms = 0;
            MLFMA_HIP(hipEventElapsedTime(&ms, ev_[16], ev_[17]));
            prof_ms[3] += ms;
        }
    }
    // FP32 far field with the matrix-core disaggregation: one product
    // aggregates all R into the passes' leaf workspaces, the passes write
    // their effective matrices into Eall, and one product disaggregates all R,
    // so each pattern row is read once per matvec in each direction
    // (with profiling on, the two all-R products are timed into the
    // aggregation and disaggregation groups)
    const bool all = f32 && mfma_near && disagg_mfma_;
    auto timed = [&](int group, int launches, auto&& body) {
        if (prof_on_) MLFMA_HIP(hipEventRecord(ev_[16], st));
        body();
        if (prof_on_) {
            MLFMA_HIP(hipEventRecord(ev_[17], st));
            MLFMA_HIP(hipEventSynchronize(ev_[17]));
            float ms = 0;
            MLFMA_HIP(hipEventElapsedTime(&ms, ev_[16], ev_[17]));
            prof_ms[group] += ms;
            prof_launch[group] += launches;
        }
    };
    if (all) {
        ensure_eall(R);
        const int L = tree.max_level;
        timed(0, nw, [&] {
            for (int w = 0; w < nw; ++w) {
                if (d_Va_[w] && xsplit)   // split-pair BF16 (near_precision = MLFMA_BF16), X split above
                    launch_aggregate_bf3(st, ncur, nleaf, wd_[w].K[L], N, R, cap, Qa_[w], Np_, d_leaf_ptr_, d_joff_,
                                         d_perm_, d_Va_[w], reinterpret_cast<const uint2*>(d_xf_), d_Sall_[w]);
                else if (d_Va_[w])
                    launch_aggregate_bf3(st, ncur, nleaf, wd_[w].K[L], N, R, cap, Qa_[w], Np_, d_leaf_ptr_, d_joff_,
                                         d_perm_, d_Va_[w], X, d_Sall_[w]);
                else
                    launch_aggregate_all(st, ncur, nleaf, wd_[w].K[L], N, R, cap, d_leaf_ptr_, d_perm_,
                                         far<float>(w).Vs, X, d_Sall_[w]);
            }
        });
        eall_R_ = R;
    }
    try {
        for (int r0 = 0; r0 < R;) {
            int rb = 1;
            while (rb * 2 <= cap && r0 + rb * 2 <= R) rb *= 2;
            apply_pass(X, Y, R, r0, rb, prec, !mfma_near);
            if (prof_on_) collect_profile(rb);
            r0 += rb;
        }
    } catch (...) {
        eall_R_ = 0;
        throw;
    }
    eall_R_ = 0;
    if (all) {
        const int L = tree.max_level;
        const bool pm = ie == MLFMA_PMCHWT;
        timed(4, 1, [&] {
            if (d_Vb_[0])   // split-pair BF16 (near_precision = MLFMA_BF16)
                launch_disagg_bf3(st, R, nd_, pm, wd_[0].K[L], d_Vb_[0], reinterpret_cast<const uint2*>(d_Eall_[0]),
                                  pm ? wd_[1].K[L] : 0, pm ? d_Vb_[1] : nullptr,
                                  reinterpret_cast<const uint2*>(d_Eall_[1]), Y);
            else
                launch_disagg_all(st, R, nd_, pm, wd_[0].K[L], far<float>(0).Vf, d_Eall_[0], pm ? wd_[1].K[L] : 0,
                                  pm ? far<float>(1).Vf : nullptr, d_Eall_[1], Y);
        });
    }
}

// All-RHS effective matrices of the FP32 far field (sized before any
// recording: the recorded sequences hold their addresses)
void Engine::ensure_eall(int R) {
    if (P.far_precision != MLFMA_FP32 || R <= eall_cap_) return;
    clear_graphs();
    const int L = tree.max_level, no = ie == MLFMA_PMCHWT ? 2 : 1;
    for (int w = 0; w < nw; ++w) {
        dfree(d_Eall_[w]);
        dfree(d_Sall_[w]);
        d_Eall_[w] = dalloc<float2>((size_t)nleaf * no * 3 * wd_[w].K[L] * R);
        d_Sall_[w] = dalloc<float2>((size_t)nleaf * ncur * 3 * wd_[w].K[L] * R);
    }
    eall_cap_ = R;
}

// Column-major device vectors (vector j at x + j * nunk), as the C ABI passes them.
void Engine::apply_device(const double2* x, double2* y, int nrhs) {
    if (nrhs == 1) {
        apply_one(x, y, apply_prec);
        if (prof_on_) collect_profile(1);
    } else {
        ensure_block(nrhs);
        launch_transpose(st, nrhs, nunk, x, d_bx_);
        if (use_graphs_ && !prof_on_) {
            run_graph(GraphKey{0, nrhs, apply_prec}, [&] { apply_block(d_bx_, d_by_, nrhs, apply_prec); });
        } else {
            apply_block(d_bx_, d_by_, nrhs, apply_prec);
        }
        launch_transpose(st, nunk, nrhs, d_by_, y);
    }
    MLFMA_HIP(hipGetLastError());
}

void Engine::ensure_block(int R) {
    if (R > 256) throw Error(MLFMA_ERR_ARG, "at most 256 right-hand sides per call");
    ensure_xf(R);
    ensure_eall(R);
    if (R <= block_cap_) return;
    clear_graphs();
    dfree(d_bx_);
    dfree(d_by_);
    d_bx_ = dalloc<double2>((size_t)nunk * R);
    d_by_ = dalloc<double2>((size_t)nunk * R);
    block_cap_ = R;
}

// ---------------------------------------------------------------- hipGraph persistence
// A graph per (kind, R, precision) records the kernel sequence of a batched
// matvec (kind 0) or of a preconditioned batched matvec (kind 1) between fixed
// buffers, and is replayed on the handle's stream.  Nothing in the recorded
// sequence allocates, synchronises or clears memory with hipMemsetAsync.
void Engine::run_graph(const GraphKey& key, const std::function<void()>& body) {
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        hipGraph_t g = nullptr;
        MLFMA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
        try {
            body();
        } catch (...) {
            capturing_ = false;
            hipStreamEndCapture(st, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        capturing_ = false;
        MLFMA_HIP(hipStreamEndCapture(st, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        MLFMA_HIP(e);
        it = graphs_.emplace(key, ex).first;
    }
    MLFMA_HIP(hipGraphLaunch(it->second, st));
}

void Engine::clear_graphs() {
    if (capturing_) throw Error(MLFMA_ERR_STATE, "workspace resize inside a recorded matvec");
    if (!graphs_.empty()) hipStreamSynchronize(st);
    for (auto& kv : graphs_) hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

void Engine::set_graphs(bool on) {
    if (!on) clear_graphs();
    use_graphs_ = on;
}

// ---------------------------------------------------------------- autotuner
// Launch-configuration search for the batched matvec of R (at most 32)
// right-hand sides in precision prec, by coordinate descent from the widest
// configuration (the most right-hand sides per far-field pass, every
// available matrix-core and cluster path): each choice (bit 0 near product on
// the matrix cores, FP32 or FP64 entries; bit 1 disaggregation on the matrix
// cores; bit 2 translation by sibling clusters) is
// flipped in turn and kept when faster, then narrower passes are tried.  The
// choices act on separate kernels, so they barely interact, and the search
// times about six configurations instead of all of them.  Each configuration
// runs one untimed matvec and two timed ones; one whose untimed matvec already
// took twice the best time so far is not timed further (it counts with that
// time, so it cannot be chosen).  out[0] = chosen rb,
// out[1] = choice bits, out[2] = milliseconds per matvec of the choice, out[3]
// = of the slowest configuration timed.
void Engine::autotune(int R, int prec, double* out) {
    if (R < 1 || R > 256) throw Error(MLFMA_ERR_ARG, "autotune: 1..256 right-hand sides");
    R = std::min(R, 32);   // the per-pass choices do not depend on larger batches; keeps tuning short
    const bool f32 = prec == MLFMA_FP32 && P.far_precision == MLFMA_FP32;
    if (prec == MLFMA_FP32 && !f32 && !d_Z32_) throw Error(MLFMA_ERR_ARG, "no FP32 tables built");
    ensure_block(R);
    launch_probe(st, (int)std::min<size_t>((size_t)nunk * R, 0x7fffffff), 99u, d_bx_);
    const int rbmax = f32 ? rb_max<float>() : rb_max<double>();
    const bool nm = R >= 2, dm = R >= 2 && (f32 ? wd_[0].E != nullptr : wd_[0].E64 != nullptr);
    const int valid = (nm ? 1 : 0) | (dm ? 2 : 0) | 4 | (f32 ? 16 : 0);
    hipEvent_t a, b;
    MLFMA_HIP(hipEventCreate(&a));
    MLFMA_HIP(hipEventCreate(&b));
    double worst = 0, best = 0;
    auto time = [&](int rb, int m) {
        near_mfma_ = (m & 1) != 0;
        disagg_mfma_ = (m & 2) != 0;
        m2l_cluster_ = (m & 4) != 0;
        m2l_variant_ = (m & 16) ? 2 : ((m & 8) ? 1 : 0);
        rb_tuned_[f32 ? 1 : 0] = rb;
        float ms = 0;
        MLFMA_HIP(hipEventRecord(a, st));
        apply_block(d_bx_, d_by_, R, prec);   // warm-up
        MLFMA_HIP(hipEventRecord(b, st));
        MLFMA_HIP(hipEventSynchronize(b));
        MLFMA_HIP(hipEventElapsedTime(&ms, a, b));
        double per = ms;
        if (best == 0 || per < 2.0 * best) {
            MLFMA_HIP(hipEventRecord(a, st));
            for (int k = 0; k < 2; ++k) apply_block(d_bx_, d_by_, R, prec);
            MLFMA_HIP(hipEventRecord(b, st));
            MLFMA_HIP(hipEventSynchronize(b));
            MLFMA_HIP(hipEventElapsedTime(&ms, a, b));
            per = ms / 2.0;
        }
        if (P.verbose) std::fprintf(stderr, "[mlfma] autotune R=%d prec=%d rb=%d bits=%d: %.3f ms\n", R, prec, rb, m, per);
        worst = std::max(worst, per);
        return per;
    };
    int best_rb = rbmax, best_m = valid;
    best = time(best_rb, best_m);
    for (int bit = 1; bit <= 4; bit <<= 1) {
        if (!(valid & bit)) continue;
        const double t = time(best_rb, best_m ^ bit);
        if (t < best) best = t, best_m ^= bit;
    }
    if (f32 && (best_m & 4))   // the cluster translation's forms: two samples (start), twelve columns, six
        for (int v : {8, 0}) {
            const int m = (best_m & ~24) | v;
            const double t = time(best_rb, m);
            if (t < best) best = t, best_m = m;
        }
    for (int rb = rbmax / 2; rb >= std::max(1, rbmax / 4); rb /= 2) {
        const double t = time(rb, best_m);
        if (t < best) best = t, best_rb = rb;
        else break;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    rb_tuned_[f32 ? 1 : 0] = best_rb;
    near_mfma_ = (best_m & 1) != 0;
    disagg_mfma_ = (best_m & 2) != 0;
    m2l_cluster_ = (best_m & 4) != 0;
    m2l_variant_ = (best_m & 16) ? 2 : ((best_m & 8) ? 1 : 0);
    clear_graphs();   // recorded sequences follow the tuned configuration
    if (out) out[0] = best_rb, out[1] = best_m, out[2] = best, out[3] = worst;
}

void Engine::set_tuning(int prec, int rb, int bits) {
    const bool f32 = prec == MLFMA_FP32 && P.far_precision == MLFMA_FP32;
    if (prec == MLFMA_FP32 && !f32 && !d_Z32_) throw Error(MLFMA_ERR_ARG, "no FP32 tables built");
    rb_tuned_[f32 ? 1 : 0] = rb;
    near_mfma_ = (bits & 1) != 0;
    disagg_mfma_ = (bits & 2) != 0;
    m2l_cluster_ = (bits & 4) != 0;
    m2l_variant_ = (bits & 16) ? 2 : ((bits & 8) ? 1 : 0);
    clear_graphs();
}

void Engine::dfree(void* p) {
    if (!p) return;
    if (capturing_) throw Error(MLFMA_ERR_STATE, "device free inside a recorded matvec");
    auto it = allocs_.find(p);
    if (it != allocs_.end()) {
        dev_bytes -= it->second;
        allocs_.erase(it);
    }
    MLFMA_HIP(hipStreamSynchronize(st));
    MLFMA_HIP(hipFree(p));
}

// Column-major device blocks (vector j at b + j * nunk), as the C ABI passes them.
void Engine::solve_device(const mlfma_solve_params& sp, const double2* d_b, double2* d_x, int R,
                          mlfma_solve_stats* stats) {
    block_load(d_b, d_x, R);
    solve_block(sp, d_bx_, d_by_, R, stats);
    block_store(d_x, R);
}

void Engine::block_load(const double2* d_b, const double2* d_x, int R) {
    ensure_block(R);
    launch_transpose(st, R, nunk, d_b, d_bx_);
    launch_transpose(st, R, nunk, d_x, d_by_);
}

void Engine::block_store(double2* d_x, int R) {
    launch_transpose(st, nunk, R, d_by_, d_x);
    MLFMA_HIP(hipStreamSynchronize(st));
}

void Engine::set_apply_precision(int prec) {
    if (prec != MLFMA_FP64 && prec != MLFMA_FP32) throw Error(MLFMA_ERR_ARG, "unknown precision");
    if (prec == MLFMA_FP32 && P.far_precision != MLFMA_FP32 && P.near_precision == MLFMA_FP64)
        throw Error(MLFMA_ERR_ARG, "no FP32 tables built (create with far_precision or near_precision = MLFMA_FP32)");
    apply_prec = prec;
}

// Single stages of the FP64 operator (parity inspection).
void Engine::stage(int which, const double2* dx, double2* dy) {
    const int L = tree.max_level;
    switch (which) {
        case MLFMA_STAGE_AGGREGATE:
            for (int w = 0; w < nw; ++w)
                launch_aggregate<double>(st, ncur, nleaf, wd_[w].K[L], N, d_leaf_order_, d_leaf_ptr_, d_perm_,
                                         wd_[w].f64.Vs, dx, wd_[w].f64.Sm[L]);
            break;
        case MLFMA_STAGE_M2M:
            for (int w = 0; w < nw; ++w) {
                WaveDev& D = wd_[w];
                const FarDev<double>& F = D.f64;
                (void)D;
                (void)F;
                for (int l = L - 1; l >= 2; --l) interpolate<double>(w, l, ncur);
            }
            break;
        case MLFMA_STAGE_M2L_L2L:
            for (int w = 0; w < nw; ++w)
                for (int l = 2; l <= L; ++l) translate<double>(w, l, ncur);
            break;
        case MLFMA_STAGE_FAR:
            MLFMA_HIP(hipMemsetAsync(dy, 0, sizeof(double2) * nunk, st));
            disaggregate<double>(dy, 1, 0, 1);
            break;
        case MLFMA_STAGE_NEAR:
            launch_near<double>(st, nd_, d_Z_, dx, dy);
            break;
        default: throw Error(MLFMA_ERR_ARG, "unknown stage");
    }
    MLFMA_HIP(hipGetLastError());
    MLFMA_HIP(hipStreamSynchronize(st));
}

double Engine::norm(const double2* v, int n) {
    launch_norm2_partial(st, n, v, d_rpart_, kRedBlocks);
    std::vector<double> h(kRedBlocks);
    MLFMA_HIP(hipMemcpyAsync(h.data(), d_rpart_, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
    MLFMA_HIP(hipStreamSynchronize(st));
    double s = 0;
    for (double x : h) s += x;
    return std::sqrt(s);
}

// Restarted GMRES(m) on the device.  Same outer logic as Code/GMRES.h:65-217
// (Givens QR of the Hessenberg matrix, exact residual at every restart,
// divergence guard); the Arnoldi step uses classical Gram-Schmidt applied twice
// (two block reductions per pass instead of one per basis vector), which has
// the orthogonality of the reference's MGS + re-orthogonalisation.  With a
// preconditioner the Krylov space is built for A M^{-1} (right preconditioning)
// in flexible form: z_k = M^-1 v_k is kept and the update is x += Z y, so a
// preconditioner applied in reduced precision (not exactly linear) enters the
// solution only through the z_k themselves.  The reported residual is the true
// ||b - A x|| / ||b||.
void Engine::solve(const mlfma_solve_params& sp, const double2* d_b, double2* d_x, mlfma_solve_stats& stats) {
    const double t0 = wall();
    const int n = nunk;
    const int m = std::max(1, sp.restart);
    const double eps = 1e-16;
    if (V_cols_ < std::max(m + 1, 10)) {
        V_cols_ = std::max(m + 1, 10);
        d_V_ = dalloc<double2>((size_t)V_cols_ * n);
        d_w_ = dalloc<double2>(n);
        d_r_ = dalloc<double2>(n);
        d_t_ = dalloc<double2>(n);
        d_part_ = dalloc<double2>((size_t)V_cols_ * kRedBlocks);
        d_small_ = dalloc<double2>(V_cols_);
        d_rpart_ = dalloc<double>(kRedBlocks);
    }
    if (sp.precond < 0 || sp.precond > 3) throw Error(MLFMA_ERR_ARG, "precond must be 0, 1, 2 or 3");
    stats_fallbacks_ = 0;
    if (sp.precond == 2 && poly_alpha_.empty()) throw Error(MLFMA_ERR_STATE, "polynomial preconditioner coefficients not set");
    bool pc = sp.precond != 0;
    precond_kind_ = sp.precond;
    if (sp.precond == 1) precond_setup();
    if (pc && sp.solver == MLFMA_SOLVER_GMRES && Zf_cols_ < m) {
        dfree(d_Zf_);
        d_Zf_ = dalloc<double2>((size_t)m * n);
        Zf_cols_ = m;
    }
    int matvecs = 0, fp64_matvecs = 0, total = 0;
    auto A = [&](const double2* in, double2* out) {
        if (pc) {
            precond_apply(in, d_t_);
            apply_one(d_t_, out, apply_prec);
        } else {
            apply_one(in, out, apply_prec);
        }
        ++matvecs;
    };
    auto true_residual = [&]() {
        apply_one(d_x, d_r_, MLFMA_FP64);
        ++matvecs;
        ++fp64_matvecs;
        launch_residual(st, n, d_b, d_r_);
        return norm(d_r_, n);
    };
    auto dots = [&](int cnt, const double2* w, std::vector<cplx>& h) {
        launch_multidot(st, n, cnt, d_V_, (size_t)n, w, d_part_, kRedBlocks);
        launch_reduce_partials(st, cnt, kRedBlocks, d_part_, d_small_);
        std::vector<double2> tmp(cnt);
        MLFMA_HIP(hipMemcpyAsync(tmp.data(), d_small_, cnt * sizeof(double2), hipMemcpyDeviceToHost, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        h.resize(cnt);
        for (int j = 0; j < cnt; ++j) h[j] = cx(tmp[j]);
    };
    auto put_small = [&](const std::vector<cplx>& v, int cnt) {
        std::vector<double2> tmp(cnt);
        for (int j = 0; j < cnt; ++j) tmp[j] = d2(v[j]);
        MLFMA_HIP(hipMemcpyAsync(d_small_, tmp.data(), cnt * sizeof(double2), hipMemcpyHostToDevice, st));
        MLFMA_HIP(hipStreamSynchronize(st));
    };

    double bnorm = norm(d_b, n);
    if (bnorm < eps) bnorm = 1.0;
    double rnorm = true_residual();
    double err = rnorm / bnorm;
    bool converged = (err < sp.tol_rel || err < sp.tol_abs);
    if (!std::isfinite(err)) throw Error(MLFMA_ERR_NONFINITE, "non-finite initial residual");

    if (!converged && sp.solver == MLFMA_SOLVER_GMRES) {
        std::vector<cplx> H((size_t)(m + 1) * m), cs(m), sn(m), s(m + 1), y(m), h1, h2;
        auto Hm = [&](int i, int j) -> cplx& { return H[i + (size_t)j * (m + 1)]; };
        double err_cycle = err;
        for (int cyc = 1; cyc <= sp.itr_max && !converged; ++cyc) {
            launch_scale_copy(st, n, 1.0 / rnorm, d_r_, d_V_);
            std::fill(s.begin(), s.end(), cplx(0, 0));
            s[0] = rnorm;
            std::fill(H.begin(), H.end(), cplx(0, 0));
            int kused = 0;
            for (int k = 0; k < m; ++k) {
                if (total >= sp.itr_max) break;
                ++total;
                kused = k + 1;
                if (pc) {   // flexible: keep z_k = M^-1 v_k, so x += Z y needs no further application
                    double2* z = d_Zf_ + (size_t)k * n;
                    precond_apply(d_V_ + (size_t)k * n, z);
                    apply_one(z, d_w_, apply_prec);
                    ++matvecs;
                } else {
                    A(d_V_ + (size_t)k * n, d_w_);
                }
                dots(k + 1, d_w_, h1);
                put_small(h1, k + 1);
                launch_sub_combination(st, n, k + 1, d_V_, (size_t)n, d_small_, d_w_);
                dots(k + 1, d_w_, h2);
                put_small(h2, k + 1);
                launch_sub_combination(st, n, k + 1, d_V_, (size_t)n, d_small_, d_w_);
                for (int j = 0; j <= k; ++j) Hm(j, k) = h1[j] + h2[j];
                const double hn = norm(d_w_, n);
                if (!std::isfinite(hn)) throw Error(MLFMA_ERR_NONFINITE, "non-finite Arnoldi vector");
                Hm(k + 1, k) = hn;
                if (hn < eps * rnorm) break;
                launch_scale_copy(st, n, 1.0 / hn, d_w_, d_V_ + (size_t)(k + 1) * n);
                // Code/GMRES.h:22-48, 144-152
                for (int j = 0; j < k; ++j) {
                    cplx tmp = std::conj(cs[j]) * Hm(j, k) + std::conj(sn[j]) * Hm(j + 1, k);
                    Hm(j + 1, k) = -sn[j] * Hm(j, k) + cs[j] * Hm(j + 1, k);
                    Hm(j, k) = tmp;
                }
                cplx a = Hm(k, k), b = Hm(k + 1, k);
                if (std::abs(b) < eps) {
                    cs[k] = 1.0, sn[k] = 0.0;
                } else if (std::abs(a) < eps) {
                    cs[k] = 0.0, sn[k] = 1.0;
                } else {
                    double r = std::sqrt(std::real(a * std::conj(a) + b * std::conj(b)));
                    cs[k] = a / r, sn[k] = b / r;
                }
                cplx tmp = std::conj(cs[k]) * a + std::conj(sn[k]) * b;
                Hm(k, k) = tmp;
                Hm(k + 1, k) = 0.0;
                tmp = std::conj(cs[k]) * s[k] + std::conj(sn[k]) * s[k + 1];
                s[k + 1] = -sn[k] * s[k] + cs[k] * s[k + 1];
                s[k] = tmp;
                err = std::abs(s[k + 1]) / bnorm;
                if (err < sp.tol_rel || err < sp.tol_abs) break;
            }
            const int mm = std::min(kused, m);
            for (int i = mm - 1; i >= 0; --i) {
                y[i] = s[i];
                for (int j = i + 1; j < mm; ++j) y[i] -= Hm(i, j) * y[j];
                y[i] = (std::abs(Hm(i, i)) > eps) ? y[i] / Hm(i, i) : cplx(0, 0);
            }
            put_small(y, mm);
            if (pc) {
                launch_add_combination(st, n, mm, d_Zf_, (size_t)n, d_small_, d_x);
            } else {
                launch_add_combination(st, n, mm, d_V_, (size_t)n, d_small_, d_x);
            }
            const double err_prev = err_cycle;   // true residual at the start of this cycle
            const double err_est = std::abs(s[std::min(kused, m)]) / bnorm;
            rnorm = true_residual();
            err = rnorm / bnorm;
            if (!std::isfinite(err)) throw Error(MLFMA_ERR_NONFINITE, "non-finite residual");
            if (P.verbose)
                std::fprintf(stderr, "[mlfma] gmres cycle %d: %d iterations, residual estimate %.3e, true %.3e\n", cyc,
                             total, err_est, err);
            if (err < sp.tol_rel || err < sp.tol_abs) converged = true;
            if (total >= sp.itr_max) break;
            if (err > 1000.0 * sp.tol_rel && cyc > 8) break;   // Code/GMRES.h:208-214
            // fallback: a preconditioner that does not reduce the true residual by
            // 10% over a whole cycle is replaced by the next plainer one: the
            // polynomial by the near-field block-diagonal, that by none
            if (!converged && pc && err > 0.9 * err_prev) {
                precond_kind_ = precond_kind_ >= 2 ? 1 : 0;
                if (precond_kind_ == 1) precond_setup();
                pc = precond_kind_ != 0;
                ++stats_fallbacks_;
                if (P.verbose)
                    std::fprintf(stderr, "[mlfma] preconditioner stalled: continuing with %s\n",
                                 pc ? "the block-diagonal one" : "none");
            }
            err_cycle = err;
        }
    } else if (!converged) {
        // Code/CGS.h:104-234 (bilinear inner products); columns of V as vectors
        double2 *r = d_r_, *Rt = d_V_, *u = d_V_ + (size_t)n, *p = d_V_ + 2 * (size_t)n, *Vv = d_V_ + 3 * (size_t)n,
                *q = d_V_ + 4 * (size_t)n, *uq = d_V_ + 5 * (size_t)n, *Q = d_V_ + 6 * (size_t)n,
                *tmp = d_V_ + 7 * (size_t)n;
        auto bidot = [&](const double2* a, const double2* b) {
            launch_bidot_partial(st, n, a, b, d_part_, kRedBlocks);
            launch_reduce_partials(st, 1, kRedBlocks, d_part_, d_small_);
            double2 h;
            MLFMA_HIP(hipMemcpyAsync(&h, d_small_, sizeof(double2), hipMemcpyDeviceToHost, st));
            MLFMA_HIP(hipStreamSynchronize(st));
            return cx(h);
        };
        MLFMA_HIP(hipMemcpyAsync(Rt, r, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
        cplx rho_prev(0, 0);
        for (int it = 1; it <= sp.itr_max; ++it) {
            const cplx rho = bidot(r, Rt);
            if (std::abs(rho) < eps) break;
            if (it == 1) {
                MLFMA_HIP(hipMemcpyAsync(u, r, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
                MLFMA_HIP(hipMemcpyAsync(p, u, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
            } else {
                const cplx beta = rho / rho_prev;
                launch_axpy(st, n, d2(beta), q, r, u);      // u = r + beta q
                launch_axpy(st, n, d2(beta), p, q, tmp);    // tmp = q + beta p
                launch_axpy(st, n, d2(beta), tmp, u, p);    // p = u + beta tmp
            }
            A(p, Vv);
            const cplx den = bidot(Rt, Vv);
            if (std::abs(den) < eps) break;
            const cplx al = rho / den;
            launch_axpy(st, n, d2(-al), Vv, u, q);                       // q = u - al V
            launch_axpy(st, n, make_double2(1.0, 0.0), q, u, uq);        // uq = u + q
            if (pc) {
                precond_apply(uq, tmp);
                launch_axpy(st, n, d2(al), tmp, d_x, d_x);
                apply_one(tmp, Q, apply_prec);
                ++matvecs;
            } else {
                launch_axpy(st, n, d2(al), uq, d_x, d_x);
                A(uq, Q);
            }
            launch_axpy(st, n, d2(-al), Q, r, r);
            total = it;
            err = norm(r, n) / bnorm;
            if (!std::isfinite(err)) throw Error(MLFMA_ERR_NONFINITE, "non-finite residual");
            if (err < sp.tol_rel) break;
            rho_prev = rho;
        }
        rnorm = true_residual();
        err = rnorm / bnorm;
        converged = err < sp.tol_rel || err < sp.tol_abs;
    }
    stats.iterations = total;
    stats.matvecs = matvecs;
    stats.fp64_matvecs = fp64_matvecs;
    stats.precond_fallbacks = stats_fallbacks_;
    stats_fallbacks_ = 0;
    stats.converged = converged ? 1 : 0;
    stats.rel_residual = err;
    stats.seconds = wall() - t0;
}

// Restarted GMRES(m) for R right-hand sides in lockstep: one batched matvec
// per Arnoldi step serves every right-hand side, while the Gram-Schmidt
// coefficients, Givens rotations and convergence tests stay per right-hand
// side (the same recurrences as Engine::solve).  Block vectors are
// RHS-fastest.  A right-hand side that has converged inside a cycle keeps its
// Krylov column count fixed; its later columns are not used.  Krylov steps run
// with the apply precision, every restart starts from the true FP64 residual;
// the polynomial preconditioner is applied in flexible form as in solve().
// Once a quarter of the block has converged, the rest continues as a narrower
// block (a recursive call), so converged columns cost no further matvecs.
// ---------------------------------------------------------------- batched GMRES
// Lockstep flexible GMRES(m) over the R right-hand sides of one operator, as
// phases on a BlockState: start, per restart cycle begin / steps / end, and
// per step the (preconditioned) matvec (pre, body between fixed buffers, post)
// and four orthogonalisation phases.  The host reads the device between the
// orthogonalisation phases only after synchronising st, so several operators
// can advance their phases together with one synchronisation per phase
// (solve_multi), and their matvec bodies can be recorded into one graph.
// Reductions go through pinned host buffers (h_get_ device -> host, h_put_
// host -> device), rewritten only after such a synchronisation.
void Engine::ensure_pinned(size_t count) {
    if (count <= pin_cap_) return;
    MLFMA_HIP(hipStreamSynchronize(st));
    put_pending_ = false;
    if (h_get_) hipHostFree(h_get_);
    if (h_put_) hipHostFree(h_put_);
    h_get_ = h_put_ = nullptr;
    pin_cap_ = 0;
    MLFMA_HIP(hipHostMalloc((void**)&h_get_, count * sizeof(double2), hipHostMallocDefault));
    MLFMA_HIP(hipHostMalloc((void**)&h_put_, count * sizeof(double2), hipHostMallocDefault));
    pin_cap_ = count;
}

void Engine::bs_issue_norms(BlockState& s, const double2* v) {
    launch_norm2_partial_b(st, s.n, s.R, v, d_bpart_, kRedBlocks);
    launch_reduce_partials_b(st, 1, s.R, kRedBlocks, d_bpart_, d_bsmall_);
    MLFMA_HIP(hipMemcpyAsync(h_get_, d_bsmall_, s.R * sizeof(double2), hipMemcpyDeviceToHost, st));
}

void Engine::bs_read_norms(const BlockState& s, std::vector<double>& out) const {
    out.resize(s.R);
    for (int r = 0; r < s.R; ++r) out[r] = std::sqrt(h_get_[r].x);
}

// the host -> device staging buffer, once the copy last issued from it has run
double2* Engine::put_buffer() {
    if (put_pending_) MLFMA_HIP(hipStreamSynchronize(st));
    put_pending_ = false;
    return h_put_;
}

void Engine::bs_put(const std::vector<cplx>& v, double2* dst) {
    double2* hp = put_buffer();
    for (size_t e = 0; e < v.size(); ++e) hp[e] = d2(v[e]);
    MLFMA_HIP(hipMemcpyAsync(dst, hp, v.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    put_pending_ = true;
}

void Engine::bs_put_scale(const std::vector<double>& a) {
    double* hp = reinterpret_cast<double*>(put_buffer());
    for (size_t r = 0; r < a.size(); ++r) hp[r] = a[r];
    MLFMA_HIP(hipMemcpyAsync(d_bscale_, hp, a.size() * sizeof(double), hipMemcpyHostToDevice, st));
    put_pending_ = true;
}

void Engine::bs_true_residual(BlockState& s, std::vector<double>& rn) {
    apply_block(s.X, d_bRr_, s.R, MLFMA_FP64);
    s.matvecs += s.R;
    s.fp64_matvecs += s.R;
    launch_residual(st, (int)s.nR, s.B, d_bRr_);
    bs_issue_norms(s, d_bRr_);
    MLFMA_HIP(hipStreamSynchronize(st));
    bs_read_norms(s, rn);
}

void Engine::bs_start(BlockState& s, const mlfma_solve_params& sp, const double2* d_B, double2* d_X, int R) {
    s = BlockState();
    s.t0 = wall();
    s.sp = sp;
    s.B = d_B;
    s.X = d_X;
    s.R = R;
    s.n = nunk;
    s.m = std::max(1, sp.restart);
    const int m = s.m;
    if (sp.solver != MLFMA_SOLVER_GMRES) throw Error(MLFMA_ERR_ARG, "batched solve supports GMRES only");
    if (sp.precond < 0 || sp.precond > 3) throw Error(MLFMA_ERR_ARG, "batched solve: precond must be 0, 1, 2 or 3");
    if (sp.precond == 2 && poly_alpha_.empty()) throw Error(MLFMA_ERR_STATE, "polynomial preconditioner coefficients not set");
    s.pkind = sp.precond;   // 0 none, 1 near-field block-diagonal, 2 polynomial, 3 inner near-field solve
    if (s.pkind == 1) precond_setup();
    s.pc = s.pkind != 0;
    if (R < 1 || R > 256) throw Error(MLFMA_ERR_ARG, "batched solve: 1..256 right-hand sides");
    s.nR = (size_t)s.n * R;
    if (s.nR > 0x7fffffff) throw Error(MLFMA_ERR_ARG, "batched solve: n * R exceeds 2^31");
    // every workspace a batched matvec of R may use is sized here, before any
    // recorded matvec could reach it (a resize inside a capture is an error)
    ensure_xf(R);
    ensure_eall(R);
    ensure_pinned((size_t)(2 * m + 3) * R);
    const size_t nR = s.nR;
    if (bV_cols_ < m + 1 || bV_R_ != R) {
        for (void* p : {(void*)d_bV_, (void*)d_bW_, (void*)d_bRr_, (void*)d_bpart_, (void*)d_bsmall_, (void*)d_bscale_,
                        (void*)d_bZ_, (void*)d_bG_})
            dfree(p);
        bV_cols_ = m + 1;
        bV_R_ = R;
        d_bV_ = dalloc<double2>(nR * (m + 1));
        d_bW_ = dalloc<double2>(nR);
        d_bRr_ = dalloc<double2>(nR);
        d_bpart_ = dalloc<double2>((size_t)(m + 1) * kRedBlocks * R);
        d_bsmall_ = dalloc<double2>((size_t)(m + 1) * R);
        d_bscale_ = dalloc<double>(R);
        d_bZ_ = dalloc<double2>(nR);
        d_bG_ = dalloc<double2>(nR);
        clear_graphs();   // recorded sequences refer to the freed block buffers
    }
    // the flexible basis z_k = M^-1 v_k is sized whenever a preconditioner may run
    // (a stalling polynomial falls back to the block-diagonal one mid-solve)
    if (s.pc && (bZf_cols_ < m || bZf_nR_ != nR)) {
        dfree(d_bZf_);
        d_bZf_ = dalloc<double2>(nR * m);
        bZf_cols_ = m;
        bZf_nR_ = nR;
    }
    // FP32-mode solves keep the Krylov bases in FP32 (half the traffic of the
    // Gram-Schmidt passes); the reductions, W and the solution stay FP64
    s.vf = apply_prec == MLFMA_FP32;
    if (s.vf && (bVf_cols_ < m + 1 || bVf_nR_ != nR)) {
        dfree(d_bVf_);
        d_bVf_ = dalloc<float2>(nR * (m + 1));
        bVf_cols_ = m + 1;
        bVf_nR_ = nR;
    }
    if (s.vf && s.pc && (bZff_cols_ < m || bZff_nR_ != nR)) {
        dfree(d_bZff_);
        d_bZff_ = dalloc<float2>(nR * m);
        bZff_cols_ = m;
        bZff_nR_ = nR;
    }
    if (s.pkind == 2) ensure_poly(R);
    const double eps = 1e-16;
    s.iters.assign(R, 0);
    s.kused.assign(R, 0);
    s.err.assign(R, 0.0);
    bs_issue_norms(s, d_B);
    MLFMA_HIP(hipStreamSynchronize(st));
    bs_read_norms(s, s.bnorm);
    for (double& b : s.bnorm)
        if (b < eps) b = 1.0;
    bs_true_residual(s, s.rnorm);
    s.done.assign(R, 0);
    for (int r = 0; r < R; ++r) {
        s.err[r] = s.rnorm[r] / s.bnorm[r];
        if (!std::isfinite(s.err[r])) throw Error(MLFMA_ERR_NONFINITE, "non-finite initial residual");
        s.done[r] = s.err[r] < sp.tol_rel || s.err[r] < sp.tol_abs;
    }
    s.err_cycle = s.err;
    s.H.assign((size_t)R * (m + 1) * m, cplx(0, 0));
    s.cs.assign((size_t)R * m, cplx(0, 0));
    s.sn.assign((size_t)R * m, cplx(0, 0));
    s.sv.assign((size_t)R * (m + 1), cplx(0, 0));
}

bool Engine::bs_active(const BlockState& s) const {
    return !s.stopped && s.total < s.sp.itr_max &&
           std::any_of(s.done.begin(), s.done.end(), [](char c) { return c == 0; });
}

void Engine::bs_cycle_begin(BlockState& s) {
    ++s.cyc;
    const int R = s.R;
    std::vector<double> sc(R);
    for (int r = 0; r < R; ++r) sc[r] = (s.done[r] || s.rnorm[r] == 0.0) ? 0.0 : 1.0 / s.rnorm[r];
    bs_put_scale(sc);
    if (s.vf) launch_scale_copy_b(st, s.n, R, d_bscale_, d_bRr_, d_bVf_);
    else launch_scale_copy_b(st, s.n, R, d_bscale_, d_bRr_, d_bV_);
    std::fill(s.H.begin(), s.H.end(), cplx(0, 0));
    std::fill(s.sv.begin(), s.sv.end(), cplx(0, 0));
    s.live.assign(R, 0);
    for (int r = 0; r < R; ++r) {
        s.sv[(size_t)r * (s.m + 1)] = s.rnorm[r];
        s.live[r] = !s.done[r];
        s.kused[r] = 0;
    }
    s.k = 0;
}

bool Engine::bs_step_open(const BlockState& s) const {
    return s.k < s.m && s.total < s.sp.itr_max &&
           std::any_of(s.live.begin(), s.live.end(), [](char c) { return c != 0; });
}

int Engine::bs_graph_kind(const BlockState& s) const { return s.pc && s.pkind != 3 ? (s.pkind == 2 ? 1 : 3) : 2; }

// The matvec of step k.  recorded: the input goes to the fixed buffer G
// (through the inner near-field solve first with precond 3, which synchronises
// with the host and so is never recorded), and bs_body runs between fixed
// buffers.  Otherwise the whole step runs here.
void Engine::bs_pre(BlockState& s, bool recorded) {
    ++s.total;
    const size_t nR = s.nR;
    if (poly_f32(s)) {   // FP32 bases and the FP32-vector polynomial: no FP64 copy of v_k or z_k
        const float2* Vk = d_bVf_ + (size_t)s.k * nR;
        ensure_poly(s.R);
        if (recorded) {
            MLFMA_HIP(hipMemcpyAsync(d_bG_, Vk, nR * sizeof(float2), hipMemcpyDeviceToDevice, st));
        } else {
            float2* Zk = d_bZff_ + (size_t)s.k * nR;
            poly_apply_bf16f(Vk, Zk, s.R);
            apply_block_f(Zk, d_bW_, s.R, apply_prec, d_bG_);
        }
        return;
    }
    if (s.vf) {   // FP32 bases: v_k widened into G; z_k = M^-1 v_k rounded into the FP32 basis afterwards
        launch_widen(st, nR, d_bVf_ + (size_t)s.k * nR, d_bG_);
        if (s.pc && (s.pkind == 3 || !recorded)) {
            precond_block(s.pkind, d_bG_, d_bZ_, s.R, apply_prec);
            launch_narrow(st, nR, d_bZ_, d_bZff_ + (size_t)s.k * nR);
            if (recorded) {
                MLFMA_HIP(hipMemcpyAsync(d_bG_, d_bZ_, nR * sizeof(double2), hipMemcpyDeviceToDevice, st));
            } else {
                apply_block(d_bZ_, d_bW_, s.R, apply_prec);
            }
        } else if (s.pc && s.pkind == 2) {
            ensure_poly(s.R);
        } else if (!recorded) {
            apply_block(d_bG_, d_bW_, s.R, apply_prec);
        }
        return;
    }
    double2* Vk = d_bV_ + (size_t)s.k * nR;
    double2* Zk = s.pc ? d_bZf_ + (size_t)s.k * nR : nullptr;
    if (recorded) {
        if (s.pc && s.pkind == 3) {
            precond_block(3, Vk, Zk, s.R, apply_prec);
            MLFMA_HIP(hipMemcpyAsync(d_bG_, Zk, nR * sizeof(double2), hipMemcpyDeviceToDevice, st));
        } else {
            if (s.pc && s.pkind == 2) ensure_poly(s.R);
            MLFMA_HIP(hipMemcpyAsync(d_bG_, Vk, nR * sizeof(double2), hipMemcpyDeviceToDevice, st));
        }
    } else if (s.pc) {   // flexible: keep z_k = M^-1 v_k
        precond_block(s.pkind, Vk, Zk, s.R, apply_prec);
        apply_block(Zk, d_bW_, s.R, apply_prec);
    } else {
        apply_block(Vk, d_bW_, s.R, apply_prec);
    }
}

void Engine::bs_body(BlockState& s) {
    if (poly_f32(s)) {   // FP32 v_k and z_k in the first halves of G and Z (G is then free as scratch)
        float2* Gf = reinterpret_cast<float2*>(d_bG_);
        float2* Zf = reinterpret_cast<float2*>(d_bZ_);
        poly_apply_bf16f(Gf, Zf, s.R);
        apply_block_f(Zf, d_bW_, s.R, apply_prec, d_bG_);
        return;
    }
    if (s.pc && s.pkind != 3) {
        precond_block(s.pkind, d_bG_, d_bZ_, s.R, apply_prec);
        apply_block(d_bZ_, d_bW_, s.R, apply_prec);
    } else {
        apply_block(d_bG_, d_bW_, s.R, apply_prec);
    }
}

void Engine::bs_post(BlockState& s, bool recorded) {
    if (recorded && poly_f32(s)) {
        MLFMA_HIP(hipMemcpyAsync(d_bZff_ + (size_t)s.k * s.nR, d_bZ_, s.nR * sizeof(float2), hipMemcpyDeviceToDevice,
                                 st));
    } else if (recorded && s.pc && s.pkind != 3) {
        if (s.vf)
            launch_narrow(st, s.nR, d_bZ_, d_bZff_ + (size_t)s.k * s.nR);
        else
            MLFMA_HIP(hipMemcpyAsync(d_bZf_ + (size_t)s.k * s.nR, d_bZ_, s.nR * sizeof(double2),
                                     hipMemcpyDeviceToDevice, st));
    }
    s.matvecs += s.R;
}

// Classical Gram-Schmidt twice (CGS2) and the Givens update of step k.
// Phase 0 runs both passes on the device (each pass's coefficients <V_j, W>
// stay in d_bsmall_ for the subtraction) and copies both passes'
// coefficients and ||W|| to the host; after the caller has synchronised st,
// phase 1 updates every live right-hand side's Hessenberg column and scales
// W into V_{k+1}.
void Engine::bs_orth(BlockState& s, int phase) {
    const int R = s.R, m = s.m, k = s.k;
    const double eps = 1e-16;
    const size_t nh = (size_t)(k + 1) * R;
    if (phase == 0) {
        for (int pass = 0; pass < 2; ++pass) {
            if (s.vf) launch_multidot_b(st, s.n, R, k + 1, d_bVf_, s.nR, d_bW_, d_bpart_, kRedBlocks);
            else launch_multidot_b(st, s.n, R, k + 1, d_bV_, s.nR, d_bW_, d_bpart_, kRedBlocks);
            launch_reduce_partials_b(st, k + 1, R, kRedBlocks, d_bpart_, d_bsmall_);
            MLFMA_HIP(hipMemcpyAsync(h_get_ + pass * nh, d_bsmall_, nh * sizeof(double2), hipMemcpyDeviceToHost, st));
            if (s.vf) launch_sub_combination_b(st, s.n, R, k + 1, d_bVf_, s.nR, d_bsmall_, d_bW_);
            else launch_sub_combination_b(st, s.n, R, k + 1, d_bV_, s.nR, d_bsmall_, d_bW_);
        }
        launch_norm2_partial_b(st, s.n, R, d_bW_, d_bpart_, kRedBlocks);
        launch_reduce_partials_b(st, 1, R, kRedBlocks, d_bpart_, d_bsmall_);
        MLFMA_HIP(hipMemcpyAsync(h_get_ + 2 * nh, d_bsmall_, R * sizeof(double2), hipMemcpyDeviceToHost, st));
        return;
    }
    put_pending_ = false;   // the caller synchronised st
    auto Hm = [&](int r, int i, int j) -> cplx& { return s.H[((size_t)r * m + j) * (m + 1) + i]; };
    const double2* h1 = h_get_;
    const double2* h2 = h_get_ + nh;
    const double2* hn2 = h_get_ + 2 * nh;
    std::vector<double> inv(R, 0.0);
    for (int r = 0; r < R; ++r) {
        if (!s.live[r]) continue;
        s.kused[r] = k + 1;
        ++s.iters[r];
        for (int j = 0; j <= k; ++j) Hm(r, j, k) = cx(h1[(size_t)j * R + r]) + cx(h2[(size_t)j * R + r]);
        const double hn = std::sqrt(hn2[r].x);
        if (!std::isfinite(hn)) throw Error(MLFMA_ERR_NONFINITE, "non-finite Arnoldi vector");
        Hm(r, k + 1, k) = hn;
        const bool breakdown = hn < eps * s.rnorm[r];
        if (!breakdown) inv[r] = 1.0 / hn;
        cplx* c_ = &s.cs[(size_t)r * m];
        cplx* s_ = &s.sn[(size_t)r * m];
        cplx* g = &s.sv[(size_t)r * (m + 1)];
        // Code/GMRES.h:22-48, 144-152
        for (int j = 0; j < k; ++j) {
            cplx tmp = std::conj(c_[j]) * Hm(r, j, k) + std::conj(s_[j]) * Hm(r, j + 1, k);
            Hm(r, j + 1, k) = -s_[j] * Hm(r, j, k) + c_[j] * Hm(r, j + 1, k);
            Hm(r, j, k) = tmp;
        }
        cplx a = Hm(r, k, k), b = Hm(r, k + 1, k);
        if (std::abs(b) < eps) {
            c_[k] = 1.0, s_[k] = 0.0;
        } else if (std::abs(a) < eps) {
            c_[k] = 0.0, s_[k] = 1.0;
        } else {
            double rr = std::sqrt(std::real(a * std::conj(a) + b * std::conj(b)));
            c_[k] = a / rr, s_[k] = b / rr;
        }
        Hm(r, k, k) = std::conj(c_[k]) * a + std::conj(s_[k]) * b;
        Hm(r, k + 1, k) = 0.0;
        cplx tmp = std::conj(c_[k]) * g[k] + std::conj(s_[k]) * g[k + 1];
        g[k + 1] = -s_[k] * g[k] + c_[k] * g[k + 1];
        g[k] = tmp;
        s.err[r] = std::abs(g[k + 1]) / s.bnorm[r];
        if (breakdown || s.err[r] < s.sp.tol_rel || s.err[r] < s.sp.tol_abs) s.live[r] = 0;
    }
    bs_put_scale(inv);
    if (s.vf) launch_scale_copy_b(st, s.n, R, d_bscale_, d_bW_, d_bVf_ + (size_t)(k + 1) * s.nR);
    else launch_scale_copy_b(st, s.n, R, d_bscale_, d_bW_, d_bV_ + (size_t)(k + 1) * s.nR);
    ++s.k;
}

// End of a restart cycle: x_r += Z y_r (V y_r without preconditioner), the
// true FP64 residuals, the preconditioner fallback, the reference's
// divergence guard and (single operator only) the compaction of the block.
void Engine::bs_cycle_end(BlockState& s, bool compaction) {
    const int R = s.R, m = s.m, n = s.n;
    const double eps = 1e-16;
    const mlfma_solve_params& sp = s.sp;
    auto Hm = [&](int r, int i, int j) -> cplx& { return s.H[((size_t)r * m + j) * (m + 1) + i]; };
    // x_r += V y_r with y_r from the triangular Hessenberg of each right-hand side
    int kmax = 0;
    for (int r = 0; r < R; ++r) kmax = std::max(kmax, s.kused[r]);
    std::vector<cplx> yv((size_t)std::max(kmax, 1) * R, cplx(0, 0));
    for (int r = 0; r < R; ++r) {
        const int mm = s.kused[r];
        const cplx* g = &s.sv[(size_t)r * (m + 1)];
        std::vector<cplx> y(mm);
        for (int i = mm - 1; i >= 0; --i) {
            y[i] = g[i];
            for (int j = i + 1; j < mm; ++j) y[i] -= Hm(r, i, j) * y[j];
            y[i] = (std::abs(Hm(r, i, i)) > eps) ? y[i] / Hm(r, i, i) : cplx(0, 0);
        }
        for (int i = 0; i < mm; ++i) yv[(size_t)i * R + r] = y[i];
    }
    if (kmax > 0) {
        bs_put(yv, d_bsmall_);
        if (s.pc) {   // x += Z y with the kept z_k = M^-1 v_k
            if (s.vf) launch_add_combination_b(st, n, R, kmax, d_bZff_, s.nR, d_bsmall_, s.X);
            else launch_add_combination_b(st, n, R, kmax, d_bZf_, s.nR, d_bsmall_, s.X);
        } else {
            if (s.vf) launch_add_combination_b(st, n, R, kmax, d_bVf_, s.nR, d_bsmall_, s.X);
            else launch_add_combination_b(st, n, R, kmax, d_bV_, s.nR, d_bsmall_, s.X);
        }
    }
    const std::vector<double> err_prev = s.err_cycle;   // true residuals at the start of this cycle
    bs_true_residual(s, s.rnorm);
    bool progress = false;
    for (int r = 0; r < R; ++r) {
        s.err[r] = s.rnorm[r] / s.bnorm[r];
        if (!std::isfinite(s.err[r])) throw Error(MLFMA_ERR_NONFINITE, "non-finite residual");
        if (s.err[r] < sp.tol_rel || s.err[r] < sp.tol_abs) s.done[r] = 1;
        if (!s.done[r] && s.err[r] <= 0.9 * err_prev[r]) progress = true;
    }
    const bool all_done = std::all_of(s.done.begin(), s.done.end(), [](char c) { return c != 0; });
    if (P.verbose) {
        int nd = 0;
        double emax = 0;
        for (int r = 0; r < R; ++r) nd += s.done[r] ? 1 : 0, emax = std::max(emax, s.err[r]);
        std::fprintf(stderr, "[mlfma] batched gmres cycle: %d of %d converged, max true residual %.3e\n", nd, R, emax);
    }
    // fallback as in solve(): a stalling preconditioner is replaced by the
    // next plainer one (polynomial -> block-diagonal -> none)
    if (s.pc && !all_done && !progress) {
        s.pkind = s.pkind >= 2 ? 1 : 0;
        if (s.pkind == 1) precond_setup();
        s.pc = s.pkind != 0;
        ++s.fallbacks;
        if (P.verbose)
            std::fprintf(stderr, "[mlfma] preconditioner stalled: continuing with %s\n",
                         s.pc ? "the block-diagonal one" : "none");
    }
    s.err_cycle = s.err;
    bool stalled = true;   // Code/GMRES.h:208-214, all right-hand sides
    for (int r = 0; r < R; ++r)
        if (!s.done[r] && !(s.err[r] > 1000.0 * sp.tol_rel)) stalled = false;
    if (stalled && s.cyc > 8) {
        s.stopped = true;
        return;
    }
    if (!compaction) return;
    // compaction: once a quarter of the block has converged, the remaining
    // right-hand sides continue as a narrower block (same recurrences; each
    // restarts from its own current iterate), so no matvec work is spent on
    // converged columns
    std::vector<int> idx;
    for (int r = 0; r < R; ++r)
        if (!s.done[r]) idx.push_back(r);
    const int Rc = (int)idx.size();
    if (Rc > 0 && 4 * Rc <= 3 * R && s.total < sp.itr_max) {
        int* d_idx = upload(idx);
        double2* Bc = dalloc<double2>((size_t)n * Rc);
        double2* Xc = dalloc<double2>((size_t)n * Rc);
        launch_gather_cols(st, n, R, Rc, d_idx, s.B, Bc);
        launch_gather_cols(st, n, R, Rc, d_idx, s.X, Xc);
        mlfma_solve_params sp2 = sp;
        sp2.itr_max = sp.itr_max - s.total;
        sp2.precond = s.pkind;   // a preconditioner replaced above stays replaced
        std::vector<mlfma_solve_stats> st2(Rc);
        if (P.verbose) std::fprintf(stderr, "[mlfma] batched gmres: continuing with %d of %d right-hand sides\n", Rc, R);
        solve_block(sp2, Bc, Xc, Rc, st2.data());
        launch_scatter_cols(st, n, R, Rc, d_idx, Xc, s.X);
        MLFMA_HIP(hipStreamSynchronize(st));
        dfree(Bc);
        dfree(Xc);
        dfree(d_idx);
        s.fallbacks += st2[0].precond_fallbacks;
        for (int j = 0; j < Rc; ++j) {
            const int r = idx[j];
            s.iters[r] += st2[j].iterations;
            s.done[r] = st2[j].converged ? 1 : 0;
            s.err[r] = st2[j].rel_residual;
        }
        s.matvecs += st2[0].matvecs * Rc;
        s.fp64_matvecs += st2[0].fp64_matvecs * Rc;
        s.stopped = true;
    }
}

bool Engine::bs_narrow(BlockState& s, BlockState& child, NarrowRecord& rec) {
    if (s.stopped) return false;
    std::vector<int> idx;
    for (int r = 0; r < s.R; ++r)
        if (!s.done[r]) idx.push_back(r);
    const int Rc = (int)idx.size();
    if (!(Rc > 0 && 4 * Rc <= 3 * s.R && s.total < s.sp.itr_max)) return false;
    const int n = s.n;
    rec.idx = idx;
    rec.d_idx = upload(idx);
    rec.Bc = dalloc<double2>((size_t)n * Rc);
    rec.Xc = dalloc<double2>((size_t)n * Rc);
    launch_gather_cols(st, n, s.R, Rc, rec.d_idx, s.B, rec.Bc);
    launch_gather_cols(st, n, s.R, Rc, rec.d_idx, s.X, rec.Xc);
    mlfma_solve_params sp2 = s.sp;
    sp2.itr_max = s.sp.itr_max - s.total;
    sp2.precond = s.pkind;   // a preconditioner replaced before stays replaced
    if (P.verbose) std::fprintf(stderr, "[mlfma] batched gmres: continuing with %d of %d right-hand sides\n", Rc, s.R);
    bs_start(child, sp2, rec.Bc, rec.Xc, Rc);
    s.stopped = true;
    return true;
}

void Engine::bs_unnarrow(BlockState& s, const BlockState& child, NarrowRecord& rec) {
    const int Rc = child.R;
    std::vector<mlfma_solve_stats> st2(Rc);
    bs_stats(child, st2.data());
    launch_scatter_cols(st, s.n, s.R, Rc, rec.d_idx, rec.Xc, s.X);
    MLFMA_HIP(hipStreamSynchronize(st));
    s.fallbacks += st2[0].precond_fallbacks;
    for (int j = 0; j < Rc; ++j) {
        const int r = rec.idx[j];
        s.iters[r] += st2[j].iterations;
        s.done[r] = st2[j].converged ? 1 : 0;
        s.err[r] = st2[j].rel_residual;
    }
    s.matvecs += st2[0].matvecs * Rc;
    s.fp64_matvecs += st2[0].fp64_matvecs * Rc;
    dfree(rec.Bc);
    dfree(rec.Xc);
    dfree(rec.d_idx);
    rec = NarrowRecord();
}

void Engine::bs_stats(const BlockState& s, mlfma_solve_stats* stats) const {
    const double secs = wall() - s.t0;
    for (int r = 0; r < s.R; ++r) {
        stats[r].iterations = s.iters[r];
        stats[r].matvecs = s.matvecs / s.R;
        stats[r].fp64_matvecs = s.fp64_matvecs / s.R;
        stats[r].converged = s.done[r] ? 1 : 0;
        stats[r].rel_residual = s.err[r];
        stats[r].seconds = secs;
        stats[r].precond_fallbacks = s.fallbacks;
    }
}

void Engine::solve_block(const mlfma_solve_params& sp, const double2* d_B, double2* d_X, int R,
                         mlfma_solve_stats* stats) {
    BlockState s;
    bs_start(s, sp, d_B, d_X, R);
    while (bs_active(s)) {
        bs_cycle_begin(s);
        while (bs_step_open(s)) {
            const bool recorded = use_graphs_ && !prof_on_;
            bs_pre(s, recorded);
            if (recorded) run_graph(GraphKey{bs_graph_kind(s), R, apply_prec}, [&] { bs_body(s); });
            bs_post(s, recorded);
            bs_orth(s, 0);
            MLFMA_HIP(hipStreamSynchronize(st));
            bs_orth(s, 1);
        }
        bs_cycle_end(s, true);
    }
    bs_stats(s, stats);
}

// Near-field block-diagonal preconditioner: exact inverse of every leaf's self
// block (no reference counterpart; the reference solves unpreconditioned).
void Engine::precond_setup() {
    if (d_Dinv_) return;
    const bool pm = ie == MLFMA_PMCHWT;
    std::vector<long long> off(nleaf + 1, 0);
    for (int t = 0; t < nleaf; ++t) {
        const int nt = tree.leaf_ptr[t + 1] - tree.leaf_ptr[t];
        const int m = pm ? 2 * nt : nt;
        off[t + 1] = off[t] + (long long)m * m;
    }
    std::vector<double2> Dinv((size_t)off[nleaf]);
    std::vector<std::vector<double2>> self(nleaf);
    for (int t = 0; t < nleaf; ++t) self[t] = near_block(pair_ptr[t + 1] - 1);   // self block: last pair of the row
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < nleaf; ++t) {
        const int nt = tree.leaf_ptr[t + 1] - tree.leaf_ptr[t];
        const int m = pm ? 2 * nt : nt;
        std::vector<double2> blk;
        blk.swap(self[t]);
        std::vector<cplx> A((size_t)m * m), I((size_t)m * m, cplx(0, 0));
        for (int j = 0; j < nt; ++j)
            for (int i = 0; i < nt; ++i) {
                if (!pm) {
                    A[i + (size_t)j * m] = cx(blk[(size_t)j * nt + i]);
                } else {
                    const cplx z11 = cx(blk[(size_t)j * nt + i]), zk = cx(blk[(size_t)(nt + j) * nt + i]),
                               z22 = cx(blk[(size_t)(2 * nt + j) * nt + i]);
                    A[i + (size_t)j * m] = z11;
                    A[i + (size_t)(nt + j) * m] = -zk;
                    A[(nt + i) + (size_t)j * m] = zk;
                    A[(nt + i) + (size_t)(nt + j) * m] = z22;
                }
            }
        for (int i = 0; i < m; ++i) I[i + (size_t)i * m] = 1.0;
        // Gauss-Jordan with partial pivoting
        for (int c = 0; c < m; ++c) {
            int piv = c;
            for (int r = c + 1; r < m; ++r)
                if (std::abs(A[r + (size_t)c * m]) > std::abs(A[piv + (size_t)c * m])) piv = r;
            if (piv != c)
                for (int j = 0; j < m; ++j) {
                    std::swap(A[c + (size_t)j * m], A[piv + (size_t)j * m]);
                    std::swap(I[c + (size_t)j * m], I[piv + (size_t)j * m]);
                }
            const cplx d = A[c + (size_t)c * m];
            for (int j = 0; j < m; ++j) {
                A[c + (size_t)j * m] /= d;
                I[c + (size_t)j * m] /= d;
            }
            for (int r = 0; r < m; ++r) {
                if (r == c) continue;
                const cplx f = A[r + (size_t)c * m];
                if (f == cplx(0, 0)) continue;
                for (int j = 0; j < m; ++j) {
                    A[r + (size_t)j * m] -= f * A[c + (size_t)j * m];
                    I[r + (size_t)j * m] -= f * I[c + (size_t)j * m];
                }
            }
        }
        for (size_t e = 0; e < I.size(); ++e) Dinv[off[t] + e] = d2(I[e]);
    }
    d_Dinv_ = upload(Dinv);
    d_Doff_ = upload(off);
}

// Single application of preconditioner `kind` (1 block-diagonal, 2 polynomial)
// with the apply precision, for inspection.
void Engine::precond_apply_kind(int kind, const double2* x, double2* y) {
    if (kind == 1) precond_setup();
    const int keep = precond_kind_;
    precond_kind_ = kind;
    precond_apply(x, y);
    precond_kind_ = keep;
    MLFMA_HIP(hipStreamSynchronize(st));
}

// Polynomial preconditioner on nrhs column-major device vectors (apply precision).
void Engine::poly_apply_device(const double2* x, double2* y, int nrhs) {
    ensure_block(nrhs);
    ensure_poly(nrhs);
    launch_transpose(st, nrhs, nunk, x, d_bx_);
    poly_apply(d_bx_, d_by_, nrhs, apply_prec);
    launch_transpose(st, nunk, nrhs, d_by_, y);
    MLFMA_HIP(hipStreamSynchronize(st));
}

void Engine::precond_apply(const double2* x, double2* y) {
    if (precond_kind_ == 2) {
        poly_apply(x, y, 1, apply_prec);
        return;
    }
    if (precond_kind_ == 3) {
        inner_near_solve(x, y, 1, apply_prec);
        return;
    }
    launch_block_diag(st, ie == MLFMA_PMCHWT ? 3 : 1, nleaf, N, d_leaf_ptr_, d_perm_, d_Doff_, d_Dinv_, max_leaf, x, y);
}

// Preconditioner `kind` on an RHS-fastest block of R vectors (apply precision for
// the polynomial's near products; the block-diagonal inverses are FP64).
void Engine::precond_block(int kind, const double2* x, double2* y, int R, int prec) {
    if (kind == 2) {
        poly_apply(x, y, R, prec);
        return;
    }
    if (kind == 3) {
        inner_near_solve(x, y, R, prec);
        return;
    }
    launch_block_diag_b(st, ie == MLFMA_PMCHWT ? 3 : 1, nleaf, N, R, d_leaf_ptr_, d_perm_, d_Doff_, d_Dinv_, max_leaf,
                        x, y);
}

// ---------------------------------------------------------------- inner near-field solve
// Preconditioner 3: z = M^-1 v is `inner_steps_` steps of GMRES on the stored near
// field Z_near z = v, from z = 0, for an RHS-fastest block of R vectors (each
// right-hand side its own Arnoldi recurrence and Givens rotations; classical
// Gram-Schmidt with one re-orthogonalisation).  The near products use the
// preconditioner's storage of the apply precision (BF16 on the matrix cores in
// FP32-mode solves).  It is a different operator for every v, so the outer GMRES
// keeps z_k (flexible form).  No reference counterpart: the reference solves
// unpreconditioned (Code/MLFMM.cpp:370-397).
void Engine::set_inner_steps(int k) {
    if (k < 1 || k > 64) throw Error(MLFMA_ERR_ARG, "inner near-field steps must be 1..64");
    inner_steps_ = k;
}

void Engine::ensure_inner(int R) {
    ensure_poly(R);   // the BF16 copy of the near product's input
    const int k = inner_steps_;
    if (R <= inner_cap_R_ && k <= inner_cap_k_) return;
    clear_graphs();
    for (void* p : {(void*)d_iU_, (void*)d_iW_, (void*)d_ipart_, (void*)d_ismall_, (void*)d_iscale_}) dfree(p);
    const int Rc = std::max(R, inner_cap_R_), kc = std::max(k, inner_cap_k_);
    d_iU_ = dalloc<double2>((size_t)nunk * Rc * (kc + 1));
    d_iW_ = dalloc<double2>((size_t)nunk * Rc);
    d_ipart_ = dalloc<double2>((size_t)(kc + 1) * kRedBlocks * Rc);
    d_ismall_ = dalloc<double2>((size_t)(kc + 1) * Rc);
    d_iscale_ = dalloc<double>(Rc);
    inner_cap_R_ = Rc;
    inner_cap_k_ = kc;
}

void Engine::inner_near_solve(const double2* V, double2* Z, int R, int prec) {
    const int k = inner_steps_, n = nunk;
    const size_t nR = (size_t)n * R;
    if (nR > 0x7fffffff) throw Error(MLFMA_ERR_ARG, "inner near-field solve: n * R exceeds 2^31");
    ensure_inner(R);
    const double eps = 1e-16;
    auto norms = [&](const double2* v, std::vector<double>& out) {
        launch_norm2_partial_b(st, n, R, v, d_ipart_, kRedBlocks);
        launch_reduce_partials_b(st, 1, R, kRedBlocks, d_ipart_, d_ismall_);
        std::vector<double2> h(R);
        MLFMA_HIP(hipMemcpyAsync(h.data(), d_ismall_, R * sizeof(double2), hipMemcpyDeviceToHost, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        out.resize(R);
        for (int r = 0; r < R; ++r) out[r] = std::sqrt(h[r].x);
    };
    auto dots = [&](int cnt, std::vector<cplx>& h) {   // h[j * R + r] = <U_j, W>_r
        launch_multidot_b(st, n, R, cnt, d_iU_, nR, d_iW_, d_ipart_, kRedBlocks);
        launch_reduce_partials_b(st, cnt, R, kRedBlocks, d_ipart_, d_ismall_);
        std::vector<double2> tmp((size_t)cnt * R);
        MLFMA_HIP(hipMemcpyAsync(tmp.data(), d_ismall_, tmp.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        h.resize(tmp.size());
        for (size_t e = 0; e < tmp.size(); ++e) h[e] = cx(tmp[e]);
    };
    auto put = [&](const std::vector<cplx>& v) {
        std::vector<double2> tmp(v.size());
        for (size_t e = 0; e < v.size(); ++e) tmp[e] = d2(v[e]);
        MLFMA_HIP(hipMemcpyAsync(d_ismall_, tmp.data(), tmp.size() * sizeof(double2), hipMemcpyHostToDevice, st));
        MLFMA_HIP(hipStreamSynchronize(st));
    };
    auto scale_into = [&](const std::vector<double>& a, const double2* src, double2* dst) {
        MLFMA_HIP(hipMemcpyAsync(d_iscale_, a.data(), R * sizeof(double), hipMemcpyHostToDevice, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        launch_scale_copy_b(st, n, R, d_iscale_, src, dst);
    };
    std::vector<double> beta;
    norms(V, beta);
    std::vector<double> inv(R);
    for (int r = 0; r < R; ++r) inv[r] = beta[r] > 0 ? 1.0 / beta[r] : 0.0;
    scale_into(inv, V, d_iU_);
    std::vector<cplx> H((size_t)R * (k + 1) * k, cplx(0, 0)), cs((size_t)R * k), sn((size_t)R * k), g((size_t)R * (k + 1));
    auto Hm = [&](int r, int i, int j) -> cplx& { return H[((size_t)r * k + j) * (k + 1) + i]; };
    std::vector<int> used(R, 0);
    std::vector<char> live(R);
    for (int r = 0; r < R; ++r) {
        g[(size_t)r * (k + 1)] = beta[r];
        live[r] = beta[r] > 0;
    }
    std::vector<cplx> h1, h2;
    int steps = 0;
    for (int j = 0; j < k; ++j) {
        if (std::none_of(live.begin(), live.end(), [](char c) { return c != 0; })) break;
        near_block(d_iU_ + (size_t)j * nR, d_iW_, R, prec, true);
        dots(j + 1, h1);
        put(h1);
        launch_sub_combination_b(st, n, R, j + 1, d_iU_, nR, d_ismall_, d_iW_);
        dots(j + 1, h2);
        put(h2);
        launch_sub_combination_b(st, n, R, j + 1, d_iU_, nR, d_ismall_, d_iW_);
        std::vector<double> hn;
        norms(d_iW_, hn);
        for (int r = 0; r < R; ++r) {
            if (!live[r]) continue;
            used[r] = j + 1;
            for (int i = 0; i <= j; ++i) Hm(r, i, j) = h1[(size_t)i * R + r] + h2[(size_t)i * R + r];
            if (!std::isfinite(hn[r])) throw Error(MLFMA_ERR_NONFINITE, "non-finite inner Arnoldi vector");
            Hm(r, j + 1, j) = hn[r];
            cplx* c_ = &cs[(size_t)r * k];
            cplx* s_ = &sn[(size_t)r * k];
            cplx* gr = &g[(size_t)r * (k + 1)];
            for (int i = 0; i < j; ++i) {
                const cplx t = std::conj(c_[i]) * Hm(r, i, j) + std::conj(s_[i]) * Hm(r, i + 1, j);
                Hm(r, i + 1, j) = -s_[i] * Hm(r, i, j) + c_[i] * Hm(r, i + 1, j);
                Hm(r, i, j) = t;
            }
            const cplx a = Hm(r, j, j), b = Hm(r, j + 1, j);
            const double den = std::sqrt(std::norm(a) + std::norm(b));
            if (den < eps) {
                c_[j] = 1.0, s_[j] = 0.0;
            } else {
                c_[j] = a / den, s_[j] = b / den;
            }
            Hm(r, j, j) = std::conj(c_[j]) * a + std::conj(s_[j]) * b;
            Hm(r, j + 1, j) = 0.0;
            gr[j + 1] = -s_[j] * gr[j];
            gr[j] = std::conj(c_[j]) * gr[j];
            if (hn[r] < eps * beta[r]) live[r] = 0;   // breakdown: the inner solution is exact
        }
        ++steps;
        if (j + 1 < k) {
            for (int r = 0; r < R; ++r) inv[r] = (live[r] && hn[r] > 0) ? 1.0 / hn[r] : 0.0;
            scale_into(inv, d_iW_, d_iU_ + (size_t)(j + 1) * nR);
        }
    }
    // z_r = U y_r with y_r from each right-hand side's triangular Hessenberg
    std::vector<cplx> yv((size_t)std::max(steps, 1) * R, cplx(0, 0));
    for (int r = 0; r < R; ++r) {
        const int mm = used[r];
        const cplx* gr = &g[(size_t)r * (k + 1)];
        std::vector<cplx> y(mm);
        for (int i = mm - 1; i >= 0; --i) {
            y[i] = gr[i];
            for (int jj = i + 1; jj < mm; ++jj) y[i] -= Hm(r, i, jj) * y[jj];
            y[i] = (std::abs(Hm(r, i, i)) > eps) ? y[i] / Hm(r, i, i) : cplx(0, 0);
        }
        for (int i = 0; i < mm; ++i) yv[(size_t)i * R + r] = y[i];
    }
    MLFMA_HIP(hipMemsetAsync(Z, 0, nR * sizeof(double2), st));
    if (steps > 0) {
        put(yv);
        launch_add_combination_b(st, n, R, steps, d_iU_, nR, d_ismall_, Z);
    }
}

// ---------------------------------------------------------------- polynomial preconditioner
// M^-1 r = sum_k alpha_k t_k, t_0 = D^-1 r, t_{k+1} = t_k - D^-1 Z_near t_k on an
// RHS-fastest block of R vectors; the near products use the apply precision's
// near-field storage.  No reference counterpart.
void Engine::set_poly(const std::vector<cplx>& alpha) {
    if (alpha.empty() || alpha.size() > 16) throw Error(MLFMA_ERR_ARG, "polynomial degree must be 0..15");
    for (const cplx& a : alpha)
        if (!std::isfinite(a.real()) || !std::isfinite(a.imag())) throw Error(MLFMA_ERR_NONFINITE, "non-finite coefficient");
    poly_alpha_ = alpha;
    clear_graphs();   // the coefficients are kernel arguments of recorded sequences
}

// FP32 copy of the matrix-core near product's input (sized before any
// recording: the recorded sequences hold its address)
void Engine::ensure_xf(int R) {
    if (!d_Z32_ || R <= xf_cap_) return;
    clear_graphs();
    dfree(d_xf_);
    d_xf_ = dalloc<float2>((size_t)nunk * R);
    xf_cap_ = R;
}

// (the FP32-vector polynomial of launch_poly_init_bf16 / launch_near_bf16_poly
// keeps its basis vector in d_pt_ and the second BF16 input buffer in d_pu_,
// both large enough, as the FP64 polynomial does not run at the same time)
void Engine::ensure_poly(int R) {
    ensure_xf(R);
    ensure_eall(R);
    if (R <= poly_cap_) return;
    clear_graphs();
    dfree(d_pt_);
    dfree(d_pu_);
    d_pt_ = dalloc<double2>((size_t)nunk * R);
    d_pu_ = dalloc<double2>((size_t)nunk * R);
    if (d_Zbf_) {
        dfree(d_xb_);
        d_xb_ = dalloc<unsigned>((size_t)nunk * R);
    }
    poly_cap_ = R;
}

// Y = Z_near X for an RHS-fastest block (near storage of precision prec; the
// preconditioner's products in FP32 mode use the BF16 copy when there is one)
void Engine::near_block(const double2* X, double2* Y, int R, int prec, bool precond) {
    if (precond && prec == MLFMA_FP32 && d_Zbf_) {
        if (R > poly_cap_) throw Error(MLFMA_ERR_STATE, "preconditioner buffers smaller than the batch");
        launch_near_bf16(st, R, R, 0, nd_, d_Zbf_, d_bfoff_, X, d_xb_, Y);
        return;
    }
    if (R >= 2 && near_mfma_) {
        if (prec == MLFMA_FP32 && d_Zbs_) {
            ensure_xf(R);
            launch_near_bf16x3(st, R, R, 0, nd_, d_Zbs_, d_bsoff_, X, reinterpret_cast<uint2*>(d_xf_), Y);
        } else if (prec == MLFMA_FP32 && d_Z32_) {
            ensure_xf(R);
            launch_near_mfma(st, R, R, 0, nd_, d_Z32_, X, d_xf_, Y);
        } else {
            launch_near_mfma64(st, R, R, 0, nd_, d_Z_, X, Y);
        }
        return;
    }
    const int cap = (ie == MLFMA_PMCHWT) ? 4 : 8;
    for (int r0 = 0; r0 < R;) {
        int rb = 1;
        while (rb * 2 <= cap && r0 + rb * 2 <= R) rb *= 2;
        if (prec == MLFMA_FP32 && d_Z32_) launch_near_b<float>(st, rb, R, r0, nd_, d_Z32_, X, Y);
        else launch_near_b<double>(st, rb, R, r0, nd_, d_Z_, X, Y);
        r0 += rb;
    }
}

void Engine::poly_apply(const double2* X, double2* Y, int R, int prec) {
    if (poly_alpha_.empty()) throw Error(MLFMA_ERR_STATE, "polynomial preconditioner coefficients not set");
    ensure_poly(R);
    launch_poly_init(st, nunk, R, d_dinv_, X, d2(poly_alpha_[0]), d_pt_, Y);
    for (size_t k = 1; k < poly_alpha_.size(); ++k) {
        near_block(d_pt_, d_pu_, R, prec, true);
        launch_poly_step(st, nunk, R, d_dinv_, d_pu_, d2(poly_alpha_[k]), d_pt_, Y);
    }
}

// The same polynomial with FP32 vectors (x, y: nunk x R, RHS-fastest) around
// the single-BF16 products: each product's epilogue takes the polynomial's
// step and rounds the next input, so between products nothing else runs.
// Allocates nothing (ensure_poly(R) first).
void Engine::poly_apply_bf16f(const float2* X, float2* Y, int R) {
    if (poly_alpha_.empty()) throw Error(MLFMA_ERR_STATE, "polynomial preconditioner coefficients not set");
    if (!d_Zbf_ || !d_dinvf_ || R > poly_cap_)
        throw Error(MLFMA_ERR_STATE, "FP32 polynomial: BF16 preconditioner records and sized buffers needed");
    auto f2 = [](std::complex<double> a) { return make_float2((float)a.real(), (float)a.imag()); };
    float2* t = reinterpret_cast<float2*>(d_pt_);
    unsigned* xb[2] = {d_xb_, reinterpret_cast<unsigned*>(d_pu_)};
    launch_poly_init_bf16(st, R, nd_, d_dinvf_, X, f2(poly_alpha_[0]), t, Y, xb[0]);
    const size_t d = poly_alpha_.size() - 1;
    for (size_t k = 1; k <= d; ++k) {
        const PolyEpi pe{d_dinvf_, f2(poly_alpha_[k]), t, Y, k < d ? xb[k & 1] : nullptr};
        launch_near_bf16_poly(st, R, nd_, d_Zbf_, d_bfoff_, xb[(k - 1) & 1], pe);
    }
}

// Problem features for the coefficient network (kPrecondFeatures values):
// spectral moments mu_k = z^H B^k z / n of B = D^-1 Z_near for k = 1..8 with a
// seeded unit-modulus probe z (re, im), then eps_r (re, im), PMCHWT flag and
// log10 of the near-field entries per row.  FP64 near field.
void Engine::precond_features(float* f) {
    ensure_poly(1);
    double2* z = d_pu_;
    launch_probe(st, nunk, 12345u, d_pt_);
    std::vector<double2> hz(nunk), hv(nunk);
    MLFMA_HIP(hipMemcpyAsync(hz.data(), d_pt_, nunk * sizeof(double2), hipMemcpyDeviceToHost, st));
    double2* v = dalloc<double2>(nunk);
    MLFMA_HIP(hipMemcpyAsync(v, d_pt_, nunk * sizeof(double2), hipMemcpyDeviceToDevice, st));
    for (int k = 1; k <= 8; ++k) {
        near_block(v, z, 1, MLFMA_FP64);
        launch_diag_scale(st, nunk, 1, d_dinv_, z, v);
        MLFMA_HIP(hipMemcpyAsync(hv.data(), v, nunk * sizeof(double2), hipMemcpyDeviceToHost, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        cplx mu(0, 0);
        for (int i = 0; i < nunk; ++i) mu += std::conj(cx(hz[i])) * cx(hv[i]);
        mu /= (double)nunk;
        f[k - 1] = (float)mu.real();
        f[8 + k - 1] = (float)mu.imag();
    }
    dfree(v);
    f[16] = (float)P.eps_r_re;
    f[17] = (float)P.eps_r_im;
    f[18] = ie == MLFMA_PMCHWT ? 1.0f : 0.0f;
    // discretisation density, not object size: log10 of the near-field entries per row of the
    // full near matrix (PMCHWT stores 3 of its 4 blocks), so a family of smaller objects meshed
    // as densely covers a large member's value
    const double full = (ie == MLFMA_PMCHWT) ? (double)near_entries / 3.0 * 4.0 : (double)near_entries;
    f[19] = (float)std::log10(full / (double)nunk);
}

// Least-squares statistics of the polynomial coefficients from the solver's
// own Krylov pairs: `steps` Arnoldi steps of the unpreconditioned operator
// (apply precision) from rhs give pairs (r_k = A v_k, e_k = v_k); with the basis
// a_k = [t_0(r_k) .. t_d(r_k)], G = sum_k a_k^H a_k ((d+1) x (d+1), row-major),
// h = sum_k a_k^H e_k and c = sum_k |e_k|^2, so that
// sum_k |sum_j alpha_j t_j(r_k) - e_k|^2 = alpha^H G alpha - 2 Re alpha^H h + c.
void Engine::precond_statistics(const double2* d_b, int steps, int degree, double2* G, double2* h, double* c) {
    if (steps < 1 || steps > 256 || degree < 0 || degree > 15) throw Error(MLFMA_ERR_ARG, "bad steps / degree");
    const int n = nunk, m = degree + 1;
    const size_t nS = (size_t)n * steps;
    double2* V = dalloc<double2>((size_t)n * (steps + 1));   // column-major Arnoldi basis
    double2* W = dalloc<double2>(nS);                        // A v_k, column-major
    double2* part = dalloc<double2>((size_t)(steps + 1) * kRedBlocks);
    double2* small = dalloc<double2>(std::max(steps + 1, m * steps));
    auto nrm = [&](const double2* x) {
        launch_norm2_partial(st, n, x, d_rpart_ ? d_rpart_ : (double*)part, kRedBlocks);
        std::vector<double> hp(kRedBlocks);
        MLFMA_HIP(hipMemcpyAsync(hp.data(), d_rpart_ ? d_rpart_ : (double*)part, kRedBlocks * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        double s2 = 0;
        for (double x2 : hp) s2 += x2;
        return std::sqrt(s2);
    };
    const double bn = nrm(d_b);
    if (!(bn > 0)) throw Error(MLFMA_ERR_ARG, "zero right-hand side");
    launch_scale_copy(st, n, 1.0 / bn, d_b, V);
    int kept = 0;
    for (int k = 0; k < steps; ++k) {
        double2* w = W + (size_t)k * n;
        apply_one(V + (size_t)k * n, w, apply_prec);
        ++kept;
        // CGS2 into the next basis vector
        double2* nx = V + (size_t)(k + 1) * n;
        MLFMA_HIP(hipMemcpyAsync(nx, w, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
        for (int pass = 0; pass < 2; ++pass) {
            launch_multidot(st, n, k + 1, V, (size_t)n, nx, part, kRedBlocks);
            launch_reduce_partials(st, k + 1, kRedBlocks, part, small);
            launch_sub_combination(st, n, k + 1, V, (size_t)n, small, nx);
        }
        const double hn = nrm(nx);
        if (!(hn > 1e-14)) break;
        launch_scale_copy(st, n, 1.0 / hn, nx, nx);
    }
    // RHS-fastest blocks of the pairs, then the basis t_j of every pair
    const int R = kept;
    ensure_poly(R);
    double2* E = dalloc<double2>((size_t)n * R);
    double2* T = dalloc<double2>((size_t)n * R * m);
    launch_transpose(st, R, n, V, E);
    launch_transpose(st, R, n, W, d_pt_);
    launch_diag_scale(st, n, R, d_dinv_, d_pt_, T);
    for (int j = 1; j < m; ++j) {
        double2* Tj = T + (size_t)j * n * R;
        MLFMA_HIP(hipMemcpyAsync(Tj, Tj - (size_t)n * R, (size_t)n * R * sizeof(double2), hipMemcpyDeviceToDevice, st));
        near_block(Tj, d_pu_, R, MLFMA_FP64);
        launch_poly_step(st, n, R, d_dinv_, d_pu_, make_double2(0, 0), Tj, nullptr);
    }
    double2* bpart = dalloc<double2>((size_t)m * kRedBlocks * R);
    double2* bsmall = dalloc<double2>((size_t)m * R);
    std::vector<double2> hs((size_t)m * R);
    auto block_dots = [&](const double2* w, std::vector<cplx>& out) {   // out[j] = sum_r <T_j, w>_r
        launch_multidot_b(st, n, R, m, T, (size_t)n * R, w, bpart, kRedBlocks);
        launch_reduce_partials_b(st, m, R, kRedBlocks, bpart, bsmall);
        MLFMA_HIP(hipMemcpyAsync(hs.data(), bsmall, hs.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
        MLFMA_HIP(hipStreamSynchronize(st));
        out.assign(m, cplx(0, 0));
        for (int j = 0; j < m; ++j)
            for (int r = 0; r < R; ++r) out[j] += cx(hs[(size_t)j * R + r]);
    };
    std::vector<cplx> col;
    for (int l = 0; l < m; ++l) {
        block_dots(T + (size_t)l * n * R, col);
        for (int j = 0; j < m; ++j) G[(size_t)j * m + l] = d2(col[j]);
    }
    block_dots(E, col);
    for (int j = 0; j < m; ++j) h[j] = d2(col[j]);
    *c = (double)R;   // unit basis vectors
    for (void* q : {(void*)V, (void*)W, (void*)part, (void*)small, (void*)E, (void*)T, (void*)bpart, (void*)bsmall})
        dfree(q);
}

void Engine::excitation(double pol, double th, double ph, double2* d_v) {
    ExcitationArgs a;
    double vth[3], vph[3];
    plane_wave(pol, th, ph, a.kinc, a.einc, a.hinc, vth, vph);
    a.k = d2(med.k1);
    a.alpha = P.cfie_alpha;
    launch_excitation(st, N, ie, d_geo_, d_perm_, gauss_, a, d_v);
    MLFMA_HIP(hipGetLastError());
    MLFMA_HIP(hipStreamSynchronize(st));
}

// Right-hand sides of n_inc plane waves into column-major d_V (column j at d_V + j * nunk).
void Engine::excitation_batch(int n_inc, const double* pol, const double* th, const double* ph, double2* d_V) {
    for (int j = 0; j < n_inc; ++j) {
        ExcitationArgs a;
        double vth[3], vph[3];
        plane_wave(pol[j], th[j], ph[j], a.kinc, a.einc, a.hinc, vth, vph);
        a.k = d2(med.k1);
        a.alpha = P.cfie_alpha;
        launch_excitation(st, N, ie, d_geo_, d_perm_, gauss_, a, d_V + (size_t)j * nunk);
    }
    MLFMA_HIP(hipGetLastError());
    MLFMA_HIP(hipStreamSynchronize(st));
}

// Monostatic sweep (Code/RCS.h:272-325 PEC, :428-497 PMCHWT): for every angle the
// plane wave incident from it, solved for all angles together in batched GMRES
// (chunks of at most 256), then the backscatter RCS of each solution at its own
// angle.  out[3 a + {0,1,2}] = total, V, H in dBsm.
void Engine::monostatic(const mlfma_solve_params& sp, double pol, int nang, const double* th, const double* ph,
                        double* out, mlfma_solve_stats* stats) {
    if (nang < 1) throw Error(MLFMA_ERR_ARG, "monostatic: at least one angle");
    const int chunk = std::min(nang, 256);
    double2* B = dalloc<double2>((size_t)nunk * chunk);
    double2* X = dalloc<double2>((size_t)nunk * chunk);
    double2* Bt = dalloc<double2>((size_t)nunk * chunk);
    double2* Xt = dalloc<double2>((size_t)nunk * chunk);
    std::vector<double> pols(chunk, pol);
    for (int a0 = 0; a0 < nang; a0 += chunk) {
        const int R = std::min(chunk, nang - a0);
        excitation_batch(R, pols.data(), th + a0, ph + a0, B);
        launch_transpose(st, R, nunk, B, Bt);   // column-major -> RHS-fastest
        MLFMA_HIP(hipMemsetAsync(Xt, 0, (size_t)nunk * R * sizeof(double2), st));
        solve_block(sp, Bt, Xt, R, stats + a0);
        launch_transpose(st, nunk, R, Xt, X);
        rcs_multi(X, (size_t)nunk, R, th + a0, ph + a0, pol, out + 3 * (size_t)a0);
    }
    dfree(B);
    dfree(X);
    dfree(Bt);
    dfree(Xt);
}

void Engine::rcs(const double2* d_I, int nang, const double* th, const double* ph, double pol, double* out) {
    rcs_multi(d_I, 0, nang, th, ph, pol, out);
}

// RCS at nang angles; angle a uses the current vector d_I + a * istride.
void Engine::rcs_multi(const double2* d_I, size_t istride, int nang, const double* th, const double* ph, double pol,
                       double* out) {
    std::vector<double> dirs(3 * (size_t)nang);
    for (int a = 0; a < nang; ++a) {
        const double t = th[a] / 180.0 * kPi, p = ph[a] / 180.0 * kPi;
        dirs[3 * a] = std::sin(t) * std::cos(p);
        dirs[3 * a + 1] = std::sin(t) * std::sin(p);
        dirs[3 * a + 2] = std::cos(t);
    }
    double* d_dirs = nullptr;
    double2* d_part = nullptr;
    const int nblk = (N + 255) / 256;
    MLFMA_HIP(hipMalloc(&d_dirs, dirs.size() * sizeof(double)));
    MLFMA_HIP(hipMalloc(&d_part, (size_t)nang * nblk * 3 * sizeof(double2)));
    MLFMA_HIP(hipMemcpy(d_dirs, dirs.data(), dirs.size() * sizeof(double), hipMemcpyHostToDevice));
    launch_farfield(st, N, ie == MLFMA_PMCHWT, d_geo_, d_perm_, gauss_, d2(med.k1), nang, d_dirs, d_I, istride, d_part,
                    nblk);
    std::vector<double2> part((size_t)nang * nblk * 3);
    MLFMA_HIP(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
    MLFMA_HIP(hipStreamSynchronize(st));
    hipFree(d_dirs);
    hipFree(d_part);
    // PEC: jk eta F.I (Code/RCS.h:109-112); PMCHWT: jk1 (F_J I_J - F_M I_M) (:186-190)
    const cplx coeff = (ie == MLFMA_PMCHWT) ? cplx(0, 1) * med.k1 : cplx(0, 1) * med.k1 * med.eta1;
    for (int a = 0; a < nang; ++a) {
        cplx F[3] = {0, 0, 0};
        for (int b = 0; b < nblk; ++b)
            for (int c = 0; c < 3; ++c) F[c] += cx(part[((size_t)a * nblk + b) * 3 + c]);
        for (int c = 0; c < 3; ++c) F[c] *= coeff;
        double k[3], e[3], h[3], vth[3], vph[3];
        plane_wave(pol, th[a], ph[a], k, e, h, vth, vph);
        const cplx Ev = vth[0] * F[0] + vth[1] * F[1] + vth[2] * F[2];
        const cplx Eh = vph[0] * F[0] + vph[1] * F[1] + vph[2] * F[2];
        // Code/RCS.h:255-257
        out[3 * a] = 10.0 * std::log10((std::norm(Ev) + std::norm(Eh)) / (4.0 * kPi));
        out[3 * a + 1] = 10.0 * std::log10(std::norm(Ev) / (4.0 * kPi));
        out[3 * a + 2] = 10.0 * std::log10(std::norm(Eh) / (4.0 * kPi));
    }
}

// The same far field and dB values as rcs(), with every buffer either the
// caller's (currents, angles, output: device memory) or the file-scope scratch
// array of k_rcs_noalloc.hip; no allocation of any kind.
void Engine::rcs_noalloc(const double2* d_I, int nang, const double* d_th, const double* d_ph, double* d_out) {
    const cplx coeff = (ie == MLFMA_PMCHWT) ? cplx(0, 1) * med.k1 : cplx(0, 1) * med.k1 * med.eta1;
    launch_rcs_noalloc(st, P.device, N, ie == MLFMA_PMCHWT, d_geo_, d_perm_, gauss_, d2(med.k1), d2(coeff), nang,
                       d_th, d_ph, d_I, d_out);
    MLFMA_HIP(hipGetLastError());
    MLFMA_HIP(hipStreamSynchronize(st));
}

void Engine::get_workspace(int wn, int cur, int which, int level, int node, double* out) {
    if (wn < 1 || wn > nw || cur < 0 || cur >= ncur || level < 2 || level > tree.max_level || node < 0 ||
        node >= tree.lv[level].n)
        throw Error(MLFMA_ERR_ARG, "workspace index out of range");
    const WaveDev& D = wd_[wn - 1];
    const int K = D.K[level];
    const double2* src = (which == 0 ? D.f64.Sm[level] : D.f64.Bm[level]) + ((size_t)node * ncur + cur) * 3 * K;
    std::vector<double2> h(3 * (size_t)K);
    MLFMA_HIP(hipMemcpy(h.data(), src, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
    for (int q = 0; q < K; ++q)
        for (int c = 0; c < 3; ++c) {
            out[6 * q + 2 * c] = h[(size_t)c * K + q].x;
            out[6 * q + 2 * c + 1] = h[(size_t)c * K + q].y;
        }
}

void Engine::get_pattern(int wn, int which, int rwg, double* out) {
    if (wn < 1 || wn > nw || rwg < 0 || rwg >= N) throw Error(MLFMA_ERR_ARG, "pattern index out of range");
    const WaveDev& D = wd_[wn - 1];
    const int K = D.K[tree.max_level];
    const int i = inv_perm[rwg];
    std::vector<double2> h(3 * (size_t)K);
    MLFMA_HIP(hipMemcpy(h.data(), (which == 0 ? D.f64.Vs : D.f64.Vf) + (size_t)i * 3 * K, h.size() * sizeof(double2),
                        hipMemcpyDeviceToHost));
    for (int q = 0; q < K; ++q)
        for (int c = 0; c < 3; ++c) {
            out[6 * q + 2 * c] = h[(size_t)c * K + q].x;
            out[6 * q + 2 * c + 1] = h[(size_t)c * K + q].y;
        }
}

void Engine::get_near_row(int row, int* len, int* cols, double* vals) {
    if (row < 0 || row >= nunk) throw Error(MLFMA_ERR_ARG, "row out of range");
    const bool pm = ie == MLFMA_PMCHWT;
    const bool mrow = pm && row >= N;
    const int i = inv_perm[mrow ? row - N : row];
    const int t = leaf_of[i];
    const int nt = tree.leaf_ptr[t + 1] - tree.leaf_ptr[t];
    const int rr = i - tree.leaf_ptr[t];
    int cnt = 0;
    for (int p = pair_ptr[t]; p < pair_ptr[t + 1]; ++p) {
        const int s = pair_src[p];
        const int j0 = tree.leaf_ptr[s], ns = tree.leaf_ptr[s + 1] - j0;
        if (!cols) {
            cnt += pm ? 2 * ns : ns;
            continue;
        }
        const std::vector<double2> blk = near_block(p);
        for (int j = 0; j < ns; ++j) {
            const int sid = tree.perm[j0 + j];
            auto put = [&](double2 v, int col) {
                cols[cnt] = col;
                vals[2 * cnt] = v.x;
                vals[2 * cnt + 1] = v.y;
                ++cnt;
            };
            const double2 a = blk[(size_t)j * nt + rr];
            if (!pm) {
                put(a, sid);
            } else {
                const double2 k = blk[(size_t)(ns + j) * nt + rr], c = blk[(size_t)(2 * ns + j) * nt + rr];
                if (!mrow) {
                    put(a, sid);
                    put(make_double2(-k.x, -k.y), N + sid);
                } else {
                    put(k, sid);
                    put(c, N + sid);
                }
            }
        }
    }
    *len = cnt;
}

}  // namespace mlfma
