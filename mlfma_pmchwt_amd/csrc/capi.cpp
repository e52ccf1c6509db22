// C ABI of libmlfma.so (declared in include/mlfma.h).
#include <cstring>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"

// device staging buffers of the host-vector entry points
struct HostIO {
    double2* x = nullptr;
    double2* y = nullptr;
    size_t n = 0;
    ~HostIO() {
        if (x) hipFree(x);
        if (y) hipFree(y);
    }
    void ensure(size_t count) {
        if (count <= n) return;
        if (x) hipFree(x);
        if (y) hipFree(y);
        x = y = nullptr;
        n = 0;
        MLFMA_HIP(hipMalloc(&x, count * sizeof(double2)));
        MLFMA_HIP(hipMalloc(&y, count * sizeof(double2)));
        n = count;
    }
};

struct mlfma_handle {
    std::unique_ptr<mlfma::Engine> eng;
    std::string err;
    HostIO io;
};

namespace {
thread_local std::string g_create_error;

template <typename F>
int guarded(mlfma_handle* h, F&& f) {
    try {
        f();
        return MLFMA_OK;
    } catch (const mlfma::Error& e) {
        if (h) h->err = e.what();
        else g_create_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (h) h->err = "host allocation failed";
        else g_create_error = "host allocation failed";
        return MLFMA_ERR_OOM;
    } catch (const std::exception& e) {
        if (h) h->err = e.what();
        else g_create_error = e.what();
        return MLFMA_ERR_ARG;
    }
}

inline const double2* D2(const void* p) { return static_cast<const double2*>(p); }
inline double2* D2(void* p) { return static_cast<double2*>(p); }

int check_params(const mlfma_params* p, std::string& why) {
    if (!p) return why = "params is NULL", MLFMA_ERR_ARG;
    if (p->integral_equation < 0 || p->integral_equation > 2) return why = "integral_equation must be 0, 1 or 2", MLFMA_ERR_ARG;
    if (!(p->frequency_hz > 0)) return why = "frequency must be positive", MLFMA_ERR_ARG;
    if (!(p->leaf_coeff > 0)) return why = "leaf_coeff must be positive", MLFMA_ERR_ARG;
    if (p->far_precision != MLFMA_FP64 && p->far_precision != MLFMA_FP32) return why = "bad far_precision", MLFMA_ERR_ARG;
    if (p->near_precision != MLFMA_FP64 && p->near_precision != MLFMA_FP32 && p->near_precision != MLFMA_BF16)
        return why = "bad near_precision", MLFMA_ERR_ARG;
    if (p->precond_precision != MLFMA_FP64 && p->precond_precision != MLFMA_BF16)
        return why = "bad precond_precision", MLFMA_ERR_ARG;
    if (p->max_rhs < 1) return why = "max_rhs must be >= 1", MLFMA_ERR_ARG;
    const int g = p->gauss_points;
    if (g != 1 && g != 3 && g != 4 && g != 6 && g != 7 && g != 12 && g != 13 && g != 16)
        return why = "Unsupported Gauss point count: " + std::to_string(g), MLFMA_ERR_ARG;
    return MLFMA_OK;
}

}  // namespace

extern "C" {

void mlfma_params_default(mlfma_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->integral_equation = MLFMA_PMCHWT;
    p->gauss_points = 7;
    p->frequency_hz = 299792458.0;
    p->eps_r_re = 1.0;
    p->mu_r_re = 1.0;
    p->cfie_alpha = 0.2;
    p->leaf_coeff = 0.3;
    p->far_precision = MLFMA_FP64;
    p->near_precision = MLFMA_FP64;
    p->precond_precision = MLFMA_FP64;
    p->tf_mode = 0;
    p->max_rhs = 1;
    p->device = 0;
    p->verbose = 0;
}

int mlfma_create(const mlfma_params* p, int n_points, const int* point_ids, const double* xyz, int n_faces,
                 const int* face_ids, const int* face_vertex_ids, mlfma_handle** out) {
    if (!out) return MLFMA_ERR_ARG;
    *out = nullptr;
    std::string why;
    int rc = check_params(p, why);
    if (rc != MLFMA_OK) {
        g_create_error = why;
        return rc;
    }
    if (n_points < 3 || n_faces < 2 || !point_ids || !xyz || !face_ids || !face_vertex_ids) {
        g_create_error = "empty or NULL mesh arrays";
        return MLFMA_ERR_ARG;
    }
    auto h = std::make_unique<mlfma_handle>();
    rc = guarded(nullptr, [&] {
        h->eng = std::make_unique<mlfma::Engine>(*p);
        h->eng->build(n_points, point_ids, xyz, n_faces, face_ids, face_vertex_ids);
    });
    if (rc != MLFMA_OK) return rc;
    *out = h.release();
    return MLFMA_OK;
}

int mlfma_create_nas(const mlfma_params* p, const char* nas_path, mlfma_handle** out) {
    if (!out || !nas_path) return MLFMA_ERR_ARG;
    *out = nullptr;
    std::vector<int> ids, fids, fvid;
    std::vector<double> xyz;
    std::string err;
    if (mlfma::read_nas(nas_path, ids, xyz, fids, fvid, err) != 0) {
        g_create_error = err;
        return MLFMA_ERR_MESH;
    }
    return mlfma_create(p, (int)ids.size(), ids.data(), xyz.data(), (int)fids.size(), fids.data(), fvid.data(), out);
}

void mlfma_destroy(mlfma_handle* h) { delete h; }

const char* mlfma_last_error(const mlfma_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int mlfma_dims(const mlfma_handle* h, long long* o) {
    if (!h || !o) return MLFMA_ERR_ARG;
    const mlfma::Engine& e = *h->eng;
    o[0] = e.N;
    o[1] = e.nunk;
    o[2] = e.tree.max_level;
    o[3] = e.nleaf;
    o[4] = e.near_entries;
    o[5] = (long long)e.dev_bytes;
    return MLFMA_OK;
}

int mlfma_apply_device(mlfma_handle* h, int n, int nrhs, const void* x_dev, void* y_dev) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        if (n != e.nunk || nrhs < 1 || !x_dev || !y_dev) throw mlfma::Error(MLFMA_ERR_ARG, "MLFMM::applyAx invalid input.");
        e.apply_device(D2(x_dev), D2(y_dev), nrhs);
        MLFMA_HIP(hipStreamSynchronize(e.st));
    });
}

int mlfma_set_graphs(mlfma_handle* h, int on) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->set_graphs(on != 0); });
}

int mlfma_autotune(mlfma_handle* h, int nrhs, int precision, double* out4) {
    if (!h || !out4) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        if (precision != MLFMA_FP64 && precision != MLFMA_FP32) throw mlfma::Error(MLFMA_ERR_ARG, "bad precision");
        h->eng->autotune(nrhs, precision, out4);
        MLFMA_HIP(hipStreamSynchronize(h->eng->st));
    });
}

int mlfma_set_tuning(mlfma_handle* h, int precision, int rhs_per_pass, int bits) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        if (precision != MLFMA_FP64 && precision != MLFMA_FP32) throw mlfma::Error(MLFMA_ERR_ARG, "bad precision");
        if (!(rhs_per_pass == 0 || rhs_per_pass == 1 || rhs_per_pass == 2 || rhs_per_pass == 4 || rhs_per_pass == 8) ||
            bits < 0 || bits > 31)
            throw mlfma::Error(MLFMA_ERR_ARG, "rhs_per_pass must be 0, 1, 2, 4 or 8 and bits in 0..31");
        h->eng->set_tuning(precision, rhs_per_pass, bits);
    });
}

int mlfma_set_apply_precision(mlfma_handle* h, int precision) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->set_apply_precision(precision); });
}

int mlfma_apply_batch(mlfma_handle* h, int n, int nrhs, const double* x, double* y) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        if (n != e.nunk || nrhs < 1 || !x || !y) throw mlfma::Error(MLFMA_ERR_ARG, "MLFMM::applyAx invalid input.");
        const size_t cnt = (size_t)n * nrhs;
        h->io.ensure(cnt);
        MLFMA_HIP(hipMemcpy(h->io.x, x, cnt * sizeof(double2), hipMemcpyHostToDevice));
        e.apply_device(h->io.x, h->io.y, nrhs);
        MLFMA_HIP(hipStreamSynchronize(e.st));
        MLFMA_HIP(hipMemcpy(y, h->io.y, cnt * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int mlfma_apply(mlfma_handle* h, int n, const double* x, double* y) { return mlfma_apply_batch(h, n, 1, x, y); }

int mlfma_solve(mlfma_handle* h, const mlfma_solve_params* sp, int n, const double* rhs, double* x,
                mlfma_solve_stats* stats) {
    if (!h) return MLFMA_ERR_ARG;
    int conv = 0;
    int rc = guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        if (!sp || n != e.nunk || !rhs || !x) throw mlfma::Error(MLFMA_ERR_ARG, "invalid solve arguments");
        if (sp->solver != MLFMA_SOLVER_GMRES && sp->solver != MLFMA_SOLVER_CGS)
            throw mlfma::Error(MLFMA_ERR_ARG, "Invalid matrixSolverType_ in MLFMM (must be 0 for GMRES or 1 for CGS)");
        if (sp->itr_max < 0 || sp->restart < 1 || sp->tol_rel < 0) throw mlfma::Error(MLFMA_ERR_ARG, "bad solver parameters");
        h->io.ensure((size_t)n);
        MLFMA_HIP(hipMemcpy(h->io.x, rhs, (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
        MLFMA_HIP(hipMemcpy(h->io.y, x, (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
        mlfma_solve_stats s{};
        e.solve(*sp, h->io.x, h->io.y, s);
        MLFMA_HIP(hipMemcpy(x, h->io.y, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost));
        if (stats) *stats = s;
        conv = s.converged;
    });
    if (rc != MLFMA_OK) return rc;
    return conv ? MLFMA_OK : MLFMA_ERR_NOT_CONVERGED;
}

int mlfma_solve_batch(mlfma_handle* h, const mlfma_solve_params* sp, int n, int nrhs, const double* rhs, double* x,
                      mlfma_solve_stats* stats) {
    if (!h) return MLFMA_ERR_ARG;
    int conv = 1;
    int rc = guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        if (!sp || n != e.nunk || nrhs < 1 || nrhs > 256 || !rhs || !x)
            throw mlfma::Error(MLFMA_ERR_ARG, "invalid batched solve arguments");
        if (sp->solver != MLFMA_SOLVER_GMRES) throw mlfma::Error(MLFMA_ERR_ARG, "batched solve supports GMRES only");
        if (sp->itr_max < 0 || sp->restart < 1 || sp->tol_rel < 0) throw mlfma::Error(MLFMA_ERR_ARG, "bad solver parameters");
        const size_t cnt = (size_t)n * nrhs;
        h->io.ensure(cnt);
        MLFMA_HIP(hipMemcpy(h->io.x, rhs, cnt * sizeof(double2), hipMemcpyHostToDevice));
        MLFMA_HIP(hipMemcpy(h->io.y, x, cnt * sizeof(double2), hipMemcpyHostToDevice));
        std::vector<mlfma_solve_stats> s(nrhs);
        e.solve_device(*sp, h->io.x, h->io.y, nrhs, s.data());
        MLFMA_HIP(hipMemcpy(x, h->io.y, cnt * sizeof(double2), hipMemcpyDeviceToHost));
        for (int r = 0; r < nrhs; ++r) {
            if (stats) stats[r] = s[r];
            conv = conv && s[r].converged;
        }
    });
    if (rc != MLFMA_OK) return rc;
    return conv ? MLFMA_OK : MLFMA_ERR_NOT_CONVERGED;
}

int mlfma_solve_multi(int count, mlfma_handle* const* hs, const mlfma_solve_params* sp, const int* nrhs,
                      const double* const* rhs, double* const* x, mlfma_solve_stats* const* stats) {
    if (count < 1 || !hs || !sp || !nrhs || !rhs || !x) return MLFMA_ERR_ARG;
    for (int k = 0; k < count; ++k)
        if (!hs[k]) return MLFMA_ERR_ARG;
    int conv = 1;
    int rc = guarded(hs[0], [&] {
        if (sp->solver != MLFMA_SOLVER_GMRES) throw mlfma::Error(MLFMA_ERR_ARG, "lockstep solve supports GMRES only");
        if (sp->itr_max < 0 || sp->restart < 1 || sp->tol_rel < 0) throw mlfma::Error(MLFMA_ERR_ARG, "bad solver parameters");
        std::vector<mlfma::Engine*> es(count);
        std::vector<const double2*> B(count);
        std::vector<double2*> X(count);
        std::vector<int> R(count);
        std::vector<std::vector<mlfma_solve_stats>> st(count);
        std::vector<mlfma_solve_stats*> sp_out(count);
        for (int k = 0; k < count; ++k) {
            mlfma::Engine& e = *hs[k]->eng;
            if (nrhs[k] < 1 || nrhs[k] > 256 || !rhs[k] || !x[k])
                throw mlfma::Error(MLFMA_ERR_ARG, "lockstep solve: 1..256 right-hand sides and buffers per object");
            const size_t cnt = (size_t)e.nunk * nrhs[k];
            hs[k]->io.ensure(cnt);
            MLFMA_HIP(hipMemcpy(hs[k]->io.x, rhs[k], cnt * sizeof(double2), hipMemcpyHostToDevice));
            MLFMA_HIP(hipMemcpy(hs[k]->io.y, x[k], cnt * sizeof(double2), hipMemcpyHostToDevice));
            e.block_load(hs[k]->io.x, hs[k]->io.y, nrhs[k]);
            es[k] = &e;
            B[k] = e.block_b();
            X[k] = e.block_x();
            R[k] = nrhs[k];
            st[k].resize(nrhs[k]);
            sp_out[k] = st[k].data();
        }
        mlfma::solve_multi(es, *sp, B, X, R, sp_out);
        for (int k = 0; k < count; ++k) {
            mlfma::Engine& e = *hs[k]->eng;
            const size_t cnt = (size_t)e.nunk * nrhs[k];
            e.block_store(hs[k]->io.y, nrhs[k]);
            MLFMA_HIP(hipMemcpy(x[k], hs[k]->io.y, cnt * sizeof(double2), hipMemcpyDeviceToHost));
            for (int r = 0; r < nrhs[k]; ++r) {
                if (stats && stats[k]) stats[k][r] = st[k][r];
                conv = conv && st[k][r].converged;
            }
        }
    });
    if (rc != MLFMA_OK) return rc;
    return conv ? MLFMA_OK : MLFMA_ERR_NOT_CONVERGED;
}

int mlfma_precond_features(mlfma_handle* h, float* features) {
    if (!h || !features) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->precond_features(features); });
}

int mlfma_precond_statistics(mlfma_handle* h, const double* rhs, int steps, int degree, double* G, double* hv,
                             double* c) {
    if (!h || !rhs || !G || !hv || !c) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        h->io.ensure((size_t)e.nunk);
        MLFMA_HIP(hipMemcpy(h->io.x, rhs, (size_t)e.nunk * sizeof(double2), hipMemcpyHostToDevice));
        e.precond_statistics(h->io.x, steps, degree, reinterpret_cast<double2*>(G), reinterpret_cast<double2*>(hv), c);
    });
}

int mlfma_set_precond_inner(mlfma_handle* h, int steps) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->set_inner_steps(steps); });
}

int mlfma_set_precond_poly(mlfma_handle* h, int degree, const double* alpha) {
    if (!h || !alpha || degree < 0 || degree > 15) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        std::vector<mlfma::cplx> a(degree + 1);
        for (int k = 0; k <= degree; ++k) a[k] = mlfma::cplx(alpha[2 * k], alpha[2 * k + 1]);
        h->eng->set_poly(a);
    });
}

int mlfma_precond_apply(mlfma_handle* h, int kind, int n, const double* x, double* y) {
    if (!h || !x || !y) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        if (n != e.nunk || (kind != 1 && kind != 2)) throw mlfma::Error(MLFMA_ERR_ARG, "bad preconditioner apply");
        h->io.ensure((size_t)n);
        MLFMA_HIP(hipMemcpy(h->io.x, x, (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
        e.precond_apply_kind(kind, h->io.x, h->io.y);
        MLFMA_HIP(hipMemcpy(y, h->io.y, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int mlfma_precond_apply_batch(mlfma_handle* h, int n, int nrhs, const double* x, double* y) {
    if (!h || !x || !y) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        if (n != e.nunk || nrhs < 1 || nrhs > 256) throw mlfma::Error(MLFMA_ERR_ARG, "bad preconditioner apply");
        const size_t cnt = (size_t)n * nrhs;
        h->io.ensure(cnt);
        MLFMA_HIP(hipMemcpy(h->io.x, x, cnt * sizeof(double2), hipMemcpyHostToDevice));
        e.poly_apply_device(h->io.x, h->io.y, nrhs);
        MLFMA_HIP(hipMemcpy(y, h->io.y, cnt * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int mlfma_excitation(mlfma_handle* h, double pol, double th, double ph, double* v) {
    if (!h || !v) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        h->io.ensure((size_t)e.nunk);
        e.excitation(pol, th, ph, h->io.x);
        MLFMA_HIP(hipMemcpy(v, h->io.x, (size_t)e.nunk * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int mlfma_rcs(mlfma_handle* h, const double* I, int nang, const double* th, const double* ph, double pol,
              double* out) {
    if (!h || !I || nang < 1 || !th || !ph || !out) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        h->io.ensure((size_t)e.nunk);
        MLFMA_HIP(hipMemcpy(h->io.x, I, (size_t)e.nunk * sizeof(double2), hipMemcpyHostToDevice));
        e.rcs(h->io.x, nang, th, ph, pol, out);
    });
}

int mlfma_rcs_device_noalloc(mlfma_handle* h, const void* I_dev, int n_angles, const double* theta_dev,
                             const double* phi_dev, double* out_dev) {
    if (!h || !I_dev || n_angles < 1 || !theta_dev || !phi_dev || !out_dev) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        h->eng->rcs_noalloc(static_cast<const double2*>(I_dev), n_angles, theta_dev, phi_dev, out_dev);
    });
}

int mlfma_excitation_batch(mlfma_handle* h, int n_inc, const double* pol, const double* th, const double* ph,
                           double* v) {
    if (!h || n_inc < 1 || !pol || !th || !ph || !v) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        const size_t cnt = (size_t)e.nunk * n_inc;
        h->io.ensure(cnt);
        e.excitation_batch(n_inc, pol, th, ph, h->io.x);
        MLFMA_HIP(hipMemcpy(v, h->io.x, cnt * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int mlfma_monostatic(mlfma_handle* h, const mlfma_solve_params* sp, double pol, int n_angles, const double* th,
                     const double* ph, double* out3, mlfma_solve_stats* stats) {
    if (!h || !sp || n_angles < 1 || !th || !ph || !out3) return MLFMA_ERR_ARG;
    int conv = 1;
    const int rc = guarded(h, [&] {
        if (sp->solver != MLFMA_SOLVER_GMRES) throw mlfma::Error(MLFMA_ERR_ARG, "monostatic sweep: GMRES only");
        if (sp->itr_max < 0 || sp->restart < 1 || sp->tol_rel < 0) throw mlfma::Error(MLFMA_ERR_ARG, "bad solver parameters");
        std::vector<mlfma_solve_stats> s(n_angles);
        h->eng->monostatic(*sp, pol, n_angles, th, ph, out3, s.data());
        for (int a = 0; a < n_angles; ++a) {
            if (stats) stats[a] = s[a];
            conv = conv && s[a].converged;
        }
    });
    if (rc != MLFMA_OK) return rc;
    return conv ? MLFMA_OK : MLFMA_ERR_NOT_CONVERGED;
}

int mlfma_scatter_grid(double th_s, double th_f, double ph_s, double ph_f, double step, int* n, double* theta,
                       double* phi) {
    if (!n || !(step > 0)) return MLFMA_ERR_ARG;
    // Code/RCS.h:46-51 (point count) and :239-246 (theta sweep unless the theta range is empty)
    const bool along_phi = std::abs(th_f - th_s) < 1e-6;
    const int cnt = (int)(std::abs(along_phi ? ph_f - ph_s : th_f - th_s) / step) + 1;
    if (theta && phi) {
        if (*n < cnt) return MLFMA_ERR_ARG;
        for (int i = 0; i < cnt; ++i) {
            theta[i] = along_phi ? th_s : th_s + i * step;
            phi[i] = along_phi ? ph_s + i * step : ph_s;
        }
    }
    *n = cnt;
    return MLFMA_OK;
}

int mlfma_stage(mlfma_handle* h, int stage, const double* x, double* y) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] {
        mlfma::Engine& e = *h->eng;
        h->io.ensure((size_t)e.nunk);
        if (x) MLFMA_HIP(hipMemcpy(h->io.x, x, (size_t)e.nunk * sizeof(double2), hipMemcpyHostToDevice));
        e.stage(stage, h->io.x, h->io.y);
        if (y) MLFMA_HIP(hipMemcpy(y, h->io.y, (size_t)e.nunk * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int mlfma_get_workspace(mlfma_handle* h, int wn, int cur, int which, int level, int node, double* out) {
    if (!h || !out) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->get_workspace(wn, cur, which, level, node, out); });
}

int mlfma_set_profiling(mlfma_handle* h, int enable) {
    if (!h) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->set_profiling(enable != 0); });
}

int mlfma_get_profile(mlfma_handle* h, double* ms, long long* launches, int* napply) {
    if (!h) return MLFMA_ERR_ARG;
    const mlfma::Engine& e = *h->eng;
    for (int g = 0; g < MLFMA_PROFILE_GROUPS; ++g) {
        if (ms) ms[g] = e.prof_ms[g];
        if (launches) launches[g] = e.prof_launch[g];
    }
    if (napply) *napply = e.prof_apply;
    return MLFMA_OK;
}

int mlfma_get_rwg(mlfma_handle* h, int* ints6, double* len, double* mid3) {
    if (!h) return MLFMA_ERR_ARG;
    const mlfma::Basis& b = h->eng->basis;
    for (int i = 0; i < b.n; ++i) {
        if (ints6) {
            const int v[6] = {b.fpos[i], b.fneg[i], b.e0[i], b.e1[i], b.vpos[i], b.vneg[i]};
            std::memcpy(ints6 + 6 * i, v, sizeof(v));
        }
        if (len) len[i] = b.len[i];
        if (mid3) mid3[3 * i] = b.mid[i].x, mid3[3 * i + 1] = b.mid[i].y, mid3[3 * i + 2] = b.mid[i].z;
    }
    return MLFMA_OK;
}

int mlfma_get_level(mlfma_handle* h, int level, int* n, unsigned long long* codes, double* centres3, double* len,
                    int* parent) {
    if (!h || !n) return MLFMA_ERR_ARG;
    const mlfma::Octree& t = h->eng->tree;
    if (level < 0 || level > t.max_level) return MLFMA_ERR_ARG;
    const mlfma::Level& v = t.lv[level];
    *n = v.n;
    for (int i = 0; i < v.n; ++i) {
        if (codes) codes[i] = v.code[i];
        if (centres3) centres3[3 * i] = v.centre[i].x, centres3[3 * i + 1] = v.centre[i].y, centres3[3 * i + 2] = v.centre[i].z;
        if (len) len[i] = v.len;
        if (parent) parent[i] = v.parent.empty() ? -1 : v.parent[i];
    }
    return MLFMA_OK;
}

int mlfma_get_level_list(mlfma_handle* h, int level, int which, int* total, int* ptr, int* idx) {
    if (!h || !total) return MLFMA_ERR_ARG;
    const mlfma::Engine& e = *h->eng;
    const mlfma::Octree& t = e.tree;
    if (level < 0 || level > t.max_level || which < 0 || which > 4) return MLFMA_ERR_ARG;
    const mlfma::Level& v = t.lv[level];
    std::vector<int> P(v.n + 1, 0), I;
    for (int i = 0; i < v.n; ++i) {
        switch (which) {
            case 0: I.insert(I.end(), v.near_idx.begin() + v.near_ptr[i], v.near_idx.begin() + v.near_ptr[i + 1]); break;
            case 1:
                if (!v.dr_ptr.empty()) I.insert(I.end(), v.dr_idx.begin() + v.dr_ptr[i], v.dr_idx.begin() + v.dr_ptr[i + 1]);
                break;
            case 2:
                if (!v.dr_ptr.empty()) I.insert(I.end(), v.dr_vec.begin() + v.dr_ptr[i], v.dr_vec.begin() + v.dr_ptr[i + 1]);
                break;
            case 3:
                if (level < t.max_level)
                    I.insert(I.end(), v.son_idx.begin() + v.son_ptr[i], v.son_idx.begin() + v.son_ptr[i + 1]);
                break;
            default: {
                // a node holds the RWG functions of all leaves below it, in son
                // order (Code/OCTree.cpp:359-363): one contiguous leaf range
                int lo = i, hi = i + 1;
                for (int l = level; l < t.max_level; ++l) {
                    const mlfma::Level& c = t.lv[l];
                    lo = c.son_idx[c.son_ptr[lo]];
                    hi = c.son_idx[c.son_ptr[hi] - 1] + 1;
                }
                for (int k = t.leaf_ptr[lo]; k < t.leaf_ptr[hi]; ++k) I.push_back(t.perm[k]);
                break;
            }
        }
        P[i + 1] = (int)I.size();
    }
    *total = (int)I.size();
    if (ptr) std::memcpy(ptr, P.data(), P.size() * sizeof(int));
    if (idx && !I.empty()) std::memcpy(idx, I.data(), I.size() * sizeof(int));
    return MLFMA_OK;
}

int mlfma_get_drvec(mlfma_handle* h, int level, int* n, double* xyz) {
    if (!h || !n) return MLFMA_ERR_ARG;
    const mlfma::Octree& t = h->eng->tree;
    if (level < 2 || level > t.max_level) return MLFMA_ERR_ARG;
    const auto& v = t.lv[level].drvec;
    *n = (int)v.size();
    if (xyz)
        for (size_t i = 0; i < v.size(); ++i) xyz[3 * i] = v[i].x, xyz[3 * i + 1] = v[i].y, xyz[3 * i + 2] = v[i].z;
    return MLFMA_OK;
}

int mlfma_get_grid(mlfma_handle* h, int wn, int level, int* L, double* theta, double* phi, double* wgl,
                   double* dphi) {
    if (!h || !L) return MLFMA_ERR_ARG;
    const mlfma::Engine& e = *h->eng;
    if (wn < 1 || wn > e.nw || level < 2 || level > e.tree.max_level) return MLFMA_ERR_ARG;
    const mlfma::Grid& g = e.W[wn - 1].grid[level];
    *L = g.L;
    if (theta) std::memcpy(theta, g.theta.data(), g.theta.size() * sizeof(double));
    if (phi) std::memcpy(phi, g.phi.data(), g.phi.size() * sizeof(double));
    if (wgl) std::memcpy(wgl, g.wgl.data(), g.wgl.size() * sizeof(double));
    if (dphi) *dphi = g.dphi;
    return MLFMA_OK;
}

int mlfma_get_interp(mlfma_handle* h, int wn, int level, int* taps, double* weights) {
    if (!h) return MLFMA_ERR_ARG;
    const mlfma::Engine& e = *h->eng;
    if (wn < 1 || wn > e.nw || level < 2 || level >= e.tree.max_level) return MLFMA_ERR_ARG;
    const mlfma::InterpTable& it = e.W[wn - 1].interp[level + 1];
    for (int q = 0; q < it.Kp; ++q)
        for (int s = 0; s < 9; ++s) {
            if (taps) taps[9 * q + s] = it.tap[(size_t)s * it.Kp + q];
            if (weights) weights[9 * q + s] = it.w[(size_t)s * it.Kp + q];
        }
    return MLFMA_OK;
}

int mlfma_get_tf(mlfma_handle* h, int wn, int level, double* out) {
    if (!h || !out) return MLFMA_ERR_ARG;
    const mlfma::Engine& e = *h->eng;
    if (wn < 1 || wn > e.nw || level < 2 || level > e.tree.max_level) return MLFMA_ERR_ARG;
    const auto& tf = e.W[wn - 1].tf[level];
    for (size_t i = 0; i < tf.size(); ++i) out[2 * i] = tf[i].real(), out[2 * i + 1] = tf[i].imag();
    return MLFMA_OK;
}

int mlfma_get_pattern(mlfma_handle* h, int wn, int which, int rwg, double* out) {
    if (!h || !out) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->get_pattern(wn, which, rwg, out); });
}

int mlfma_get_near_row(mlfma_handle* h, int row, int* len, int* cols, double* vals) {
    if (!h || !len) return MLFMA_ERR_ARG;
    return guarded(h, [&] { h->eng->get_near_row(row, len, cols, vals); });
}

}  // extern "C"
