// The coefficient network of the learned preconditioner: a 2-hidden-layer
// tanh MLP (n_in -> width -> width -> n_out), FP32, trained on the device with
// the kernels of k_mlp.hip.  Inputs are standardised with a stored mean / std;
// outputs are the real and imaginary parts of the polynomial coefficients.
// No reference counterpart.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/mlfma.h"
#include "engine.hpp"
#include "precond.hpp"

namespace mlfma {

class Mlp {
public:
    Mlp(int nin, int width, int nout, unsigned seed) : nin_(nin), w_(width), nout_(nout) {
        MLFMA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        off_[0] = 0;                              // W1 (w x nin)
        off_[1] = off_[0] + (size_t)w_ * nin_;    // b1
        off_[2] = off_[1] + w_;                   // W2 (w x w)
        off_[3] = off_[2] + (size_t)w_ * w_;      // b2
        off_[4] = off_[3] + w_;                   // W3 (nout x w)
        off_[5] = off_[4] + (size_t)nout_ * w_;   // b3
        np_ = off_[5] + nout_;
        std::vector<float> p(np_, 0.0f);
        std::mt19937 rng(seed);
        auto xavier = [&](size_t o, int fo, int fi, float gain) {
            const float a = gain * std::sqrt(6.0f / (float)(fi + fo));
            std::uniform_real_distribution<float> u(-a, a);
            for (size_t e = 0; e < (size_t)fo * fi; ++e) p[o + e] = u(rng);
        };
        xavier(off_[0], w_, nin_, 1.0f);
        xavier(off_[2], w_, w_, 1.0f);
        xavier(off_[4], nout_, w_, 0.01f);   // starts close to the output bias
        d_p_ = alloc<float>(np_);
        d_g_ = alloc<float>(np_);
        d_m1_ = alloc<float>(np_);
        d_m2_ = alloc<float>(np_);
        set_params(p.data());
        mean_.assign(nin_, 0.0f);
        stdv_.assign(nin_, 1.0f);
    }
    ~Mlp() {
        if (st_) hipStreamSynchronize(st_);
        for (void* a : allocs_) hipFree(a);
        if (st_) hipStreamDestroy(st_);
    }

    size_t nparams() const { return np_; }
    void set_params(const float* p) {
        MLFMA_HIP(hipMemcpy(d_p_, p, np_ * sizeof(float), hipMemcpyHostToDevice));
        MLFMA_HIP(hipMemset(d_m1_, 0, np_ * sizeof(float)));
        MLFMA_HIP(hipMemset(d_m2_, 0, np_ * sizeof(float)));
        step_ = 0;
    }
    void get_params(float* p) { MLFMA_HIP(hipMemcpy(p, d_p_, np_ * sizeof(float), hipMemcpyDeviceToHost)); }
    void set_normalization(const float* mean, const float* stdv) {
        for (int i = 0; i < nin_; ++i) {
            if (!(stdv[i] > 0.0f) || !std::isfinite(mean[i])) throw Error(MLFMA_ERR_ARG, "bad normalisation");
            mean_[i] = mean[i];
            stdv_[i] = stdv[i];
        }
    }

    // Y = net(X) for B samples (host arrays)
    void predict(int B, const float* X, float* Y) {
        stage_inputs(B, X);
        forward(B);
        MLFMA_HIP(hipMemcpyAsync(Y, d_Y_, (size_t)B * nout_ * sizeof(float), hipMemcpyDeviceToHost, st_));
        MLFMA_HIP(hipStreamSynchronize(st_));
    }

    // mean quadratic loss over B samples and its gradient w.r.t. the parameters
    // (m = n_out / 2 complex coefficients; G row-major m x m per sample)
    double loss_grad(int B, const float* X, const double* G, const double* h, const double* c, float* grad) {
        stage_inputs(B, X);
        stage_targets(B, G, h, c);
        const double l = forward_backward(B);
        if (grad) MLFMA_HIP(hipMemcpy(grad, d_g_, np_ * sizeof(float), hipMemcpyDeviceToHost));
        return l;
    }

    // `steps` full-batch Adam steps; loss history written per step
    void train(int B, const float* X, const double* G, const double* h, const double* c, int steps, float lr,
               float* history) {
        stage_inputs(B, X);
        stage_targets(B, G, h, c);
        for (int s = 0; s < steps; ++s) {
            const double l = forward_backward(B);
            if (history) history[s] = (float)l;
            ++step_;
            launch_adam(st_, (int)np_, d_p_, d_g_, d_m1_, d_m2_, lr, 0.9f, 0.999f, 1e-8f, step_);
        }
        MLFMA_HIP(hipStreamSynchronize(st_));
        MLFMA_HIP(hipGetLastError());
    }

    int nin_, w_, nout_;

private:
    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        MLFMA_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        allocs_.push_back(p);
        return static_cast<T*>(p);
    }
    void ensure(int B) {
        if (B <= cap_) return;
        for (void* a : bufs_) {
            auto it = std::find(allocs_.begin(), allocs_.end(), a);
            if (it != allocs_.end()) allocs_.erase(it);
            hipFree(a);
        }
        bufs_.clear();
        auto take = [&](size_t n) {
            float* p = alloc<float>(n);
            bufs_.push_back(p);
            return p;
        };
        d_X_ = take((size_t)B * nin_);
        d_Z1_ = take((size_t)B * w_);
        d_H1_ = take((size_t)B * w_);
        d_Z2_ = take((size_t)B * w_);
        d_H2_ = take((size_t)B * w_);
        d_Z3_ = take((size_t)B * nout_);
        d_Y_ = take((size_t)B * nout_);
        d_dY_ = take((size_t)B * nout_);
        d_dH2_ = take((size_t)B * w_);
        d_dZ2_ = take((size_t)B * w_);
        d_dH1_ = take((size_t)B * w_);
        d_dZ1_ = take((size_t)B * w_);
        d_loss_ = take((size_t)B);
        const int m = nout_ / 2;
        double2* g = alloc<double2>((size_t)B * m * m);
        double2* hh = alloc<double2>((size_t)B * m);
        double* cc = alloc<double>((size_t)B);
        bufs_.push_back(g);
        bufs_.push_back(hh);
        bufs_.push_back(cc);
        d_G_ = g;
        d_h_ = hh;
        d_c_ = cc;
        cap_ = B;
    }
    void stage_inputs(int B, const float* X) {
        if (B < 1) throw Error(MLFMA_ERR_ARG, "empty batch");
        ensure(B);
        std::vector<float> xs((size_t)B * nin_);
        for (int s = 0; s < B; ++s)
            for (int i = 0; i < nin_; ++i) xs[(size_t)s * nin_ + i] = (X[(size_t)s * nin_ + i] - mean_[i]) / stdv_[i];
        MLFMA_HIP(hipMemcpy(d_X_, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    void stage_targets(int B, const double* G, const double* h, const double* c) {
        if (nout_ % 2) throw Error(MLFMA_ERR_ARG, "n_out must be even (re, im)");
        const int m = nout_ / 2;
        if (m > 16) throw Error(MLFMA_ERR_ARG, "at most 16 coefficients");
        MLFMA_HIP(hipMemcpy(d_G_, G, (size_t)B * m * m * sizeof(double2), hipMemcpyHostToDevice));
        MLFMA_HIP(hipMemcpy(d_h_, h, (size_t)B * m * sizeof(double2), hipMemcpyHostToDevice));
        MLFMA_HIP(hipMemcpy(d_c_, c, (size_t)B * sizeof(double), hipMemcpyHostToDevice));
    }
    float* P(int k) { return d_p_ + off_[k]; }
    float* Gr(int k) { return d_g_ + off_[k]; }
    void forward(int B) {
        launch_dense_fwd(st_, B, nin_, w_, d_X_, P(0), P(1), d_Z1_, d_H1_, 1);
        launch_dense_fwd(st_, B, w_, w_, d_H1_, P(2), P(3), d_Z2_, d_H2_, 1);
        launch_dense_fwd(st_, B, w_, nout_, d_H2_, P(4), P(5), d_Z3_, d_Y_, 0);
    }
    double forward_backward(int B) {
        forward(B);
        const int m = nout_ / 2;
        launch_quad_loss(st_, B, m, d_Y_, d_G_, d_h_, d_c_, d_loss_, d_dY_);
        launch_dense_bwd_weights(st_, B, w_, nout_, d_dY_, d_H2_, Gr(4), Gr(5));
        launch_dense_bwd_input(st_, B, w_, nout_, d_dY_, P(4), d_dH2_);
        launch_act_bwd(st_, B * w_, d_H2_, d_dH2_, d_dZ2_, 1);
        launch_dense_bwd_weights(st_, B, w_, w_, d_dZ2_, d_H1_, Gr(2), Gr(3));
        launch_dense_bwd_input(st_, B, w_, w_, d_dZ2_, P(2), d_dH1_);
        launch_act_bwd(st_, B * w_, d_H1_, d_dH1_, d_dZ1_, 1);
        launch_dense_bwd_weights(st_, B, nin_, w_, d_dZ1_, d_X_, Gr(0), Gr(1));
        std::vector<float> l(B);
        MLFMA_HIP(hipMemcpyAsync(l.data(), d_loss_, B * sizeof(float), hipMemcpyDeviceToHost, st_));
        MLFMA_HIP(hipStreamSynchronize(st_));
        MLFMA_HIP(hipGetLastError());
        double s = 0;
        for (float v : l) s += v;
        return s / B;
    }

    hipStream_t st_ = nullptr;
    size_t off_[6], np_ = 0;
    float *d_p_ = nullptr, *d_g_ = nullptr, *d_m1_ = nullptr, *d_m2_ = nullptr;
    float *d_X_ = nullptr, *d_Z1_ = nullptr, *d_H1_ = nullptr, *d_Z2_ = nullptr, *d_H2_ = nullptr, *d_Z3_ = nullptr,
          *d_Y_ = nullptr;
    float *d_dY_ = nullptr, *d_dH2_ = nullptr, *d_dZ2_ = nullptr, *d_dH1_ = nullptr, *d_dZ1_ = nullptr,
          *d_loss_ = nullptr;
    double2* d_G_ = nullptr;
    double2* d_h_ = nullptr;
    double* d_c_ = nullptr;
    std::vector<float> mean_, stdv_;
    std::vector<void*> allocs_, bufs_;
    int cap_ = 0, step_ = 0;
};

}  // namespace mlfma

struct mlfma_mlp {
    std::unique_ptr<mlfma::Mlp> net;
    std::string err;
};

namespace {
template <typename F>
int mlp_guarded(mlfma_mlp* m, F&& f) {
    try {
        f();
        return MLFMA_OK;
    } catch (const mlfma::Error& e) {
        if (m) m->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (m) m->err = "host allocation failed";
        return MLFMA_ERR_OOM;
    } catch (const std::exception& e) {
        if (m) m->err = e.what();
        return MLFMA_ERR_HIP;
    }
}
std::string g_mlp_create_error;
}  // namespace

extern "C" {

int mlfma_mlp_create(int n_in, int width, int n_out, unsigned seed, mlfma_mlp** out) {
    if (!out || n_in < 1 || width < 1 || n_out < 2 || n_out > 32 || (n_out & 1)) {
        g_mlp_create_error = "bad network shape";
        return MLFMA_ERR_ARG;
    }
    *out = nullptr;
    auto m = std::make_unique<mlfma_mlp>();
    const int rc = mlp_guarded(m.get(), [&] { m->net = std::make_unique<mlfma::Mlp>(n_in, width, n_out, seed); });
    if (rc != MLFMA_OK) {
        g_mlp_create_error = m->err;
        return rc;
    }
    *out = m.release();
    return MLFMA_OK;
}

void mlfma_mlp_destroy(mlfma_mlp* m) { delete m; }

const char* mlfma_mlp_last_error(const mlfma_mlp* m) { return m ? m->err.c_str() : g_mlp_create_error.c_str(); }

int mlfma_mlp_num_params(const mlfma_mlp* m, long long* count) {
    if (!m || !count) return MLFMA_ERR_ARG;
    *count = (long long)m->net->nparams();
    return MLFMA_OK;
}

int mlfma_mlp_get_params(mlfma_mlp* m, float* p) {
    if (!m || !p) return MLFMA_ERR_ARG;
    return mlp_guarded(m, [&] { m->net->get_params(p); });
}

int mlfma_mlp_set_params(mlfma_mlp* m, const float* p) {
    if (!m || !p) return MLFMA_ERR_ARG;
    return mlp_guarded(m, [&] { m->net->set_params(p); });
}

int mlfma_mlp_set_normalization(mlfma_mlp* m, const float* mean, const float* stdv) {
    if (!m || !mean || !stdv) return MLFMA_ERR_ARG;
    return mlp_guarded(m, [&] { m->net->set_normalization(mean, stdv); });
}

int mlfma_mlp_predict(mlfma_mlp* m, int batch, const float* x, float* y) {
    if (!m || batch < 1 || !x || !y) return MLFMA_ERR_ARG;
    return mlp_guarded(m, [&] { m->net->predict(batch, x, y); });
}

int mlfma_mlp_loss_grad(mlfma_mlp* m, int batch, const float* x, const double* G, const double* h, const double* c,
                        double* loss, float* grad) {
    if (!m || batch < 1 || !x || !G || !h || !c) return MLFMA_ERR_ARG;
    return mlp_guarded(m, [&] {
        const double l = m->net->loss_grad(batch, x, G, h, c, grad);
        if (loss) *loss = l;
    });
}

int mlfma_mlp_train(mlfma_mlp* m, int batch, const float* x, const double* G, const double* h, const double* c,
                    int steps, float lr, float* loss_history) {
    if (!m || batch < 1 || !x || !G || !h || !c || steps < 0 || !(lr > 0)) return MLFMA_ERR_ARG;
    return mlp_guarded(m, [&] { m->net->train(batch, x, G, h, c, steps, lr, loss_history); });
}

}  // extern "C"
