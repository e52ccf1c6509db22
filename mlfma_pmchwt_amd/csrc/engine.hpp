// Engine: one MLFMA operator resident on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mlfma.h"
#include "geo.hpp"
#include "kernels.hpp"
#include "model.hpp"

namespace mlfma {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MLFMA_HIP(call)                                                                                     \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            throw ::mlfma::Error(e_ == hipErrorOutOfMemory ? MLFMA_ERR_OOM : MLFMA_ERR_HIP,                  \
                                 std::string(#call) + ": " + hipGetErrorString(e_));                        \
    } while (0)

struct LevelDev {
    int n = 0;
    int* dr_ptr = nullptr;
    int* dr_idx = nullptr;
    int* dr_vec = nullptr;
    int* son_ptr = nullptr;
    int* son_idx = nullptr;
    int* parent = nullptr;
    int* oct = nullptr;   // Morton octant of each node inside its parent
    // sibling clusters for the translation: up to 4 children of one parent,
    // the union of their distant relatives (ascending), and per union entry the
    // translation vector of each member (-1: not its distant relative)
    int ncl = 0;
    int* cl_child = nullptr;   // [ncl][4], -1 for an empty slot
    int* cl_ptr = nullptr;     // [ncl + 1] into cl_src
    int* cl_src = nullptr;
    int* cl_vec = nullptr;     // [entries][4]
    long long cl_pairs = 0, cl_entries = 0;   // sum of DR list lengths, sum of union sizes
    // the same with every child of a parent in one cluster (up to 8 slots)
    int ncl8 = 0;
    int* cl8_child = nullptr;   // [ncl8][8]
    int* cl8_ptr = nullptr;
    int* cl8_src = nullptr;
    int* cl8_vec = nullptr;     // [entries][8]
};

// Far-field tables and workspaces of one wavenumber in real type R.
template <typename R>
struct FarDev {
    cplx_t<R>* Vs = nullptr;                 // [N][3][K0], internal RWG order
    cplx_t<R>* Vf = nullptr;
    R* kp0 = nullptr;                    // [3][K0] finest-level directions
    std::vector<cplx_t<R>*> Sm, Bm, TFw;     // per tree level
    std::vector<R*> tw, t_w;             // per child tree level
    std::vector<R*> th_w, ph_w;          // per child tree level: the stencil's weights by direction
    std::vector<R*> at_w, ap_w;          // per child tree level: the transposed stencils' weights
    std::vector<cplx_t<R>*> sh_up, sh_dn;
};

struct WaveDev {
    std::vector<int> K;                  // per tree level
    std::vector<int*> tap, t_ptr, t_kp;  // per child tree level
    std::vector<int*> th_tap, ph_tap;    // per child tree level: the stencil's taps by direction
    std::vector<int*> at_ptr, at_idx, ap_ptr, ap_idx;   // per child tree level: transposed, by direction
    std::vector<int> Lp, Lc;             // per child tree level: theta rows of the parent and child grids
    FarDev<double> f64;                  // always built (true residuals, stage API)
    FarDev<float> f32;                   // built when far_precision = MLFMA_FP32
    float2* E = nullptr;                 // batched matrix-core disaggregation: effective receive-side matrices,
    double2* E64 = nullptr;              // FP32 and FP64 far field
};

// Host state of one batched GMRES (lockstep over the R right-hand sides of
// one operator), advanced phase by phase by Engine::solve_block and by the
// lockstep multi-object solve (solve_multi).
struct BlockState {
    mlfma_solve_params sp{};
    const double2* B = nullptr;
    double2* X = nullptr;
    int R = 0, n = 0, m = 1;
    size_t nR = 0;
    int pkind = 0;        // 0 none, 1 block-diagonal, 2 polynomial, 3 inner near-field solve
    bool pc = false;
    bool vf = false;      // Krylov bases in FP32 (FP32-mode solves)
    int matvecs = 0, fp64_matvecs = 0, total = 0, cyc = 0, k = 0, fallbacks = 0;
    bool stopped = false;   // the divergence guard fired, or the block was handed on narrower
    double t0 = 0;
    std::vector<int> iters, kused;
    std::vector<double> bnorm, rnorm, err, err_cycle;
    std::vector<char> done, live;
    std::vector<cplx> H, cs, sn, sv;
};

// A block narrowed to its unconverged columns (the compaction of the batched
// solve, as a child state instead of a nested call): the columns' indices and
// their own right-hand-side and solution blocks.
struct NarrowRecord {
    std::vector<int> idx;
    int* d_idx = nullptr;
    double2* Bc = nullptr;
    double2* Xc = nullptr;
};

class Engine {
public:
    explicit Engine(const mlfma_params& p);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void build(int np, const int* ids, const double* xyz, int nf, const int* fids, const int* fvid);

    // y = A x for nrhs device vectors (column-major, n_unknowns each), far
    // field in the apply precision
    void apply_device(const double2* x, double2* y, int nrhs);
    void set_apply_precision(int prec);
    int apply_prec = MLFMA_FP64;
    void stage(int which, const double2* dx, double2* dy);
    void solve(const mlfma_solve_params& sp, const double2* d_b, double2* d_x, mlfma_solve_stats& st);
    // R right-hand sides at once, RHS-fastest device blocks; stats[R]
    void solve_block(const mlfma_solve_params& sp, const double2* d_B, double2* d_X, int R, mlfma_solve_stats* stats);
    // same with column-major blocks (vector j at b + j * n_unknowns)
    void solve_device(const mlfma_solve_params& sp, const double2* d_b, double2* d_x, int R,
                      mlfma_solve_stats* stats);
    void ensure_block(int R);
    // phases of the batched GMRES (the caller synchronises st between the
    // orthogonalisation phases 0 and 1)
    void bs_start(BlockState& s, const mlfma_solve_params& sp, const double2* d_B, double2* d_X, int R);
    bool bs_active(const BlockState& s) const;
    void bs_cycle_begin(BlockState& s);
    bool bs_step_open(const BlockState& s) const;
    int bs_graph_kind(const BlockState& s) const;
    // FP32 bases with the polynomial preconditioner on the BF16 records: the
    // polynomial runs in FP32 vectors (poly_apply_bf16f) and the matvec reads
    // z_k in FP32
    bool poly_f32(const BlockState& s) const { return s.vf && s.pc && s.pkind == 2 && d_Zbf_ != nullptr; }
    void bs_pre(BlockState& s, bool recorded);
    void bs_body(BlockState& s);
    void bs_post(BlockState& s, bool recorded);
    void bs_orth(BlockState& s, int phase);
    void bs_cycle_end(BlockState& s, bool compaction);
    void bs_stats(const BlockState& s, mlfma_solve_stats* stats) const;
    // compaction without nesting: start `child` on the unconverged columns of a
    // block whose cycle ended with a quarter or more converged (s stops), and
    // later fold the child's solutions and counts back into s
    bool bs_narrow(BlockState& s, BlockState& child, NarrowRecord& rec);
    void bs_unnarrow(BlockState& s, const BlockState& child, NarrowRecord& rec);
    // column-major device blocks (vector j at b + j * n_unknowns) to and from
    // the RHS-fastest blocks the batched solver works on
    void block_load(const double2* d_b, const double2* d_x, int R);
    void block_store(double2* d_x, int R);
    const double2* block_b() const { return d_bx_; }
    double2* block_x() { return d_by_; }
    // a capture of several operators' matvecs into one graph (solve_multi)
    void set_capturing(bool on) { capturing_ = on; }
    bool graphs_on() const { return use_graphs_ && !prof_on_; }
    // hipGraph replay of batched matvecs; launch-configuration autotuner
    void set_graphs(bool on);
    void autotune(int R, int prec, double* out);
    void set_tuning(int prec, int rb, int bits);
    // polynomial near-field preconditioner (precond = 2)
    void set_inner_steps(int k);
    void set_poly(const std::vector<cplx>& alpha);
    void poly_apply(const double2* X, double2* Y, int R, int prec);
    void poly_apply_bf16f(const float2* X, float2* Y, int R);
    void precond_features(float* f);
    void poly_apply_device(const double2* x, double2* y, int nrhs);
    void precond_apply_kind(int kind, const double2* x, double2* y);
    void precond_statistics(const double2* d_b, int steps, int degree, double2* G, double2* h, double* c);
    void excitation(double pol, double th, double ph, double2* d_v);
    void rcs(const double2* d_I, int nang, const double* th, const double* ph, double pol, double* out);
    // RCS of device currents at device angles into device output, allocating nothing
    void rcs_noalloc(const double2* d_I, int nang, const double* d_th, const double* d_ph, double* d_out);
    void rcs_multi(const double2* d_I, size_t istride, int nang, const double* th, const double* ph, double pol,
                   double* out);
    void excitation_batch(int n_inc, const double* pol, const double* th, const double* ph, double2* d_V);
    void monostatic(const mlfma_solve_params& sp, double pol, int nang, const double* th, const double* ph,
                    double* out, mlfma_solve_stats* stats);

    // per-group kernel timing with hipEvents on st
    void set_profiling(bool on);
    double prof_ms[5] = {0, 0, 0, 0, 0};
    long long prof_launch[5] = {0, 0, 0, 0, 0};
    int prof_apply = 0;

    // inspection
    void get_workspace(int wn, int cur, int which, int level, int node, double* out);
    void get_pattern(int wn, int which, int rwg, double* out);
    void get_near_row(int row, int* len, int* cols, double* vals);

    mlfma_params P;
    int ie = 2, nw = 2, ncur = 2;
    int N = 0, nunk = 0, nleaf = 0, max_leaf = 0;
    Mesh mesh;
    Basis basis;
    Octree tree;
    Medium med;
    WaveTables W[2];
    std::vector<int> inv_perm, leaf_of;
    hipStream_t st = nullptr;
    size_t dev_bytes = 0;
    long long near_entries = 0;
    std::string last_error;

    // near-field structure (host copies kept for inspection): pairs of leaves,
    // and per target leaf its block row (layout of NearDev, geo.hpp)
    std::vector<int> pair_ptr, pair_tgt, pair_src, pair_col;
    std::vector<long long> tgt_off;
    std::vector<int> tgt_cols;

private:
    template <typename T>
    T* dalloc(size_t count);
    template <typename T>
    T* upload(const std::vector<T>& v);
    void dfree(void* p);
    struct GraphKey {
        int kind, R, prec;
        bool operator<(const GraphKey& o) const {
            return kind != o.kind ? kind < o.kind : (R != o.R ? R < o.R : prec < o.prec);
        }
    };
    void run_graph(const GraphKey& key, const std::function<void()>& body);
    void clear_graphs();
    std::map<GraphKey, hipGraphExec_t> graphs_;
    bool use_graphs_ = false;
    int rb_tuned_[2] = {0, 0};   // tuned right-hand sides per far pass (FP64, FP32); 0: widest
    void build_device_tree();
    void build_wave(int w);
    void build_near();
    template <typename R>
    FarDev<R>& far(int w);
    template <typename R>
    void build_far(int w, const WaveTables& T);
    template <typename R>
    int rb_max() const;
    template <typename R>
    void far_pass(int w, const double2* x, int ldr, int r0, int rb);
    template <typename R>
    void disaggregate(double2* y, int ldr, int r0, int rb);
    template <typename R>
    void translate(int w, int l, int nc);
    template <typename R>
    void interpolate(int w, int l, int nc);
    template <typename R>
    cplx_t<R>* beff_buf(int w);
    void apply_pass(const double2* x, double2* y, int ldr, int r0, int rb, int prec, bool with_near);
    void mark(int slot);
    void collect_profile(int count);
    void apply_one(const double2* x, double2* y, int prec);
    void apply_block(const double2* X, double2* Y, int R, int prec, const float2* Xf = nullptr);
    void apply_block_f(const float2* X, double2* Y, int R, int prec, double2* scratch);
    bool fp32_input_ok(int R, int prec) const;
    std::vector<double2> near_block(int p);   // planes x (ns x nt column-major) of pair p
    void precond_setup();
    void precond_apply(const double2* x, double2* y);
    void precond_block(int kind, const double2* x, double2* y, int R, int prec);
    void inner_near_solve(const double2* V, double2* Z, int R, int prec);
    void ensure_inner(int R);
    void ensure_poly(int R);
    void ensure_xf(int R);
    void ensure_eall(int R);
    void near_block(const double2* X, double2* Y, int R, int prec, bool precond = false);
    double norm(const double2* v, int n);

    std::unordered_map<void*, size_t> allocs_;   // device allocation -> bytes
    bool capturing_ = false;                      // inside run_graph's stream capture
    std::vector<LevelDev> lv_;
    WaveDev wd_[2];
    int* d_perm_ = nullptr;
    int* d_leaf_ptr_ = nullptr;
    int* d_leaf_of_ = nullptr;
    int* d_leaf_order_ = nullptr;   // leaves by decreasing size (aggregation launch order)
    double* d_leaf_c_ = nullptr;
    RwgGeo* d_geo_ = nullptr;
    GaussRule gauss_{};
    NearDev nd_{};
    double2* d_Z_ = nullptr;
    float2* d_Z32_ = nullptr;   // FP32 copy (near_precision = MLFMA_FP32)
    uint4* d_Zbs_ = nullptr;    // split-pair BF16 records of the batched products (near_precision = MLFMA_BF16)
    long long* d_bsoff_ = nullptr;
    uint4* d_Zbf_ = nullptr;    // BF16 record copy for the preconditioner (precond_precision = MLFMA_BF16)
    long long* d_bfoff_ = nullptr;
    float2* d_Eall_[2] = {nullptr, nullptr};   // all-RHS effective matrices (FP32), nleaf x no x 3K x eall_cap_
    unsigned short* d_Vb_[2] = {nullptr, nullptr};   // split-pair BF16 receive-pattern planes (near_precision = MLFMA_BF16)
    unsigned short* d_Va_[2] = {nullptr, nullptr};   // split-pair BF16 radiation-pattern planes, transposed [4][Qa][Np]
    int Qa_[2] = {0, 0};                             // their sample rows (3 K rounded up to 64)
    int Np_ = 0;                                     // their RWG columns (every leaf rounded up to 32)
    int* d_joff_ = nullptr;                          // first column of each leaf
    int* d_jmap_ = nullptr;                          // RWG of each column, -1 for padding
    float2* d_Sall_[2] = {nullptr, nullptr};   // all-RHS leaf radiation workspaces of the passes, nleaf x ncur x 3K x eall_cap_
    int eall_cap_ = 0;
    int eall_R_ = 0;   // > 0 while a batched FP32 matvec defers its disaggregation to one all-RHS product
    float2* d_xf_ = nullptr;    // FP32 copy of the matrix-core near product's input, nunk x xf_cap_
    int xf_cap_ = 0;
    unsigned* d_xb_ = nullptr;  // BF16 copy of the preconditioner's near-product input, nunk x poly_cap_
    // RHS-fastest block vectors of the batched matvec
    double2* d_bx_ = nullptr;
    double2* d_by_ = nullptr;
    int block_cap_ = 0;
    // solver scratch
    double2* d_V_ = nullptr;
    int V_cols_ = 0;
    double2* d_w_ = nullptr;
    double2* d_r_ = nullptr;
    double2* d_t_ = nullptr;
    double2* d_part_ = nullptr;
    double2* d_small_ = nullptr;
    double* d_rpart_ = nullptr;
    // batched solver scratch
    double2* d_bV_ = nullptr;
    double2* d_bW_ = nullptr;
    double2* d_bRr_ = nullptr;
    double2* d_bpart_ = nullptr;
    double2* d_bsmall_ = nullptr;
    double* d_bscale_ = nullptr;
    double2* d_bZ_ = nullptr;
    double2* d_bG_ = nullptr;   // fixed input of recorded Arnoldi matvecs
    int bV_cols_ = 0, bV_R_ = 0;
    int stats_fallbacks_ = 0;
    // pinned host staging of the batched solver's reductions and coefficients
    double2* h_get_ = nullptr;
    double2* h_put_ = nullptr;
    size_t pin_cap_ = 0;
    bool put_pending_ = false;   // a host -> device copy from h_put_ may not have run yet
    void ensure_pinned(size_t count);
    double2* put_buffer();
    void bs_put(const std::vector<cplx>& v, double2* dst);
    void bs_put_scale(const std::vector<double>& a);
    void bs_issue_norms(BlockState& s, const double2* v);
    void bs_read_norms(const BlockState& s, std::vector<double>& out) const;
    void bs_true_residual(BlockState& s, std::vector<double>& rn);
    // profiling events: per wave 3 groups x 2, near 2, far 2
    bool prof_on_ = false;
    hipEvent_t ev_[18] = {};   // 16..17: the all-RHS near product of a batched matvec
    // block-diagonal preconditioner
    double2* d_dinv_ = nullptr;   // 1 / diagonal of the operator, caller order
    float2* d_dinvf_ = nullptr;   // its FP32 copy (with the BF16 preconditioner records)
    std::vector<cplx> poly_alpha_;
    double2* d_Zf_ = nullptr;    // flexible GMRES: preconditioned basis z_k = M^-1 v_k (n x Zf_cols_)
    int Zf_cols_ = 0;
    double2* d_bZf_ = nullptr;   // same for the batched solver (n R x bZf_cols_)
    int bZf_cols_ = 0;
    size_t bZf_nR_ = 0;
    float2* d_bVf_ = nullptr;    // FP32 Krylov bases of FP32-mode batched solves
    float2* d_bZff_ = nullptr;
    int bVf_cols_ = 0, bZff_cols_ = 0;
    size_t bVf_nR_ = 0, bZff_nR_ = 0;
    double2* d_pt_ = nullptr;
    double2* d_pu_ = nullptr;
    int poly_cap_ = 0;
    int precond_kind_ = 0;
    // inner near-field solve (preconditioner 3): Krylov basis and reduction scratch
    int inner_steps_ = 10;
    double2* d_iU_ = nullptr;
    double2* d_iW_ = nullptr;
    double2* d_ipart_ = nullptr;
    double2* d_ismall_ = nullptr;
    double* d_iscale_ = nullptr;
    int inner_cap_R_ = 0, inner_cap_k_ = 0;
    bool near_mfma_ = true;     // batched near product on the matrix cores (FP32 copy or FP64 entries)
    bool disagg_mfma_ = true;   // batched disaggregation on the matrix cores
    bool m2l_cluster_ = true;   // translation by sibling clusters
    bool m2m_separable_ = true;  // interpolation: separable (theta, then phi) form where it applies
    bool l2l_separable_ = true;  // anterpolation: separable pass after the translation where it applies
    int m2l_variant_ = 0;       // FP32 cluster translation: 1 twelve columns per thread, 2 all-sibling clusters, LDS-staged samples
    double2* d_Dinv_ = nullptr;
    long long* d_Doff_ = nullptr;
};

// Batched GMRES of several operators in lockstep, one recorded graph per
// Arnoldi step over all their matvecs (multi.cpp).  B[k], X[k]: RHS-fastest
// device blocks of R[k] columns (block_b / block_x after block_load); stats[k]
// has R[k] entries.
void solve_multi(const std::vector<Engine*>& es, const mlfma_solve_params& sp, const std::vector<const double2*>& B,
                 const std::vector<double2*>& X, const std::vector<int>& R,
                 const std::vector<mlfma_solve_stats*>& stats);

}  // namespace mlfma
