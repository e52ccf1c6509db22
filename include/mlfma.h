/*
 * mlfma.h -- C ABI of the MI355X MLFMA solver engine (libmlfma.so).
 *
 * Plain C types only: pointers, sizes, doubles.  Complex vectors are
 * interleaved (re, im) doubles.  A handle owns every device buffer of one
 * operator (mesh + octree + translation/interpolation tables + radiation
 * patterns + near-field blocks) and one HIP stream; a handle is not
 * thread-safe, distinct handles are independent.
 *
 * Unknown ordering is the reference's: RWG functions numbered by their
 * sorted (min vertex id, max vertex id) edge key (Code/RWGGenerator.cpp:39-113);
 * PMCHWT vectors are x[0:N) = eta0*J and x[N:2N) = M (Code/MLFMM.cpp:238-245).
 *
 * Every entry point returns MLFMA_OK (0) or an error code; the message of the
 * last failure is available from mlfma_last_error().
 */
#ifndef MLFMA_H
#define MLFMA_H

#ifdef __cplusplus
extern "C" {
#endif

#define MLFMA_OK 0
#define MLFMA_ERR_ARG 1          /* invalid argument (reference: std::invalid_argument) */
#define MLFMA_ERR_HIP 2          /* HIP runtime failure */
#define MLFMA_ERR_OOM 3          /* device or host allocation failed */
#define MLFMA_ERR_NONFINITE 4    /* NaN/Inf met in a residual */
#define MLFMA_ERR_NOT_CONVERGED 5
#define MLFMA_ERR_MESH 6         /* unreadable / degenerate mesh */
#define MLFMA_ERR_STATE 7        /* call not valid for this handle */

typedef struct mlfma_handle mlfma_handle;

/* Integral equation ids are the reference's selectIntegralEqu values
 * (Code/RCSExportConfig.h:17). */
#define MLFMA_EFIE 0
#define MLFMA_CFIE 1
#define MLFMA_PMCHWT 2

/* Arithmetic of the far-field stages (aggregation, interpolation, translation,
 * anterpolation, disaggregation): storage and accumulation.  Near field,
 * translation-operator generation, patterns (computed in FP64, then rounded)
 * and Krylov steps are FP64 in both modes.  With MLFMA_FP32 the FP64 far
 * field is built as well: it gives the true residual of every solve
 * (restarted GMRES then refines an FP32-operator solution against the FP64
 * operator). */
#define MLFMA_FP64 0
#define MLFMA_FP32 1
#define MLFMA_BF16 2

/* mlfma_params.tf_mode bits.  MLFMA_TF_COMPLEX_ARG is the accuracy mode for lossy media: the
 * spherical Hankel functions of the translation series are evaluated at the complex kR, as the
 * Green's function exp(-jkR)/R of a lossy medium requires; the reference's real |kR| drops the
 * medium's decay between distant boxes (25% matvec error on a 6-level lossy PMCHWT tree, 1.3% in
 * this mode).  MLFMA_TF_FULL_ORDER keeps all L terms (unstable where L > |kR|; study only). */
#define MLFMA_TF_FULL_ORDER 1
#define MLFMA_TF_COMPLEX_ARG 2

typedef struct mlfma_params {
    int integral_equation;  /* MLFMA_EFIE / MLFMA_CFIE / MLFMA_PMCHWT */
    int gauss_points;       /* Dunavant rule: 1,3,4,6,7,12,13,16 (Code/GaussPoints.cpp:286-301) */
    double frequency_hz;    /* Code/EMSource.cpp:6-14 */
    double eps_r_re, eps_r_im;   /* interior medium, eps' - j eps'' (Code/EMSource.cpp:19-32) */
    double mu_r_re, mu_r_im;
    double cfie_alpha;      /* reference compiles 0.2 (Code/PhyConstant.h:15) */
    double leaf_coeff;      /* finest box = leaf_coeff * lambda0; reference 0.3 (Code/PhyConstant.h:14) */
    int far_precision;      /* MLFMA_FP64 or MLFMA_FP32 */
    int near_precision;     /* storage of the near-field entries used by FP32-mode matvecs: MLFMA_FP64,
                               MLFMA_FP32 (entries computed in FP64, rounded once; the FP64 copy is
                               kept for true residuals) or MLFMA_BF16 (FP32 storage for single-vector
                               products; products of 2 or more right-hand sides on the BF16 matrix
                               cores from split pairs hi + lo of entries and vectors, three partial
                               products, FP32 accumulation: about 2^-17 relative, within the FP64
                               refinement of FP32-mode solves; with far_precision = MLFMA_FP32 the
                               batched disaggregation runs on split pairs of the patterns too) */
    int tf_mode;            /* translation operators, a bit set of MLFMA_TF_*: 0 = the reference's series
                               (parity mode, Code/AlgoMLFMM.h:262-265): length min(L, ceil|kR|), Hankel
                               functions at the real |kR| even for a lossy medium */
    int max_rhs;            /* largest batch a single apply/solve will be given */
    int device;             /* HIP device ordinal */
    int verbose;            /* 1: setup timings on stderr */
    int precond_precision;  /* near-field products of the polynomial preconditioner in FP32-mode
                               solves: MLFMA_FP64 (same storage as the matvec's near field) or
                               MLFMA_BF16 (a BF16 copy built at setup, matrix-core products with
                               FP32 accumulation; residuals and the operator are unaffected) */
} mlfma_params;

/* Reference defaults: PMCHWT, 7 points, 299792458 Hz (lambda = 1 m), eps_r 1,
 * alpha 0.2, leaf 0.3, FP64, reference TF truncation, 1 RHS, device 0. */
void mlfma_params_default(mlfma_params* p);

/* Builds the whole operator on the device.  Replaces the setup half of
 * MLFMM::MLFMM (Code/MLFMM.cpp:15-104): readNasData + GenerateRWG + OCTree +
 * AlgoMLFMM::computeBase_MLFMM / computeTransferFactorMatrix /
 * computeInterpolationMatrix + ProcessFMM::computeVsmi_Vfmj_* +
 * Zmartix::computeZ_fmm_* .  face_vertex_ids holds 3 point ids per face,
 * in file order (Code/NASReader.cpp:98-119). */
int mlfma_create(const mlfma_params* p, int n_points, const int* point_ids, const double* xyz,
                 int n_faces, const int* face_ids, const int* face_vertex_ids, mlfma_handle** out);

/* Same, reading the reference's NAS format (Code/NASReader.cpp:60-144). */
int mlfma_create_nas(const mlfma_params* p, const char* nas_path, mlfma_handle** out);

void mlfma_destroy(mlfma_handle* h);

/* Message of the last failure on this handle (h may be NULL for create errors). */
const char* mlfma_last_error(const mlfma_handle* h);

/* n_rwg (N), n_unknowns (N or 2N), max tree level, leaves, near-field complex
 * entries stored, device bytes held. */
int mlfma_dims(const mlfma_handle* h, long long* out6);

/* y = A x on host vectors of length n (= n_unknowns).  Replaces the ax(n, x, w)
 * lambda of MLFMM::matrix_solver (Code/MLFMM.cpp:189-369) and the proposed
 * MLFMM::applyAx (README/矩阵向量乘法导出头文件设计.md:392-437): y is
 * overwritten; n must match. */
int mlfma_apply(mlfma_handle* h, int n, const double* x, double* y);

/* nrhs vectors at once, column-major (vector j at x + 2*n*j). */
int mlfma_apply_batch(mlfma_handle* h, int n, int nrhs, const double* x, double* y);

/* Device-resident form: x_dev / y_dev are device pointers (interleaved complex
 * double, column-major n x nrhs), work is queued on the handle's stream and
 * the call returns after it completes. */
int mlfma_apply_device(mlfma_handle* h, int n, int nrhs, const void* x_dev, void* y_dev);

/* Far-field precision used by mlfma_apply* and by the Krylov iterations of
 * mlfma_solve: MLFMA_FP64 always, MLFMA_FP32 when the handle was created with
 * far_precision = MLFMA_FP32 (its default then).  No reference counterpart:
 * the reference's operator is FP64 only. */
int mlfma_set_apply_precision(mlfma_handle* h, int precision);

/* Persistent batched matvec: with on = 1, the kernel sequence of a batched
 * matvec (mlfma_apply_device / mlfma_apply_batch with nrhs > 1, and every
 * Arnoldi step of mlfma_solve_batch, preconditioner included) is recorded
 * once per (right-hand-side count, precision) as a hipGraph and replayed.
 * Changing coefficients, buffers or tuning drops the recordings.  No reference
 * counterpart. */
int mlfma_set_graphs(mlfma_handle* h, int on);

/* Launch-configuration autotuner: times the batched matvec of min(nrhs, 32)
 * right-hand sides in `precision` and picks, by coordinate descent from the
 * widest configuration, the right-hand sides per far-field pass and whether
 * the near product and the disaggregation run on the matrix cores or on the
 * vector units and the translation by sibling clusters or per box; keeps the
 * choice for this handle.  out4 = {rhs per pass, choice bits (bit 0 near on
 * the matrix cores, bit 1 disaggregation on the matrix cores, bit 2 cluster
 * translation, bit 3 cluster translation with 12 workspace columns per thread
 * (FP32), bit 4 cluster translation over every child of a parent (up to 8)
 * with the members' translation samples staged in LDS (FP32; takes precedence
 * over bit 3)), ms per batched matvec of the choice, ms of the slowest
 * configuration timed}.  No reference counterpart. */
int mlfma_autotune(mlfma_handle* h, int nrhs, int precision, double* out4);

/* Sets the launch configuration directly (e.g. one that mlfma_autotune chose
 * in an earlier process): rhs_per_pass in {1, 2, 4, 8} (clamped to what the
 * far-field precision allows; 0 keeps the default) for `precision`, and the
 * choice bits of mlfma_autotune's out4[1]: bit 0 near field on the matrix
 * cores, bit 1 disaggregation on the matrix cores, bit 2 translation by
 * sibling clusters (up to 4 children of one parent share each source box,
 * read once), bit 3 12 instead of 6 workspace columns per thread in the FP32
 * cluster translation, bit 4 FP32 cluster translation over every child of a
 * parent with LDS-staged translation samples (takes precedence over bit 3).
 * No reference counterpart. */
int mlfma_set_tuning(mlfma_handle* h, int precision, int rhs_per_pass, int bits);

/* Krylov solve on the device.  Replaces MGMRES::mgmres (Code/GMRES.h:65-217)
 * and MCGS::CGS::solve (Code/CGS.h:72-234) as driven by MLFMM::matrix_solver
 * (Code/MLFMM.cpp:370-397): x holds the initial guess on entry and the
 * solution on exit; returns MLFMA_ERR_NOT_CONVERGED (x still written) when
 * itr_max is exhausted. */
#define MLFMA_SOLVER_GMRES 0
#define MLFMA_SOLVER_CGS 1
typedef struct mlfma_solve_params {
    int solver;         /* MLFMA_SOLVER_GMRES / MLFMA_SOLVER_CGS (Code/MLFMM.cpp:372-395) */
    int itr_max;        /* total inner iterations */
    int restart;        /* GMRES(mr) */
    double tol_rel;     /* stop when ||r|| / ||b|| < tol_rel (or < tol_abs) */
    double tol_abs;
    int precond;        /* 0 none (reference), 1 near-field block-diagonal, 2 near-field polynomial
                           (coefficients from mlfma_set_precond_poly, e.g. predicted by the network),
                           3 inner near-field solve (mlfma_set_precond_inner steps of GMRES on the
                           stored near field per application); mlfma_solve_batch accepts all four;
                           a solve replaces a stalling preconditioner by the block-diagonal one,
                           and a stalling block-diagonal one by none */
} mlfma_solve_params;

typedef struct mlfma_solve_stats {
    int iterations;
    int matvecs;
    int converged;
    double rel_residual;      /* true ||b - A x|| / ||b|| with the FP64 operator */
    double seconds;
    int fp64_matvecs;         /* of matvecs, those with the FP64 operator (true residuals) */
    int precond_fallbacks;    /* 1 if a stalling polynomial preconditioner was dropped mid-solve */
} mlfma_solve_stats;

int mlfma_solve(mlfma_handle* h, const mlfma_solve_params* sp, int n, const double* rhs, double* x,
                mlfma_solve_stats* stats);

/* nrhs (<= 256) independent GMRES solves in lockstep, sharing one batched
 * matvec per Arnoldi step (each pass through the operator's tables serves up
 * to 8 right-hand sides).  rhs / x are column-major (vector j at 2*n*j
 * doubles); stats has nrhs entries; returns MLFMA_ERR_NOT_CONVERGED if any
 * right-hand side did not converge.  Replaces the sequential per-incidence
 * MLFMM::matrix_solver calls of the monostatic sweeps (Code/RCS.h:272-325,
 * 428-497) and of batch mode (Code/BatchRunner.h:422-484); precond may be 0, 1 or 2. */
int mlfma_solve_batch(mlfma_handle* h, const mlfma_solve_params* sp, int n, int nrhs, const double* rhs, double* x,
                      mlfma_solve_stats* stats);

/* Several objects solved together (no reference counterpart; replaces the
 * one-operator-per-task, one-task-after-another loop of
 * Code/BatchRunner.h:470-480).  Object k is handle hs[k] with nrhs[k] (1..256)
 * right-hand sides rhs[k] (n_k x nrhs[k] complex, column-major) and initial
 * guesses / solutions x[k]; stats[k] (may be NULL, as may stats) receives
 * nrhs[k] entries.  The objects' batched GMRES runs (as mlfma_solve_batch,
 * same parameters for all) advance in lockstep, and every Arnoldi step
 * launches the preconditioned matvecs of all objects still iterating as one
 * recorded graph with one branch per object, so objects of uneven size run
 * side by side from one launch.  Each object's iterates are those of
 * mlfma_solve_batch on it alone, except that a block whose columns have
 * partly converged keeps its width.  Every handle needs mlfma_set_graphs(h, 1)
 * and profiling off, all on one device, each handle at most once.  Errors are
 * reported through hs[0]. */
int mlfma_solve_multi(int count, mlfma_handle* const* hs, const mlfma_solve_params* sp, const int* nrhs,
                      const double* const* rhs, double* const* x, mlfma_solve_stats* const* stats);

/* ---- Learned near-field preconditioner (no reference counterpart; the
 * reference solves unpreconditioned, Code/MLFMM.cpp:370-397).
 *
 * precond = 2 applies M^-1 r = sum_k alpha_k t_k with t_0 = D^-1 r and
 * t_{k+1} = t_k - D^-1 Z_near t_k (D the operator's diagonal, Z_near the stored
 * near field): one near-field product per degree.  Its coefficients are
 * either fitted from the solver's own Krylov pairs (mlfma_precond_statistics)
 * or predicted for an unseen problem by the coefficient network
 * (mlfma_mlp_*) from mlfma_precond_features.  A solve replaces a polynomial
 * preconditioner that stalls (true residual not 10% lower after a restart
 * cycle) by the block-diagonal one, and that, if it stalls too, by none. */
#define MLFMA_PRECOND_FEATURES 20

/* 20 problem features: spectral moments mu_k = z^H (D^-1 Z_near)^k z / n,
 * k = 1..8 (8 real parts, 8 imaginary parts) for a seeded unit-modulus probe
 * z, then eps_r (re, im), PMCHWT flag, log10(near-field entries per row of the full near matrix):
 * a discretisation-density feature, so smaller objects meshed as densely cover a large one. */
int mlfma_precond_features(mlfma_handle* h, float* features);

/* Least-squares statistics of the coefficients from `steps` Arnoldi steps of
 * the unpreconditioned operator started at rhs: pairs (r_k = A v_k, e_k = v_k),
 * basis a_k = [t_0(r_k) .. t_degree(r_k)]; G = sum a_k^H a_k ((degree+1)^2
 * complex, row-major), hv = sum a_k^H e_k, c = sum |e_k|^2. */
int mlfma_precond_statistics(mlfma_handle* h, const double* rhs, int steps, int degree, double* G, double* hv,
                             double* c);

/* Sets the degree + 1 complex coefficients (interleaved re, im). */
int mlfma_set_precond_poly(mlfma_handle* h, int degree, const double* alpha);

/* Steps (1..64, default 10) of the inner near-field solve of preconditioner 3: each
 * application runs that many GMRES steps on Z_near z = v from z = 0 (per right-hand side,
 * in lockstep for batched solves; near products in the preconditioner's storage, BF16 on the
 * matrix cores in FP32-mode solves with precond_precision = MLFMA_BF16).  No reference
 * counterpart; the outer GMRES is flexible, so the preconditioner may change between steps. */
int mlfma_set_precond_inner(mlfma_handle* h, int steps);

/* y = M^-1 x for kind 1 (block-diagonal) or 2 (polynomial), host vectors. */
int mlfma_precond_apply(mlfma_handle* h, int kind, int n, const double* x, double* y);

/* The polynomial preconditioner (precond = 2) on nrhs <= 256 column-major
 * vectors in one batched pass, in the apply precision (with
 * precond_precision = MLFMA_BF16 and FP32 apply precision, its near products
 * use the BF16 copy).  No reference counterpart. */
int mlfma_precond_apply_batch(mlfma_handle* h, int n, int nrhs, const double* x, double* y);

/* The coefficient network: n_in -> width -> width -> n_out, tanh hidden
 * layers, FP32, trained on the device.  Loss per sample s (features x_s,
 * statistics G_s, h_s, c_s of one problem): with a = net(x_s) read as
 * n_out / 2 complex coefficients (real parts first),
 *   l_s = (a^H G_s a - 2 Re a^H h_s + c_s) / c_s,   loss = mean_s l_s. */
typedef struct mlfma_mlp mlfma_mlp;
int mlfma_mlp_create(int n_in, int width, int n_out, unsigned seed, mlfma_mlp** out);
void mlfma_mlp_destroy(mlfma_mlp* m);
const char* mlfma_mlp_last_error(const mlfma_mlp* m);
/* parameter vector: W1 (width x n_in), b1, W2 (width x width), b2, W3 (n_out x width), b3 */
int mlfma_mlp_num_params(const mlfma_mlp* m, long long* count);
int mlfma_mlp_get_params(mlfma_mlp* m, float* params);
int mlfma_mlp_set_params(mlfma_mlp* m, const float* params);
/* inputs are standardised as (x - mean) / std before the first layer */
int mlfma_mlp_set_normalization(mlfma_mlp* m, const float* mean, const float* stdv);
int mlfma_mlp_predict(mlfma_mlp* m, int batch, const float* x, float* y);
/* loss and its gradient with respect to the parameter vector (grad may be NULL) */
int mlfma_mlp_loss_grad(mlfma_mlp* m, int batch, const float* x, const double* G, const double* h, const double* c,
                        double* loss, float* grad);
/* `steps` full-batch Adam steps (beta 0.9 / 0.999), loss per step in loss_history */
int mlfma_mlp_train(mlfma_mlp* m, int batch, const float* x, const double* G, const double* h, const double* c,
                    int steps, float lr, float* loss_history);

/* Plane-wave excitation for the handle's equation; replaces RHS::computeV_E,
 * computeV_C and computeV_PMCHWT (Code/Vm.h:10-203) with EMSource::initPW
 * (Code/EMSource.cpp:34-72).  pol 0 = "V" (E along -theta_hat), 90 = "H". */
int mlfma_excitation(mlfma_handle* h, double pol_deg, double theta_deg, double phi_deg, double* v);

/* RCS (dBsm) of solved currents I at n_angles observation angles:
 * out[3*i] = total, out[3*i+1] = V part, out[3*i+2] = H part.  Replaces
 * RCSUtils::calculateRCS / calculateRCS_PMCHWT and the dB formula of
 * compute{Dual,Mono}Static[_PMCHWT] (Code/RCS.h:54-192, :247-259). */
int mlfma_rcs(mlfma_handle* h, const double* I, int n_angles, const double* theta_deg, const double* phi_deg,
              double pol_deg, double* out);

/* Allocation-free variant of mlfma_rcs.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).  Every buffer is
 * the caller's device memory: I_dev the solved currents (n interleaved complex), theta_dev and
 * phi_dev the n_angles observation angles in degrees, out_dev 3 doubles per angle (total, V, H in
 * dBsm, as mlfma_rcs).  The directions and partial sums live in a file-scope device array of
 * 8 MiB per GPU, held until the process ends; a call needs 3*8 + 48*ceil(N/256) bytes per angle
 * (about 300 angles per call at 150k RWG functions) and returns MLFMA_ERR_ARG beyond that.  All
 * handles on one GPU share that array: calls that overlap on one GPU race.  The call completes
 * before it returns (the handle's stream is synchronised). */
int mlfma_rcs_device_noalloc(mlfma_handle* h, const void* I_dev, int n_angles, const double* theta_dev,
                             const double* phi_dev, double* out_dev);

/* n_inc plane-wave right-hand sides at once, column-major into v (column j at
 * v + 2*n*j, interleaved complex); pol/theta/phi per column, as
 * mlfma_excitation.  The batched form of RHS::computeV_* (Code/Vm.h:150-203)
 * for a monostatic or multi-incidence run. */
int mlfma_excitation_batch(mlfma_handle* h, int n_inc, const double* pol_deg, const double* theta_deg,
                           const double* phi_deg, double* v);

/* Monostatic sweep: for each of n_angles directions the plane wave incident
 * from it is solved (all directions together in lockstep batched GMRES,
 * chunks of at most 256; solver must be GMRES) and the backscatter RCS of its
 * solution taken at the same direction: out3[3*a + {0,1,2}] = total, V, H in
 * dBsm, stats[a] per direction (may be NULL).  Replaces
 * RCSUtils::computeMonoStatic / computeMonoStatic_PMCHWT (Code/RCS.h:272-325,
 * :428-497), which solve the directions one at a time.  Returns
 * MLFMA_ERR_NOT_CONVERGED if any direction did not converge (out3 written). */
int mlfma_monostatic(mlfma_handle* h, const mlfma_solve_params* sp, double pol_deg, int n_angles,
                     const double* theta_deg, const double* phi_deg, double* out3, mlfma_solve_stats* stats);

/* The reference's scattering-angle grid (Code/RCS.h:46-51, :239-246): a theta
 * sweep th_s + i*step at phi = ph_s, or, when the theta range is empty, a phi
 * sweep ph_s + i*step at theta = th_s; floor(range / step) + 1 points.  With
 * theta or phi NULL only *n is set; otherwise *n is the capacity on entry and
 * the count on return.  Host only. */
int mlfma_scatter_grid(double th_s, double th_f, double ph_s, double ph_f, double step, int* n, double* theta,
                       double* phi);

/* ---- Batch runner (replaces BatchConfigParser::parse, BatchConfig::generateTasks
 * and BatchRunner::run, Code/BatchRunner.h:55-484, on the configuration format
 * of Code/batch_config.txt).  Tasks expand as mesh x frequency x incidence x
 * algorithm x material, each with polarisations "h" (pol 0 here) and "v"
 * (pol 1).  One operator is built per (mesh, frequency, material) and all its
 * (incidence, polarisation) runs are solved together (dual: one batched GMRES
 * and the RCS of every run on the scattering grid; mono: one monostatic sweep
 * per polarisation).  MoM / FMM tasks (algorithm 0 / 1) are not run and report
 * MLFMA_ERR_ARG.  Results are kept in memory (no files). */
typedef struct mlfma_batch mlfma_batch;
int mlfma_batch_load(const char* config_path, mlfma_batch** out);
int mlfma_batch_load_text(const char* config_text, mlfma_batch** out);
void mlfma_batch_destroy(mlfma_batch* b);
int mlfma_batch_count(const mlfma_batch* b, int* n_tasks, int* n_points);
/* d7 = {frequency, incidence theta, phi, eps_r re, im, mu_r re, im};
 * i5 = {algorithm, integral equation, matrix solver, gauss points, mono (1) / dual (0)} */
int mlfma_batch_task(const mlfma_batch* b, int task, double* d7, int* i5, char* nas_path, int nas_capacity);
int mlfma_batch_grid(const mlfma_batch* b, double* theta_deg, double* phi_deg);
/* base: operator parameters other than equation, Gauss rule, frequency and
 * medium (NULL: defaults); sp: NULL for the reference drivers' GMRES(50),
 * tol_rel 1e-4, tol_abs 1e-8, n iterations (Code/RCS.h:215-219). */
int mlfma_batch_run(mlfma_batch* b, const mlfma_params* base, const mlfma_solve_params* sp);
/* Objects (operator groups) processed concurrently by n <= 16 host threads, each
 * with its own operator, stream and graphs (default 1: one after another).
 * Results do not depend on n; device memory holds n operators at a time. */
int mlfma_batch_set_concurrency(mlfma_batch* b, int n);
/* rcs3[3*i + {0,1,2}] = total, V, H (dBsm) on the scattering grid; returns the
 * run's status (MLFMA_ERR_STATE before mlfma_batch_run). */
int mlfma_batch_result(const mlfma_batch* b, int task, int pol, double* rcs3, mlfma_solve_stats* stats);

/* ------------------------------------------------------------------ stages
 * One MLFMA stage on the handle's device workspaces, for per-stage parity and
 * profiling; mirrors ProcessMLFMM (Code/ProcessMLFMM.h:14-249).
 *   MLFMA_STAGE_AGGREGATE  clearSmBm + mexp     (x = host vector of n_unknowns)
 *   MLFMA_STAGE_M2M        m2m, all levels
 *   MLFMA_STAGE_M2L_L2L    m2l + l2l, all levels (fused per level)
 *   MLFMA_STAGE_FAR        finest-level disaggregation into y (host)
 *   MLFMA_STAGE_NEAR       near-field product into y (host)
 * Workspaces can be read back with mlfma_get_workspace. */
#define MLFMA_STAGE_AGGREGATE 0
#define MLFMA_STAGE_M2M 1
#define MLFMA_STAGE_M2L_L2L 2
#define MLFMA_STAGE_FAR 3
#define MLFMA_STAGE_NEAR 4
int mlfma_stage(mlfma_handle* h, int stage, const double* x, double* y);

/* Sm (which = 0) or Bm (which = 1) of wavenumber wn (1, 2), current cur
 * (0 = J, 1 = M; 0 for PEC), tree level `level` (2..maxLevel), node `node`
 * (Morton order): K x (x, y, z) complex. */
int mlfma_get_workspace(mlfma_handle* h, int wn, int cur, int which, int level, int node, double* out);

/* ------------------------------------------------------------------ profiling
 * No reference counterpart (the reference has no per-matvec timing).  When
 * enabled, every matvec brackets its kernel groups with hipEvents on the
 * handle's stream and accumulates their durations:
 *   group 0 aggregation, 1 interpolation (m2m), 2 translation + anterpolation,
 *   3 near-field product, 4 disaggregation.
 * mlfma_get_profile fills ms[5] (summed milliseconds), launches[5] (kernel
 * launches counted) and returns the number of matvecs profiled in *napply. */
#define MLFMA_PROFILE_GROUPS 5
int mlfma_set_profiling(mlfma_handle* h, int enable);
int mlfma_get_profile(mlfma_handle* h, double* ms, long long* launches, int* napply);

/* ------------------------------------------------------------------ inspection
 * Read-back of the setup products in the reference's layouts, for parity
 * tests.  Counts first (pass NULL buffers), then data. */

/* rwg: 6 ints per function (pos face, neg face, edge v0, edge v1, free+, free-
 * as point indices / face indices in input order), length, edge midpoint. */
int mlfma_get_rwg(mlfma_handle* h, int* ints6, double* len, double* mid3);

/* Tree level: n nodes, Morton codes, centres, box length, parent index. */
int mlfma_get_level(mlfma_handle* h, int level, int* n, unsigned long long* codes, double* centres3,
                    double* len, int* parent);

/* CSR list of a level: which = 0 near neighbours, 1 distant relatives,
 * 2 their translation-vector ids, 3 sons, 4 RWG ids (leaves only). */
int mlfma_get_level_list(mlfma_handle* h, int level, int which, int* total, int* ptr, int* idx);

/* Unique translation vectors of a level (Code/OCTree.cpp:500-551). */
int mlfma_get_drvec(mlfma_handle* h, int level, int* n, double* xyz);

/* Angular grid of wavenumber wn at tree level `level`: L, theta, phi,
 * Gauss-Legendre weights (node order), dphi. */
int mlfma_get_grid(mlfma_handle* h, int wn, int level, int* L, double* theta, double* phi, double* wgl,
                   double* dphi);

/* 3x3 Lagrange table from level+1 to level (level >= 2): per parent sample 9
 * child sample indices and 9 weights (Code/AlgoMLFMM.h:298-368). */
int mlfma_get_interp(mlfma_handle* h, int wn, int level, int* taps, double* weights);

/* Translation operators of a level without quadrature weight
 * (n_vec x K complex, Code/AlgoMLFMM.h:192-295). */
int mlfma_get_tf(mlfma_handle* h, int wn, int level, double* out);

/* Radiation (which = 0, Vsmi) or receiving (which = 1, Vfmj) pattern of one
 * RWG: K x (x, y, z) complex (Code/ProcessFMM.h:17-233). */
int mlfma_get_pattern(mlfma_handle* h, int wn, int which, int rwg, double* out);

/* One near-field row in the reference's column order (Code/Zmartix.h:188-643):
 * *len entries, column ids and complex values. */
int mlfma_get_near_row(mlfma_handle* h, int row, int* len, int* cols, double* vals);

#ifdef __cplusplus
}
#endif
#endif /* MLFMA_H */
