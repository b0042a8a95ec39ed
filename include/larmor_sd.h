/*
 * larmor_sd.h -- C-ABI of liblarmor_sd.so, the MI355X (gfx950) per-voxel spherical
 * deconvolution of larmor (larmor-dmri/larmor 2.2.0).
 *
 * The entry points replace the nanobind module larmor._core for the spherical-deconvolution
 * path; each declaration names the reference binding it replaces (paths relative to the
 * reference repository).  Conventions:
 *  - matrices are row-major float64, C-contiguous;
 *  - "host" pointers are read during the call; "device" pointers are HIP device buffers, and
 *    the work is enqueued on `stream` (a hipStream_t, or NULL for the default stream) -- the
 *    call returns before the device work ends;
 *  - every function returns 0 on success and a non-zero LSD_ERR_* code otherwise, with the
 *    message (the reference's ValueError text where it has one) from lsd_last_error(), kept per
 *    calling thread.
 * Results are bit-identical for any grid, any stream and any batch composition: each voxel's
 * arithmetic depends on its own signal only.
 */
#ifndef LARMOR_SD_H
#define LARMOR_SD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSD_ABI_VERSION 6
#define LSD_ERR_INVALID 1 /* invalid argument (the reference raises ValueError) */
#define LSD_ERR_HIP 2     /* a HIP runtime call failed */

/* Per-voxel status bits written by the fits. */
#define LSD_STATUS_NOT_CONVERGED 1 /* an NNLS stopped at its 3n inner-step limit */
#define LSD_STATUS_QR_FALLBACK 4   /* a Cholesky pivot was refused; pivoted QR finished */

typedef struct lsd_solver lsd_solver;
typedef struct lsd_tissue_kernel lsd_tissue_kernel;

/* The last error message of this thread. */
const char* lsd_last_error(void);
int lsd_abi_version(void);

/*
 * Opt-in kernel timing (no counterpart in the reference, which has no profiling hooks): with
 * profiling on, every kernel launch is bracketed by HIP events on its own stream, and the FOD
 * kernels add up the FP64 operations of the solves they complete (counted by formula from each
 * step's factor size: the algorithmic work).  lsd_profile_collect waits for them and returns per
 * kernel name (newline-separated in names) the summed milliseconds, launch counts and counted
 * work (work may be NULL; 0 for kernels that count none), then forgets them.
 */
int lsd_profile_enable(int on);
int lsd_profile_collect(char* names, int64_t names_bytes, double* ms, int64_t* counts,
                        double* work, int64_t max_entries, int64_t* n_entries);

/*
 * An L2-regularised NNLS FOD solver for kernel K (host, m x n): its device copy, the Gram matrix
 * K^T K + ridge^2 I, and whether fits use the normal equations (ridge 0, or
 * (||K||_F^2 + ridge^2) / ridge^2 <= 1e8).
 * Replaces: FodSolver.l2nnls(k, reg, tol_x)  cpp/src/python/sd.cpp:89-99 (l2nnls_solver),
 *           make_ridge_design cpp/src/sd/nnls.cpp:280-292.
 * Errors: "reg must be at least 0", "tol_x must be positive".
 */
int lsd_solver_create_l2nnls(const double* kernel, int64_t m, int64_t n, double ridge,
                             double tol_x, lsd_solver** out);
/*
 * A damped Richardson-Lucy FOD solver from k_t = K^T (c x m) and k_t_k = K^T K (c x c), host
 * arrays, with damping threshold nu, at most n_iter iterations, starting from init in every
 * direction.  Fits with it run drl_kernel: per voxel the damping strength and stopping threshold
 * from the signal, K^T s, then the multiplicative updates until the FOD changes in total by no
 * more than 1e-4 of the signal's largest value (status bit 1 when n_iter stopped the voxel).
 * Replaces: FodSolver.drl(k_t, k_t_k, nu, n_iter, init)  cpp/src/python/sd.cpp:127-146
 *           (drl_solver), DrlSolver cpp/src/sd/solver.hpp:16-33, damped_richardson_lucy
 *           cpp/src/sd/drl.cpp:39-99.
 * Errors: "shapes do not match: k_t (c, m), k_t_k (r, s)", "n_iter must be at least 0".
 */
int lsd_solver_create_drl(const double* kernel_t, int64_t c, int64_t m, const double* gram,
                          int64_t gram_rows, int64_t gram_cols, double nu, int64_t n_iter,
                          double init, lsd_solver** out);
/* 1 for a damped Richardson-Lucy solver, 0 for ridge NNLS (the variant of FodSolver's method,
 * cpp/src/sd/solver.hpp:103-105). */
int lsd_solver_is_drl(const lsd_solver* solver);
/*
 * An elastic-net FOD solver for kernel K (host, m x n): each column divided by its root mean
 * square (a column of zeros kept out), device copies of the scaled columns (column-major) and
 * scales.  Fits with it run lasso_kernel: per voxel, coordinate descent on
 * (1 / 2m) ||s - K0 b||^2 + reg (alpha ||b||_1 + (1 - alpha) / 2 ||b||^2), full sweeps each
 * followed by sweeps over the non-zero coefficients, until a full sweep moves no coefficient by
 * tol relative to 1 + its size or max_iter sweeps ran (status bit 1, and the solver's
 * unconverged count); the FOD is b over the scales.
 * Replaces: FodSolver.lasso(k, reg, alpha, tol, max_iter)  cpp/src/python/sd.cpp:102-125
 *           (elastic_net_solver), ElasticNetSolver cpp/src/sd/solver.hpp:56-64,
 *           standardize_columns / solve_elastic_net cpp/src/sd/elastic_net.cpp:25-151.
 * Errors: "reg must be at least 0", "alpha must be in (0, 1]", "tol must be positive",
 * "max_iter must be at least 1".
 */
int lsd_solver_create_lasso(const double* kernel, int64_t m, int64_t n, double reg, double alpha,
                            double tol, int64_t max_iter, lsd_solver** out);
/* 1 for an elastic-net solver. */
int lsd_solver_is_lasso(const lsd_solver* solver);
/* FodSolver.n_unconverged (cpp/src/python/sd.cpp:441-443): the solves so far that max_iter
 * stopped, always 0 except for an elastic-net solver; waits for the device. */
int lsd_solver_n_unconverged(const lsd_solver* solver, int64_t* out);
void lsd_solver_destroy(lsd_solver* solver);
/* FodSolver.n_dirs / n_meas  cpp/src/python/sd.cpp:437-440 */
int64_t lsd_solver_n_dirs(const lsd_solver* solver);
int64_t lsd_solver_n_meas(const lsd_solver* solver);
int lsd_solver_normal_equations(const lsd_solver* solver);

/*
 * One resolution of a multi-tissue kernel (host arrays): the unweighted anisotropic columns
 * (m x ca), the unweighted isotropic columns (m x ci), the weighted isotropic columns (m x ci) and
 * the anisotropic blocks as (start, stop) pairs (n_blocks x 2).
 * Replaces: DeconvKernel(k_aniso, k_iso, w_k_iso, aniso_blocks)  cpp/src/python/sd.cpp:261-294
 *           (construct_kernel), TissueKernel cpp/src/sd/fractions.hpp:27-38.
 * Errors: "aniso_blocks must be non-empty ranges within 0..ca, got [...]"; at most 3 anisotropic
 * blocks and 6 fraction columns (blocks + isotropic tissues).
 */
int lsd_tissue_kernel_create(const double* k_aniso, const double* k_iso, const double* w_k_iso,
                             int64_t m, int64_t ca, int64_t ci, const int64_t* blocks,
                             int64_t n_blocks, lsd_tissue_kernel** out);
void lsd_tissue_kernel_destroy(lsd_tissue_kernel* kernel);

/*
 * The FOD of each voxel with one solver: signals (device, n_voxels x m) -> fods (device,
 * n_voxels x n), status (device int32, n_voxels).
 * Replaces: fit_single_tissue(signals, solver, n_threads)  cpp/src/python/sd.cpp:148-169,
 *           L2NnlsSolver::solve_block cpp/src/sd/solver.cpp:14-21,
 *           solve_ridge_nnls cpp/src/sd/nnls.cpp:294-350.
 * m <= 1024 for ridge NNLS solvers (512 for the elastic-net and Richardson-Lucy ones).
 * Errors: "signals (r, c) do not have the solver's N measurements".
 */
int lsd_fit_single_tissue(const double* signals, int64_t n_voxels, int64_t m,
                          const lsd_solver* solver, double* fods, int32_t* status, void* stream);

/*
 * The multi-tissue fit: signals (device, n_voxels x m, unweighted), weights (device, m);
 * low-resolution rounds of FOD and lumped-fraction fits (at most max_rounds, a voxel leaving when
 * no fraction moves by fraction_tol), then the high-resolution FOD and the per-block fractions
 * (NNLS tolerance final_tol) normalised to sum to 1.
 * Outputs (device): fods (n_voxels x hr.ca, may be NULL), fractions (n_voxels x
 * (hr.n_blocks + ci)), rounds (int32, n_voxels: low-resolution rounds run), status (int32).
 * Replaces: fit_multi_tissue(signals, weights, lr, lr_solver, hr, hr_solver, max_outer,
 *           fraction_tol, final_tol_x, n_threads)  cpp/src/python/sd.cpp:310-358,
 *           fit_multi_tissue_block cpp/src/sd/multi_tissue.cpp:28-100,
 *           compute_fractions cpp/src/sd/fractions.cpp:45-76.
 * m <= 1024 with ridge NNLS solvers (512 with the others).
 * Errors: the reference's shape and kernel/solver mismatch messages; "max_outer must be at
 * least 0", "problem size exceeds 1024 measurements or 640 directions".
 */
int lsd_fit_multi_tissue(const double* signals, int64_t n_voxels, int64_t m,
                         const double* weights, const lsd_tissue_kernel* lr,
                         const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                         const lsd_solver* hr_solver, int64_t max_rounds, double fraction_tol,
                         double final_tol, double* fods, double* fractions, int32_t* rounds,
                         int32_t* status, void* stream);

/*
 * The correlations K^T a of each voxel, a = max(w s - W_iso p, 0) (all device: signals
 * n_voxels x m, weights m, iso n_voxels x ci the isotropic fractions p, corr n_voxels x n), with
 * K the ridge NNLS solver's kernel and W_iso the tissue kernel's weighted isotropic columns:
 * precision 0 the FP64 pass of the fits (each an in-order sum over the measurements), 1 the
 * mixed-precision form's BF16 matrix-core pass (n <= 64).
 * Replaces: the K^T s of solve_ridge_nnls cpp/src/sd/nnls.cpp:302 with anisotropic_signal
 *           cpp/src/sd/multi_tissue.cpp:20-24 (exposed on their own to test the two passes).
 */
int lsd_correlations(const double* signals, int64_t n_voxels, int64_t m, const double* weights,
                     const lsd_tissue_kernel* kernel, const lsd_solver* solver, const double* iso,
                     int precision, double* corr, void* stream);

/*
 * The mixed-precision form of lsd_fit_multi_tissue (same arguments and outputs, FP64 in and out).
 * The low-resolution rounds run in two phases.  First, the correlations K^T a of each round come
 * from the BF16 matrix cores: a and K split into BF16 halves (x = hi + lo), the products
 * hi.hi + (lo.hi + hi.lo) accumulated in FP32 (one refinement step of the split product).  Then,
 * from where that phase left every voxel, the rounds continue with the exact FP64 correlations
 * until the fractions settle again (the iterative refinement of the loop).  In both phases a
 * voxel's NNLS starts from the support of its previous solve, and the Cholesky factor and its
 * updates, the gradients and the KKT test stay FP64.  The high-resolution solve uses FP64
 * correlations.  Ridge NNLS solvers with a low-resolution design of at most 64 directions.  The
 * results match the FP64 form to the loop tolerance, not to the bit; rounds counts the rounds of
 * both phases.  No counterpart in the reference: the precision form of the benchmark.
 */
int lsd_fit_multi_tissue_mixed(const double* signals, int64_t n_voxels, int64_t m,
                               const double* weights, const lsd_tissue_kernel* lr,
                               const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                               const lsd_solver* hr_solver, int64_t max_rounds,
                               double fraction_tol, double final_tol, double* fods,
                               double* fractions, int32_t* rounds, int32_t* status, void* stream);

/*
 * The plain FP32 form of lsd_fit_multi_tissue: the same kernels and launch sequence with every
 * value in single precision (signals, weights and outputs float32; the solvers' and tissue
 * kernels' matrices are float32 copies of their FP64 ones, made at creation).  Ridge NNLS solvers
 * only.  Its results are not the FP64 form's: they differ by FP32 rounding, and active sets at
 * near-ties with it.  No counterpart in the reference (which computes in FP64): the baseline
 * form of the precision comparison.
 */
int lsd_fit_multi_tissue_fp32(const float* signals, int64_t n_voxels, int64_t m,
                              const float* weights, const lsd_tissue_kernel* lr,
                              const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                              const lsd_solver* hr_solver, int64_t max_rounds, double fraction_tol,
                              double final_tol, float* fods, float* fractions, int32_t* rounds,
                              int32_t* status, void* stream);

/*
 * lsd_fit_multi_tissue with no device allocation: every per-call buffer (voxel lists, counters,
 * the low-resolution FODs and correlations, the loop state, the overflow lists and the
 * global-memory rerun slabs) is carved from a static device workspace of
 * lsd_noalloc_workspace_bytes() (64 MiB), one per device, held until the process ends.  fods and
 * fractions must be given; rounds and status may be NULL.  Ridge NNLS solvers and at most 512
 * measurements only ("problem size exceeds 512 measurements or 640 directions").  The
 * high-resolution correlations are computed inside the FOD kernel (the same sums, so the results
 * are bit-identical to lsd_fit_multi_tissue's), and the rerun slabs fill what the per-voxel state
 * leaves.  A call whose buffers do not fit fails with LSD_ERR_INVALID ("... pass fewer voxels per
 * call"): at the 202-measurement, 44/300-direction three-tissue shape, 16,384 voxels per call fit.
 * The workspace keeps the last call's contents; calls on one device share it, so calls that
 * overlap on one device (other streams, other threads) race: serialise them.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 */
int lsd_fit_multi_tissue_noalloc(const double* signals, int64_t n_voxels, int64_t m,
                                 const double* weights, const lsd_tissue_kernel* lr,
                                 const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                                 const lsd_solver* hr_solver, int64_t max_rounds,
                                 double fraction_tol, double final_tol, double* fods,
                                 double* fractions, int32_t* rounds, int32_t* status,
                                 void* stream);
/* The size of lsd_fit_multi_tissue_noalloc's per-device workspace in bytes. */
int64_t lsd_noalloc_workspace_bytes(void);

/*
 * The device bytes of the global-memory rerun slabs one fit allocates besides its per-voxel
 * buffers, for m measurements and ridge NNLS solves of n_lr (0 for none: a single-tissue fit) and
 * n_hr directions; -1 for sizes the kernels do not take.  Host-only (no device call).  No
 * counterpart in the reference, which allocates per thread on the host: it lets a caller budget
 * device memory (the request server's admission).
 */
int64_t lsd_fit_scratch_bytes(int64_t m, int64_t n_lr, int64_t n_hr);

/*
 * One volume of a batched multi-tissue fit: its signals (n_voxels, m) and shell weights (m) on
 * the device, and its own low- and high-resolution kernels and ridge NNLS solvers.
 */
typedef struct lsd_batch_job {
    const double* signals;
    int64_t n_voxels;
    int64_t m;
    const double* weights;
    const lsd_tissue_kernel* lr;
    const lsd_solver* lr_solver;
    const lsd_tissue_kernel* hr;
    const lsd_solver* hr_solver;
} lsd_batch_job;

/*
 * The multi-tissue fit of 1 to 64 volumes in one launch sequence: each low-resolution round is
 * one correlation, one FOD and one fraction launch over every volume's voxels still in the loop,
 * then one high-resolution solve and one final fraction launch.  The outputs hold the volumes'
 * voxels one volume after the other: fods (sum n_voxels, n_hr), fractions (sum n_voxels,
 * n_blocks + ci), rounds and status (nullable) (sum n_voxels).  Each volume's rows are
 * bit-identical to lsd_fit_multi_tissue on that volume alone.  The volumes share their direction
 * counts (at most 64 low-, 320 high-resolution), isotropic tissues and block counts and their
 * row-slot class (m <= 256, <= 384 or <= 512); the solvers are ridge NNLS (l2nnls).
 * No counterpart in the reference, which fits one volume per call with its voxels spread over
 * threads (cpp/src/common/parallel.hpp:25-88): for callers with many volumes (the request
 * server's batches of acquisitions).
 * Errors: "a batch holds 1 to 64 jobs, got N", "job i: ...", "max_outer must be at least 0".
 */
int lsd_fit_multi_tissue_batched(const lsd_batch_job* jobs, int64_t n_jobs, int64_t max_rounds,
                                 double fraction_tol, double final_tol, double* fods,
                                 double* fractions, int32_t* rounds, int32_t* status,
                                 void* stream);

/*
 * One non-negative least-squares problem, min ||c x - d|| with x >= 0 (device arrays: c (m, n)
 * row-major, d (d_len = m), x (n); stream-ordered on `stream`).  A coefficient enters the
 * solution only when its gradient exceeds tol_x; has_tol 0 takes 10 eps norm(c, 1) max(m, n),
 * each column's absolute sum a wave sum.  Lawson-Hanson with a pivoted-QR least-squares fit per
 * active set and at most 3 n inner steps; status (nullable, device int32) receives the status
 * bits (1: the inner-step limit stopped it, the last candidate kept).  m <= 1024, n <= 640.
 * Replaces: lsqnonneg        cpp/src/python/sd.cpp:209-229 (module.def 471-479),
 *           solve_nnls        cpp/src/sd/nnls.cpp:258-262,
 *           solve_nnls_with_limit nnls.cpp:264-278.
 * Errors: "c must be (m, n) and d (m,): got (r, c) and N".
 */
int lsd_lsqnonneg(const double* c, int64_t m, int64_t n, const double* d, int64_t d_len,
                  double tol_x, int has_tol, double* x, int32_t* status, void* stream);

/*
 * min ||c x - d||^2 + ridge^2 ||x||^2 + 2 l1 sum(x) with x >= 0 (device arrays as
 * lsd_lsqnonneg): the ridge NNLS of the FOD fits on one problem, its correlations c^T d - l1,
 * its Gram matrix c^T c + ridge^2 I (each entry a sequential sum over rows, made on the device)
 * and the normal equations when ridge is 0 or sum(c^2) + ridge^2 <= 1e8 ridge^2, pivoted QR on
 * [c_A; ridge I] otherwise.  Default tolerance as lsd_lsqnonneg.
 * Replaces: ridge_l1_nnls     cpp/src/python/sd.cpp:231-258 (module.def 481-491),
 *           make_ridge_design cpp/src/sd/nnls.cpp:280-292,
 *           solve_ridge_nnls  nnls.cpp:294-350.
 * Errors: "c must be (m, n) and d (m,): ...", "ridge and l1 must be finite and >= 0.",
 *         "l1 > 0 needs ridge > 0.".
 */
int lsd_ridge_l1_nnls(const double* c, int64_t m, int64_t n, const double* d, int64_t d_len,
                      double ridge, double l1, double tol_x, int has_tol, double* x,
                      int32_t* status, void* stream);

/*
 * The log-signal fits of n_voxels voxels on one design (device arrays): each row of log_s
 * (n_voxels, n) fitted on the prepared design -- its columns scaled to unit norm, transposed xt
 * (p, n), their pseudo-inverse pinv (p, n) and the column norms scales (p) -- by the ordinary
 * fit pinv y and, with weighted != 0, refits weighted by the predicted signal squared
 * (exp(2 (x.p - max x.p))) by LU with partial pivoting, until no predicted log signal moves by
 * more than tol, at most max_reweightings of them.  A refit whose normal matrix has a zero pivot
 * or a non-finite result leaves the ordinary fit.  params (n_voxels, p) are the scaled
 * parameters divided by scales; refits (nullable, device int32) the refits kept, -1 when the
 * ordinary fit was kept.  n <= 512, p <= 8; no constraints.
 * Replaces: fit_log_signal        cpp/src/python/dwi.cpp:41-82 (module.def 317-331),
 *           fit_log_signal_block  cpp/src/dwi/log_fit.cpp:159-185,
 *           reweighted_fit        log_fit.cpp:28-62 (the design prepared by the caller as
 *           make_log_signal_design log_fit.cpp:134-149 prepares it).
 * Errors: "x must be (n, p) with n >= p >= 1: ...", "max_iter must be at least 0".
 */
int lsd_fit_log_signal(const double* xt, const double* pinv, const double* scales, int64_t n,
                       int64_t p, const double* log_s, int64_t n_voxels, int weighted,
                       int64_t max_reweightings, double tol, double* params, int32_t* refits,
                       void* stream);

/*
 * A constrained spherical deconvolution problem (host arrays): the linear start initial (c0 x m),
 * the forward model F (m x c), the weighted SH basis on the constraint directions C (k x c),
 * with the Gram matrix F^T F + norm_weight I made here.  1 <= c <= 496 (lmax up to 30: lmax 2-8
 * run with the system in registers, larger c with it in LDS, or in global memory past ~190
 * coefficients); k <= 512.
 * Replaces: fit_csd's problem set-up  cpp/src/python/sd.cpp:171-192,
 *           make_csd_problem cpp/src/sd/csd.cpp:11-21.
 * Errors: "shapes do not match: ...", "max_iter must be at least 0", "CSD on the GPU takes 1 to
 * 496 coefficients (lmax up to 30), got c", "CSD on the GPU takes at most 512 constraint
 * directions, got k".
 */
typedef struct lsd_csd lsd_csd;
int lsd_csd_create(const double* initial, int64_t c0, int64_t m, const double* forward,
                   int64_t c, const double* constraint, int64_t k, double norm_weight,
                   double threshold, int64_t max_iter, lsd_csd** out);
void lsd_csd_destroy(lsd_csd* problem);

/*
 * CSD of each voxel: signals (device, n_voxels x m) -> fods (device, n_voxels x c), converged
 * (device uint8, n_voxels: 1 when the constrained set repeated, or max_iter is 0) and, when not
 * NULL, solves (device int32, n_voxels: Cholesky solves run).
 * Replaces: fit_csd(signals, initial, forward, constraint, norm_weight, threshold, max_iter,
 *           n_threads)  cpp/src/python/sd.cpp:171-206, solve_csd cpp/src/sd/csd.cpp:23-65.
 * Errors: "shapes do not match: signals (r, c), initial ..., forward ..., constraint ...".
 */
int lsd_fit_csd(const double* signals, int64_t n_voxels, int64_t m, const lsd_csd* problem,
                double* fods, uint8_t* converged, int32_t* solves, void* stream);

/*
 * The sphere a peak search climbs from (host arrays): unit vertices (nv x 3, nv <= 3072), the 6
 * neighbours of each vertex (nv x 6; a vertex with 5 lists itself as the sixth) and, optionally,
 * the tournier SH basis at the vertices (nv x n_coeffs) for searches that compute the amplitudes
 * on the device.
 * Replaces: the sphere argument checks and copy of peaks_chunk  cpp/src/python/sh.cpp:118-146.
 * Errors: "a neighbour index is outside the sphere".
 */
typedef struct lsd_sphere lsd_sphere;
int lsd_sphere_create(const double* vertices, int64_t nv, const int64_t* neighbours,
                      const double* basis, int64_t n_coeffs, lsd_sphere** out);
void lsd_sphere_destroy(lsd_sphere* sphere);

/*
 * The peaks of each SH series: coeffs (device, n_voxels x n_coeffs, tournier basis; all degrees
 * when full) with its amplitudes at the sphere's vertices on_sphere (device, n_voxels x nv), or
 * NULL to compute them from the sphere's basis.  Outputs (device): directions (n_voxels x
 * n_peaks x 3) and amplitudes (n_voxels x n_peaks), largest first, NaN-padded; overflow (device
 * int32, may be NULL) counts voxels with more than 128 distinct peaks (the rest are dropped).
 * n_peaks <= 64; lmax <= 20 (above 12 symmetric or 8 full, one kernel takes the lmax at run time).
 * Replaces: peaks_chunk(coeffs, lmax, on_sphere, vertices, neighbours, n_peaks,
 *           relative_threshold, threshold, n_threads, full)  cpp/src/python/sh.cpp:112-175,
 *           voxel_peaks cpp/src/sh/peaks.cpp:79-133, climb cpp/src/sh/peaks.cpp:33-77.
 * Errors: "lmax must be at most 20, and even unless full, got L", "lmax L has N coefficients,
 * got C", "n_peaks must be at least 1".
 */
int lsd_find_peaks(const double* coeffs, int64_t n_voxels, int64_t n_coeffs, int64_t lmax,
                   int full, const double* on_sphere, const lsd_sphere* sphere, int64_t n_peaks,
                   double relative_threshold, double threshold, double* directions,
                   double* amplitudes, int32_t* overflow, void* stream);

/*
 * The value and derivatives of series i at (theta[i], phi[i]) (all device): out (n x 6) =
 * (value, d_theta, d_phi / sin, d_theta_theta, d_theta_phi / sin, d_phi_phi / sin^2).
 * Replaces: sh_derivatives(sh, lmax, theta, phi, full)  cpp/src/python/sh.cpp:70-83,
 *           derivatives cpp/src/sh/harmonics.cpp:139-185 (batched over series).
 */
int lsd_sh_derivatives(const double* coeffs, int64_t n, int64_t n_coeffs, int64_t lmax, int full,
                       const double* theta, const double* phi, double* out, void* stream);

/*
 * One climb per series from starts[i] (all device): directions (n x 3) and amplitudes (n), NaN
 * where the climb did not stop within max_iters.
 * Replaces: climb(sh, lmax, start, max_iters, tol, max_step, full)  cpp/src/python/sh.cpp:85-108
 *           (batched over series).
 * Errors: "max_iters must be at least 0".
 */
int lsd_sh_climb(const double* coeffs, int64_t n, int64_t n_coeffs, int64_t lmax, int full,
                 const double* starts, int64_t max_iters, double tol, double max_step,
                 double* directions, double* amplitudes, void* stream);

/*
 * A least-squares problem with inequality constraints, min ||a x - s||^2 + ridge^2 ||x||^2
 * subject to g x >= 0, prepared for many data vectors (host arrays a (m x n), g (k x g_cols)):
 * the column-pivoted QR of [a; ridge I] and the least-distance form E = g P R^-1 (unit rows),
 * E E^T, made on the host and kept on the device.  n <= 255 (MSMT-CSD up to WM lmax 20), k <= 512.
 * Replaces: LsiDesign(a, g, ridge)  cpp/src/python/sd.cpp:361-367,
 *           make_lsi_design cpp/src/sd/lsi.cpp:19-69.
 * Errors: "the constraints have C columns but the design has N", "the ridge must be at least 0",
 * "the design does not have full column rank: the measurements cannot determine all N
 * parameters".
 */
typedef struct lsd_lsi lsd_lsi;
int lsd_lsi_create(const double* a, int64_t m, int64_t n, const double* g, int64_t k,
                   int64_t g_cols, double ridge, lsd_lsi** out);
void lsd_lsi_destroy(lsd_lsi* design);
int64_t lsd_lsi_n_meas(const lsd_lsi* design);
int64_t lsd_lsi_n_params(const lsd_lsi* design);
int64_t lsd_lsi_n_constraints(const lsd_lsi* design);

/*
 * The constrained solution of each voxel: signals (device, n_voxels x m) -> x (device,
 * n_voxels x n), converged (device uint8: 0 when the NNLS stopped at its inner-step limit) and,
 * when not NULL, status (device int32: LSD_STATUS_QR_FALLBACK when the QR-based NNLS finished
 * the voxel).
 * Replaces: fit_lsi(signals, design, n_threads)  cpp/src/python/sd.cpp:369-396,
 *           solve_lsi cpp/src/sd/lsi.cpp:245-288.
 * Errors: "signals (r, c) do not have the design's M measurements".
 */
int lsd_fit_lsi(const double* signals, int64_t n_voxels, int64_t m, const lsd_lsi* design,
                double* x, uint8_t* converged, int32_t* status, void* stream);

/*
 * lsd_fit_lsi with the QR-based NNLS for every voxel.
 * Replaces: solve_lsi_by_qr cpp/src/sd/lsi.cpp:290-292 (the fallback, on its own).
 */
int lsd_fit_lsi_by_qr(const double* signals, int64_t n_voxels, int64_t m, const lsd_lsi* design,
                      double* x, uint8_t* converged, int32_t* status, void* stream);

/*
 * FOD amplitudes to SH coefficients (all device): coeffs (n x c) = amplitudes (n x k) times
 * projection (k x c, the least-squares projection of the basis on the amplitude directions
 * times their density, c <= 96); with exact_integral, coeffs[:, 0] = the sum of each row of
 * amplitudes / sqrt(4 pi).
 * Replaces: _to_sh  src/larmor/sd/model.py:412-440 (numpy lstsq through fit_sh, sh.py:134-177).
 */
int lsd_sh_project(const double* amplitudes, int64_t n, int64_t k, const double* projection,
                   int64_t c, int exact_integral, double* coeffs, void* stream);

/* ------------------------------------------------------------------------------------------
 * Streamline tracking through a field of FODs (the consumer of the deconvolution's output).
 */

/* Tracking algorithms of lsd_track_params.algorithm. */
#define LSD_TRACK_DET 0  /* deterministic: climb to the nearest FOD peak (SD_STREAM) */
#define LSD_TRACK_PROB 1 /* probabilistic: rejection sampling in a cone (iFOD1) */
#define LSD_TRACK_TEST 2 /* the fixed-rule algorithm that tests the step loop */

/* End codes of lsd_track_batch's status (larmor's EndStatus). */
#define LSD_END_NONE 0
#define LSD_END_MAX_LENGTH 1
#define LSD_END_LEFT_VOLUME 2
#define LSD_END_LEFT_MASK 3
#define LSD_END_LOW_FOD 4
#define LSD_END_SHARP_TURN 5
#define LSD_END_NO_DIRECTION 6

/*
 * The algorithm of a tracking call.  det and prob: cutoff (FOD amplitude), max_angle_deg (det:
 * the largest turn per step; prob: the cone's half-angle), prob's trials per step and attempts per
 * seed, and full (the FOD holds all SH degrees).  test: test_rule 0 straight, 1 refuse with
 * test_stop (1..6), 2 quarter turn about z, 3 wander (three draws per step); its first-direction
 * search draws a random unit vector and fails where the seed's x > test_refuses_past_x.
 * Replaces: DetParams, ProbParams, TestAlgorithm  cpp/src/python/tracking.cpp:180-261, 409-478.
 */
typedef struct lsd_track_params {
    int32_t algorithm;
    int32_t full;
    double cutoff;
    double max_angle_deg;
    int64_t trials;
    int64_t attempts;
    int32_t test_rule;
    int32_t test_stop;
    double test_refuses_past_x;
} lsd_track_params;

/* A tracking call's sampling counts, summed over its streamlines (prob only; 0 otherwise). */
typedef struct lsd_track_stats {
    uint64_t sampled_steps;
    uint64_t accepted_index_sum;
    uint64_t truncations;
    double max_truncation;
    uint64_t calibration_directions;
} lsd_track_stats;

/*
 * Tuning (no counterpart in the reference): the low-resolution solve's LDS factor capacity for
 * the row class of m measurements (<= 256, <= 384, <= 512, <= 768, <= 1024): 16, 20, 24 or 32
 * columns, or 0 for the default (20), process-wide, used by every later plain and batched fit of
 * that class (the batched form stops at 512 measurements).  A
 * voxel whose active set outgrows the capacity finishes in the global-memory solve with the same
 * arithmetic, so the choice changes speed only, never a bit of the results; the autotuner of
 * larmor_amd/sd/tuning.py times the capacities on a sample of a workload and keeps the fastest.
 * lsd_tuned_lr_capacity returns the capacity in force for m (-1 beyond 1024).
 * Errors: "the low-resolution factor capacity is 16, 20, 24 or 32 columns (0: the default), got N".
 */
int lsd_tune_lr_capacity(int64_t m, int64_t capacity);
int64_t lsd_tuned_lr_capacity(int64_t m);

/*
 * The rejection-sampling calibration of prob for an FOD of degree lmax (all degrees if full) and
 * a cone of half-angle max_angle_rad: the grid's size in *n (grid, host, receives min(n, cap)
 * rows of 3 unit directions around +z; may be NULL with cap 0) and the bound's ratio.
 * Replaces: calibrate, direction_grid  cpp/src/tracking/calibrate.cpp:41-108.
 * Errors: "the maximum angle must be above 0.001 rad (0.057 degrees) and at most 180, got X".
 */
int lsd_track_calibrate(int64_t lmax, int full, double max_angle_rad, double* grid, int64_t cap,
                        int64_t* n, double* ratio);

/*
 * Tracks every seed into its row of positions.  fod (device, nx x ny x nz x nc float32, channel
 * fastest, tournier basis); fod_inv_affine (host, 4 x 4 world-to-voxel); mask (device, mx x my x
 * mz float32, or NULL) with mask_inv_affine (host); seeds (device, n x 3 world mm), directions
 * (device, n x 3, or NULL for the algorithm's first-direction search), indices (device int64,
 * n, the global seed numbers that key each streamline's random numbers with run_seed);
 * positions (device float32, positions_rows x (max_steps + 1) x 3, positions_rows >= n): row i
 * receives seed i's points, rows past n and points past a row's count are not written.
 * Outputs (device): counts (n, 0 where the seed did not start), status (n x 2 uint8: the
 * LSD_END_* of the row's start and end), followed (n x 3, the direction given or found).
 * stats (host, may be NULL) receives the sampling counts; it makes the call wait for the
 * stream.  A streamline's points do not depend on the batch, the grid or the other seeds.
 * GPU limits: lmax <= 20 (above 12 symmetric or 8 full, the series lives in the lane's private
 * memory).
 * Replaces: track_batch(fod, fod_inv_affine, mask, seeds, directions, indices, run_seed,
 *           step_size, max_steps, bidirectional, algorithm, positions, n_threads, full)
 *           cpp/src/python/tracking.cpp:264-404; track_batch, track_seed, track_one_way
 *           cpp/src/tracking/tracker.cpp:42-139; det.cpp, prob.cpp, testing.cpp, grid.cpp,
 *           rng.cpp.
 * Errors: "the cutoff must be non-negative and finite, got X", "the maximum angle must be in
 * (0, 180], got X", "trials and attempts must be at least 1, got T and A", "need a positive step
 * size and max_steps, got S and M", "lmax must be at most 20, and even unless full, got L",
 * "lmax L has N coefficients, got C".
 */
int lsd_track_batch(const float* fod, int64_t nx, int64_t ny, int64_t nz, int64_t nc,
                    const double* fod_inv_affine, const float* mask, int64_t mx, int64_t my,
                    int64_t mz, const double* mask_inv_affine, const double* seeds,
                    const double* directions, const int64_t* indices, int64_t n_seeds,
                    uint64_t run_seed, double step_size, int64_t max_steps, int bidirectional,
                    const lsd_track_params* params, float* positions, int64_t positions_rows,
                    int64_t* counts, uint8_t* status, double* followed, lsd_track_stats* stats,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LARMOR_SD_H */
