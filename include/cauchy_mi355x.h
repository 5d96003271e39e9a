/*
 * cauchy_mi355x.h -- C ABI of the MI355X (gfx950) Multivariate Cauchy Estimator
 * term-propagation engine.
 *
 * The pycauchy_* entry points keep the signatures, argument meaning, output
 * ownership (outputs are malloc'd here and released by the caller with free())
 * and error behaviour of the reference's SWIG boundary,
 * scripts/swig/cauchy/pycauchy.hpp (declared for SWIG in pycauchy.i:80-423),
 * so the reference's Python module, MATLAB mex files and C++ drivers can bind
 * this library instead.  Every term of the estimator lives in GPU memory and
 * every measurement update runs as HIP kernels; the host only sequences them.
 *
 * The mce_* entry points have no counterpart in the reference; they expose the
 * per-update diagnostics (pre-reduction moments, term counts per shape) and the
 * term tables that the parity tests compare against the reference.
 */
#ifndef CAUCHY_MI355X_H
#define CAUCHY_MI355X_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as CauchyDynamicsUpdateContainer, include/dynamic_models.hpp:62-80,
 * mirrored by ctypes in scripts/swig/cauchy/cauchy_estimator.py:40-58. */
typedef struct CauchyDynamicsUpdateContainer {
    double* x;
    double* u;
    double dt;
    int step;
    int n;
    int cmcc;
    int pncc;
    int p;
    double* Phi;
    double* Gamma;
    double* B;
    double* H;
    double* beta;
    double* gamma;
    bool is_xbar_set_for_ece;
    void* other_stuff;
} CauchyDynamicsUpdateContainer;

/* Replaces pycauchy_initialize_lti, scripts/swig/cauchy/pycauchy.hpp:295-328.
 * Sizes follow the reference: n = size_b0, pncc = size_beta, cmcc = size_B / n,
 * p = size_gamma.  Matrices are row-major.  Returns NULL on failure (the
 * reference exits; this library prints the reason and returns NULL). */
void* pycauchy_initialize_lti(int num_steps, double* A0, int size_A0, double* p0, int size_p0, double* b0,
                              int size_b0, double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B,
                              int size_B, double* beta, int size_beta, double* H, int size_H, double* gamma,
                              int size_gamma, double dt, int init_step, bool debug_print);

/* Replaces pycauchy_initialize_ltv, scripts/swig/cauchy/pycauchy.hpp:330-370.  The
 * callback is invoked on the calling thread before every time step, as in the reference. */
void* pycauchy_initialize_ltv(int num_steps, double* A0, int size_A0, double* p0, int size_p0, double* b0,
                              int size_b0, double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B,
                              int size_B, double* beta, int size_beta, double* H, int size_H, double* gamma,
                              int size_gamma, void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                              double dt, int init_step, bool debug_print);

/* Replaces pycauchy_single_step_ltiv, scripts/swig/cauchy/pycauchy.hpp:461-543.
 * Processes size_msmts scalar measurement updates (rows p-size_msmts .. p-1 of H).
 * All 13 outputs are malloc'd by the library. */
void pycauchy_single_step_ltiv(void* handle, double* msmts, int size_msmts, double* controls, int size_controls,
                               double** out_Phi, int* size_out_Phi, double** out_Gamma, int* size_out_Gamma,
                               double** out_B, int* size_out_B, double** out_H, int* size_out_H, double** out_beta,
                               int* size_out_beta, double** out_gamma, int* size_out_gamma, double** out_fz,
                               int* size_out_fz, double** out_xhat, int* size_out_xhat, double** out_Phat,
                               int* size_out_Phat, double** out_cerr_fz, int* size_out_cerr_fz,
                               double** out_cerr_xhat, int* size_out_cerr_xhat, double** out_cerr_Phat,
                               int* size_out_cerr_Phat, int** out_err_code, int* size_out_err_code);

/* Replaces pycauchy_initialize_nonlin, scripts/swig/cauchy/pycauchy.hpp:365-396: the
 * extended (linearised) estimator.  n = size_b0, pncc = size_beta, p = size_gamma; x0 is the
 * initial x_bar.  The estimator itself runs on the state deviation, without a control term.
 * The three callbacks run on the calling thread: f_dyn_update_callback propagates duc->x and
 * sets Phi, Gamma, B, beta and is_xbar_set_for_ece; f_nonlinear_msmt_model writes h(duc->x) into
 * its p-vector argument; f_extended_msmt_update_callback sets H (and gamma) at duc->x.
 * Returns NULL on failure. */
void* pycauchy_initialize_nonlin(int num_steps, double* x0, int size_x0, double* A0, int size_A0, double* p0,
                                 int size_p0, double* b0, int size_b0, double* beta, int size_beta, double* gamma,
                                 int size_gamma, void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                                 void (*f_nonlinear_msmt_model)(CauchyDynamicsUpdateContainer*, double*),
                                 void (*f_extended_msmt_update_callback)(CauchyDynamicsUpdateContainer*), int cmcc,
                                 double dt, int init_step, bool debug_print);

/* Replaces pycauchy_single_step_nonlin, scripts/swig/cauchy/pycauchy.hpp:545-652.  With
 * with_propagate, the dynamics callback runs first.  Then, for rows p-size_msmts .. p-1: zbar
 * and H from the callbacks, one update with the residual z - zbar, and the extended finish
 * (terms shifted by the estimated deviation, x_bar <- Re(x_bar + deviation)).  All 15 outputs
 * are malloc'd; the dynamics outputs hold the current dynamics when the step does not
 * propagate (the reference leaves them unset). */
void pycauchy_single_step_nonlin(void* handle, double* msmts, int size_msmts, double* controls, int size_controls,
                                 bool with_propagate, double** out_Phi, int* size_out_Phi, double** out_Gamma,
                                 int* size_out_Gamma, double** out_B, int* size_out_B, double** out_H,
                                 int* size_out_H, double** out_beta, int* size_out_beta, double** out_gamma,
                                 int* size_out_gamma, double** out_fz, int* size_out_fz, double** out_xhat,
                                 int* size_out_xhat, double** out_Phat, int* size_out_Phat, double** out_xbar,
                                 int* size_out_xbar, double** out_zbar, int* size_out_zbar, double** out_cerr_fz,
                                 int* size_out_cerr_fz, double** out_cerr_xhat, int* size_out_cerr_xhat,
                                 double** out_cerr_Phat, int* size_out_cerr_Phat, int** out_err_code,
                                 int* size_out_err_code);

/* Batched stepping (no counterpart in the reference, whose windows and estimators step one
 * at a time or in forked processes).  Steps count estimators concurrently from up to
 * num_threads host threads (0: one per hardware thread); estimator i runs exactly
 * pycauchy_single_step_ltiv(handles[i], msmts[i], size_msmts[i], controls[i], size_controls[i],
 * ...) on its own HIP streams, so each result is that of the single step.  Its 13 outputs are
 * outs[13*i + j] / out_sizes[13*i + j] in the single step's order (j = 12, the error codes, is
 * an int*), malloc'd as there.  controls / size_controls may be NULL.  LTV callbacks run on
 * the worker threads.  Returns the number of estimators that reported an error (message in
 * mce_last_error()). */
int pycauchy_batch_single_step_ltiv(void** handles, int count, double** msmts, const int* size_msmts,
                                    double** controls, const int* size_controls, void** outs, int* out_sizes,
                                    int num_threads);

/* Batched stepping with merged launches (no counterpart in the reference).  Same arguments,
 * outputs and results as pycauchy_batch_single_step_ltiv, but the estimators are dealt to
 * num_groups host threads, each with one HIP stream, and the estimators of a group step
 * together: every kernel of their updates is launched once for the whole group, each
 * estimator's part a block range of that launch with its own arguments.  The kernels,
 * their order per estimator and each estimator's arithmetic are those of the single step.
 * Not for estimators in tangent mode.  Returns the number of estimators that reported an
 * error (message in mce_last_error()). */
int pycauchy_batch_single_step_ltiv_merged(void** handles, int count, double** msmts, const int* size_msmts,
                                           double** controls, const int* size_controls, void** outs, int* out_sizes,
                                           int num_groups);

/* Batched pycauchy_single_step_nonlin, as above with 15 outputs per estimator (j = 14, the
 * error codes, is an int*) and with_propagate[i] != 0 for the step's propagation. */
int pycauchy_batch_single_step_nonlin(void** handles, int count, double** msmts, const int* size_msmts,
                                     double** controls, const int* size_controls, const int* with_propagate,
                                     void** outs, int* out_sizes, int num_threads);

/* Replaces pycauchy_get_marginal_1D_pointwise_cpdf, scripts/swig/cauchy/pycauchy.hpp:890-932:
 * the marginal density of state marg_idx1 on the grid gridx_low + i * gridx_resolution
 * (clamped to gridx_high), as (x, f(x)) pairs, malloc'd; out_num_gridx points.  No points
 * after a window's last update (its g-tables are not built) or on error.  The first call on a
 * handle draws d values from glibc rand(), as the reference's density object does.  log_dir
 * is accepted and not written. */
void pycauchy_get_marginal_1D_pointwise_cpdf(void* handle, int marg_idx1, double gridx_low, double gridx_high,
                                             double gridx_resolution, char* log_dir, double** out_cpdf_data,
                                             int* size_out_cpdf_data, int* out_num_gridx);

/* No counterpart in the reference: pycauchy_get_marginal_1D_pointwise_cpdf for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The grid's x values and densities go to the caller's device arrays
 * out_x and out_f (capacity points each), launched on `stream` (a hipStream_t; NULL for the
 * null stream) without waiting: they are ready once the stream reaches this point.  Every
 * intermediate (40 B per term, 8 B per 1024-term chunk and grid point) lives in one static
 * 128 MiB device array per GPU, held until the process ends; a larger request, or a grid
 * beyond `capacity`, is refused.  Calls on one GPU share that array: calls that overlap on one
 * GPU race.  Returns 0 and the point count in out_num_gridx; 1 after a window's last update
 * (as the allocating form, no points); -1 on error (message in mce_last_error). */
int mce_cpdf1d_noalloc(void* handle, int marg_idx, double gridx_low, double gridx_high, double gridx_resolution,
                       double* out_x, double* out_f, int capacity, void* stream, int* out_num_gridx);

/* Replaces pycauchy_get_1D_pointwise_cpdf, scripts/swig/cauchy/pycauchy.hpp:934-945 (state 0). */
void pycauchy_get_1D_pointwise_cpdf(void* handle, double gridx_low, double gridx_high, double gridx_resolution,
                                    char* log_dir, double** out_cpdf_data, int* size_out_cpdf_data,
                                    int* out_num_gridx);

/* Replaces pycauchy_get_marginal_2D_pointwise_cpdf, scripts/swig/cauchy/pycauchy.hpp:822-872:
 * the joint marginal density of states marg_idx1 < marg_idx2 on the grid (x from gridx_low
 * by gridx_resolution, y likewise, each clamped to its high end; x runs fastest), as (x, y, f)
 * triples, malloc'd; out_num_gridx x out_num_gridy points.  No points after a window's last
 * update, or on error (mce_last_error()).  Every call builds its per-term cell data afresh, so
 * reset_cache changes nothing; log_dir is accepted and not written. */
void pycauchy_get_marginal_2D_pointwise_cpdf(void* handle, int marg_idx1, int marg_idx2, double gridx_low,
                                             double gridx_high, double gridx_resolution, double gridy_low,
                                             double gridy_high, double gridy_resolution, char* log_dir,
                                             bool reset_cache, double** out_cpdf_data, int* size_out_cpdf_data,
                                             int* out_num_gridx, int* out_num_gridy);

/* Replaces pycauchy_get_2D_pointwise_cpdf, scripts/swig/cauchy/pycauchy.hpp:874-888 (states 0, 1). */
void pycauchy_get_2D_pointwise_cpdf(void* handle, double gridx_low, double gridx_high, double gridx_resolution,
                                    double gridy_low, double gridy_high, double gridy_resolution, char* log_dir,
                                    bool reset_cache, double** out_cpdf_data, int* size_out_cpdf_data,
                                    int* out_num_gridx, int* out_num_gridy);

/* Replaces pycauchy_single_step_eval_2d_rsys_cpdf, scripts/swig/cauchy/pycauchy.hpp:673-750: the
 * relative system (secondary - primary) of two estimators of one state dimension n,
 * transformed by Trel (2 x n, row-major).  Every pair of their terms forms one relative
 * term; terms whose bounded contribution is below RSYS_APPROX_EPS are left out of the
 * density (cauchy_prediction.hpp:445-471), not out of the moments.  Outputs (malloc'd): the
 * relative normalisation factor, mean (2) and covariance (4) with their imaginary parts,
 * and the density on the grid [xlow, xhigh] x [ylow, yhigh] about the relative mean as
 * (x, y, f) triples with y fastest, out_num_gridx x out_num_gridy points.  Both estimators
 * must have stepped and not be past their window's last update; on error the density is
 * empty and mce_last_error() says why.  Draws n values from glibc rand() per call, as the
 * reference does. */
void pycauchy_single_step_eval_2d_rsys_cpdf(double* Trel, int size_Trel, void* _s_pcdh, void* _p_pcdh,
                                            double RSYS_APPROX_EPS, double xlow, double xhigh, double delta_x,
                                            double ylow, double yhigh, double delta_y, double** out_rsys_fz,
                                            int* size_out_rsys_fz, double** out_rsys_xhat, int* size_out_rsys_xhat,
                                            double** out_rsys_Phat, int* size_out_rsys_Phat,
                                            double** out_rsys_cerr_fz, int* size_out_rsys_cerr_fz,
                                            double** out_rsys_cerr_xhat, int* size_out_rsys_cerr_xhat,
                                            double** out_rsys_cerr_Phat, int* size_out_rsys_cerr_Phat,
                                            double** out_cpdf_data, int* size_out_cpdf_data, int* out_num_gridx,
                                            int* out_num_gridy);

/* Replaces pycauchy_single_step_reset, scripts/swig/cauchy/pycauchy.hpp:798-819. */
void pycauchy_single_step_reset(void* handle, double* A0, int size_A0, double* p0, int size_p0, double* b0,
                                int size_b0, double* xbar, int size_xbar);

/* Replaces pycauchy_single_step_set_master_step, scripts/swig/cauchy/pycauchy.hpp:767-777. */
void pycauchy_single_step_set_master_step(void* handle, int step);

/* Replaces pycauchy_single_step_set_window_number, scripts/swig/cauchy/pycauchy.hpp:779-789. */
void pycauchy_single_step_set_window_number(void* handle, int win_num);

/* Replaces pycauchy_single_step_get_number_of_terms, scripts/swig/cauchy/pycauchy.hpp:666-671. */
int pycauchy_single_step_get_number_of_terms(void* handle);

/* Replaces pycauchy_single_step_deterministic_transform, scripts/swig/cauchy/pycauchy.hpp:654-664:
 * A <- A Trans^T, b <- Trans b + bias for every term (cauchy_estimator.hpp:1312-1355). */
void pycauchy_single_step_deterministic_transform(void* handle, double* Trans, int size_Trans, double* bias,
                                                  int size_bias);

/* Replaces pycauchy_single_step_get_duc, scripts/swig/cauchy/pycauchy.hpp:791-795. */
void* pycauchy_single_step_get_duc(void* handle);

/* Replaces pycauchy_single_step_shutdown, scripts/swig/cauchy/pycauchy.hpp:752-765. */
void pycauchy_single_step_shutdown(void* handle);

/* Replaces pycauchy_set_tr_search_idxs_ordering, scripts/swig/cauchy/pycauchy.hpp:994-1006.
 * Returns 1 on error (more than 12 entries), 0 otherwise. */
int pycauchy_set_tr_search_idxs_ordering(int* ordering, int size_ordering);

/* ---- the window manager (one bank per process) -------------------------------------- */

/* Replace pycauchy_initialize_{lti,ltv,nonlin}_window_manager, scripts/swig/cauchy/
 * pycauchy.hpp:80-182: a bank of num_windows windows for num_sim_steps measurements, with
 * the reference's arguments (sizes as for pycauchy_initialize_lti; win_var_boost, n entries,
 * is added to the diagonal of the restart covariance; size_wvb == 0 for none).  The
 * reference forks one process per window (SlidingWindowManager, include/cauchy_windows.hpp:
 * 729-1705); here every window is an estimator of this library in the calling process, with
 * the same schedule, window selection, Speyer restarts and outputs, and the caller's rand()
 * state left as it was.  log_seq / log_full / log_dir are accepted and no log is written.
 * On failure no bank is made (message in mce_last_error). */
void pycauchy_initialize_lti_window_manager(int num_windows, int num_sim_steps, double* A0, int size_A0, double* p0,
                                            int size_p0, double* b0, int size_b0, double* Phi, int size_Phi,
                                            double* Gamma, int size_Gamma, double* B, int size_B, double* beta,
                                            int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                                            bool debug_print, bool log_seq, bool log_full, char* log_dir, double dt,
                                            int init_step, double* win_var_boost, int size_wvb);
void pycauchy_initialize_ltv_window_manager(int num_windows, int num_sim_steps, double* A0, int size_A0, double* p0,
                                            int size_p0, double* b0, int size_b0, double* Phi, int size_Phi,
                                            double* Gamma, int size_Gamma, double* B, int size_B, double* beta,
                                            int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                                            void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                                            bool debug_print, bool log_seq, bool log_full, char* log_dir, double dt,
                                            int init_step, double* win_var_boost, int size_wvb);
void pycauchy_initialize_nonlin_window_manager(int num_windows, int num_sim_steps, double* x0, int size_x0,
                                               double* A0, int size_A0, double* p0, int size_p0, double* b0,
                                               int size_b0, double* beta, int size_beta, double* gamma,
                                               int size_gamma,
                                               void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*),
                                               void (*f_nonlinear_msmt_model)(CauchyDynamicsUpdateContainer*,
                                                                              double*),
                                               void (*f_extended_msmt_update_callback)(CauchyDynamicsUpdateContainer*),
                                               int cmcc, bool debug_print, bool log_seq, bool log_full, char* log_dir,
                                               double dt, int init_step, double* win_var_boost, int size_wvb);

/* Replaces pycauchy_step, scripts/swig/cauchy/pycauchy.hpp:184-214: one measurement (p values,
 * cmcc controls) into the bank; the best window's fz, x_hat (malloc'd, n), P_hat (malloc'd,
 * n x n), their largest imaginary parts, its index and its error bits.  On error the arrays
 * are NULL with size 0 and the index -1. */
void pycauchy_step(double* msmts, int size_msmts, double* controls, int size_controls, double* out_swm_fz,
                   double** out_xhat, int* size_out_xhat, double** out_Phat, int* size_out_Phat,
                   double* out_swm_cerr_fz, double* out_swm_cerr_xhat, double* out_swm_cerr_Phat,
                   int* out_swm_win_idx, int* out_swm_err_code);

/* Replaces pycauchy_shutdown, scripts/swig/cauchy/pycauchy.hpp:216-221: ends the bank. */
void pycauchy_shutdown(void);

/* How the running bank steps its windows (no counterpart in the reference, whose windows are
 * forked processes): 0 one window after another; 1 the windows that take a measurement
 * without restarting step together, each kernel of their updates one merged launch; 2 (the
 * default) those windows on one host thread each, each on its own streams (extended mode
 * steps merged instead, so that its callbacks run on the calling thread).  Results are the
 * same in every mode.  Returns 0, or -1 with no bank running or another mode. */
int mce_window_bank_set_mode(int mode);

/* Replaces pycauchy_speyers_window_init, scripts/swig/cauchy/pycauchy.hpp:973-993: the
 * initial terms (A0 n x n rows, p0, b0; malloc'd) of a window restarted about the estimate
 * (xhat, Phat) after scalar measurement z = H x + v, v ~ Cauchy(0, gamma)
 * (speyers_window_init, include/cauchy_util.hpp:23-113).  Phat is not modified. */
void pycauchy_speyers_window_init(double* xhat, int size_xhat, double* Phat, int size_Phat, double* H, int size_H,
                                  double gamma, double z, double** out_A0, int* size_out_A0, double** out_p0,
                                  int* size_out_p0, double** out_b0, int* size_out_b0);

/* Replaces pycauchy_get_reinitialization_statistics, scripts/swig/cauchy/pycauchy.hpp:948-971
 * (the same computation for the estimator's state dimension). */
void pycauchy_get_reinitialization_statistics(void* handle, double z, double* xhat, int size_xhat, double* Phat,
                                               int size_Phat, double* H, int size_H, double gamma, double** out_A0,
                                               int* size_out_A0, double** out_p0, int* size_out_p0, double** out_b0,
                                               int* size_out_b0);

/* ---- the C++ CauchyEstimator interface (include/cauchy_estimator.hpp:47-1432) --------- */

/* Replaces the CauchyEstimator constructor, include/cauchy_estimator.hpp:87-185: steps time
 * steps of p scalar updates each; the root point and the cell enumerator's perturbations are
 * drawn from glibc rand() in the reference's order.  Returns NULL on failure. */
void* mce_estimator_create(double* A0, double* p0, double* b0, int steps, int d, int cmcc, int pncc, int p,
                           bool print_basic_info);

/* Replaces CauchyEstimator::step, include/cauchy_estimator.hpp:1211-1245: one scalar update
 * with row H (d) and scale gamma; Phi (d x d), Gamma (d x pncc), beta (pncc), B (d x cmcc), u
 * (cmcc) are used when the update starts a time step.  Returns numeric_moment_errors; -1 when
 * the estimator refuses to step (window full or fz negative). */
int mce_estimator_step(void* handle, double msmt, double* Phi, double* Gamma, double* beta, double* H,
                       double gamma, double* B, double* u);

/* Public members fz, conditional_mean, conditional_variance as (re, im) pairs, plus Nt. */
int mce_estimator_moments(void* handle, double* fz, double* mean, double* cov);

/* Replaces CauchyEstimator::reset, include/cauchy_estimator.hpp:1247-1300. */
void mce_estimator_reset(void* handle);

/* Releases an estimator made by mce_estimator_create (same as pycauchy_single_step_shutdown). */
void mce_estimator_destroy(void* handle);

/* Replaces CauchyEstimator::finalize_extended_moments, include/cauchy_estimator.hpp:1358-1394:
 * b -= Re(mean) for every term (not after a window's last update), mean += x_bar,
 * x_bar <- Re(mean).  Returns 0, or -1 on a device error. */
int mce_estimator_finalize_extended_moments(void* handle, double* x_bar);

/* ---- derivatives (no reference counterpart) ----------------------------------------- */

/* Tangent mode for num_directions (1..16) derivative directions, before a window's first
 * update (after mce_estimator_create or mce_estimator_reset).  Every later step carries the
 * derivatives of the terms and g-tables along the directions, with the term structure of
 * the nominal pass.  Returns 0, or -1 (message in mce_last_error). */
int mce_estimator_enable_tangents(void* handle, int num_directions);

/* mce_estimator_step in tangent mode.  Seeds of this update per direction k < P:
 * dPhi (P x d x d), dbeta (P x pncc, before the pre-coalignment of Gamma's columns),
 * dH (P x d, the derivative of this update's row H) and dgamma (P); a NULL seed is zero.
 * Writes dfz (2P doubles, (re, im) per direction): the derivative of this update's
 * measurement likelihood fz before term reduction (the moment sum the reference's
 * cache_moments forms, include/cauchy_estimator.hpp:307-338).  Returns as
 * mce_estimator_step; -2 on an error. */
int mce_estimator_step_tangent(void* handle, double msmt, double* Phi, double* Gamma, double* beta, double* H,
                               double gamma, double* B, double* u, const double* dPhi, const double* dbeta,
                               const double* dH, const double* dgamma, double* dfz);

/* ---- diagnostics (no reference counterpart) ---------------------------------------- */

/* Diagnostics of the last scalar update, in the layout of oracle/ref_harness.cpp ref_step:
 *   out_mom (6 + 2d + 2d^2): fz public (re,im), fz before reduction (re,im), last_fz (re,im),
 *                            mean (re,im) x d, covariance (re,im) x d^2
 *   out_cnt (4 + 2*shape_range): err code, terms after MU/MUC, terms after reduction,
 *                            reference thread count, per-shape counts after MU/MUC, after reduction.
 * Returns the error code, or -1 before any update. */
int mce_last_update(void* handle, double* out_mom, int* out_cnt);

/* Shape range (largest hyperplane count + 1) of the estimator. */
int mce_shape_range(void* handle);

/* Terms generated since the estimator was created: the sum over its updates of the terms
 * after MU/MUC (after MU alone on a window's last update), the numerator of terms/s. */
long long mce_terms_generated(void* handle);

/* Copies term `idx` of shape bucket `shape` (after an update that ran term reduction):
 * info = [m, g-table size, m]; A (m x d), p (m), b (d), keys / g (re,im interleaved) of its g-table.
 * Returns 0, or -1 when there is no such term. */
int mce_term(void* handle, int shape, int idx, int* info, double* A, double* p, double* b, unsigned int* keys,
             double* gvals);

/* Replaces the construction-time random draws: root point (d) and the 8 x max_shape
 * perturbation vectors of the cell enumerator. */
void mce_set_random_draws(void* handle, const double* root_point, const double* b_pert_8xmax);
void mce_get_random_draws(void* handle, double* root_point, double* b_pert_8xmax);

/* Accumulated device time (ms, HIP events on the engine's stream) since the last reset:
 * out[0..8] per phase [time-prop, B-table enumeration, measurement update, moments, gather,
 * reduction, group B-tables, group g-tables, compaction], out[9] their total, out[10] the
 * launch count of the group g-table kernel, out[11] its algorithmic bytes, out[12] its
 * kernel time, out[13] reduction groups evaluated, out[14] their members, out[15] the
 * largest group (root + members), out[16] the algorithmic bytes of the profiled updates
 * (every parent's record and g-table read once, every child written and read, every
 * surviving g-table and B-table written, every merged member's table read and added),
 * out[17] the updates profiled.  Recorded only after mce_profile_enable(h, 1); out must
 * hold 18 doubles. */
void mce_profile_enable(void* handle, int on);
void mce_profile_read(void* handle, double* out);
void mce_profile_reset(void* handle);

/* The last scalar update's fz (re, im) before the numeric checks, which replace fz by the
 * previous update's when both the mean and the covariance fail them (no counterpart in the
 * reference, whose public fz is the checked one).  The window log-likelihood sum_k log fz_k
 * and its tangent use this value. */
void mce_last_fz_raw(void* handle, double* out2);

/* Kernel shapes of one estimator (no counterpart in the reference), set by the autotuner
 * (cauchyfriendly_amd/autotune.py); results never depend on them.  enum_threads: lanes per
 * parent in the time-propagation cell enumerator (64, 128 or 256; 0 keeps it);
 * big_group_members: reduction groups with more members evaluate them one wavefront per
 * member (> 0; 0 keeps it); geval_mode: parent g-table lookup of the g-table kernel, 1..4
 * for binary search in global memory, binary search in LDS, LDS rank structure, rank
 * structure with LDS values (a mode the estimator cannot use is ignored), -1 automatic,
 * 0 keeps it.  Returns 0, or -1 with mce_last_error() set. */
int mce_estimator_set_shapes(void* handle, int enum_threads, int big_group_members, int geval_mode);

/* Profile of merged batch stepping (pycauchy_batch_single_step_ltiv_merged), summed over
 * its schedulers since mce_batch_profile_enable(1) (which also resets it): out[0] merged
 * group g-table launches, out[1] their algorithmic bytes, out[2] their kernel time (ms,
 * HIP events on the scheduler's stream), out[3] flushes, out[4] merged launches, out[5]
 * library calls run one estimator at a time.  out must hold 6 doubles. */
void mce_batch_profile_enable(int on);
void mce_batch_profile_read(double* out);

/* Last error message ("" when none). */
const char* mce_last_error(void);

/* The scalar updates of the handle's last pycauchy_single_step_* call (no counterpart in the
 * reference, whose step returns the fz left after term reduction): out_fz[i] = the measurement
 * likelihood fz of update i before term reduction (cauchy_estimator.hpp:455-460),
 * out_counts[2i], [2i+1] = its terms after MU/MUC and after reduction.  Writes at most
 * max_updates entries; returns the number of updates of that step. */
int mce_step_record(void* handle, double* out_fz, int* out_counts, int max_updates);

/* Host helper for the batched calls; no counterpart in the reference.  Copies the outputs
 * that a pycauchy_batch_* call returned (outs / sizes of count estimators x nout outputs) into
 * two caller buffers in order: every output but the last of each estimator into dbuf
 * (doubles), the last (the int error codes) into ibuf, then frees each output and sets its
 * pointer to NULL.  dbuf and ibuf must hold the summed sizes.  Returns the doubles written. */
long long mce_collect_outputs(void** outs, const int* sizes, long long count, int nout, double* dbuf, int* ibuf);

#ifdef __cplusplus
}
#endif
#endif /* CAUCHY_MI355X_H */
