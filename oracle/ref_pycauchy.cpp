// Reference build of the SWIG-level estimator functions
// (scripts/swig/cauchy/pycauchy.hpp, compiled from the reference tree where it
// lies), re-exported with C linkage under refpy_ names.  TEST INFRASTRUCTURE
// ONLY: tests/golden/make_window_golden.py binds these to run the reference's
// own Python classes on the reference backend and record fixtures.
#include "pycauchy.hpp"

extern "C" {

void* refpy_initialize_lti(int num_steps, double* A0, int size_A0, double* p0, int size_p0, double* b0, int size_b0,
                           double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B, int size_B,
                           double* beta, int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                           double dt, int init_step, bool debug_print) {
    return pycauchy_initialize_lti(num_steps, A0, size_A0, p0, size_p0, b0, size_b0, Phi, size_Phi, Gamma, size_Gamma,
                                   B, size_B, beta, size_beta, H, size_H, gamma, size_gamma, dt, init_step,
                                   debug_print);
}

void* refpy_initialize_ltv(int num_steps, double* A0, int size_A0, double* p0, int size_p0, double* b0, int size_b0,
                           double* Phi, int size_Phi, double* Gamma, int size_Gamma, double* B, int size_B,
                           double* beta, int size_beta, double* H, int size_H, double* gamma, int size_gamma,
                           void (*cb)(CauchyDynamicsUpdateContainer*), double dt, int init_step, bool debug_print) {
    return pycauchy_initialize_ltv(num_steps, A0, size_A0, p0, size_p0, b0, size_b0, Phi, size_Phi, Gamma, size_Gamma,
                                   B, size_B, beta, size_beta, H, size_H, gamma, size_gamma, cb, dt, init_step,
                                   debug_print);
}

void refpy_single_step_ltiv(void* h, double* msmts, int size_msmts, double* controls, int size_controls,
                            double** o1, int* s1, double** o2, int* s2, double** o3, int* s3, double** o4, int* s4,
                            double** o5, int* s5, double** o6, int* s6, double** o7, int* s7, double** o8, int* s8,
                            double** o9, int* s9, double** o10, int* s10, double** o11, int* s11, double** o12,
                            int* s12, int** o13, int* s13) {
    pycauchy_single_step_ltiv(h, msmts, size_msmts, controls, size_controls, o1, s1, o2, s2, o3, s3, o4, s4, o5, s5,
                              o6, s6, o7, s7, o8, s8, o9, s9, o10, s10, o11, s11, o12, s12, o13, s13);
}

void* refpy_initialize_nonlin(int num_steps, double* x0, int size_x0, double* A0, int size_A0, double* p0,
                              int size_p0, double* b0, int size_b0, double* beta, int size_beta, double* gamma,
                              int size_gamma, void (*f_dyn)(CauchyDynamicsUpdateContainer*),
                              void (*f_msmt)(CauchyDynamicsUpdateContainer*, double*),
                              void (*f_ext)(CauchyDynamicsUpdateContainer*), int cmcc, double dt, int init_step,
                              bool debug_print) {
    return pycauchy_initialize_nonlin(num_steps, x0, size_x0, A0, size_A0, p0, size_p0, b0, size_b0, beta, size_beta,
                                      gamma, size_gamma, f_dyn, f_msmt, f_ext, cmcc, dt, init_step, debug_print);
}

void refpy_single_step_nonlin(void* h, double* msmts, int size_msmts, double* controls, int size_controls,
                              bool with_propagate, double** o1, int* s1, double** o2, int* s2, double** o3, int* s3,
                              double** o4, int* s4, double** o5, int* s5, double** o6, int* s6, double** o7, int* s7,
                              double** o8, int* s8, double** o9, int* s9, double** o10, int* s10, double** o11,
                              int* s11, double** o12, int* s12, double** o13, int* s13, double** o14, int* s14,
                              int** o15, int* s15) {
    pycauchy_single_step_nonlin(h, msmts, size_msmts, controls, size_controls, with_propagate, o1, s1, o2, s2, o3, s3,
                                o4, s4, o5, s5, o6, s6, o7, s7, o8, s8, o9, s9, o10, s10, o11, s11, o12, s12, o13,
                                s13, o14, s14, o15, s15);
}

void refpy_get_marginal_1D_pointwise_cpdf(void* h, int marg_idx1, double lo, double hi, double res, char* log_dir,
                                          double** out, int* size, int* num) {
    pycauchy_get_marginal_1D_pointwise_cpdf(h, marg_idx1, lo, hi, res, log_dir, out, size, num);
}

void refpy_get_1D_pointwise_cpdf(void* h, double lo, double hi, double res, char* log_dir, double** out, int* size,
                                 int* num) {
    pycauchy_get_1D_pointwise_cpdf(h, lo, hi, res, log_dir, out, size, num);
}

void refpy_get_marginal_2D_pointwise_cpdf(void* h, int i1, int i2, double lox, double hix, double resx, double loy,
                                          double hiy, double resy, char* log_dir, bool reset_cache, double** out,
                                          int* size, int* nx, int* ny) {
    pycauchy_get_marginal_2D_pointwise_cpdf(h, i1, i2, lox, hix, resx, loy, hiy, resy, log_dir, reset_cache, out, size,
                                            nx, ny);
}

void refpy_single_step_eval_2d_rsys_cpdf(double* Trel, int sT, void* sh, void* ph, double eps, double xlo,
                                         double xhi, double dx, double ylo, double yhi, double dy, double** o1,
                                         int* s1, double** o2, int* s2, double** o3, int* s3, double** o4, int* s4,
                                         double** o5, int* s5, double** o6, int* s6, double** o7, int* s7, int* nx,
                                         int* ny) {
    pycauchy_single_step_eval_2d_rsys_cpdf(Trel, sT, sh, ph, eps, xlo, xhi, dx, ylo, yhi, dy, o1, s1, o2, s2, o3, s3,
                                           o4, s4, o5, s5, o6, s6, o7, s7, nx, ny);
}

void refpy_single_step_reset(void* h, double* A0, int sA0, double* p0, int sp0, double* b0, int sb0, double* xbar,
                             int sxbar) {
    pycauchy_single_step_reset(h, A0, sA0, p0, sp0, b0, sb0, xbar, sxbar);
}

void refpy_single_step_set_master_step(void* h, int step) { pycauchy_single_step_set_master_step(h, step); }
void refpy_single_step_set_window_number(void* h, int w) { pycauchy_single_step_set_window_number(h, w); }
int refpy_single_step_get_number_of_terms(void* h) { return pycauchy_single_step_get_number_of_terms(h); }
void refpy_single_step_deterministic_transform(void* h, double* T, int sT, double* b, int sb) {
    pycauchy_single_step_deterministic_transform(h, T, sT, b, sb);
}
void* refpy_single_step_get_duc(void* h) { return pycauchy_single_step_get_duc(h); }
void refpy_single_step_shutdown(void* h) { pycauchy_single_step_shutdown(h); }
int refpy_set_tr_search_idxs_ordering(int* ordering, int n) { return pycauchy_set_tr_search_idxs_ordering(ordering, n); }

void refpy_speyers_window_init(double* xhat, int sx, double* Phat, int sP, double* H, int sH, double gamma, double z,
                               double** oA, int* sA, double** op, int* sp, double** ob, int* sb) {
    pycauchy_speyers_window_init(xhat, sx, Phat, sP, H, sH, gamma, z, oA, sA, op, sp, ob, sb);
}

void refpy_get_reinitialization_statistics(void* h, double z, double* xhat, int sx, double* Phat, int sP, double* H,
                                           int sH, double gamma, double** oA, int* sA, double** op, int* sp,
                                           double** ob, int* sb) {
    pycauchy_get_reinitialization_statistics(h, z, xhat, sx, Phat, sP, H, sH, gamma, oA, sA, op, sp, ob, sb);
}

// The reference's forked window manager (pycauchy.hpp:80-221): every window a child process.
void refpy_initialize_lti_window_manager(int nw, int nsim, double* A0, int sA0, double* p0, int sp0, double* b0,
                                         int sb0, double* Phi, int sPhi, double* Gamma, int sGamma, double* B, int sB,
                                         double* beta, int sbeta, double* H, int sH, double* gamma, int sgamma,
                                         bool dbg, bool lseq, bool lfull, char* ldir, double dt, int init_step,
                                         double* wvb, int swvb) {
    pycauchy_initialize_lti_window_manager(nw, nsim, A0, sA0, p0, sp0, b0, sb0, Phi, sPhi, Gamma, sGamma, B, sB, beta,
                                           sbeta, H, sH, gamma, sgamma, dbg, lseq, lfull, ldir, dt, init_step, wvb,
                                           swvb);
}

void refpy_initialize_ltv_window_manager(int nw, int nsim, double* A0, int sA0, double* p0, int sp0, double* b0,
                                         int sb0, double* Phi, int sPhi, double* Gamma, int sGamma, double* B, int sB,
                                         double* beta, int sbeta, double* H, int sH, double* gamma, int sgamma,
                                         void (*cb)(CauchyDynamicsUpdateContainer*), bool dbg, bool lseq, bool lfull,
                                         char* ldir, double dt, int init_step, double* wvb, int swvb) {
    pycauchy_initialize_ltv_window_manager(nw, nsim, A0, sA0, p0, sp0, b0, sb0, Phi, sPhi, Gamma, sGamma, B, sB, beta,
                                           sbeta, H, sH, gamma, sgamma, cb, dbg, lseq, lfull, ldir, dt, init_step, wvb,
                                           swvb);
}

void refpy_initialize_nonlin_window_manager(int nw, int nsim, double* x0, int sx0, double* A0, int sA0, double* p0,
                                            int sp0, double* b0, int sb0, double* beta, int sbeta, double* gamma,
                                            int sgamma, void (*f_dyn)(CauchyDynamicsUpdateContainer*),
                                            void (*f_msmt)(CauchyDynamicsUpdateContainer*, double*),
                                            void (*f_ext)(CauchyDynamicsUpdateContainer*), int cmcc, bool dbg,
                                            bool lseq, bool lfull, char* ldir, double dt, int init_step, double* wvb,
                                            int swvb) {
    pycauchy_initialize_nonlin_window_manager(nw, nsim, x0, sx0, A0, sA0, p0, sp0, b0, sb0, beta, sbeta, gamma,
                                              sgamma, f_dyn, f_msmt, f_ext, cmcc, dbg, lseq, lfull, ldir, dt,
                                              init_step, wvb, swvb);
}

void refpy_step(double* msmts, int sm, double* controls, int sc, double* fz, double** x, int* sx, double** P, int* sP,
                double* cfz, double* cx, double* cP, int* win, int* err) {
    pycauchy_step(msmts, sm, controls, sc, fz, x, sx, P, sP, cfz, cx, cP, win, err);
}

void refpy_shutdown() { pycauchy_shutdown(); }

}  // extern "C"
