/* SWIG interface: a Python module named `pycauchy` over this library's C ABI
 * (cauchy_mi355x.h), for machines that have SWIG and numpy.i.  It gives the functions,
 * array arguments and returned arrays of the reference's SWIG module
 * (scripts/swig/cauchy/pycauchy.i), so the reference's scripts/swig/cauchy/cauchy_estimator.py
 * imports it unchanged; without SWIG, cauchyfriendly_amd/pycauchy.py is the same module over
 * ctypes.
 *
 * Build (SWIG 4; numpy.i from numpy's tools/swig directory next to this file):
 *   swig -python -o pycauchy_wrap.c include/pycauchy_mi355x.i
 *   gcc -shared -fPIC -O2 pycauchy_wrap.c -Iinclude \
 *       -I$(python3 -c 'import sysconfig; print(sysconfig.get_paths()["include"])') \
 *       -I$(python3 -c 'import numpy; print(numpy.get_include())') \
 *       -Lcauchyfriendly_amd -lcauchy_mi355x -Wl,-rpath,$PWD/cauchyfriendly_amd -o _pycauchy.so
 * then put _pycauchy.so and the generated pycauchy.py on the Python path.
 *
 * Conventions (the reference's argument names, so each typemap applies by name):
 *   (double* x, int size_x)          numpy array in
 *   (double** out_x, int* size_out_x) numpy array out, malloc'd by the library and freed by
 *                                     numpy (ARGOUTVIEWM)
 *   int* out_num_grid*               returned scalar
 *   callbacks                        ctypes CFUNCTYPE objects passed as ints (their address)
 */
%module pycauchy
%{
#define SWIG_FILE_WITH_INIT
#include "cauchy_mi355x.h"
%}

%include "typemaps.i"
%include "numpy.i"
%init %{
import_array();
%}

/* array inputs */
%apply (double* IN_ARRAY1, int DIM1) {
    (double* msmts, int size_msmts), (double* controls, int size_controls), (double* x0, int size_x0),
    (double* A0, int size_A0), (double* p0, int size_p0), (double* b0, int size_b0), (double* Phi, int size_Phi),
    (double* Gamma, int size_Gamma), (double* B, int size_B), (double* beta, int size_beta),
    (double* H, int size_H), (double* gamma, int size_gamma), (double* xbar, int size_xbar),
    (double* xhat, int size_xhat), (double* Phat, int size_Phat), (double* Trans, int size_Trans),
    (double* bias, int size_bias), (double* Trel, int size_Trel), (double* win_var_boost, int size_wvb)
};
%apply (int* IN_ARRAY1, int DIM1) { (int* ordering, int size_ordering) };

/* array outputs */
%apply (double** ARGOUTVIEWM_ARRAY1, int* DIM1) {
    (double** out_A0, int* size_out_A0), (double** out_p0, int* size_out_p0), (double** out_b0, int* size_out_b0),
    (double** out_Phi, int* size_out_Phi), (double** out_Gamma, int* size_out_Gamma),
    (double** out_B, int* size_out_B), (double** out_H, int* size_out_H), (double** out_beta, int* size_out_beta),
    (double** out_gamma, int* size_out_gamma), (double** out_fz, int* size_out_fz),
    (double** out_xhat, int* size_out_xhat), (double** out_Phat, int* size_out_Phat),
    (double** out_xbar, int* size_out_xbar), (double** out_zbar, int* size_out_zbar),
    (double** out_cerr_fz, int* size_out_cerr_fz), (double** out_cerr_xhat, int* size_out_cerr_xhat),
    (double** out_cerr_Phat, int* size_out_cerr_Phat), (double** out_cpdf_data, int* size_out_cpdf_data),
    (double** out_rsys_fz, int* size_out_rsys_fz), (double** out_rsys_xhat, int* size_out_rsys_xhat),
    (double** out_rsys_Phat, int* size_out_rsys_Phat), (double** out_rsys_cerr_fz, int* size_out_rsys_cerr_fz),
    (double** out_rsys_cerr_xhat, int* size_out_rsys_cerr_xhat),
    (double** out_rsys_cerr_Phat, int* size_out_rsys_cerr_Phat)
};
%apply (int** ARGOUTVIEWM_ARRAY1, int* DIM1) { (int** out_err_code, int* size_out_err_code) };

/* scalar outputs */
%apply int* OUTPUT { int* out_num_gridx, int* out_num_gridy };
%apply double* OUTPUT { double* out_swm_fz, double* out_swm_cerr_xhat, double* out_swm_cerr_Phat, double* out_swm_cerr_fz };
%apply int* OUTPUT { int* out_swm_win_idx, int* out_swm_err_code };

/* callbacks: the address of a ctypes function object */
%typemap(in) void (*f_dyn_update_callback)(CauchyDynamicsUpdateContainer*) {
    $1 = (void (*)(CauchyDynamicsUpdateContainer*))PyLong_AsVoidPtr($input);
}
%typemap(in) void (*f_nonlinear_msmt_model)(CauchyDynamicsUpdateContainer*, double*) {
    $1 = (void (*)(CauchyDynamicsUpdateContainer*, double*))PyLong_AsVoidPtr($input);
}
%typemap(in) void (*f_extended_msmt_update_callback)(CauchyDynamicsUpdateContainer*) {
    $1 = (void (*)(CauchyDynamicsUpdateContainer*))PyLong_AsVoidPtr($input);
}

/* This library's additions that take arrays of pointers or unsized buffers stay C-only
 * (Python reaches them through cauchyfriendly_amd); the module is the reference's surface. */
%ignore pycauchy_batch_single_step_ltiv;
%ignore pycauchy_batch_single_step_ltiv_merged;
%ignore pycauchy_batch_single_step_nonlin;
%ignore mce_estimator_create;
%ignore mce_estimator_step;
%ignore mce_estimator_moments;
%ignore mce_estimator_reset;
%ignore mce_estimator_destroy;
%ignore mce_estimator_finalize_extended_moments;
%ignore mce_cpdf1d_noalloc;
%ignore mce_batch_profile_enable;
%ignore mce_batch_profile_read;
%ignore mce_estimator_enable_tangents;
%ignore mce_estimator_step_tangent;
%ignore mce_last_update;
%ignore mce_shape_range;
%ignore mce_terms_generated;
%ignore mce_term;
%ignore mce_set_random_draws;
%ignore mce_get_random_draws;
%ignore mce_profile_enable;
%ignore mce_profile_read;
%ignore mce_profile_reset;
%ignore mce_estimator_set_shapes;
%ignore mce_last_fz_raw;
%ignore mce_window_bank_set_mode;
%ignore mce_step_record;
%ignore mce_collect_outputs;

%include "cauchy_mi355x.h"
