"""Build and load the HIP engine (libcauchy_mi355x.so, gfx950).

The product path has no CPU fallback: `lib()` raises when the shared library
is missing or fails to load.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libcauchy_mi355x.so")
SOURCES = [os.path.join(PKG_DIR, "csrc", "mce_engine.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCE_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps every multiply and add separately rounded, as in the
# reference's x86-64 build; packed FP32 instructions are disabled for gfx950.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    """Compiles the engine in-tree; returns the library path."""
    cmd = [HIPCC] + HIP_FLAGS + ["-o", LIB_PATH + ".tmp"] + SOURCES
    res = subprocess.run(cmd, cwd=REPO_DIR, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(" ".join(cmd))
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


_lib = None
_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_dpp = ct.POINTER(_dp)
_ipp = ct.POINTER(_ip)

DYN_CALLBACK = ct.CFUNCTYPE(None, ct.c_void_p)



def declare_pycauchy(L, sym=lambda name: name):
    """Declares the argument and result types of the pycauchy_* C ABI on the CDLL L;
    sym maps a pycauchy_* name to the symbol L exports for it."""
    init_args = [ct.c_int] + [_dp, ct.c_int] * 9
    getattr(L, sym("pycauchy_initialize_lti")).restype = ct.c_void_p
    getattr(L, sym("pycauchy_initialize_lti")).argtypes = init_args + [ct.c_double, ct.c_int, ct.c_bool]
    getattr(L, sym("pycauchy_initialize_ltv")).restype = ct.c_void_p
    getattr(L, sym("pycauchy_initialize_ltv")).argtypes = init_args + [ct.c_void_p, ct.c_double, ct.c_int, ct.c_bool]
    getattr(L, sym("pycauchy_single_step_ltiv")).restype = None
    getattr(L, sym("pycauchy_single_step_ltiv")).argtypes = [ct.c_void_p, _dp, ct.c_int, _dp, ct.c_int] + \
        [_dpp, _ip] * 12 + [_ipp, _ip]
    getattr(L, sym("pycauchy_initialize_nonlin")).restype = ct.c_void_p
    getattr(L, sym("pycauchy_initialize_nonlin")).argtypes = [ct.c_int] + [_dp, ct.c_int] * 6 + \
        [ct.c_void_p] * 3 + [ct.c_int, ct.c_double, ct.c_int, ct.c_bool]
    getattr(L, sym("pycauchy_single_step_nonlin")).restype = None
    getattr(L, sym("pycauchy_single_step_nonlin")).argtypes = [ct.c_void_p, _dp, ct.c_int, _dp, ct.c_int, ct.c_bool] + \
        [_dpp, _ip] * 14 + [_ipp, _ip]
    if hasattr(L, sym("pycauchy_batch_single_step_ltiv")):  # this library's batched stepping
        getattr(L, sym("pycauchy_batch_single_step_ltiv")).restype = ct.c_int
        getattr(L, sym("pycauchy_batch_single_step_ltiv")).argtypes = [ct.POINTER(ct.c_void_p), ct.c_int, _dpp, _ip,
                                                                       _dpp, _ip, ct.POINTER(ct.c_void_p), _ip,
                                                                       ct.c_int]
        if hasattr(L, sym("pycauchy_batch_single_step_ltiv_merged")):
            getattr(L, sym("pycauchy_batch_single_step_ltiv_merged")).restype = ct.c_int
            getattr(L, sym("pycauchy_batch_single_step_ltiv_merged")).argtypes = \
                getattr(L, sym("pycauchy_batch_single_step_ltiv")).argtypes
        getattr(L, sym("pycauchy_batch_single_step_nonlin")).restype = ct.c_int
        getattr(L, sym("pycauchy_batch_single_step_nonlin")).argtypes = [ct.POINTER(ct.c_void_p), ct.c_int, _dpp,
                                                                         _ip, _dpp, _ip, _ip,
                                                                         ct.POINTER(ct.c_void_p), _ip, ct.c_int]
    if hasattr(L, sym("pycauchy_step")):  # the window-manager set (pycauchy.hpp:80-221)
        tail = [ct.c_bool] * 3 + [ct.c_char_p, ct.c_double, ct.c_int, _dp, ct.c_int]
        getattr(L, sym("pycauchy_initialize_lti_window_manager")).restype = None
        getattr(L, sym("pycauchy_initialize_lti_window_manager")).argtypes = [ct.c_int, ct.c_int] + \
            [_dp, ct.c_int] * 9 + tail
        getattr(L, sym("pycauchy_initialize_ltv_window_manager")).restype = None
        getattr(L, sym("pycauchy_initialize_ltv_window_manager")).argtypes = [ct.c_int, ct.c_int] + \
            [_dp, ct.c_int] * 9 + [ct.c_void_p] + tail
        getattr(L, sym("pycauchy_initialize_nonlin_window_manager")).restype = None
        getattr(L, sym("pycauchy_initialize_nonlin_window_manager")).argtypes = [ct.c_int, ct.c_int] + \
            [_dp, ct.c_int] * 6 + [ct.c_void_p] * 3 + [ct.c_int] + tail
        getattr(L, sym("pycauchy_step")).restype = None
        getattr(L, sym("pycauchy_step")).argtypes = [_dp, ct.c_int, _dp, ct.c_int, _dp, _dpp, _ip, _dpp, _ip, _dp, _dp,
                                                     _dp, _ip, _ip]
        getattr(L, sym("pycauchy_shutdown")).restype = None
        getattr(L, sym("pycauchy_shutdown")).argtypes = []
    getattr(L, sym("pycauchy_get_marginal_1D_pointwise_cpdf")).restype = None
    getattr(L, sym("pycauchy_get_marginal_1D_pointwise_cpdf")).argtypes = [ct.c_void_p, ct.c_int] + [ct.c_double] * 3 + \
        [ct.c_char_p, _dpp, _ip, _ip]
    getattr(L, sym("pycauchy_get_marginal_2D_pointwise_cpdf")).restype = None
    getattr(L, sym("pycauchy_get_marginal_2D_pointwise_cpdf")).argtypes = [ct.c_void_p, ct.c_int, ct.c_int] + \
        [ct.c_double] * 6 + [ct.c_char_p, ct.c_bool, _dpp, _ip, _ip, _ip]
    getattr(L, sym("pycauchy_single_step_eval_2d_rsys_cpdf")).restype = None
    getattr(L, sym("pycauchy_single_step_eval_2d_rsys_cpdf")).argtypes = [_dp, ct.c_int, ct.c_void_p, ct.c_void_p] + \
        [ct.c_double] * 7 + [_dpp, _ip] * 7 + [_ip, _ip]
    getattr(L, sym("pycauchy_single_step_reset")).argtypes = [ct.c_void_p] + [_dp, ct.c_int] * 4
    getattr(L, sym("pycauchy_single_step_set_master_step")).argtypes = [ct.c_void_p, ct.c_int]
    getattr(L, sym("pycauchy_single_step_set_window_number")).argtypes = [ct.c_void_p, ct.c_int]
    getattr(L, sym("pycauchy_single_step_get_number_of_terms")).restype = ct.c_int
    getattr(L, sym("pycauchy_single_step_get_number_of_terms")).argtypes = [ct.c_void_p]
    getattr(L, sym("pycauchy_single_step_deterministic_transform")).argtypes = [ct.c_void_p, _dp, ct.c_int, _dp, ct.c_int]
    getattr(L, sym("pycauchy_single_step_get_duc")).restype = ct.c_void_p
    getattr(L, sym("pycauchy_single_step_get_duc")).argtypes = [ct.c_void_p]
    getattr(L, sym("pycauchy_single_step_shutdown")).argtypes = [ct.c_void_p]
    getattr(L, sym("pycauchy_set_tr_search_idxs_ordering")).restype = ct.c_int
    getattr(L, sym("pycauchy_set_tr_search_idxs_ordering")).argtypes = [_ip, ct.c_int]
    getattr(L, sym("pycauchy_speyers_window_init")).restype = None
    getattr(L, sym("pycauchy_speyers_window_init")).argtypes = [_dp, ct.c_int] * 3 + [ct.c_double, ct.c_double] + [_dpp, _ip] * 3
    getattr(L, sym("pycauchy_get_reinitialization_statistics")).restype = None
    getattr(L, sym("pycauchy_get_reinitialization_statistics")).argtypes = [ct.c_void_p, ct.c_double] + [_dp, ct.c_int] * 3 + \
        [ct.c_double] + [_dpp, _ip] * 3


def lib():
    """The loaded engine; raises if the HIP library is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP engine {LIB_PATH} is not built; run __graft_entry__.build()")
    L = ct.CDLL(LIB_PATH)
    declare_pycauchy(L)
    L.mce_estimator_create.restype = ct.c_void_p
    L.mce_estimator_create.argtypes = [_dp, _dp, _dp] + [ct.c_int] * 5 + [ct.c_bool]
    L.mce_estimator_step.restype = ct.c_int
    L.mce_estimator_step.argtypes = [ct.c_void_p, ct.c_double, _dp, _dp, _dp, _dp, ct.c_double, _dp, _dp]
    L.mce_estimator_moments.restype = ct.c_int
    L.mce_estimator_moments.argtypes = [ct.c_void_p, _dp, _dp, _dp]
    L.mce_estimator_reset.argtypes = [ct.c_void_p]
    L.mce_estimator_destroy.argtypes = [ct.c_void_p]
    L.mce_estimator_finalize_extended_moments.restype = ct.c_int
    L.mce_estimator_finalize_extended_moments.argtypes = [ct.c_void_p, _dp]
    L.mce_last_update.restype = ct.c_int
    L.mce_last_update.argtypes = [ct.c_void_p, _dp, _ip]
    L.mce_shape_range.restype = ct.c_int
    L.mce_terms_generated.restype = ct.c_longlong
    L.mce_terms_generated.argtypes = [ct.c_void_p]
    L.mce_shape_range.argtypes = [ct.c_void_p]
    L.mce_term.restype = ct.c_int
    L.mce_term.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, _ip, _dp, _dp, _dp, ct.POINTER(ct.c_uint), _dp]
    L.mce_set_random_draws.argtypes = [ct.c_void_p, _dp, _dp]
    L.mce_get_random_draws.argtypes = [ct.c_void_p, _dp, _dp]
    L.mce_profile_enable.argtypes = [ct.c_void_p, ct.c_int]
    L.mce_profile_read.argtypes = [ct.c_void_p, _dp]
    L.mce_profile_reset.argtypes = [ct.c_void_p]
    L.mce_last_fz_raw.restype = None
    L.mce_last_fz_raw.argtypes = [ct.c_void_p, _dp]
    L.mce_window_bank_set_mode.restype = ct.c_int
    L.mce_window_bank_set_mode.argtypes = [ct.c_int]
    L.mce_estimator_set_shapes.restype = ct.c_int
    L.mce_estimator_set_shapes.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int]
    L.mce_batch_profile_enable.restype = None
    L.mce_batch_profile_enable.argtypes = [ct.c_int]
    L.mce_batch_profile_read.restype = None
    L.mce_batch_profile_read.argtypes = [_dp]
    L.mce_last_error.restype = ct.c_char_p
    L.mce_step_record.restype = ct.c_int
    L.mce_step_record.argtypes = [ct.c_void_p, _dp, _ip, ct.c_int]
    L.mce_collect_outputs.restype = ct.c_longlong
    L.mce_collect_outputs.argtypes = [ct.POINTER(ct.c_void_p), _ip, ct.c_longlong, ct.c_int, _dp, _ip]
    L.mce_estimator_enable_tangents.restype = ct.c_int
    L.mce_estimator_enable_tangents.argtypes = [ct.c_void_p, ct.c_int]
    L.mce_estimator_step_tangent.restype = ct.c_int
    L.mce_estimator_step_tangent.argtypes = [ct.c_void_p, ct.c_double, _dp, _dp, _dp, _dp, ct.c_double, _dp, _dp,
                                             _dp, _dp, _dp, _dp, _dp]
    L.mce_cpdf1d_noalloc.restype = ct.c_int
    L.mce_cpdf1d_noalloc.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_void_p,
                                     ct.c_void_p, ct.c_int, ct.c_void_p, _ip]
    _lib = L
    return L


def libc():
    return ct.CDLL("libc.so.6")


def free(ptr):
    libc().free(ct.cast(ptr, ct.c_void_p))
