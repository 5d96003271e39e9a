"""ctypes stand-in for the reference's SWIG module `pycauchy`.

Same function names, argument order and return tuples as the SWIG bindings of
scripts/swig/cauchy/pycauchy.i (numpy.i typemaps: an input `double* x, int
size_x` pair is one numpy array; `double** out, int* size_out` pairs are
returned as numpy arrays that own malloc'd memory), so the reference's
`cauchy_estimator.py` (PyCauchyEstimator) runs unchanged on the HIP engine.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _native

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)


def _arr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    return a, a.ctypes.data_as(_dp), int(a.size)


def _take_d(ptr, n):
    """Copies a library-malloc'd double array into numpy and frees it."""
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 0),)).copy() if n > 0 else np.zeros(0)
    _native.free(ptr)
    return out


def _take_i(ptr, n):
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 0),)).copy() if n > 0 else np.zeros(0, dtype=np.int32)
    _native.free(ptr)
    return out


def _check(handle, what):
    if not handle:
        msg = _native.lib().mce_last_error().decode()
        raise RuntimeError(f"{what} failed: {msg}")
    return handle


def pycauchy_initialize_lti(num_steps, A0, p0, b0, Phi, Gamma, B, beta, H, gamma, dt, init_step, debug_print):
    L = _native.lib()
    args = [_arr(v) for v in (A0, p0, b0, Phi, Gamma, B, beta, H, gamma)]
    flat = []
    for a in args:
        flat += [a[1], a[2]]
    h = L.pycauchy_initialize_lti(int(num_steps), *flat, float(dt), int(init_step), bool(debug_print))
    return _check(h, "pycauchy_initialize_lti")


def pycauchy_initialize_ltv(num_steps, A0, p0, b0, Phi, Gamma, B, beta, H, gamma, f_duc_ptr, dt, init_step,
                            debug_print):
    L = _native.lib()
    args = [_arr(v) for v in (A0, p0, b0, Phi, Gamma, B, beta, H, gamma)]
    flat = []
    for a in args:
        flat += [a[1], a[2]]
    if not isinstance(f_duc_ptr, int):
        f_duc_ptr = ct.cast(f_duc_ptr, ct.c_void_p).value
    h = L.pycauchy_initialize_ltv(int(num_steps), *flat, ct.c_void_p(f_duc_ptr), float(dt), int(init_step),
                                  bool(debug_print))
    return _check(h, "pycauchy_initialize_ltv")


def pycauchy_single_step_ltiv(handle, msmts, controls):
    """Returns (Phi, Gamma, B, H, beta, gamma, fz, xhat, Phat, cerr_fz, cerr_xhat, cerr_Phat, err_code)."""
    L = _native.lib()
    m_ = _arr(msmts)
    c_ = _arr(controls)
    outs_d = [(_dp(), ct.c_int()) for _ in range(12)]
    out_i = (_ip(), ct.c_int())
    args = [ct.c_void_p(handle), m_[1], m_[2], c_[1] if c_[2] > 0 else None, c_[2]]
    for p, n in outs_d:
        args += [ct.byref(p), ct.byref(n)]
    args += [ct.byref(out_i[0]), ct.byref(out_i[1])]
    L.pycauchy_single_step_ltiv(*args)
    res = [_take_d(p, n.value) for p, n in outs_d]
    res.append(_take_i(out_i[0], out_i[1].value))
    return tuple(res)


def pycauchy_single_step_reset(handle, A0, p0, b0, xbar):
    L = _native.lib()
    a = [_arr(v) for v in (A0, p0, b0, xbar)]
    L.pycauchy_single_step_reset(ct.c_void_p(handle), a[0][1], a[0][2], a[1][1], a[1][2], a[2][1], a[2][2],
                                 a[3][1], a[3][2])


def pycauchy_single_step_set_master_step(handle, step):
    _native.lib().pycauchy_single_step_set_master_step(ct.c_void_p(handle), int(step))


def pycauchy_single_step_set_window_number(handle, win_num):
    _native.lib().pycauchy_single_step_set_window_number(ct.c_void_p(handle), int(win_num))


def pycauchy_single_step_get_number_of_terms(handle):
    return int(_native.lib().pycauchy_single_step_get_number_of_terms(ct.c_void_p(handle)))


def pycauchy_single_step_deterministic_transform(handle, Trans, bias):
    t = _arr(Trans)
    b = _arr(bias)
    _native.lib().pycauchy_single_step_deterministic_transform(ct.c_void_p(handle), t[1], t[2], b[1], b[2])


def pycauchy_single_step_get_duc(handle):
    return _native.lib().pycauchy_single_step_get_duc(ct.c_void_p(handle))


def pycauchy_single_step_shutdown(handle):
    _native.lib().pycauchy_single_step_shutdown(ct.c_void_p(handle))


def pycauchy_set_tr_search_idxs_ordering(ordering):
    o = np.ascontiguousarray(np.asarray(ordering, dtype=np.int32).reshape(-1))
    return int(_native.lib().pycauchy_set_tr_search_idxs_ordering(o.ctypes.data_as(_ip), int(o.size)))
