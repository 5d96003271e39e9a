"""ctypes stand-in for the reference's SWIG module `pycauchy`.

Same function names, argument order and return tuples as the SWIG bindings of
scripts/swig/cauchy/pycauchy.i (numpy.i typemaps: an input `double* x, int
size_x` pair is one numpy array; `double** out, int* size_out` pairs are
returned as numpy arrays that own malloc'd memory), so the reference's
`cauchy_estimator.py` (PyCauchyEstimator, PySlidingWindowManager) runs
unchanged on the HIP engine.

The module-level functions are bound to the MI355X library.  `bind(cdll, sym)`
returns the same functions bound to any other library that exports this C ABI
(sym maps a pycauchy_* name to the library's symbol).
"""
from __future__ import annotations

import ctypes as ct
import types

import numpy as np

from . import _native

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)


_EMPTY_D = np.zeros(0)
_EMPTY_D.flags.writeable = False


def _arr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    return a, a.ctypes.data_as(_dp), int(a.size)


def _packed(arrays):
    """One contiguous copy of the arrays and, per array, a pointer into it (NULL when empty)
    and its size; the copy must outlive the call."""
    flat = [np.asarray(a, dtype=np.float64).reshape(-1) for a in arrays]
    n = np.array([a.size for a in flat], dtype=np.int64)
    cat = np.concatenate(flat) if len(flat) else np.zeros(0)
    if cat.size == 0:
        cat = np.zeros(1)
    addr = cat.ctypes.data + 8 * (np.cumsum(n) - n)
    ptrs = (ct.c_void_p * len(flat))(*[int(a) if k else None for a, k in zip(addr.tolist(), n.tolist())])
    return cat, ct.cast(ptrs, ct.POINTER(_dp)), (ct.c_int * len(flat))(*n.tolist())


def _take_d(ptr, n):
    """Copies a library-malloc'd double array into numpy and frees it."""
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 0),)).copy() if n > 0 else np.zeros(0)
    _native.free(ptr)
    return out


def _take_i(ptr, n):
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 0),)).copy() if n > 0 else np.zeros(0, dtype=np.int32)
    _native.free(ptr)
    return out


def _three_outputs(call):
    """Calls a function whose last six arguments are three (double**, int*) outputs."""
    outs = [_dp() for _ in range(3)]
    sizes = [ct.c_int(0) for _ in range(3)]
    flat = []
    for o, n in zip(outs, sizes):
        flat += [ct.byref(o), ct.byref(n)]
    call(*flat)
    return tuple(_take_d(o, n.value) for o, n in zip(outs, sizes))


class _Bindings:
    """The SWIG module's functions over one library."""

    def __init__(self, get_lib, sym=lambda name: name, last_error=None):
        self._get_lib = get_lib
        self._sym = sym
        self._last_error = last_error

    def _f(self, name):
        return getattr(self._get_lib(), self._sym(name))

    def _check(self, handle, what):
        if not handle:
            msg = self._last_error() if self._last_error else ""
            raise RuntimeError(f"{what} failed: {msg}")
        return handle

    def pycauchy_initialize_lti(self, num_steps, A0, p0, b0, Phi, Gamma, B, beta, H, gamma, dt, init_step,
                                debug_print):
        args = [_arr(v) for v in (A0, p0, b0, Phi, Gamma, B, beta, H, gamma)]
        flat = []
        for a in args:
            flat += [a[1], a[2]]
        h = self._f("pycauchy_initialize_lti")(int(num_steps), *flat, float(dt), int(init_step), bool(debug_print))
        return self._check(h, "pycauchy_initialize_lti")

    def pycauchy_initialize_ltv(self, num_steps, A0, p0, b0, Phi, Gamma, B, beta, H, gamma, f_duc_ptr, dt,
                                init_step, debug_print):
        args = [_arr(v) for v in (A0, p0, b0, Phi, Gamma, B, beta, H, gamma)]
        flat = []
        for a in args:
            flat += [a[1], a[2]]
        if not isinstance(f_duc_ptr, int):
            f_duc_ptr = ct.cast(f_duc_ptr, ct.c_void_p).value
        h = self._f("pycauchy_initialize_ltv")(int(num_steps), *flat, ct.c_void_p(f_duc_ptr), float(dt),
                                               int(init_step), bool(debug_print))
        return self._check(h, "pycauchy_initialize_ltv")

    def pycauchy_single_step_ltiv(self, handle, msmts, controls):
        """Returns (Phi, Gamma, B, H, beta, gamma, fz, xhat, Phat, cerr_fz, cerr_xhat, cerr_Phat, err_code)."""
        m_ = _arr(msmts)
        c_ = _arr(controls)
        outs_d = [(_dp(), ct.c_int()) for _ in range(12)]
        out_i = (_ip(), ct.c_int())
        args = [ct.c_void_p(handle), m_[1], m_[2], c_[1] if c_[2] > 0 else None, c_[2]]
        for p, n in outs_d:
            args += [ct.byref(p), ct.byref(n)]
        args += [ct.byref(out_i[0]), ct.byref(out_i[1])]
        self._f("pycauchy_single_step_ltiv")(*args)
        res = [_take_d(p, n.value) for p, n in outs_d]
        res.append(_take_i(out_i[0], out_i[1].value))
        return tuple(res)

    def pycauchy_initialize_nonlin(self, num_steps, x0, A0, p0, b0, beta, gamma, f_duc_ptr1, f_duc_ptr2, f_duc_ptr3,
                                   cmcc, dt, init_step, debug_print):
        args = [_arr(v) for v in (x0, A0, p0, b0, beta, gamma)]
        flat = []
        for a in args:
            flat += [a[1], a[2]]
        cbs = [f if isinstance(f, int) else ct.cast(f, ct.c_void_p).value for f in (f_duc_ptr1, f_duc_ptr2, f_duc_ptr3)]
        h = self._f("pycauchy_initialize_nonlin")(int(num_steps), *flat, *[ct.c_void_p(f) for f in cbs], int(cmcc),
                                                  float(dt), int(init_step), bool(debug_print))
        return self._check(h, "pycauchy_initialize_nonlin")

    def pycauchy_single_step_nonlin(self, handle, msmts, controls, with_propagate):
        """Returns (Phi, Gamma, B, H, beta, gamma, fz, xhat, Phat, xbar, zbar, cerr_fz, cerr_xhat, cerr_Phat,
        err_code)."""
        m_ = _arr(msmts)
        c_ = _arr(controls)
        outs_d = [(_dp(), ct.c_int()) for _ in range(14)]
        out_i = (_ip(), ct.c_int())
        args = [ct.c_void_p(handle), m_[1], m_[2], c_[1] if c_[2] > 0 else None, c_[2], bool(with_propagate)]
        for p, n in outs_d:
            args += [ct.byref(p), ct.byref(n)]
        args += [ct.byref(out_i[0]), ct.byref(out_i[1])]
        self._f("pycauchy_single_step_nonlin")(*args)
        res = [_take_d(p, n.value) for p, n in outs_d]
        res.append(_take_i(out_i[0], out_i[1].value))
        return tuple(res)

    def _batch(self, fname, nout, handles, msmts, controls, extra, num_threads):
        count = len(handles)
        hs = (ct.c_void_p * count)(*handles)
        mcat, mp, mn = _packed(msmts)
        ccat, cp, cn = _packed(controls)
        outs = (ct.c_void_p * (nout * count))()
        sizes = (ct.c_int * (nout * count))()
        failed = self._f(fname)(hs, count, mp, mn, cp, cn, *extra, outs, sizes, int(num_threads))
        # every output copied (and freed) in one native call, then handed out as views (empty
        # outputs share one empty array)
        sz = np.frombuffer(sizes, dtype=np.int32).reshape(count, nout).clip(min=0)
        dsz = sz[:, :-1].reshape(-1)
        dbuf = np.empty(int(dsz.sum()))
        ibuf = np.empty(int(sz[:, -1].sum()), dtype=np.int32)
        _native.lib().mce_collect_outputs(outs, sizes, count, nout, dbuf.ctypes.data_as(_dp),
                                          ibuf.ctypes.data_as(_ip))
        ends = np.cumsum(dsz)
        dviews = [dbuf[a:b] if b > a else _EMPTY_D for a, b in zip((ends - dsz).tolist(), ends.tolist())]
        iends = np.cumsum(sz[:, -1]).tolist()
        istarts = [0] + iends[:-1]
        m = nout - 1
        res = [tuple(dviews[i * m:(i + 1) * m]) + (ibuf[istarts[i]:iends[i]],) for i in range(count)]
        if failed:
            msg = self._last_error() if self._last_error else ""
            raise RuntimeError(f"{fname}: {failed} estimator(s) failed: {msg}")
        return res

    def pycauchy_batch_single_step_ltiv(self, handles, msmts, controls, num_threads=0):
        """pycauchy_single_step_ltiv of every handle, concurrently; a list of its 13-tuples."""
        return self._batch("pycauchy_batch_single_step_ltiv", 13, handles, msmts, controls, (), num_threads)

    def pycauchy_batch_single_step_ltiv_merged(self, handles, msmts, controls, num_groups=4):
        """pycauchy_single_step_ltiv of every handle, with each kernel of the updates launched
        once per group of estimators (num_groups host threads); a list of its 13-tuples."""
        return self._batch("pycauchy_batch_single_step_ltiv_merged", 13, handles, msmts, controls, (), num_groups)

    def pycauchy_batch_single_step_nonlin(self, handles, msmts, controls, with_propagate, num_threads=0):
        """pycauchy_single_step_nonlin of every handle, concurrently; a list of its 15-tuples."""
        wp = (ct.c_int * len(handles))(*[1 if w else 0 for w in with_propagate])
        return self._batch("pycauchy_batch_single_step_nonlin", 15, handles, msmts, controls, (wp,), num_threads)

    def pycauchy_get_marginal_1D_pointwise_cpdf(self, handle, marg_idx, gridx_low, gridx_high, gridx_resolution,
                                                log_dir=None):
        """(cpdf_points as x, f(x) pairs, num_gridx) of the marginal density of one state."""
        out, size, num = _dp(), ct.c_int(0), ct.c_int(0)
        ld = None if log_dir is None else str(log_dir).encode()
        self._f("pycauchy_get_marginal_1D_pointwise_cpdf")(ct.c_void_p(handle), int(marg_idx), float(gridx_low),
                                                           float(gridx_high), float(gridx_resolution), ld,
                                                           ct.byref(out), ct.byref(size), ct.byref(num))
        return _take_d(out, size.value), int(num.value)

    def pycauchy_get_1D_pointwise_cpdf(self, handle, gridx_low, gridx_high, gridx_resolution, log_dir=None):
        return self.pycauchy_get_marginal_1D_pointwise_cpdf(handle, 0, gridx_low, gridx_high, gridx_resolution,
                                                            log_dir)

    def pycauchy_get_marginal_2D_pointwise_cpdf(self, handle, marg_idx1, marg_idx2, gridx_low, gridx_high,
                                                gridx_resolution, gridy_low, gridy_high, gridy_resolution,
                                                log_dir=None, reset_cache=True):
        """(cpdf_points as x, y, f triples, num_gridx, num_gridy) of the joint marginal density."""
        out, size, nx, ny = _dp(), ct.c_int(0), ct.c_int(0), ct.c_int(0)
        ld = None if log_dir is None else str(log_dir).encode()
        self._f("pycauchy_get_marginal_2D_pointwise_cpdf")(
            ct.c_void_p(handle), int(marg_idx1), int(marg_idx2), float(gridx_low), float(gridx_high),
            float(gridx_resolution), float(gridy_low), float(gridy_high), float(gridy_resolution), ld,
            bool(reset_cache), ct.byref(out), ct.byref(size), ct.byref(nx), ct.byref(ny))
        return _take_d(out, size.value), int(nx.value), int(ny.value)

    def pycauchy_get_2D_pointwise_cpdf(self, handle, gridx_low, gridx_high, gridx_resolution, gridy_low, gridy_high,
                                       gridy_resolution, log_dir=None, reset_cache=True):
        return self.pycauchy_get_marginal_2D_pointwise_cpdf(handle, 0, 1, gridx_low, gridx_high, gridx_resolution,
                                                            gridy_low, gridy_high, gridy_resolution, log_dir,
                                                            reset_cache)

    def pycauchy_single_step_eval_2d_rsys_cpdf(self, Trel, s_handle, p_handle, rsys_approx_eps, xlow, xhigh,
                                               delta_x, ylow, yhigh, delta_y):
        """(rsys_fz, rsys_xhat, rsys_Phat, rsys_cerr_fz, rsys_cerr_xhat, rsys_cerr_Phat, cpdf_points,
        num_gridx, num_gridy) of the relative system secondary - primary transformed by Trel."""
        t = _arr(Trel)
        outs = [(_dp(), ct.c_int()) for _ in range(7)]
        nx, ny = ct.c_int(0), ct.c_int(0)
        flat = []
        for p, n in outs:
            flat += [ct.byref(p), ct.byref(n)]
        self._f("pycauchy_single_step_eval_2d_rsys_cpdf")(
            t[1], t[2], ct.c_void_p(s_handle), ct.c_void_p(p_handle), float(rsys_approx_eps), float(xlow), float(xhigh),
            float(delta_x), float(ylow), float(yhigh), float(delta_y), *flat, ct.byref(nx), ct.byref(ny))
        res = [_take_d(p, n.value) for p, n in outs]
        return tuple(res) + (int(nx.value), int(ny.value))

    def pycauchy_single_step_reset(self, handle, A0, p0, b0, xbar):
        a = [_arr(v) for v in (A0, p0, b0, xbar)]
        self._f("pycauchy_single_step_reset")(ct.c_void_p(handle), a[0][1], a[0][2], a[1][1], a[1][2], a[2][1],
                                              a[2][2], a[3][1], a[3][2])

    def pycauchy_single_step_set_master_step(self, handle, step):
        self._f("pycauchy_single_step_set_master_step")(ct.c_void_p(handle), int(step))

    def pycauchy_single_step_set_window_number(self, handle, win_num):
        self._f("pycauchy_single_step_set_window_number")(ct.c_void_p(handle), int(win_num))

    def pycauchy_single_step_get_number_of_terms(self, handle):
        return int(self._f("pycauchy_single_step_get_number_of_terms")(ct.c_void_p(handle)))

    def pycauchy_single_step_deterministic_transform(self, handle, Trans, bias):
        t = _arr(Trans)
        b = _arr(bias)
        self._f("pycauchy_single_step_deterministic_transform")(ct.c_void_p(handle), t[1], t[2], b[1], b[2])

    def pycauchy_single_step_get_duc(self, handle):
        return self._f("pycauchy_single_step_get_duc")(ct.c_void_p(handle))

    def pycauchy_single_step_shutdown(self, handle):
        self._f("pycauchy_single_step_shutdown")(ct.c_void_p(handle))

    def pycauchy_set_tr_search_idxs_ordering(self, ordering):
        o = np.ascontiguousarray(np.asarray(ordering, dtype=np.int32).reshape(-1))
        return int(self._f("pycauchy_set_tr_search_idxs_ordering")(o.ctypes.data_as(_ip), int(o.size)))

    def pycauchy_speyers_window_init(self, xhat, Phat, H, gamma, z):
        """(A0, p0, b0) of a window restarted about (xhat, Phat) after measurement z
        (pycauchy.hpp:973-993)."""
        x, P, h = _arr(xhat), _arr(Phat), _arr(H)
        f = self._f("pycauchy_speyers_window_init")
        return _three_outputs(lambda *o: f(x[1], x[2], P[1], P[2], h[1], h[2], float(gamma), float(z), *o))

    def pycauchy_get_reinitialization_statistics(self, handle, z, xhat, Phat, H, gamma):
        """(A0, p0, b0) that restart an estimator about (xhat, Phat) after measurement z
        (pycauchy.hpp:948-971)."""
        x, P, h = _arr(xhat), _arr(Phat), _arr(H)
        f = self._f("pycauchy_get_reinitialization_statistics")
        return _three_outputs(lambda *o: f(ct.c_void_p(handle), float(z), x[1], x[2], P[1], P[2], h[1], h[2],
                                           float(gamma), *o))


    # ---- the window manager (pycauchy.hpp:80-221): one bank per process
    def _wm_tail(self, debug_print, log_seq, log_full, log_dir, dt, init_step, win_var_boost):
        w = _arr(win_var_boost if win_var_boost is not None else np.zeros(0))
        self._wm_keep = w
        ld = None if log_dir is None else str(log_dir).encode()
        return [bool(debug_print), bool(log_seq), bool(log_full), ld, float(dt), int(init_step), w[1], w[2]]

    @staticmethod
    def _fptr(f):
        return ct.c_void_p(f if isinstance(f, int) else ct.cast(f, ct.c_void_p).value)

    def pycauchy_initialize_lti_window_manager(self, num_windows, num_sim_steps, A0, p0, b0, Phi, Gamma, B, beta, H,
                                               gamma, debug_print, log_seq, log_full, log_dir, dt, init_step,
                                               win_var_boost):
        args = [_arr(v) for v in (A0, p0, b0, Phi, Gamma, B, beta, H, gamma)]
        flat = [x for a in args for x in (a[1], a[2])]
        self._f("pycauchy_initialize_lti_window_manager")(
            int(num_windows), int(num_sim_steps), *flat,
            *self._wm_tail(debug_print, log_seq, log_full, log_dir, dt, init_step, win_var_boost))

    def pycauchy_initialize_ltv_window_manager(self, num_windows, num_sim_steps, A0, p0, b0, Phi, Gamma, B, beta, H,
                                               gamma, f_duc_ptr, debug_print, log_seq, log_full, log_dir, dt,
                                               init_step, win_var_boost):
        args = [_arr(v) for v in (A0, p0, b0, Phi, Gamma, B, beta, H, gamma)]
        flat = [x for a in args for x in (a[1], a[2])]
        self._f("pycauchy_initialize_ltv_window_manager")(
            int(num_windows), int(num_sim_steps), *flat, self._fptr(f_duc_ptr),
            *self._wm_tail(debug_print, log_seq, log_full, log_dir, dt, init_step, win_var_boost))

    def pycauchy_initialize_nonlin_window_manager(self, num_windows, num_sim_steps, x0, A0, p0, b0, beta, gamma,
                                                  f_duc_ptr1, f_duc_ptr2, f_duc_ptr3, cmcc, debug_print, log_seq,
                                                  log_full, log_dir, dt, init_step, win_var_boost):
        args = [_arr(v) for v in (x0, A0, p0, b0, beta, gamma)]
        flat = [x for a in args for x in (a[1], a[2])]
        self._f("pycauchy_initialize_nonlin_window_manager")(
            int(num_windows), int(num_sim_steps), *flat, self._fptr(f_duc_ptr1), self._fptr(f_duc_ptr2),
            self._fptr(f_duc_ptr3), int(cmcc),
            *self._wm_tail(debug_print, log_seq, log_full, log_dir, dt, init_step, win_var_boost))

    def pycauchy_step(self, msmts, controls):
        """(fz, xhat, Phat, cerr_fz, cerr_xhat, cerr_Phat, win_idx, err_code) of the bank's best
        window after this measurement (pycauchy.hpp:184-214)."""
        m_, c_ = _arr(msmts), _arr(controls if controls is not None else np.zeros(0))
        fz, cfz, cx, cP = ct.c_double(), ct.c_double(), ct.c_double(), ct.c_double()
        x, P, sx, sP = _dp(), _dp(), ct.c_int(0), ct.c_int(0)
        win, err = ct.c_int(0), ct.c_int(0)
        self._f("pycauchy_step")(m_[1], m_[2], c_[1] if c_[2] > 0 else None, c_[2], ct.byref(fz), ct.byref(x),
                                 ct.byref(sx), ct.byref(P), ct.byref(sP), ct.byref(cfz), ct.byref(cx), ct.byref(cP),
                                 ct.byref(win), ct.byref(err))
        if sx.value == 0:
            msg = self._last_error() if self._last_error else ""
            raise RuntimeError(f"pycauchy_step failed: {msg}")
        return (fz.value, _take_d(x, sx.value), _take_d(P, sP.value), cfz.value, cx.value, cP.value, win.value,
                err.value)

    def pycauchy_shutdown(self):
        self._f("pycauchy_shutdown")()


_NAMES = [n for n in dir(_Bindings) if n.startswith("pycauchy_")]


def bind(cdll, sym=lambda name: name, name="pycauchy"):
    """A module with the SWIG module's functions over another library exporting this
    C ABI (argument and result types are declared on it)."""
    _native.declare_pycauchy(cdll, sym)
    b = _Bindings(lambda: cdll, sym)
    mod = types.ModuleType(name)
    mod.__doc__ = __doc__
    for n in _NAMES:
        setattr(mod, n, getattr(b, n))
    return mod


_default = _Bindings(_native.lib, last_error=lambda: _native.lib().mce_last_error().decode())
for _n in _NAMES:
    globals()[_n] = getattr(_default, _n)
del _n
