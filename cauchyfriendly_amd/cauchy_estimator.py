"""Python estimator classes over the MI355X engine, mirroring the reference's
scripts/swig/cauchy/cauchy_estimator.py interface (modes lti, ltv and nonlin):

* PyCauchyEstimator (cauchy_estimator.py:535-950): one estimator window, stepped
  with the p scalar measurements of a time step; keeps `moment_info` and the
  per-measurement outputs of its last step (`_fz`, `_x`, `_P`, `_err_code`, ...);
  `reset_about_estimator` restarts it about another window's estimate (Speyer's
  initialisation) and re-processes that window's last measurement(s).
* PySlidingWindowManager (cauchy_estimator.py:1092-1335): a bank of windows,
  each restarted about the best full window when the oldest one completes.
* speyers_window_init, set_tr_search_idxs_ordering (cauchy_estimator.py:18-37).
* Py_CauchyDynamicsUpdateContainer (cauchy_estimator.py:60-136): numpy access to the
  C container that the ltv / nonlin callbacks receive.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import pycauchy


def set_tr_search_idxs_ordering(ordering):
    return pycauchy.pycauchy_set_tr_search_idxs_ordering(np.asarray(ordering, dtype=np.int32).reshape(-1))


def speyers_window_init(x1_hat, Var, H, gamma, z):
    """Initial terms (A0 rows, p0, b0) of a window restarted about (x1_hat, Var)
    after the scalar measurement z with row H and noise scale gamma."""
    n = np.asarray(x1_hat).size
    if np.asarray(Var).size != n * n or np.asarray(H).size != n:
        raise ValueError("speyers_window_init: x1_hat (n), Var (n x n) and H (n) sizes disagree")
    A0, p0, b0 = pycauchy.pycauchy_speyers_window_init(np.asarray(x1_hat, dtype=np.float64).reshape(-1),
                                                       np.asarray(Var, dtype=np.float64).reshape(-1),
                                                       np.asarray(H, dtype=np.float64).reshape(-1),
                                                       float(gamma), float(z))
    return A0.reshape(n, n), p0, b0


class C_CauchyDynamicsUpdateContainer(ct.Structure):
    """CauchyDynamicsUpdateContainer as laid out by the C ABI (include/cauchy_mi355x.h)."""
    _fields_ = [("x", ct.POINTER(ct.c_double)), ("u", ct.POINTER(ct.c_double)), ("dt", ct.c_double),
                ("step", ct.c_int), ("n", ct.c_int), ("cmcc", ct.c_int), ("pncc", ct.c_int), ("p", ct.c_int),
                ("Phi", ct.POINTER(ct.c_double)), ("Gamma", ct.POINTER(ct.c_double)),
                ("B", ct.POINTER(ct.c_double)), ("H", ct.POINTER(ct.c_double)),
                ("beta", ct.POINTER(ct.c_double)), ("gamma", ct.POINTER(ct.c_double)),
                ("is_xbar_set_for_ece", ct.c_bool), ("other_stuff", ct.c_void_p)]


class Py_CauchyDynamicsUpdateContainer:
    """The callbacks' view of a C_CauchyDynamicsUpdateContainer: cget_* return numpy
    copies, cset_* write through to the C arrays (same method names as the reference)."""

    def __init__(self, cduc):
        self.cduc = cduc
        c = cduc.contents
        self.n, self.pncc, self.cmcc, self.p = c.n, c.pncc, c.cmcc, c.p

    def _get(self, field, size, shape=None):
        ptr = getattr(self.cduc.contents, field)
        a = np.ctypeslib.as_array(ptr, shape=(size,)).copy() if size > 0 else np.zeros(0)
        return a if shape is None else a.reshape(shape)

    def _set(self, field, value, size):
        v = np.asarray(value, dtype=np.float64).reshape(-1)
        if v.size != size:
            raise ValueError(f"{field} must have {size} entries, got {v.size}")
        if size > 0:
            np.ctypeslib.as_array(getattr(self.cduc.contents, field), shape=(size,))[:] = v

    def cget_step(self):
        return self.cduc.contents.step

    def cget_dt(self):
        return self.cduc.contents.dt

    def cset_dt(self, dt):
        self.cduc.contents.dt = float(dt)

    def cget_x(self):
        return self._get("x", self.n)

    def cset_x(self, x):
        self._set("x", x, self.n)

    def cget_u(self):
        return self._get("u", self.cmcc)

    def cget_Phi(self):
        return self._get("Phi", self.n * self.n, (self.n, self.n))

    def cset_Phi(self, Phi):
        self._set("Phi", Phi, self.n * self.n)

    def cget_Gamma(self):
        return self._get("Gamma", self.n * self.pncc, (self.n, self.pncc))

    def cset_Gamma(self, Gamma):
        self._set("Gamma", Gamma, self.n * self.pncc)

    def cget_B(self):
        return self._get("B", self.n * self.cmcc, (self.n, self.cmcc))

    def cset_B(self, B):
        self._set("B", B, self.n * self.cmcc)

    def cget_beta(self):
        return self._get("beta", self.pncc)

    def cset_beta(self, beta):
        self._set("beta", beta, self.pncc)

    def cget_H(self):
        return self._get("H", self.p * self.n, (self.p, self.n))

    def cset_H(self, H):
        self._set("H", H, self.p * self.n)

    def cget_gamma(self):
        return self._get("gamma", self.p)

    def cset_gamma(self, gamma):
        self._set("gamma", gamma, self.p)

    def cset_is_xbar_set_for_ece(self):
        self.cduc.contents.is_xbar_set_for_ece = True

    def cset_zbar(self, c_zbar, zbar):
        z = np.asarray(zbar, dtype=np.float64).reshape(-1)
        if z.size != self.p:
            raise ValueError(f"zbar must have {self.p} entries, got {z.size}")
        for i in range(self.p):
            c_zbar[i] = float(z[i])


_EMPTY = np.array([], dtype=np.float64)
_MOMENT_KEYS = ("fz", "x", "P", "cerr_x", "cerr_P", "cerr_fz", "err_code")


def _flat(a):
    return _EMPTY if a is None else np.asarray(a, dtype=np.float64).reshape(-1).copy()


class PyCauchyEstimator:
    MODES = ("lti", "ltv", "nonlin")

    def __init__(self, mode, num_steps, debug_print=False, backend=None):
        """backend: a module with the pycauchy functions (default: this library's)."""
        self._pc = backend if backend is not None else pycauchy
        mode = str(mode).lower()
        if mode not in self.MODES:
            raise ValueError(f"mode {mode!r} is not supported by this engine (modes: {self.MODES})")
        self.mode = mode
        self.num_steps = int(num_steps)
        self.debug_print = bool(debug_print)
        self.is_initialized = False
        self.step_count = 0
        self.py_handle = None
        self.moment_info = {k: [] for k in _MOMENT_KEYS}

    # -- set-up ---------------------------------------------------------------
    def _dimensions(self, A0, p0, b0, Phi, B, Gamma, beta, H, gamma):
        n = np.asarray(p0).size
        if np.asarray(A0).size != n * n or np.asarray(b0).size != n or np.asarray(Phi).size != n * n:
            raise ValueError("A0 (n x n), p0 (n), b0 (n) and Phi (n x n) sizes disagree")
        G = None if Gamma is None else np.asarray(Gamma)
        pncc = 0 if G is None else (1 if G.ndim == 1 else G.shape[1])
        Bm = None if B is None else np.asarray(B)
        cmcc = 0 if Bm is None else (1 if Bm.ndim == 1 else Bm.shape[1])
        Hm = np.asarray(H)
        p = 1 if Hm.ndim == 1 else Hm.shape[0]
        if np.asarray(gamma).size != p:
            raise ValueError("gamma must hold one scale per measurement row of H")
        self.n, self.pncc, self.cmcc, self.p = n, pncc, cmcc, p

    def _store_inputs(self, A0, p0, b0, Phi, B, Gamma, beta, H, gamma):
        self._dimensions(A0, p0, b0, Phi, B, Gamma, beta, H, gamma)
        self._A0, self._p0, self._b0 = _flat(A0), _flat(p0), _flat(b0)
        self._Phi, self._Gamma, self._B = _flat(Phi), _flat(Gamma), _flat(B)
        self._beta, self._H, self._gamma = _flat(beta), _flat(H), _flat(gamma)

    def initialize_lti(self, A0, p0, b0, Phi, B, Gamma, beta, H, gamma, init_step=0, dt=0):
        if self.mode != "lti":
            raise RuntimeError(f"initialize_lti called on an estimator in mode {self.mode}")
        self._store_inputs(A0, p0, b0, Phi, B, Gamma, beta, H, gamma)
        self.py_handle = self._pc.pycauchy_initialize_lti(self.num_steps, self._A0, self._p0, self._b0, self._Phi,
                                                          self._Gamma, self._B, self._beta, self._H, self._gamma,
                                                          float(dt), int(init_step), self.debug_print)
        self.is_initialized = True

    def initialize_ltv(self, A0, p0, b0, Phi, B, Gamma, beta, H, gamma, dynamics_update_callback, init_step=0, dt=0):
        if self.mode != "ltv":
            raise RuntimeError(f"initialize_ltv called on an estimator in mode {self.mode}")
        self._store_inputs(A0, p0, b0, Phi, B, Gamma, beta, H, gamma)
        cb_type = ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer))
        self._f_duc = cb_type(dynamics_update_callback)  # kept alive as long as the estimator
        self.f_duc_ptr1 = ct.cast(self._f_duc, ct.c_void_p).value
        self.py_handle = self._pc.pycauchy_initialize_ltv(self.num_steps, self._A0, self._p0, self._b0, self._Phi,
                                                          self._Gamma, self._B, self._beta, self._H, self._gamma,
                                                          self.f_duc_ptr1, float(dt), int(init_step),
                                                          self.debug_print)
        self.is_initialized = True

    def initialize_nonlin(self, x0, A0, p0, b0, beta, gamma, dynamics_update_callback, nonlinear_msmt_model,
                          extended_msmt_update_callback, cmcc, dt=0, step=0):
        """Extended estimator: x0 is the initial x_bar; the estimator runs on the deviation
        from x_bar, with Phi, Gamma (dynamics_update_callback) and H (extended_msmt_update_callback)
        the Jacobians at x_bar, and the residual z - h(x_bar) (nonlinear_msmt_model)."""
        if self.mode != "nonlin":
            raise RuntimeError(f"initialize_nonlin called on an estimator in mode {self.mode}")
        n = np.asarray(p0).size
        if np.asarray(A0).size != n * n or np.asarray(b0).size != n or np.asarray(x0).size != n:
            raise ValueError("x0 (n), A0 (n x n), p0 (n) and b0 (n) sizes disagree")
        self.n = n
        self.pncc = 0 if beta is None else np.asarray(beta).size
        self.p = np.asarray(gamma).size
        self.cmcc = int(cmcc)
        t1 = ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer))
        t2 = ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer), ct.POINTER(ct.c_double))
        # the ctypes callbacks live as long as the estimator
        self._f_duc1 = t1(dynamics_update_callback)
        self._f_duc2 = t2(nonlinear_msmt_model)
        self._f_duc3 = t1(extended_msmt_update_callback)
        self.f_duc_ptr1 = ct.cast(self._f_duc1, ct.c_void_p).value
        self.f_duc_ptr2 = ct.cast(self._f_duc2, ct.c_void_p).value
        self.f_duc_ptr3 = ct.cast(self._f_duc3, ct.c_void_p).value
        self._x0, self._A0, self._p0, self._b0 = _flat(x0), _flat(A0), _flat(p0), _flat(b0)
        self._beta, self._gamma = _flat(beta), _flat(gamma)
        self.py_handle = self._pc.pycauchy_initialize_nonlin(self.num_steps, self._x0, self._A0, self._p0, self._b0,
                                                             self._beta, self._gamma, self.f_duc_ptr1,
                                                             self.f_duc_ptr2, self.f_duc_ptr3, self.cmcc, float(dt),
                                                             int(step), self.debug_print)
        self.is_initialized = True

    # -- stepping -------------------------------------------------------------
    def _checked_inputs(self, msmts, controls):
        z = np.asarray(msmts, dtype=np.float64).reshape(-1).copy()
        if z.size != self.p:
            raise ValueError(f"expected {self.p} measurements, got {z.size}")
        if self.cmcc == 0:
            if controls is not None and np.asarray(controls).size != 0:
                raise ValueError("this estimator has no control input")
            u = _EMPTY
        else:
            u = np.asarray(controls, dtype=np.float64).reshape(-1).copy()
            if u.size != self.cmcc:
                raise ValueError(f"expected {self.cmcc} controls, got {u.size}")
        return z, u

    def _call_step(self, z, u, full_info):
        if self.mode == "nonlin":
            # the first step of a window only updates; every later one propagates first
            out = self._pc.pycauchy_single_step_nonlin(self.py_handle, z, u, self.step_count != 0)
        else:
            out = self._pc.pycauchy_single_step_ltiv(self.py_handle, z, u)
        return self._absorb_step(out, z, full_info)

    def _absorb_step(self, out, z, full_info):
        """Keeps the outputs of one step (a single or batched pycauchy step) as step() does."""
        if self.mode == "nonlin":
            (Phi, Gamma, B, H, beta, gamma, self._fz, self._x, self._P, self._xbar, self._zbar, self._cerr_fz,
             self._cerr_x, self._cerr_P, self._err_code) = out
        else:
            (Phi, Gamma, B, H, beta, gamma, self._fz, self._x, self._P, self._cerr_fz, self._cerr_x, self._cerr_P,
             self._err_code) = out
        if self.mode != "lti":  # the dynamics of this step, as the callbacks left them
            self._Phi, self._Gamma, self._B, self._H, self._beta, self._gamma = Phi, Gamma, B, H, beta, gamma
        n = self.n
        k = z.size
        if full_info:
            xs = [self._x[i * n:(i + 1) * n].copy() for i in range(k)]
            Ps = [self._P[i * n * n:(i + 1) * n * n].reshape(n, n).copy() for i in range(k)]
        else:
            xs = self._x[-n:].copy()
            Ps = self._P[-n * n:].reshape(n, n).copy()
        # the history holds views of this step's outputs: every step's outputs are new
        # arrays, and nothing writes into them
        mi = self.moment_info
        mi["fz"].append(self._fz[-1])
        mi["x"].append(self._x[-n:])
        mi["P"].append(self._P[-n * n:].reshape(n, n))
        mi["cerr_x"].append(self._cerr_x[-1])
        mi["cerr_P"].append(self._cerr_P[-1])
        mi["cerr_fz"].append(self._cerr_fz[-1])
        mi["err_code"].append(self._err_code[-1])
        self.step_count += 1
        return xs, Ps

    def step(self, msmts, controls=None, full_info=False):
        if not self.is_initialized:
            raise RuntimeError("estimator is not initialized")
        if self.step_count == self.num_steps:
            raise RuntimeError(f"the estimator has already processed its {self.num_steps} steps; reset it first")
        z, u = self._checked_inputs(msmts, controls)
        self._msmts = z.copy()
        self._controls = u.copy()
        return self._call_step(z, u, bool(full_info))

    def deterministic_transform(self, Trans, bias):
        T = np.asarray(Trans, dtype=np.float64).reshape(-1)
        b = np.asarray(bias, dtype=np.float64).reshape(-1)
        if T.size != self.n * self.n or b.size != self.n:
            raise ValueError("Trans must be n x n and bias n")
        self._pc.pycauchy_single_step_deterministic_transform(self.py_handle, T, b)

    def terms_generated(self):
        """Terms generated by this estimator so far (this library's backend only)."""
        from . import _native
        return int(_native.lib().mce_terms_generated(self.py_handle)) if self.py_handle else 0

    def get_num_CF_terms(self):
        if not self.is_initialized:
            return 0
        if self.step_count == 0:
            return 1
        return self._pc.pycauchy_single_step_get_number_of_terms(self.py_handle)

    def get_marginal_1D_pointwise_cpdf(self, marg_idx, gridx_low, gridx_high, gridx_resolution, log_dir=None):
        """Marginal density of state marg_idx on the grid gridx_low, gridx_low + gridx_resolution, ...
        (cauchy_estimator.py:1003-1031); returns (X, Y), or (None, None) before the first step."""
        if not self.is_initialized or self.step_count < 1:
            return None, None
        if not (gridx_high > gridx_low and gridx_resolution > 0 and 0 <= int(marg_idx) < self.n):
            raise ValueError("the grid needs high > low, resolution > 0 and a state index below n")
        pts, num = self._pc.pycauchy_get_marginal_1D_pointwise_cpdf(self.py_handle, int(marg_idx), float(gridx_low),
                                                                    float(gridx_high), float(gridx_resolution),
                                                                    log_dir)
        pts = pts.reshape(num, 2)
        return pts[:, 0], pts[:, 1]

    def get_marginal_2D_pointwise_cpdf(self, marg_idx1, marg_idx2, gridx_low, gridx_high, gridx_resolution,
                                       gridy_low, gridy_high, gridy_resolution, log_dir=None, reset_cache=True):
        """Joint marginal density of states marg_idx1 < marg_idx2 on a grid
        (cauchy_estimator.py:954-989); returns meshgrid-shaped (X, Y, Z), or (None, None, None)
        before the first step."""
        if not self.is_initialized or self.step_count < 1:
            return None, None, None
        if not (gridx_high > gridx_low and gridy_high > gridy_low and gridx_resolution > 0 and gridy_resolution > 0
                and 0 <= int(marg_idx1) < int(marg_idx2) < self.n):
            raise ValueError("the grid needs high > low, resolutions > 0 and state indices i1 < i2 < n")
        pts, nx, ny = self._pc.pycauchy_get_marginal_2D_pointwise_cpdf(
            self.py_handle, int(marg_idx1), int(marg_idx2), float(gridx_low), float(gridx_high),
            float(gridx_resolution), float(gridy_low), float(gridy_high), float(gridy_resolution), log_dir,
            bool(reset_cache))
        pts = pts.reshape(ny, nx, 3)
        return pts[:, :, 0], pts[:, :, 1], pts[:, :, 2]

    def get_2D_pointwise_cpdf(self, gridx_low, gridx_high, gridx_resolution, gridy_low, gridy_high, gridy_resolution,
                              log_dir=None):
        if self.n != 2:
            raise ValueError("get_2D_pointwise_cpdf is for a 2-state system; use get_marginal_2D_pointwise_cpdf")
        return self.get_marginal_2D_pointwise_cpdf(0, 1, gridx_low, gridx_high, gridx_resolution, gridy_low,
                                                   gridy_high, gridy_resolution, log_dir)

    def get_1D_pointwise_cpdf(self, gridx_low, gridx_high, gridx_resolution, log_dir=None):
        if self.n != 1:
            raise ValueError("get_1D_pointwise_cpdf is for a 1-state system; use get_marginal_1D_pointwise_cpdf")
        return self.get_marginal_1D_pointwise_cpdf(0, gridx_low, gridx_high, gridx_resolution, log_dir)

    def get_last_mean_cov(self):
        return self.moment_info["x"][-1].copy(), self.moment_info["P"][-1].copy()

    # -- re-initialisation ----------------------------------------------------
    def _msmt_index(self, msmt_idx):
        msmt_idx = int(msmt_idx)
        if msmt_idx >= self.p or msmt_idx < -self.p:
            raise IndexError(f"msmt_idx {msmt_idx} outside [-{self.p}, {self.p})")
        return msmt_idx + self.p if msmt_idx < 0 else msmt_idx

    def get_reinitialization_statistics(self, msmt_idx=-1):
        """(A0, p0, b0) restarting a window about this estimator's estimate after its
        msmt_idx-th measurement of the last step; in nonlin mode the restart is about the
        deviation from x_bar (residual z - zbar), and x_bar is returned as a fourth value."""
        if not self.is_initialized or self.step_count == 0:
            raise RuntimeError("the estimator has not processed a measurement yet")
        i = self._msmt_index(msmt_idx)
        n = self.n
        z = self._msmts[i]
        x = self._x[i * n:(i + 1) * n].copy()
        P = self._P[i * n * n:(i + 1) * n * n].copy()
        H = self._H[i * n:(i + 1) * n].copy()
        if self.mode != "nonlin":
            A0, p0, b0 = self._pc.pycauchy_get_reinitialization_statistics(self.py_handle, z, x, P, H, self._gamma[i])
            return A0.reshape(n, n), p0, b0
        xbar = self._xbar[i * n:(i + 1) * n].copy()
        A0, p0, b0 = self._pc.pycauchy_get_reinitialization_statistics(self.py_handle, z - self._zbar[i], x - xbar, P,
                                                                       H, self._gamma[i])
        return A0.reshape(n, n), p0, b0, xbar

    def reset(self, A0=None, p0=None, b0=None, xbar=None):
        if not self.is_initialized:
            raise RuntimeError("estimator is not initialized")
        self.step_count = 0
        n = self.n
        for name, v, size in (("A0", A0, n * n), ("p0", p0, n), ("b0", b0, n), ("xbar", xbar, n)):
            if v is not None and np.asarray(v).size != size:
                raise ValueError(f"{name} must have {size} entries")
        self._A0 = _flat(A0)
        self._p0 = _flat(p0)
        self._b0 = _flat(b0)
        self._xbar = _flat(xbar)
        self._pc.pycauchy_single_step_reset(self.py_handle, self._A0, self._p0, self._b0, self._xbar)

    def reset_about_estimator(self, other_estimator, msmt_idx=-1):
        """Restarts this window about other_estimator's estimate after its msmt_idx-th
        measurement (Speyer's initialisation), then processes that window's remaining
        measurements of its last step."""
        if other_estimator is self:
            raise ValueError("an estimator cannot be restarted about itself")
        if not other_estimator.is_initialized or other_estimator.step_count == 0:
            raise RuntimeError("the other estimator has not processed a measurement yet")
        if other_estimator.p != self.p or other_estimator.mode != self.mode:
            raise ValueError("both estimators must have the same mode and measurement count")
        i = self._msmt_index(msmt_idx)
        z = other_estimator._msmts[i:].copy()
        self._msmts = z.copy()
        self.reset(*other_estimator.get_reinitialization_statistics(i))
        xs, Ps = self._call_step(z, _EMPTY, False)
        self._pc.pycauchy_single_step_set_master_step(self.py_handle, self.p)
        return xs, Ps

    def reset_with_last_measurement(self, z_scalar, A0, p0, b0, xbar=None):
        z = np.asarray(z_scalar, dtype=np.float64).reshape(-1).copy()
        self.reset(A0, p0, b0, xbar)
        xs, Ps = self._call_step(z, _EMPTY, False)
        self._pc.pycauchy_single_step_set_master_step(self.py_handle, self.p)
        return xs, Ps

    def set_window_number(self, win_num):
        self._pc.pycauchy_single_step_set_window_number(self.py_handle, int(win_num))

    def shutdown(self):
        if not self.is_initialized:
            return
        self._pc.pycauchy_single_step_shutdown(self.py_handle)
        self.py_handle = None
        self.is_initialized = False

    def __del__(self):
        try:
            self.shutdown()
        except Exception:
            pass


def eval_2d_relative_system_cpdf(Trel, secondary, primary, rsys_approx_eps, xlow, xhigh, delta_x, ylow, yhigh,
                                 delta_y):
    """Moments and density of the relative system secondary - primary transformed by Trel
    (2 x n), as pycauchy_single_step_eval_2d_rsys_cpdf returns them; the grid limits are
    relative to the relative mean (the reference calls this from
    scripts/swig/leo/gmat_mce.py:950-975)."""
    if secondary._pc is not primary._pc:
        raise ValueError("the two estimators must share one backend")
    return primary._pc.pycauchy_single_step_eval_2d_rsys_cpdf(
        np.asarray(Trel, dtype=np.float64).reshape(-1), secondary.py_handle, primary.py_handle,
        float(rsys_approx_eps), float(xlow), float(xhigh), float(delta_x), float(ylow), float(yhigh), float(delta_y))


def step_estimators(estimators, msmts, controls=None, full_info=False, num_threads=0, merged_groups=0):
    """Steps every estimator once, as est.step(msmts[i], controls[i]) would, but concurrently:
    the estimators run on their own HIP streams from up to num_threads host threads (0: one
    per hardware thread).  With merged_groups > 0 (LTI / LTV estimators) they are dealt to
    that many host threads instead, and each group's updates launch every kernel once for
    the whole group.  All must be initialized estimators of one mode on this library.
    Returns the list of est.step return values."""
    ests = list(estimators)
    if not ests:
        return []
    mode = ests[0].mode
    pc = ests[0]._pc
    if any(e.mode != mode or e._pc is not pc for e in ests):
        raise ValueError("the estimators must share one mode and one backend")
    if not hasattr(pc, "pycauchy_batch_single_step_ltiv"):
        raise RuntimeError("the backend has no batched stepping")
    if controls is None:
        controls = [None] * len(ests)
    zs, us = [], []
    for e, z, u in zip(ests, msmts, controls):
        if not e.is_initialized:
            raise RuntimeError("estimator is not initialized")
        if e.step_count == e.num_steps:
            raise RuntimeError(f"an estimator has already processed its {e.num_steps} steps; reset it first")
        z_, u_ = e._checked_inputs(z, u)  # new arrays, read only by the step
        e._msmts = z_
        e._controls = u_
        zs.append(z_)
        us.append(u_)
    handles = [e.py_handle for e in ests]
    if mode == "nonlin":
        outs = pc.pycauchy_batch_single_step_nonlin(handles, zs, us, [e.step_count != 0 for e in ests], num_threads)
    elif merged_groups > 0:
        outs = pc.pycauchy_batch_single_step_ltiv_merged(handles, zs, us, int(merged_groups))
    else:
        outs = pc.pycauchy_batch_single_step_ltiv(handles, zs, us, num_threads)
    return [e._absorb_step(o, z, bool(full_info)) for e, o, z in zip(ests, outs, zs)]


class PySlidingWindowManager:
    """A bank of num_windows estimators, each spanning num_windows time steps: window
    i starts i steps after window 0; when the oldest window completes it is reset,
    and the newest one is restarted about the best window's estimate."""

    COV_UNSTABLE = 2
    COV_DNE = 8

    def __init__(self, mode, num_windows, swm_debug_print=True, win_debug_print=False, backend=None,
                 concurrent=False):
        """concurrent: step the active windows together (step_estimators) instead of one
        after another; the results are the same, and LTV / nonlin callbacks then run on
        worker threads."""
        self.concurrent = bool(concurrent)
        self.num_windows = int(num_windows)
        if self.num_windows >= 20:
            raise ValueError("at most 19 windows")
        self.mode = str(mode).lower()
        if self.mode not in PyCauchyEstimator.MODES:
            raise ValueError(f"mode {mode!r} is not supported by this engine")
        self.debug_print = bool(swm_debug_print)
        self.is_initialized = False
        self.step_idx = 0
        self.win_idxs = np.arange(self.num_windows)
        self.win_counts = np.zeros(self.num_windows, dtype=np.int64)
        self.cauchyEsts = [PyCauchyEstimator(self.mode, self.num_windows, bool(win_debug_print), backend)
                           for _ in range(self.num_windows)]
        self.moment_info = {k: [] for k in _MOMENT_KEYS + ("win_idx",)}
        self.avg_moment_info = {k: [] for k in _MOMENT_KEYS + ("win_idx",)}

    def _dims(self):
        e = self.cauchyEsts[0]
        self.n, self.cmcc, self.pncc, self.p = e.n, e.cmcc, e.pncc, e.p

    def initialize_lti(self, A0, p0, b0, Phi, B, Gamma, beta, H, gamma, dt=0, step=0, reinit_func=None):
        if np.asarray(p0).size == 1:
            raise ValueError("the window bank is for systems of dimension > 1")
        self.reinit_func = reinit_func
        for i, est in enumerate(self.cauchyEsts):
            est.initialize_lti(A0, p0, b0, Phi, B, Gamma, beta, H, gamma, int(step) + i, float(dt))
            est.set_window_number(i + 1)
        self.is_initialized = True
        self._dims()

    def initialize_ltv(self, A0, p0, b0, Phi, B, Gamma, beta, H, gamma, dynamics_update_callback, dt=0, step=0,
                       reinit_func=None):
        if np.asarray(p0).size == 1:
            raise ValueError("the window bank is for systems of dimension > 1")
        self.reinit_func = reinit_func
        for i, est in enumerate(self.cauchyEsts):
            est.initialize_ltv(A0, p0, b0, Phi, B, Gamma, beta, H, gamma, dynamics_update_callback, int(step) + i,
                               float(dt))
            est.set_window_number(i + 1)
        self.is_initialized = True
        self._dims()

    def initialize_nonlin(self, x0, A0, p0, b0, beta, gamma, dynamics_update_callback, nonlinear_msmt_model,
                          extended_msmt_update_callback, cmcc, dt=0, step=0, reinit_func=None):
        if np.asarray(p0).size == 1:
            raise ValueError("the window bank is for systems of dimension > 1")
        self.reinit_func = reinit_func
        for i, est in enumerate(self.cauchyEsts):
            est.initialize_nonlin(x0, A0, p0, b0, beta, gamma, dynamics_update_callback, nonlinear_msmt_model,
                                  extended_msmt_update_callback, cmcc, float(dt), int(step) + i)
            est.set_window_number(i + 1)
        self.is_initialized = True
        self._dims()

    def _usable(self, est):
        err = est._err_code[self.p - 1]
        return not (err & self.COV_UNSTABLE) and not (err & self.COV_DNE)

    def _best_window_est(self):
        """The window with most measurements among those whose covariance is usable."""
        W = self.num_windows
        okays = np.zeros(W, dtype=bool)
        candidates = []
        for i in range(W):
            if self.win_counts[i] > 0 and self._usable(self.cauchyEsts[i]):
                candidates.append((i, self.win_counts[i]))
                okays[i] = True
        if self.step_idx == 0:
            best = 0
            okays[0] = True
        else:
            if not candidates:
                raise RuntimeError("no window is available without an error code")
            # largest count; among equal counts the highest window index (a stable sort
            # by count, reversed)
            best = sorted(candidates, key=lambda c: c[1])[-1][0]
        est = self.cauchyEsts[best]
        last = self.p - 1
        n = self.n
        mi = self.moment_info
        mi["fz"].append(est._fz[last])
        mi["x"].append(est._x[last * n:].copy())
        mi["P"].append(est._P[last * n * n:].copy().reshape(n, n))
        mi["cerr_x"].append(est._cerr_x[last])
        mi["cerr_P"].append(est._cerr_P[last])
        mi["cerr_fz"].append(est._cerr_fz[last])
        mi["win_idx"].append(best)
        mi["err_code"].append(est._err_code[last])
        return best, okays

    def _weighted_average_win_est(self, usable):
        """Estimates averaged over the usable windows, weighted by their measurement counts."""
        n = self.n
        last = self.p - 1
        mean = np.zeros(n)
        cov = np.zeros((n, n))
        fz = cerr_fz = cerr_x = cerr_P = 0.0
        err = 0
        norm = 0.0
        for i in range(self.num_windows):
            if self.win_counts[i] > 0 and usable[i]:
                est = self.cauchyEsts[i]
                w = self.win_counts[i] / self.num_windows
                norm += w
                x, P = est.get_last_mean_cov()
                mean += x * w
                cov += P * w
                fz += est._fz[last] * w
                cerr_fz += est._cerr_fz[last] * w
                cerr_x += est._cerr_x[last] * w
                cerr_P += est._cerr_P[last] * w
                err |= est._err_code[last]
        mean /= norm
        cov /= norm
        ami = self.avg_moment_info
        ami["fz"].append(fz / norm)
        ami["x"].append(mean.copy())
        ami["P"].append(cov.copy())
        ami["cerr_x"].append(cerr_x / norm)
        ami["cerr_P"].append(cerr_P / norm)
        ami["cerr_fz"].append(cerr_fz / norm)
        ami["win_idx"].append(-1)
        ami["err_code"].append(err)
        return mean, cov

    def step(self, msmts, controls=None, reinit_args=None):
        if not self.is_initialized:
            raise RuntimeError("the window bank is not initialized")
        if self.step_idx == 0:
            self.cauchyEsts[0].step(msmts, controls)
            self.win_counts[0] += 1
            self.min_idx = None
        else:
            max_idx = int(np.argmax(self.win_counts))
            min_idx = int(np.argmin(self.win_counts))
            active = [i for i, count in zip(self.win_idxs, self.win_counts) if count > 0]
            if self.concurrent:
                step_estimators([self.cauchyEsts[i] for i in active], [msmts] * len(active),
                                [controls] * len(active))
            else:
                for i in active:
                    self.cauchyEsts[i].step(msmts, controls)
            for i in active:
                self.win_counts[i] += 1
            self.min_idx = min_idx
        best, usable = self._best_window_est()
        wx, wP = self._weighted_average_win_est(usable)
        xhat, Phat = self.cauchyEsts[best].get_last_mean_cov()
        self.best_idx = best
        self.usable_wins = usable
        if self.step_idx > 0:
            if self.reinit_func is not None:
                self.reinit_func(self.cauchyEsts, msmts, best, min_idx, self.step_idx + 1, reinit_args)
            else:
                # Speyer's restart of the emptiest window about the best one, after the
                # last measurement of the step
                self.cauchyEsts[min_idx].reset_about_estimator(self.cauchyEsts[best], msmt_idx=self.p - 1)
            self.win_counts[min_idx] += 1
            if self.win_counts[max_idx] == self.num_windows:
                self.cauchyEsts[max_idx].reset()
                self.win_counts[max_idx] = 0
        self.step_idx += 1
        return xhat, Phat, wx, wP

    def shutdown(self):
        if not self.is_initialized:
            return
        for est in self.cauchyEsts:
            est.shutdown()
        self.win_counts = np.zeros(self.num_windows, dtype=np.int64)
        self.is_initialized = False
        self.step_idx = 0
