"""Window log-likelihood L = sum_k log fz_k of an estimator and its gradient.

fz_k is the measurement likelihood that scalar update k computes before term reduction
(the normalisation of the characteristic function, include/cauchy_estimator.hpp:340-461);
the sum runs over every scalar update of the window (p per time step).  Gradients are
taken by central differences of this library's forward path, every perturbed estimator of
every sequence stepped together (step_estimators): 1 + 2 P estimators per sequence for P
parameters.
* loglik_gradient: single-noise LTI systems, over (beta, gamma, Phi, H);
* scale_gradient: LTI or LTV systems with any number of noises and measurements, over the
  Cauchy scales (beta_1 .. beta_pncc, gamma_1 .. gamma_p).
A component is reported as valid only when the term tree's discrete structure (term counts
after MU/MUC and after reduction at every update) is identical at theta + h e_i and
theta - h e_i, the condition under which the likelihood is a smooth function of theta.
This is a finite-difference gradient, not an adjoint: it costs 2 P forward passes.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _native
from .cauchy_estimator import PyCauchyEstimator, step_estimators

PARAM_NAMES = ["beta", "gamma"] + [f"Phi{i}{j}" for i in range(3) for j in range(3)] + [f"H{j}" for j in range(3)]


def theta_of(s) -> np.ndarray:
    """[beta, gamma, Phi (row-major), H] of a single-noise LTI system."""
    return np.concatenate([np.ravel(s.beta)[:1], np.ravel(s.gamma)[:1], np.ravel(s.Phi(0)), np.ravel(s.H)])


def _unpack(theta, n):
    beta = np.array([theta[0]])
    gamma = np.array([theta[1]])
    Phi = np.asarray(theta[2:2 + n * n]).reshape(n, n)
    H = np.asarray(theta[2 + n * n:2 + n * n + n]).reshape(1, n)
    return beta, gamma, Phi, H


def _step_record(handle, p):
    """(fz before reduction, (terms after MU/MUC, after reduction)) of each update of the last step."""
    fz = np.zeros(p)
    cnt = np.zeros(2 * p, dtype=np.int32)
    n = _native.lib().mce_step_record(handle, fz.ctypes.data_as(ct.POINTER(ct.c_double)),
                                      cnt.ctypes.data_as(ct.POINTER(ct.c_int)), p)
    n = min(n, p)
    return fz[:n], [(int(cnt[2 * i]), int(cnt[2 * i + 1])) for i in range(n)]


def _dynamics_callback(s):
    """LTV dynamics callback: Phi_k of system s before the propagation into step k."""
    n = s.n

    def dyn(c_duc):
        duc = c_duc.contents
        k = min(int(duc.step), s.num_steps - 1)
        Phi = s.Phi(k).reshape(-1)
        for i in range(n * n):
            duc.Phi[i] = float(Phi[i])
        duc.step = k + 1
    return dyn


def make_estimator(s, beta=None, gamma=None, Phi=None, H=None):
    """An estimator of system s (its own scales / Phi / H unless given), constructed after
    srand(s.srand_seed); LTV (Phi_k through a dynamics callback) when s has one Phi per step."""
    beta = np.ravel(s.beta) if beta is None else np.ravel(beta)
    gamma = np.ravel(s.gamma) if gamma is None else np.ravel(gamma)
    Phi = s.Phi(0) if Phi is None else np.asarray(Phi)
    H = s.H if H is None else np.asarray(H)
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    if len(s.Phis) > 1:
        e = PyCauchyEstimator("ltv", s.num_steps, False)
        cb = _dynamics_callback(s)
        e.initialize_ltv(s.A0, s.p0, s.b0, Phi, None, s.Gamma, beta, H, gamma, cb)
        e._likelihood_callback = cb
    else:
        e = PyCauchyEstimator("lti", s.num_steps, False)
        e.initialize_lti(s.A0, s.p0, s.b0, Phi, None, s.Gamma.reshape(-1), beta, H.reshape(-1), gamma)
    return e


def run_likelihoods(systems, make, num_threads=16, max_live=None):
    """L for estimator i = make(i) on systems[i]'s measurements, and its structure (the
    (terms after MU/MUC, after reduction) of every update).  The estimators are stepped
    together, at most max_live of them at a time (all when None)."""
    n_est = len(systems)
    L = np.zeros(n_est)
    structure = [[] for _ in range(n_est)]
    chunk = n_est if not max_live else int(max_live)
    for c0 in range(0, n_est, max(chunk, 1)):
        idx = list(range(c0, min(n_est, c0 + chunk)))
        ests = {i: make(i) for i in idx}
        try:
            for k in range(max(systems[i].num_steps for i in idx)):
                live = [i for i in idx if k < systems[i].num_steps]
                step_estimators([ests[i] for i in live], [systems[i].zs[k] for i in live], num_threads=num_threads)
                for i in live:
                    fz, counts = _step_record(ests[i].py_handle, systems[i].p)
                    L[i] += float(np.sum(np.log(fz)))
                    structure[i].extend(counts)
        finally:
            for e in ests.values():
                e.shutdown()
    return L, structure


def log_likelihoods(systems, thetas, num_threads=16):
    """L(theta_i) for estimator i on systems[i]'s measurements (all stepped together), and
    the per-update (terms after MU/MUC, after reduction) that fix its discrete structure."""
    def make(i):
        s = systems[i]
        beta, gamma, Phi, H = _unpack(thetas[i], s.n)
        return make_estimator(s, beta, gamma, Phi, H)
    return run_likelihoods(systems, make, num_threads)


def _central(L, structure, n, P, h):
    stride = 1 + 2 * P
    L0 = np.zeros(n)
    grad = np.zeros((n, P))
    valid = np.zeros((n, P), dtype=bool)
    for i in range(n):
        b = i * stride
        L0[i] = L[b]
        for j in range(P):
            lp, lm = b + 1 + 2 * j, b + 2 + 2 * j
            grad[i, j] = (L[lp] - L[lm]) / (2 * h)
            valid[i, j] = structure[lp] == structure[lm] == structure[b]
    return L0, grad, valid


def scale_gradient(systems, h=1e-6, num_threads=16, max_live=None, with_terms=False):
    """Central-difference gradient of L for every system over its Cauchy scales
    (beta_1 .. beta_pncc, gamma_1 .. gamma_p; the systems must share pncc and p).
    Returns (L, grad, valid), and with_terms the terms generated by all the runs."""
    sys_all, scales = [], []
    P = None
    for s in systems:
        sc = np.concatenate([np.ravel(s.beta), np.ravel(s.gamma)]).astype(np.float64)
        if P is None:
            P = sc.size
        elif sc.size != P:
            raise ValueError("the systems must share the number of noise scales")
        sys_all.append(s)
        scales.append(sc)
        for j in range(P):
            for sign in (1.0, -1.0):
                t = sc.copy()
                t[j] += sign * h
                sys_all.append(s)
                scales.append(t)

    def make(i):
        s = sys_all[i]
        return make_estimator(s, scales[i][:s.pncc], scales[i][s.pncc:])
    L, structure = run_likelihoods(sys_all, make, num_threads, max_live)
    out = _central(L, structure, len(systems), P or 0, h)
    if with_terms:
        return out + (sum(c[0] for st in structure for c in st),)
    return out


def loglik_gradient(systems, h=1e-6, params=None, num_threads=16):
    """Central-difference gradient of L for every system, over the parameters `params`
    (indices into PARAM_NAMES; all 14 by default).  Returns (L, grad, valid)."""
    P = list(range(len(PARAM_NAMES))) if params is None else list(params)
    sys_all, th_all = [], []
    for s in systems:
        th = theta_of(s)
        sys_all.append(s)
        th_all.append(th)
        for p in P:
            for sign in (1.0, -1.0):
                t = th.copy()
                t[p] += sign * h
                sys_all.append(s)
                th_all.append(t)
    L, structure = log_likelihoods(sys_all, th_all, num_threads)
    return _central(L, structure, len(systems), len(P), h)
