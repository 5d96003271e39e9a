"""Window log-likelihood L = sum_k log fz_k of an LTI estimator and its gradient with
respect to the Cauchy scales (beta, gamma) and the Phi / H entries.

fz_k is the measurement likelihood the update at step k computes before term reduction
(the normalisation of the characteristic function, include/cauchy_estimator.hpp:340-461).
The gradient is taken by central differences of this library's forward path, every
perturbed estimator of every sequence stepped together (step_estimators): 1 + 2 P
estimators per sequence for P parameters.  A component is reported as valid only when the
term tree's discrete structure (term counts after MU/MUC and after reduction at every
update) is identical at theta + h e_i and theta - h e_i, the condition under which the
likelihood is a smooth function of theta.  This is a finite-difference gradient, not an
adjoint: it costs 2 P forward passes.
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


def _last_update(handle, d, sr):
    mom = np.zeros(6 + 2 * d + 2 * d * d)
    cnt = np.zeros(4 + 2 * sr, dtype=np.int32)
    L = _native.lib()
    L.mce_last_update(handle, mom.ctypes.data_as(ct.POINTER(ct.c_double)), cnt.ctypes.data_as(ct.POINTER(ct.c_int)))
    return mom[2], (int(cnt[1]), int(cnt[2]))


def log_likelihoods(systems, thetas, num_threads=16):
    """L(theta_i) for estimator i on systems[i]'s measurements (all stepped together), and
    the per-update (terms after MU/MUC, after reduction) that fix its discrete structure."""
    n_est = len(systems)
    ests = []
    for s, th in zip(systems, thetas):
        beta, gamma, Phi, H = _unpack(th, s.n)
        ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
        e = PyCauchyEstimator("lti", s.num_steps, False)
        e.initialize_lti(s.A0, s.p0, s.b0, Phi, None, s.Gamma.reshape(-1), beta, H.reshape(-1), gamma)
        ests.append(e)
    sr = _native.lib().mce_shape_range(ests[0].py_handle) if ests else 0
    L = np.zeros(n_est)
    structure = [[] for _ in range(n_est)]
    steps = max(s.num_steps for s in systems) if systems else 0
    for k in range(steps):
        live = [i for i, s in enumerate(systems) if k < s.num_steps]
        step_estimators([ests[i] for i in live], [systems[i].zs[k] for i in live], num_threads=num_threads)
        for i in live:
            fz, counts = _last_update(ests[i].py_handle, systems[i].n, sr)
            L[i] += np.log(fz)
            structure[i].append(counts)
    for e in ests:
        e.shutdown()
    return L, structure


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
    stride = 1 + 2 * len(P)
    n = len(systems)
    L0 = np.zeros(n)
    grad = np.zeros((n, len(P)))
    valid = np.zeros((n, len(P)), dtype=bool)
    for i in range(n):
        b = i * stride
        L0[i] = L[b]
        for j in range(len(P)):
            lp, lm = b + 1 + 2 * j, b + 2 + 2 * j
            grad[i, j] = (L[lp] - L[lm]) / (2 * h)
            valid[i, j] = structure[lp] == structure[lm] == structure[b]
    return L0, grad, valid
