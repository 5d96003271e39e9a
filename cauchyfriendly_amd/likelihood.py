"""Window log-likelihood L = sum_k log fz_k of an estimator and its gradient.

fz_k is the measurement likelihood that scalar update k computes before term reduction
(the normalisation of the characteristic function, include/cauchy_estimator.hpp:340-461),
taken before the numeric checks (which would replace it by the previous update's when both
the mean and the covariance fail them); the sum runs over every scalar update of the window
(p per time step).

Gradients:
* window_loglik_tangent / tangent_gradients / window_loglikelihood: the tangent kernels.
  One forward pass carries P directions; every discrete decision of the term tree (the
  coalignments, the reduction groups, the negligible terms) is the nominal forward's, so
  this is the derivative of L with the structure held fixed.  window_loglikelihood wraps it
  as a torch.autograd.Function.
* loglik_gradient / scale_gradient: central differences of this library's forward path,
  every perturbed estimator of every sequence stepped together (step_estimators): 1 + 2 P
  estimators per sequence for P parameters; loglik_gradient over (beta, gamma, Phi, H) of
  single-noise LTI systems, scale_gradient over the Cauchy scales (beta_1 .. beta_pncc,
  gamma_1 .. gamma_p) of LTI or LTV systems.  A component is reported as valid only when the
  term tree's discrete structure (term counts after MU/MUC and after reduction at every
  update) is identical at theta + h e_i and theta - h e_i.  They are the check of the
  tangent gradient.
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


SHRINK = (1e-7, 1e-8, 1e-9)


def _fd_gradient(systems, base, make_for, h, shrink, num_threads, max_live):
    """Central differences of L at base[i] (a parameter vector of systems[i]) over every
    component.  A component whose term structure differs at base +/- h from base's is run
    again at each step of `shrink` in turn, until the structure holds; it stays invalid
    otherwise.  Returns (L0, grad, valid, steps, terms generated, estimator runs)."""
    n = len(systems)
    P = len(base[0]) if n else 0

    def run(jobs):
        Ls, sts = run_likelihoods([systems[i] for i, _, _ in jobs], lambda k: make_for(systems[jobs[k][0]], jobs[k][2]),
                                  num_threads, max_live)
        return Ls, sts, sum(c[0] for st in sts for c in st)

    jobs = []
    for i in range(n):
        jobs.append((i, -1, base[i]))
        for j in range(P):
            for sg in (1.0, -1.0):
                t = base[i].copy()
                t[j] += sg * h
                jobs.append((i, j, t))
    L, st, terms = run(jobs)
    runs = len(jobs)
    L0 = np.zeros(n)
    S0 = [None] * n
    grad = np.zeros((n, P))
    valid = np.zeros((n, P), dtype=bool)
    steps = np.full((n, P), h)
    stride = 1 + 2 * P
    for i in range(n):
        b = i * stride
        L0[i], S0[i] = L[b], st[b]
        for j in range(P):
            lp, lm = b + 1 + 2 * j, b + 2 + 2 * j
            grad[i, j] = (L[lp] - L[lm]) / (2 * h)
            valid[i, j] = st[lp] == st[lm] == S0[i]
    for h2 in shrink:
        todo = [(i, j) for i in range(n) for j in range(P) if not valid[i, j]]
        if not todo:
            break
        jobs = []
        for i, j in todo:
            for sg in (1.0, -1.0):
                t = base[i].copy()
                t[j] += sg * h2
                jobs.append((i, j, t))
        L2, st2, t2 = run(jobs)
        terms += t2
        runs += len(jobs)
        for q, (i, j) in enumerate(todo):
            if st2[2 * q] == st2[2 * q + 1] == S0[i]:
                grad[i, j] = (L2[2 * q] - L2[2 * q + 1]) / (2 * h2)
                valid[i, j] = True
                steps[i, j] = h2
    return L0, grad, valid, steps, terms, runs


def scale_gradient(systems, h=1e-6, num_threads=16, max_live=None, shrink=SHRINK, details=False):
    """Central-difference gradient of L for every system over its Cauchy scales
    (beta_1 .. beta_pncc, gamma_1 .. gamma_p; the systems must share pncc and p), with the
    step shrunk where the term structure does not hold.  Returns (L, grad, valid), or with
    details a dict that adds the step of each component and the terms generated."""
    base = [np.concatenate([np.ravel(s.beta), np.ravel(s.gamma)]).astype(np.float64) for s in systems]
    if len({b.size for b in base}) > 1:
        raise ValueError("the systems must share the number of noise scales")
    L0, grad, valid, steps, terms, runs = _fd_gradient(
        systems, base, lambda s, t: make_estimator(s, t[:s.pncc], t[s.pncc:]), h, shrink, num_threads, max_live)
    if details:
        return {"L": L0, "grad": grad, "valid": valid, "h": steps, "terms": terms, "runs": runs}
    return L0, grad, valid


def loglik_gradient(systems, h=1e-6, params=None, num_threads=16, shrink=SHRINK, details=False):
    """Central-difference gradient of L for every single-noise LTI system over the
    parameters `params` (indices into PARAM_NAMES; all 14 by default), with the step shrunk
    where the term structure does not hold.  Returns (L, grad, valid), or with details a
    dict that adds the step of each component and the terms generated."""
    P = list(range(len(PARAM_NAMES))) if params is None else list(params)
    full = [theta_of(s) for s in systems]

    def make_for(s, sub):
        th = theta_of(s)
        th[P] = sub
        beta, gamma, Phi, H = _unpack(th, s.n)
        return make_estimator(s, beta, gamma, Phi, H)
    L0, grad, valid, steps, terms, runs = _fd_gradient(systems, [f[P].copy() for f in full], make_for, h, shrink,
                                                       num_threads, None)
    if details:
        return {"L": L0, "grad": grad, "valid": valid, "h": steps, "terms": terms, "runs": runs}
    return L0, grad, valid


# ---------------------------------------------------------------- tangent-kernel gradient
# The gradient of L by the engine's tangent kernels: one forward pass per window carries P
# derivative directions (mce_estimator_step_tangent), with the term tree's discrete
# structure that of the nominal pass.  Nothing is stored between steps: every step
# recomputes the forward with its tangents.  Directions are ("beta", j), ("gamma", i),
# ("Phi", r, c) (the same perturbation of Phi at every step) and ("H", i, j).

def theta_directions(s):
    """Directions of PARAM_NAMES for a single-noise LTI system: beta, gamma, Phi, H."""
    n = s.n
    return [("beta", 0), ("gamma", 0)] + [("Phi", r, c) for r in range(n) for c in range(n)] + \
        [("H", 0, j) for j in range(n)]


def scale_directions(s):
    """The Cauchy scales: beta_1 .. beta_pncc, gamma_1 .. gamma_p."""
    return [("beta", j) for j in range(s.pncc)] + [("gamma", i) for i in range(s.p)]


def _seeds(s, dirs, row):
    """(dPhi, dbeta, dH, dgamma) of measurement row `row` for the directions."""
    P, n = len(dirs), s.n
    dPhi = np.zeros((P, n, n))
    dbeta = np.zeros((P, s.pncc))
    dH = np.zeros((P, n))
    dgamma = np.zeros(P)
    for k, dr in enumerate(dirs):
        if dr[0] == "beta":
            dbeta[k, dr[1]] = 1.0
        elif dr[0] == "gamma":
            dgamma[k] = 1.0 if dr[1] == row else 0.0
        elif dr[0] == "Phi":
            dPhi[k, dr[1], dr[2]] = 1.0
        elif dr[0] == "H":
            if dr[1] == row:
                dH[k, dr[2]] = 1.0
        else:
            raise ValueError(f"unknown direction {dr}")
    return dPhi, dbeta, dH, dgamma


def window_loglik_tangent(s, dirs, beta=None, gamma=None, Phi=None, H=None):
    """(L, dL/dtheta along dirs, structure) of one window of system s: L = sum over every
    scalar update of log Re fz (fz before term reduction), computed on the GPU with the
    tangent kernels.  beta / gamma / Phi / H replace the system's own (Phi: one matrix for
    every step); structure is the (terms after MU/MUC, after reduction) of each update."""
    from .estimator import GpuEstimator
    beta = np.ravel(s.beta) if beta is None else np.ravel(beta)
    gamma = np.ravel(s.gamma) if gamma is None else np.ravel(gamma)
    H = s.H if H is None else np.asarray(H).reshape(s.p, s.n)
    est = GpuEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=s.cmcc, seed=s.srand_seed)
    try:
        est.enable_tangents(len(dirs))
        seeds = [_seeds(s, dirs, i) for i in range(s.p)]
        Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
        L, dL, structure = 0.0, np.zeros(len(dirs)), []
        for k in range(s.num_steps):
            Phik = s.Phi(k) if Phi is None else np.asarray(Phi)
            for i in range(s.p):
                r = est.step(s.zs[k, i], Phik.reshape(-1), Gt, beta, H[i], gamma[i], tangents=seeds[i])
                fz = r["fz_raw"].real
                L += float(np.log(fz))
                dL += r["dfz"].real / fz
                structure.append((r["nt_mu"], r["nt_ftr"]))
        return L, dL, structure
    finally:
        est.close()


def tangent_gradients(systems, thetas=None, dirs_of=theta_directions, num_threads=16):
    """L and dL/dtheta of every window by the tangent kernels, the windows stepped from
    num_threads host threads.  thetas[i] (optional) gives system i's (beta, gamma, Phi, H)
    vector in PARAM_NAMES order (single-noise LTI systems).  Returns (L, grad)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        s = systems[i]
        dirs = dirs_of(s)
        if thetas is None:
            L, g, _ = window_loglik_tangent(s, dirs)
        else:
            beta, gamma, Phi, H = _unpack(np.asarray(thetas[i], dtype=np.float64), s.n)
            L, g, _ = window_loglik_tangent(s, dirs, beta, gamma, Phi, H)
        return L, g

    with ThreadPoolExecutor(max_workers=max(1, num_threads)) as ex:
        res = list(ex.map(one, range(len(systems))))
    return np.array([r[0] for r in res]), np.array([r[1] for r in res])


def window_loglikelihood(theta, systems, num_threads=16):
    """L(theta) of a batch of single-noise LTI windows as a differentiable torch function:
    theta (batch, 2 + n^2 + n) float64 in PARAM_NAMES order -> L (batch,).  Its backward
    multiplies by the gradient that the forward's tangent kernels computed."""
    import torch

    class _WindowLogLikelihood(torch.autograd.Function):
        @staticmethod
        def forward(ctx, th):
            L, G = tangent_gradients(systems, th.detach().cpu().numpy(), num_threads=num_threads)
            ctx.save_for_backward(torch.from_numpy(G).to(th.dtype))
            return torch.from_numpy(L).to(th.dtype)

        @staticmethod
        def backward(ctx, gL):
            (G,) = ctx.saved_tensors
            return gL[:, None] * G

    return _WindowLogLikelihood.apply(theta)
