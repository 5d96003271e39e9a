"""1D marginal density of the estimator (pycauchy_get_marginal_1D_pointwise_cpdf).

Fixtures (tests/golden/make_cpdf_golden.py) hold the reference's density grids after a
few steps of a 3-state LTI, a 5-state LTV (two measurements per step) and a scalar
system.  The CPU tests restate the density over the CPU restatement's terms
(include/cpdf_ndim.hpp:1233-1368: the arrangement cut along the state, two numerator
lookups per term, the direct form at the first grid point and the cached form after
it) and pin it twice: against the fixtures, and by the closed form of its integral,
sum_t (a_t s_t + b_t) pi / w_t = Re fz with sum_t a_t = 0, which any wrong sign vector
or conjugation breaks.  The GPU tests compare this library's density with the
fixtures through the C ABI."""
import ctypes as ct
import json
import os

import numpy as np
import pytest

from cauchyfriendly_amd import workloads as W
from oracle import driver as dr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TOL = 1e-9
CASES = ["cpdf_lti3", "cpdf_ltv5_p2", "cpdf_scalar1"]
RAND_MAX = 2147483647


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        return json.load(fh)


def _lookup(keys, g, enc, m, conj_rule=True):
    top, rev = 1 << (m - 1), (1 << m) - 1
    cj = bool(enc & top)
    key = (rev ^ enc) if cj else enc
    i = int(np.searchsorted(keys, key))
    if i < len(keys) and keys[i] == key:
        return np.conj(g[i]) if (cj and conj_rule) else g[i]
    return 0j


def _density(terms, marg, X, norm, bar_nu, conj_rule=True):
    """Density on the grid X and the closed-form integral terms (sum (a s + b) pi / w, sum a)."""
    direct = 0.0
    cache = []
    for t in terms:
        m, A, p, bm = t["m"], t["A"], t["p"], t["b"][marg]
        w = 0.0
        lhs = rhs = 0
        for j in range(m):
            acj = A[j, marg]
            w += p[j] * abs(acj)
            pos = (acj > 0) if abs(acj) > 1e-15 else (float(A[j] @ bar_nu) > 0)
            if pos:
                lhs |= 1 << j
            else:
                rhs |= 1 << j
        gl = _lookup(t["keys"], t["g"], lhs, m, conj_rule)
        gr = _lookup(t["keys"], t["g"], rhs, m, conj_rule)
        bc = bm - X[0]
        direct += (gl / complex(w, bc) - gr / complex(-w, bc)).real
        cache.append((gr.imag / np.pi, (gr.real * w - gr.imag * bm) / np.pi, w, bm))
    c = np.array(cache)
    Y = np.empty(len(X))
    Y[0] = direct / (2 * np.pi) / norm
    x = np.asarray(X[1:])[:, None]
    Y[1:] = ((c[:, 0] * x + c[:, 1]) / (c[:, 2] ** 2 + (x - c[:, 3]) ** 2)).sum(axis=1) / norm
    integral = np.sum((c[:, 0] * c[:, 3] + c[:, 1]) * np.pi / c[:, 2]) / norm
    return Y, integral, np.sum(c[:, 0]) / norm


def _oracle_state(rec):
    """The CPU restatement stepped as the fixture's estimator was; its terms, public fz
    and the d draws the reference's density object takes from rand() on first use."""
    s = W.System.from_dict(rec["system"])
    libc = ct.CDLL("libc.so.6")
    libc.srand(ct.c_uint(s.srand_seed))
    est = dr.RefEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=s.cmcc, kind="oracle")
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    for k in range(rec["steps"]):
        for i in range(s.p):
            r = est.step(s.zs[k, i], s.Phi(k).reshape(-1), Gt, s.beta, s.H[i], s.gamma[i])
    terms = est.terms()
    est.close()
    bar_nu = np.array([2 * ((libc.rand() + 1.0) / (RAND_MAX + 2.0)) for _ in range(s.n)])
    return terms, r["fz_public"].real, bar_nu


@pytest.mark.parametrize("name", CASES)
def test_cpdf_fixture_matches_restatement_over_oracle_terms(name):
    rec = _load(name)
    terms, norm, bar_nu = _oracle_state(rec)
    assert len(terms) == rec["Nt"]
    for g in rec["grids"]:
        Y, _, _ = _density(terms, g["marg"], g["X"], norm, bar_nu)
        E = np.array(g["Y"])
        assert np.max(np.abs(Y - E)) / np.max(np.abs(E)) <= TOL, f"{name} state {g['marg']}"


@pytest.mark.parametrize("name", CASES)
def test_cpdf_integrates_to_one(name):
    """Integral of (a x + b) / (w^2 + (x - s)^2) over the line is (a s + b) pi / w when the
    a's cancel: the marginal density must integrate to one."""
    rec = _load(name)
    terms, norm, bar_nu = _oracle_state(rec)
    for marg in range(W.System.from_dict(rec["system"]).n):
        _, integral, suma = _density(terms, marg, [0.0, 1.0], norm, bar_nu)
        assert abs(integral - 1.0) < 1e-9 and abs(suma) < 1e-9, f"{name} state {marg}: {integral}, {suma}"


def test_cpdf_integral_pin_catches_a_dropped_conjugation():
    rec = _load("cpdf_lti3")
    terms, norm, bar_nu = _oracle_state(rec)
    _, integral, suma = _density(terms, 0, [0.0, 1.0], norm, bar_nu, conj_rule=False)
    assert abs(integral - 1.0) > 1e-6 or abs(suma) > 1e-6


def _gpu_estimator(rec):
    from cauchyfriendly_amd.cauchy_estimator import PyCauchyEstimator
    s = W.System.from_dict(rec["system"])
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    est = PyCauchyEstimator(rec["mode"], s.num_steps, False)
    Gamma = s.Gamma if s.pncc > 1 else s.Gamma.reshape(-1)
    H = s.H if s.p > 1 else s.H.reshape(-1)
    if rec["mode"] == "lti":
        est.initialize_lti(s.A0, s.p0, s.b0, s.Phi(0), None, Gamma, s.beta, H, s.gamma)
    else:
        def dyn(c_duc):
            duc = c_duc.contents
            k = int(duc.step)
            Phi = s.Phi(k).reshape(-1)
            for i in range(s.n * s.n):
                duc.Phi[i] = float(Phi[i])
            duc.step = k + 1
        est.initialize_ltv(s.A0, s.p0, s.b0, s.Phi(0), None, Gamma, s.beta, H, s.gamma, dyn)
    for k in range(rec["steps"]):
        est.step(s.zs[k], None)
    return est, s


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_cpdf_matches_reference(name):
    rec = _load(name)
    est, s = _gpu_estimator(rec)
    for g in rec["grids"]:
        if s.n == 1:
            X, Y = est.get_1D_pointwise_cpdf(g["low"], g["high"], g["res"])
        else:
            X, Y = est.get_marginal_1D_pointwise_cpdf(g["marg"], g["low"], g["high"], g["res"])
        E = np.array(g["Y"])
        assert np.array_equal(X, np.array(g["X"])), f"{name} state {g['marg']}: grid differs"
        assert np.max(np.abs(Y - E)) / np.max(np.abs(E)) <= TOL, f"{name} state {g['marg']}: density differs"
    est.shutdown()


@pytest.mark.gpu
def test_cpdf_after_last_update_has_no_points():
    rec = _load("cpdf_scalar1")
    est, s = _gpu_estimator(rec)
    for k in range(rec["steps"], s.num_steps):
        est.step(s.zs[k], None)
    X, Y = est.get_1D_pointwise_cpdf(-1.0, 1.0, 0.1)
    assert X.size == 0 and Y.size == 0
    est.shutdown()
