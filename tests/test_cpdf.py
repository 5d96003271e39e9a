"""Marginal densities of the estimator (pycauchy_get_marginal_1D_pointwise_cpdf and
pycauchy_get_marginal_2D_pointwise_cpdf).

Fixtures (tests/golden/make_cpdf_golden.py) hold the reference's density grids after a
few steps of a 3-state LTI, a 5-state LTV (two measurements per step), a 2-state and a
scalar system.  The CPU tests restate both densities over the CPU restatement's terms
and pin them:
* 1D (include/cpdf_ndim.hpp:1233-1368: the arrangement cut along the state, two
  numerator lookups per term, the direct form at the first grid point and the cached
  form after it): against the fixtures, and by the closed form of its integral,
  sum_t (a_t s_t + b_t) pi / w_t = Re fz with sum_t a_t = 0, which any wrong sign vector
  or conjugation breaks;
* 2D (cpdf_ndim.hpp:31-204 and 1400-1735: per term the two columns, zero rows and +/-
  equal rows merged, the cells of a half turn between sorted wall angles, one numerator
  per cell and the closed-form angular integral): against the fixtures, and by
  integrating it over the second state (Gauss-Legendre in t, y = tan t) to the 1D
  marginal density of the first, which ties its cells, lookups and 2 / (2 pi)^2 factor
  to the pinned 1D density.
The GPU tests compare this library's densities with the fixtures through the C ABI."""
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
CASES = ["cpdf_lti3", "cpdf_ltv5_p2", "cpdf_scalar1", "cpdf_lti2"]
CASES_2D = ["cpdf_lti3", "cpdf_ltv5_p2", "cpdf_lti2"]
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


def _cells2d(t, i1, i2, bar_nu, mu_eps=1e-10, co_eps=1e-8):
    """Cells of a half turn of one term's cut arrangement: (theta1, theta2, Re gamma1,
    Re gamma2, g numerator) each, and the term's b for the two states."""
    m, A = t["m"], t["A"]
    wa = np.zeros((m, 2))
    wp = np.array(t["p"], dtype=float)
    cmap, csm, fidx, F = [0] * m, [0] * m, [0] * m, [True] * m
    for r in range(m):
        a0, a1 = A[r, i1], A[r, i2]
        if abs(a0) < mu_eps and abs(a1) < mu_eps:
            cmap[r] = csm[r] = fidx[r] = 32
            F[r] = False
        else:
            sa = abs(a0) + abs(a1)
            wa[r] = (a0 / sa, a1 / sa)
            wp[r] *= sa
    u, wp2 = 0, []
    for r in range(m):
        if not F[r]:
            continue
        cmap[r], csm[r], fidx[r] = u, 1, r
        wp2.append(wp[r])
        for j in range(r + 1, m):
            if not F[j]:
                continue
            for sign in (1, -1):
                if abs(wa[r, 0] - sign * wa[j, 0]) < co_eps and abs(wa[r, 1] - sign * wa[j, 1]) < co_eps:
                    cmap[j], csm[j], F[j], fidx[j] = u, sign, False, r
                    wp2[u] += wp[j]
                    break
        u += 1
    wa2 = np.array([wa[r] for r in range(m) if fidx[r] == r]).reshape(-1, 2)
    wp2 = np.array(wp2)
    th = []
    for a0, a1 in wa2:
        px, py = (1.0, -a0 / a1) if abs(a0) < abs(a1) else (-a1 / a0, 1.0)
        t1 = np.arctan2(py, px)
        t1 = t1 + np.pi if t1 < 0 else t1
        th += [t1, t1 + np.pi]
    th = np.sort(np.array(th))
    cells = []
    for c in range(u):
        tm = (th[c + 1] + th[c]) / 2.0
        sv = np.where(wa2 @ np.array([np.cos(tm), np.sin(tm)]) > 0, 1.0, -1.0)
        enc = 0
        for j in range(m):
            if cmap[j] == 32:
                enc |= (1 << j) if float(A[j] @ bar_nu) < 0 else 0
            elif sv[cmap[j]] * csm[j] < 0:
                enc |= 1 << j
        g = _lookup(t["keys"], t["g"], enc, m)
        cells.append((th[c], th[c + 1], -np.sum(wa2[:, 0] * wp2 * sv), -np.sum(wa2[:, 1] * wp2 * sv), g))
    return cells, t["b"][i1], t["b"][i2]


def _density2d(terms, i1, i2, X, Y, norm, bar_nu):
    """2 Re sum over terms and cells of g * [sin t / (g1 (g1 cos t + g2 sin t))] from t1 to t2,
    over (2 pi)^2 Re fz, with g_k = Re g_k + i (b_k - x_k)."""
    X, Y = np.asarray(X, dtype=float), np.asarray(Y, dtype=float)
    tot = np.zeros(X.shape)
    for t in terms:
        cells, bx, by = _cells2d(t, i1, i2, bar_nu)
        for th1, th2, g1, g2, g in cells:
            ga1, ga2 = g1 + 1j * (bx - X), g2 + 1j * (by - Y)
            p2, p1 = ga2 * ga1, ga1 * ga1
            v = (np.sin(th2) / (p1 * np.cos(th2) + p2 * np.sin(th2)) -
                 np.sin(th1) / (p1 * np.cos(th1) + p2 * np.sin(th1))) * g
            tot += v.real
    return 2 * tot * (1 / (2 * np.pi)) ** 2 / norm


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


@pytest.mark.parametrize("name", CASES_2D)
def test_cpdf2d_fixture_matches_restatement_over_oracle_terms(name):
    rec = _load(name)
    terms, norm, bar_nu = _oracle_state(rec)
    for g in rec["grids2d"]:
        Z = _density2d(terms, g["i1"], g["i2"], g["X"], g["Y"], norm, bar_nu)
        E = np.array(g["Z"])
        assert np.max(np.abs(Z - E)) / np.max(np.abs(E)) <= TOL, f"{name} states {g['i1']},{g['i2']}"


@pytest.mark.parametrize("name", CASES_2D)
def test_cpdf2d_integrates_to_the_1d_marginal(name):
    rec = _load(name)
    terms, norm, bar_nu = _oracle_state(rec)
    g = rec["grids2d"][0]
    i1, i2 = g["i1"], g["i2"]
    xh = np.array(rec["xhat"])
    tn, wn = np.polynomial.legendre.leggauss(1500)
    scale = (np.pi / 2) * (1 - 1e-9)
    tt, ww = tn * scale, wn * scale
    ys = xh[i2] + np.tan(tt)
    jac = 1 / np.cos(tt) ** 2
    for x in (xh[i1] - 0.3, xh[i1] + 0.1):
        Z = _density2d(terms, i1, i2, np.full(ys.shape, x), ys, norm, bar_nu)
        f1, _, _ = _density(terms, i1, [x, x], norm, bar_nu)
        assert abs(np.sum(Z * jac * ww) - f1[0]) <= 1e-7 * f1[0], f"{name} at x = {x}"


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
@pytest.mark.parametrize("name", CASES_2D)
def test_cpdf2d_matches_reference(name):
    rec = _load(name)
    est, s = _gpu_estimator(rec)
    for g in rec["grids"]:   # the same call sequence as the fixture (the draws on first use)
        est.get_marginal_1D_pointwise_cpdf(g["marg"], g["low"], g["high"], g["res"])
    for g in rec["grids2d"]:
        if s.n == 2:
            X, Y, Z = est.get_2D_pointwise_cpdf(g["xlow"], g["xhigh"], g["xres"], g["ylow"], g["yhigh"], g["yres"])
        else:
            X, Y, Z = est.get_marginal_2D_pointwise_cpdf(g["i1"], g["i2"], g["xlow"], g["xhigh"], g["xres"],
                                                         g["ylow"], g["yhigh"], g["yres"])
        E = np.array(g["Z"]).reshape(g["shape"])
        assert list(Z.shape) == g["shape"], f"{name}: grid shape"
        assert np.array_equal(X.reshape(-1), np.array(g["X"])) and np.array_equal(Y.reshape(-1), np.array(g["Y"]))
        assert np.max(np.abs(Z - E)) / np.max(np.abs(E)) <= TOL, f"{name} states {g['i1']},{g['i2']}: density"
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
