"""Relative system of two estimators (pycauchy_single_step_eval_2d_rsys_cpdf).

Fixtures (tests/golden/make_rsys_golden.py) hold the reference's relative moments and
density for two 3-state LTI estimators (two transformations, term approximation off and
on), two 5-state LTV estimators (p = 2) and two 2-state estimators.  The CPU tests pin
them:
* by the mathematics of a difference of independent states: the relative mean is
  Trel (x_s - x_p) and the relative covariance Trel (P_s + P_p) Trel^T, from the two
  estimators' own moments (to 1e-9 where no term of either was approximated away; the
  5-state pair carries the reduction's approximations, so 1e-4 there), with a
  normalisation factor of 1;
* by a numpy restatement of the pair construction (include/cauchy_prediction.hpp:207-471,
  497-650, 973-1072) over the CPU restatement's terms, for the two small systems.
The GPU tests compare this library's relative system with the fixtures."""
import ctypes as ct
import json
import os

import numpy as np
import pytest

from cauchyfriendly_amd import workloads as W
from oracle import driver as dr
from test_cpdf import _lookup

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TOL = 1e-9
CASES = ["rsys_lti3", "rsys_ltv5_p2", "rsys_lti2"]
RAND_MAX = 2147483647


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        return json.load(fh)


def _rel(a, b):
    a, b = np.asarray(a, dtype=float).ravel(), np.asarray(b, dtype=float).ravel()
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_rsys_moments_are_those_of_a_difference(name):
    rec = _load(name)
    P, S = rec["primary"], rec["secondary"]
    n = len(P["xhat"])
    tol = 1e-4 if name == "rsys_ltv5_p2" else TOL
    for e in rec["evals"]:
        T = np.array(e["Trel"]).reshape(2, n)
        assert abs(e["fz"][0] - 1.0) < 1e-9
        assert _rel(e["xhat"], T @ (np.array(S["xhat"]) - np.array(P["xhat"]))) <= tol
        Pc = np.array(S["Phat"]).reshape(n, n) + np.array(P["Phat"]).reshape(n, n)
        assert _rel(e["Phat"], T @ Pc @ T.T) <= tol


def _oracle(sysobj, steps):
    ct.CDLL("libc.so.6").srand(ct.c_uint(sysobj.srand_seed))
    est = dr.RefEstimator(sysobj.num_steps, sysobj.n, sysobj.pncc, sysobj.p, sysobj.A0, sysobj.p0, sysobj.b0,
                          cmcc=sysobj.cmcc, kind="oracle")
    Gt = np.ascontiguousarray(sysobj.Gamma.reshape(-1))
    for k in range(steps):
        for i in range(sysobj.p):
            r = est.step(sysobj.zs[k, i], sysobj.Phi(k).reshape(-1), Gt, sysobj.beta, sysobj.H[i], sysobj.gamma[i])
    terms = est.terms()
    est.close()
    return terms, r["fz_public"].real


def _remove_zeros_and_coalign(a2, p2, mu_eps=1e-10, co_eps=1e-8):
    m = len(p2)
    cmap, csm, fidx, F = [0] * m, [0] * m, [0] * m, [True] * m
    for r in range(m):
        if abs(a2[r, 0]) < mu_eps and abs(a2[r, 1]) < mu_eps:
            cmap[r] = csm[r] = fidx[r] = 32
            F[r] = False
        else:
            sa = abs(a2[r, 0]) + abs(a2[r, 1])
            a2[r] /= sa
            p2[r] *= sa
    u, wp2 = 0, []
    for r in range(m):
        if not F[r]:
            continue
        cmap[r], csm[r], fidx[r] = u, 1, r
        wp2.append(p2[r])
        for j in range(r + 1, m):
            if not F[j]:
                continue
            for sign in (1, -1):
                if abs(a2[r, 0] - sign * a2[j, 0]) < co_eps and abs(a2[r, 1] - sign * a2[j, 1]) < co_eps:
                    cmap[j], csm[j], F[j], fidx[j] = u, sign, False, r
                    wp2[u] += p2[j]
                    break
        u += 1
    wa2 = np.array([a2[r] for r in range(m) if fidx[r] == r]).reshape(-1, 2)
    return wa2, np.array(wp2), cmap, csm


def _rsys(tp_terms, ts_terms, Trel, nu, eps, norm_fact, grid):
    """Relative moments (nf, mean, cov) and the kept terms' cells; then the density on grid."""
    T = np.asarray(Trel).reshape(2, -1)
    nu2 = T @ nu
    nf, mean, cov = 0j, np.zeros(2, complex), np.zeros((2, 2), complex)
    kept = []
    for tp in tp_terms:
        for ts in ts_terms:
            mp, ms = tp["m"], ts["m"]
            Arel = np.vstack([-tp["A"], ts["A"]])
            prel = np.concatenate([tp["p"], ts["p"]]).astype(float)
            a2 = Arel @ T.T
            b2 = T @ (ts["b"] - tp["b"])
            p2 = prel.copy()
            wa2, wp2, cmap, csm = _remove_zeros_and_coalign(a2, p2)
            # moments at the cell of nu
            ty = np.zeros(2)
            encp = encs = 0
            for i in range(mp + ms):
                ds = a2[i] @ nu2 if cmap[i] != 32 else Arel[i] @ nu
                s = float(np.sign(ds))
                ty += a2[i] * (p2[i] * s)
                if s < 0:
                    if i < mp:
                        encp |= 1 << i
                    else:
                        encs |= 1 << (i - mp)
            g = _lookup(tp["keys"], tp["g"], encp, mp) * _lookup(ts["keys"], ts["g"], encs, ms)
            y = -ty + 1j * b2
            nf += g
            mean += g * y
            cov -= np.outer(g * y, y)
            # cells of a half turn
            th = []
            for x0, x1 in wa2:
                px, py = (1.0, -x0 / x1) if abs(x0) < abs(x1) else (-x1 / x0, 1.0)
                t1 = np.arctan2(py, px)
                t1 = t1 + np.pi if t1 < 0 else t1
                th += [t1, t1 + np.pi]
            th = np.sort(np.array(th))
            u = len(wa2)
            cells, total, keep = [], 0.0, False
            for c in range(u):
                tm = (th[c + 1] + th[c]) / 2.0
                sv = np.where(wa2 @ np.array([np.cos(tm), np.sin(tm)]) > 0, 1.0, -1.0)
                enc = 0
                for j in range(mp + ms):
                    if cmap[j] == 32:
                        enc |= (1 << j) if float(Arel[j] @ nu) < 0 else 0
                    elif sv[cmap[j]] * csm[j] < 0:
                        enc |= 1 << j
                gc = _lookup(tp["keys"], tp["g"], enc & ((1 << mp) - 1), mp) * \
                    _lookup(ts["keys"], ts["g"], enc >> mp, ms)
                g1, g2 = -np.sum(wa2[:, 0] * wp2 * sv), -np.sum(wa2[:, 1] * wp2 * sv)
                cells.append((th[c], th[c + 1], g1, g2, gc))
                s1, c1, s2, c2 = np.sin(th[c]), np.cos(th[c]), np.sin(th[c + 1]), np.cos(th[c + 1])
                den1, den2 = g1 * (g1 * c1 + g2 * s1), g1 * (g1 * c2 + g2 * s2)
                if abs(den1) < 1e-14 or abs(den2) < 1e-14:
                    keep = True
                total += abs(abs(gc) * s2 / den2) + abs(abs(gc) * s1 / den1)
            if keep or not abs(total * 2 * norm_fact) < eps:
                kept.append((cells, b2))
    mean = mean / (1j * nf.real)
    cov = cov / nf - np.outer(mean, mean)
    xlo, xhi, dx, ylo, yhi, dy = grid
    x0, x1, y0, y1 = xlo + mean[0].real, xhi + mean[0].real, ylo + mean[1].real, yhi + mean[1].real
    nx, ny = int(np.ceil((x1 - x0) / dx + 1)), int(np.ceil((y1 - y0) / dy + 1))
    xs = x0 + dx * np.arange(nx)
    xs[-1] = x1
    ys = y0 + dy * np.arange(ny)
    ys[-1] = y1
    X, Y = np.repeat(xs, ny), np.tile(ys, nx)
    Z = np.zeros(nx * ny)
    for cells, b2 in kept:
        for th1, th2, g1, g2, gc in cells:
            ga1, ga2 = g1 + 1j * (b2[0] - X), g2 + 1j * (b2[1] - Y)
            q2, q1 = ga2 * ga1, ga1 * ga1
            Z += ((np.sin(th2) / (q1 * np.cos(th2) + q2 * np.sin(th2)) -
                   np.sin(th1) / (q1 * np.cos(th1) + q2 * np.sin(th1))) * gc).real
    return nf, mean, cov, np.stack([X, Y, Z], axis=1).ravel(), nx, ny


@pytest.mark.parametrize("name", ["rsys_lti3", "rsys_lti2"])
def test_rsys_fixture_matches_restatement_over_oracle_terms(name):
    rec = _load(name)
    sp, ss = W.System.from_dict(rec["primary"]["system"]), W.System.from_dict(rec["secondary"]["system"])
    tp_terms, fzp = _oracle(sp, rec["primary"]["steps"])
    ts_terms, fzs = _oracle(ss, rec["secondary"]["steps"])
    assert len(tp_terms) == rec["primary"]["Nt"] and len(ts_terms) == rec["secondary"]["Nt"]
    libc = ct.CDLL("libc.so.6")
    n = sp.n
    for e in rec["evals"]:
        nu = np.array([2 * ((libc.rand() + 1.0) / (RAND_MAX + 2.0)) - 1 for _ in range(n)])
        norm_fact = (1 / (2 * np.pi)) ** 2 / fzp / fzs
        nf, mean, cov, pts, nx, ny = _rsys(tp_terms, ts_terms, e["Trel"], nu, e["eps"], norm_fact, e["grid"])
        assert (nx, ny) == (e["nx"], e["ny"])
        assert _rel([nf.real], e["fz"]) <= TOL and _rel(mean.real, e["xhat"]) <= TOL
        assert _rel(cov.real.ravel(), e["Phat"]) <= TOL
        got = pts.reshape(-1, 3)
        exp = np.array(e["points"]).reshape(-1, 3)
        # the grid sits on the computed relative mean, so its points agree to rounding
        assert np.allclose(got[:, :2], exp[:, :2], rtol=0, atol=1e-12)
        z = 2 * got[:, 2] * (1 / (2 * np.pi)) ** 2 / (fzp * fzs)
        assert _rel(z, exp[:, 2]) <= TOL, f"{name} eps {e['eps']}: density"


def _gpu_pair(rec):
    from cauchyfriendly_amd.cauchy_estimator import PyCauchyEstimator
    ests = []
    for key in ("primary", "secondary"):
        s = W.System.from_dict(rec[key]["system"])
        ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
        est = PyCauchyEstimator(rec["mode"], s.num_steps, False)
        Gamma = s.Gamma if s.pncc > 1 else s.Gamma.reshape(-1)
        H = s.H if s.p > 1 else s.H.reshape(-1)
        if rec["mode"] == "lti":
            est.initialize_lti(s.A0, s.p0, s.b0, s.Phi(0), None, Gamma, s.beta, H, s.gamma)
        else:
            def dyn(c_duc, s=s):
                duc = c_duc.contents
                k = int(duc.step)
                Phi = s.Phi(k).reshape(-1)
                for i in range(s.n * s.n):
                    duc.Phi[i] = float(Phi[i])
                duc.step = k + 1
            est.initialize_ltv(s.A0, s.p0, s.b0, s.Phi(0), None, Gamma, s.beta, H, s.gamma, dyn)
        for k in range(rec[key]["steps"]):
            est.step(s.zs[k], None)
        ests.append(est)
    return ests


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_rsys_matches_reference(name):
    from cauchyfriendly_amd.cauchy_estimator import eval_2d_relative_system_cpdf
    rec = _load(name)
    P, S = _gpu_pair(rec)
    for e in rec["evals"]:
        fz, xh, Ph, cfz, cxh, cPh, pts, nx, ny = eval_2d_relative_system_cpdf(e["Trel"], S, P, e["eps"], *e["grid"])
        assert (nx, ny) == (e["nx"], e["ny"]), f"{name}: grid"
        assert _rel(fz, e["fz"]) <= TOL and _rel(xh, e["xhat"]) <= TOL and _rel(Ph, e["Phat"]) <= TOL
        got, exp = pts.reshape(-1, 3), np.array(e["points"]).reshape(-1, 3)
        assert np.allclose(got[:, :2], exp[:, :2], rtol=0, atol=1e-12), f"{name}: grid points"
        assert _rel(got[:, 2], exp[:, 2]) <= TOL, f"{name} eps {e['eps']}: density"
    P.shutdown()
    S.shutdown()
