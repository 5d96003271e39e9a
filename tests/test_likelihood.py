"""Window log-likelihood and its central-difference gradient (cauchyfriendly_amd.likelihood).

L = sum_k log fz_k over the updates of a 6-step window of the 3-state LTI system, for
seeded measurement sequences; the gradient is over (beta, gamma, Phi, H).  CPU tests pin
the oracle side: fz of the first update has the closed form s / (pi (s^2 + r^2)) with
s = sum_l p0_l |A0_l . H| + gamma and r = z - H b0, so d log fz / d gamma = 1/s - 2 s / (s^2 + r^2),
which the oracle's central difference must reproduce.  GPU tests compare this library's
likelihoods with the oracle's (1e-9) and its gradient with the oracle's central
differences at the same step (1e-6), component by component where the discrete
structure allows it."""
import ctypes as ct

import numpy as np
import pytest

from cauchyfriendly_amd import workloads as W
from oracle import driver as dr

H_STEP = 1e-6


def _oracle_L(s, theta, steps=None):
    from cauchyfriendly_amd.likelihood import _unpack
    beta, gamma, Phi, H = _unpack(theta, s.n)
    steps = s.num_steps if steps is None else steps
    est = dr.RefEstimator(steps, s.n, 1, 1, s.A0, s.p0, s.b0, seed=s.srand_seed, kind="oracle")
    L = 0.0
    structure = []
    for k in range(steps):
        r = est.step(s.zs[k, 0], Phi.reshape(-1), s.Gamma.reshape(-1), beta, H.reshape(-1), gamma[0])
        L += np.log(r["fz_raw"].real)
        structure.append((r["nt_mu"], r["nt_ftr"]))
    est.close()
    return L, structure


def _oracle_grad(s, params, steps=None):
    from cauchyfriendly_amd.likelihood import theta_of
    th = theta_of(s)
    g = []
    for p in params:
        tp, tm = th.copy(), th.copy()
        tp[p] += H_STEP
        tm[p] -= H_STEP
        g.append((_oracle_L(s, tp, steps)[0] - _oracle_L(s, tm, steps)[0]) / (2 * H_STEP))
    return np.array(g)


def test_first_update_likelihood_gradient_closed_form():
    s = W.lti_sequences(count=1)[0]
    from cauchyfriendly_amd.likelihood import theta_of
    th = theta_of(s)
    H = th[11:14]
    sc = float(np.sum(s.p0 * np.abs(s.A0 @ H)) + th[1])
    r = float(s.zs[0, 0] - H @ s.b0)
    L1, _ = _oracle_L(s, th, steps=1)
    assert abs(L1 - np.log(sc / (np.pi * (sc * sc + r * r)))) < 1e-12
    g = _oracle_grad(s, [1], steps=1)[0]
    assert abs(g - (1 / sc - 2 * sc / (sc * sc + r * r))) < 1e-6 * abs(g)


def test_sequences_are_seeded_and_distinct():
    a, b = W.lti_sequences(count=2), W.lti_sequences(count=2)
    assert np.array_equal(a[0].zs, b[0].zs) and not np.array_equal(a[0].zs, a[1].zs)


@pytest.mark.gpu
def test_likelihoods_match_oracle():
    from cauchyfriendly_amd.likelihood import log_likelihoods, theta_of
    ss = W.lti_sequences(count=6)
    L, structure = log_likelihoods(ss, [theta_of(s) for s in ss], num_threads=4)
    for s, l, st in zip(ss, L, structure):
        lo, so = _oracle_L(s, theta_of(s))
        assert st == so, f"{s.name}: structure"
        assert abs(l - lo) <= 1e-9 * abs(lo), f"{s.name}: L {l} != {lo}"


def _fd_tol(L, h, go, rel):
    """Bound on the difference of two central differences whose likelihoods agree to ~1e-12
    relative: round-off in L amplified by 1 / (2h), plus `rel` of the derivative."""
    return 1e-12 * abs(L) / h + rel * max(1.0, abs(go))


@pytest.mark.gpu
def test_gradient_matches_oracle_central_differences():
    """Each valid component is compared with the restatement's central difference at the
    step the component used (the step shrinks where the term structure changes)."""
    from cauchyfriendly_amd.likelihood import loglik_gradient, theta_of
    ss = W.lti_sequences(count=3)
    params = [0, 1, 2, 6, 11, 13]   # beta, gamma, two Phi entries, two H entries
    d = loglik_gradient(ss, h=H_STEP, params=params, num_threads=4, details=True)
    assert d["valid"].mean() > 0.9
    for i, s in enumerate(ss):
        th = theta_of(s)
        for j, p in enumerate(params):
            if not d["valid"][i, j]:
                continue
            hj = d["h"][i, j]
            tp, tm = th.copy(), th.copy()
            tp[p] += hj
            tm[p] -= hj
            go = (_oracle_L(s, tp)[0] - _oracle_L(s, tm)[0]) / (2 * hj)
            assert abs(d["grad"][i, j] - go) <= _fd_tol(d["L"][i], hj, go, 1e-6), f"{s.name} param {p} (h {hj})"


def _truncated(s, steps):
    import dataclasses
    return dataclasses.replace(s, zs=s.zs[:steps], xs=s.xs[:steps], Phis=list(s.Phis[:steps]))


def _oracle_L_scales(s, beta, gamma):
    """L over every scalar update of an LTI or LTV system on the CPU restatement."""
    est = dr.RefEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed, kind="oracle")
    L = 0.0
    structure = []
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    for k in range(s.num_steps):
        for i in range(s.p):
            r = est.step(s.zs[k, i], s.Phi(k).reshape(-1), Gt, np.asarray(beta), s.H[i], gamma[i])
            L += np.log(r["fz_raw"].real)
            structure.append((r["nt_mu"], r["nt_ftr"]))
    est.close()
    return L, structure


@pytest.mark.gpu
def test_scale_gradient_of_orbit_window_matches_oracle():
    """Two range measurements per step on the 7-state orbit (LTV through the dynamics callback):
    L summed over both updates of every step, and its (beta, gamma_1, gamma_2) central
    differences, against the CPU restatement's at the same step h."""
    from cauchyfriendly_amd.likelihood import scale_gradient
    ws = [_truncated(s, 3) for s in W.leo7_windows(count=2)]
    d = scale_gradient(ws, h=H_STEP, num_threads=2, max_live=4, details=True)
    for i, s in enumerate(ws):
        sc = np.concatenate([s.beta, s.gamma])
        lo, so = _oracle_L_scales(s, s.beta, s.gamma)
        assert abs(d["L"][i] - lo) <= 1e-9 * abs(lo), f"{s.name}: L {d['L'][i]} != {lo}"
        for j in range(sc.size):
            hj = d["h"][i, j]
            tp, tm = sc.copy(), sc.copy()
            tp[j] += hj
            tm[j] -= hj
            lp, sp = _oracle_L_scales(s, tp[:1], tp[1:])
            lm, sm = _oracle_L_scales(s, tm[:1], tm[1:])
            assert bool(d["valid"][i, j]) == (sp == sm == so), f"{s.name} scale {j}: validity at h {hj}"
            if d["valid"][i, j]:
                go = (lp - lm) / (2 * hj)
                assert abs(d["grad"][i, j] - go) <= _fd_tol(lo, hj, go, 1e-5), \
                    f"{s.name} scale {j} (h {hj}): {d['grad'][i, j]} != {go}"
    assert d["valid"].any()


@pytest.mark.gpu
def test_likelihood_sums_every_update_of_a_step():
    """p = 2: run_likelihoods adds log fz of both updates of each step (the oracle's L)."""
    from cauchyfriendly_amd.likelihood import make_estimator, run_likelihoods
    s = _truncated(W.ltv_system(d=5, steps=4, seed_offset=61, name="ltv5_lik"), 4)
    L, structure = run_likelihoods([s], lambda i: make_estimator(s), num_threads=1)
    lo, so = _oracle_L_scales(s, s.beta, s.gamma)
    assert structure[0] == so and len(so) == 2 * s.num_steps
    assert abs(L[0] - lo) <= 1e-9 * abs(lo)
