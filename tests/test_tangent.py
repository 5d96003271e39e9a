"""Gradient of the window log-likelihood by the engine's tangent kernels
(mce_estimator_step_tangent; cauchyfriendly_amd.likelihood.window_loglik_tangent).

L = sum over the scalar updates of log Re fz, fz being each update's measurement likelihood
before term reduction.  The tangent kernels carry P directions through time propagation,
the measurement update, coalignment, the moment sum and the g-table recursion, with the
nominal pass's discrete structure.  Pins:
* the first update has a closed form, fz = s / (pi (s^2 + r^2)) with
  s = sum_l p0_l |A0_l . H| + gamma and r = z - H b0, so its tangent along gamma, H, beta
  and Phi is fixed by calculus (beta and Phi do not enter before the first propagation);
* the tangent path's L equals the plain forward's and the CPU restatement's;
* on the 3-state LTI windows of 6 steps (all 14 parameters beta, gamma, Phi, H), every
  component whose discrete structure holds at theta +/- h agrees with the central
  difference of the GPU forward within 1e-5 (relative, plus the difference's own round-off
  bound) and, on a sample, with the CPU restatement's central difference;
* with two process noises and two measurements per step (5-state LTV), the scale
  gradient agrees with central differences the same way.
"""
import numpy as np
import pytest

from cauchyfriendly_amd import workloads as W

pytestmark = pytest.mark.gpu
H_STEP = 1e-6


def _fd_tol(L, h, g, rel):
    """Round-off of a central difference of L (~1e-12 relative per likelihood) amplified by
    1 / (2h), plus `rel` of the derivative."""
    return 1e-12 * abs(L) / h + rel * max(1.0, abs(g))


def test_first_update_tangent_matches_closed_form():
    from cauchyfriendly_amd.likelihood import theta_directions, _seeds
    from cauchyfriendly_amd.estimator import GpuEstimator
    s = W.lti_sequences(count=1)[0]
    dirs = theta_directions(s)
    print(f"case {s.name}: first update, {len(dirs)} directions", flush=True)
    est = GpuEstimator(s.num_steps, s.n, 1, 1, s.A0, s.p0, s.b0, seed=s.srand_seed)
    est.enable_tangents(len(dirs))
    r = est.step(s.zs[0, 0], s.Phi(0).reshape(-1), s.Gamma.reshape(-1), s.beta, s.H[0], s.gamma[0],
                 tangents=_seeds(s, dirs, 0))
    est.close()
    H = s.H[0]
    a = s.A0 @ H
    sc = float(np.sum(s.p0 * np.abs(a)) + s.gamma[0])
    res = float(s.zs[0, 0] - H @ s.b0)
    den = np.pi * (sc * sc + res * res) ** 2
    dfz_ds = (res * res - sc * sc) / den
    dfz_dr = -2 * sc * res / den
    assert abs(r["fz"].real - sc / (np.pi * (sc * sc + res * res))) <= 1e-14 * r["fz"].real
    expect = np.zeros(len(dirs))
    expect[1] = dfz_ds                                    # gamma
    for j in range(s.n):                                  # H_j: s through |A0_l . H|, r through H b0
        ds = float(np.sum(s.p0 * np.sign(a) * s.A0[:, j]))
        expect[2 + s.n * s.n + j] = dfz_ds * ds + dfz_dr * (-s.b0[j])
    got = r["dfz"]
    assert np.all(np.abs(got.imag) <= 1e-12 * np.max(np.abs(expect)))
    assert np.max(np.abs(got.real - expect)) <= 1e-12 * np.max(np.abs(expect)), (got.real, expect)


def test_tangent_likelihood_equals_forward_and_restatement():
    from cauchyfriendly_amd.likelihood import window_loglik_tangent, theta_directions, log_likelihoods, theta_of
    from test_likelihood import _oracle_L
    ss = W.lti_sequences(count=3)
    L, st = log_likelihoods(ss, [theta_of(s) for s in ss], num_threads=3)
    for i, s in enumerate(ss):
        print(f"case {s.name}", flush=True)
        lt, g, stt = window_loglik_tangent(s, theta_directions(s))
        assert stt == st[i]
        assert lt == L[i]
        lo, so = _oracle_L(s, theta_of(s))
        assert so == stt and abs(lt - lo) <= 1e-9 * abs(lo)
        assert np.all(np.isfinite(g))


def test_tangent_gradient_matches_central_differences_on_lti_windows():
    """All 14 components on all 64 of the 3-state windows against the GPU's central
    differences (components whose structure holds at theta +/- h, or at a smaller step)."""
    from cauchyfriendly_amd.likelihood import loglik_gradient, tangent_gradients
    ss = W.lti_sequences(count=64)
    print("case 64 windows x 14 parameters", flush=True)
    Lt, Gt = tangent_gradients(ss, num_threads=8)
    d = loglik_gradient(ss, h=H_STEP, num_threads=8, details=True)
    diff = np.flatnonzero(Lt != d["L"])
    assert diff.size == 0, ("windows whose forward log-likelihoods differ", diff.tolist(),
                            (Lt[diff] - d["L"][diff]).tolist())
    assert d["valid"].mean() > 0.9
    worst = 0.0
    for i in range(len(ss)):
        for j in range(Gt.shape[1]):
            if not d["valid"][i, j]:
                continue
            go, hj = d["grad"][i, j], d["h"][i, j]
            err = abs(Gt[i, j] - go)
            assert err <= _fd_tol(d["L"][i], hj, go, 1e-5), f"{ss[i].name} param {j}: {Gt[i, j]} vs {go} (h {hj})"
            worst = max(worst, err / max(1.0, abs(go)))
    print(f"  worst relative difference {worst:.2e}", flush=True)


def test_windows_built_on_several_threads_equal_one_thread():
    """Estimators created from 8 host threads at once draw their construction-time random
    values from their own seeds: L and the gradient equal the one-thread run bit for bit,
    on every one of three repeats."""
    from cauchyfriendly_amd.likelihood import tangent_gradients
    ss = W.lti_sequences(count=64)
    print("case 64 windows, 1 thread against 8 threads x 3", flush=True)
    L1, G1 = tangent_gradients(ss, num_threads=1)
    for rep in range(3):
        L8, G8 = tangent_gradients(ss, num_threads=8)
        diff = np.flatnonzero((L8 != L1) | np.any(G8 != G1, axis=1))
        assert diff.size == 0, ("repeat", rep, "windows that differ", diff.tolist())


def test_tangent_gradient_matches_restatement_central_differences():
    """A sample of components against the CPU restatement's central differences at h = 1e-6
    where the restatement's structure is the same at theta +/- h."""
    from cauchyfriendly_amd.likelihood import tangent_gradients, theta_of
    from test_likelihood import _oracle_L
    ss = W.lti_sequences(count=2)
    Lt, Gt = tangent_gradients(ss, num_threads=2)
    checked = 0
    for i, s in enumerate(ss):
        th = theta_of(s)
        L0, st0 = _oracle_L(s, th)
        for p in (0, 1, 2, 6, 10, 11, 13):
            tp, tm = th.copy(), th.copy()
            tp[p] += H_STEP
            tm[p] -= H_STEP
            (lp, sp), (lm, sm) = _oracle_L(s, tp), _oracle_L(s, tm)
            if not (sp == sm == st0):
                continue
            go = (lp - lm) / (2 * H_STEP)
            assert abs(Gt[i, p] - go) <= _fd_tol(L0, H_STEP, go, 1e-5), f"{s.name} param {p}: {Gt[i, p]} vs {go}"
            checked += 1
    assert checked >= 10


def test_scale_tangent_gradient_with_two_noises_and_two_measurements():
    """5-state LTV, two process noises and two measurements per step, 3 steps: the gradient
    over (beta_1, beta_2, gamma_1, gamma_2) against the GPU's central differences."""
    import dataclasses
    from cauchyfriendly_amd.likelihood import scale_gradient, tangent_gradients, scale_directions
    base = W.config1(steps=3)
    ws = [base, dataclasses.replace(base, zs=base.zs * 1.1, name="ltv5_p2_z11")]
    Lt, Gt = tangent_gradients(ws, dirs_of=scale_directions, num_threads=2)
    d = scale_gradient(ws, h=H_STEP, num_threads=2, details=True)
    diff = np.flatnonzero(Lt != d["L"])
    assert diff.size == 0, ("windows whose forward log-likelihoods differ", diff.tolist(),
                            (Lt[diff] - d["L"][diff]).tolist())
    for i, s in enumerate(ws):
        print(f"case {s.name}: tangent {Gt[i]} fd {d['grad'][i]}", flush=True)
        for j in range(Gt.shape[1]):
            if d["valid"][i, j]:
                go = d["grad"][i, j]
                assert abs(Gt[i, j] - go) <= _fd_tol(d["L"][i], d["h"][i, j], go, 1e-5), f"{s.name} scale {j}"


def test_autograd_function_returns_the_tangent_gradient():
    import torch
    from cauchyfriendly_amd.likelihood import window_loglikelihood, tangent_gradients, theta_of
    ss = W.lti_sequences(count=4)
    theta = torch.tensor(np.stack([theta_of(s) for s in ss]), dtype=torch.float64, requires_grad=True)
    L = window_loglikelihood(theta, ss, num_threads=4)
    w = torch.arange(1, 5, dtype=torch.float64)
    (L * w).sum().backward()
    Lt, Gt = tangent_gradients(ss, num_threads=4)
    assert np.array_equal(L.detach().numpy(), Lt)
    assert np.allclose(theta.grad.numpy(), w.numpy()[:, None] * Gt, rtol=0, atol=0)
