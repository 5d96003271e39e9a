"""Sliding-window bank and window re-initialisation (Speyer's initialisation).

Fixtures (tests/golden/make_window_golden.py) hold the reference's outputs: its
speyers_window_init on seeded inputs, and its PySlidingWindowManager run on the
reference C++ backend.  The CPU tests pin those fixtures with the mathematics of
the initialisation; the GPU tests compare this library's host initialisation
bit for bit and its window bank (over the MI355X engine) within 1e-9."""
import ctypes as ct
import json
import os

import numpy as np
import pytest

from conftest import rel_err
from cauchyfriendly_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
COV_EIG_TOL = -1e-5
TOL = 1e-9


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        return json.load(fh)


def test_speyer_fixture_is_the_restart_of_the_estimate():
    """A0 rows are the orthonormal eigenvectors a_i of M = V + V H^T H V / s with
    s = gamma^2 + (z - H x)^2 (V boosted on its diagonal when an eigenvalue is below
    the tolerance), p0_i = lambda_i |H a_i| gamma / (s + H V H^T) and
    b0 = x - V H^T (z - H x) / s."""
    for c in _load("speyer_init")["cases"]:
        n = c["n"]
        x, V, H = np.array(c["xhat"]), np.array(c["Phat"]).reshape(n, n), np.array(c["H"])
        g, z = c["gamma"], c["z"]
        A0, p0, b0 = np.array(c["A0"]).reshape(n, n), np.array(c["p0"]), np.array(c["b0"])
        w = np.linalg.eigvalsh(V)
        if w.min() < COV_EIG_TOL:
            V = V + (COV_EIG_TOL - w.min()) * np.eye(n)
        s = g * g + (z - H @ x) ** 2
        M = V + np.outer(V @ H, H @ V) / s
        assert np.allclose(A0 @ A0.T, np.eye(n), atol=1e-12)
        D = A0 @ M @ A0.T
        lam = np.diag(D)
        assert np.allclose(D - np.diag(lam), 0, atol=1e-10 * np.abs(M).max())
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(M), rtol=1e-10)
        ha = A0 @ H
        assert np.allclose(p0, lam * np.abs(ha) * g / (s + H @ V @ H), rtol=1e-10)
        assert np.allclose(b0, x - V @ H * (z - H @ x) / s, rtol=1e-12, atol=1e-14)


def test_speyer_fixture_covers_the_boost():
    cases = _load("speyer_init")["cases"]
    assert any(np.linalg.eigvalsh(np.array(c["Phat"]).reshape(c["n"], c["n"])).min() < COV_EIG_TOL for c in cases)


@pytest.mark.gpu
def test_speyer_init_matches_reference():
    from cauchyfriendly_amd.cauchy_estimator import speyers_window_init
    for c in _load("speyer_init")["cases"]:
        n = c["n"]
        A0, p0, b0 = speyers_window_init(np.array(c["xhat"]), np.array(c["Phat"]).reshape(n, n), np.array(c["H"]),
                                         c["gamma"], c["z"])
        assert np.array_equal(A0.reshape(-1), np.array(c["A0"])), f"n={n}: A0 differs"
        assert np.array_equal(p0, np.array(c["p0"])), f"n={n}: p0 differs"
        assert np.array_equal(b0, np.array(c["b0"])), f"n={n}: b0 differs"


def _run_bank(rec, concurrent=False):
    from cauchyfriendly_amd.cauchy_estimator import PySlidingWindowManager
    s = W.System.from_dict(rec["system"])
    n = s.n
    ct.CDLL("libc.so.6").srand(ct.c_uint(rec["srand"]))
    swm = PySlidingWindowManager(rec["mode"], rec["num_windows"], swm_debug_print=False, concurrent=concurrent)
    Gamma = s.Gamma if s.pncc > 1 else s.Gamma.reshape(-1)
    H = s.H if s.p > 1 else s.H.reshape(-1)
    beta = np.asarray(s.beta, dtype=np.float64).reshape(-1)
    gamma = np.asarray(s.gamma, dtype=np.float64).reshape(-1)
    if rec["mode"] == "lti":
        swm.initialize_lti(s.A0, s.p0, s.b0, s.Phi(0), None, Gamma, beta, H, gamma)
    else:
        def dyn(c_duc):
            duc = c_duc.contents
            k = int(duc.step)
            Phi = s.Phi(k).reshape(-1)
            for i in range(n * n):
                duc.Phi[i] = float(Phi[i])
            duc.step = k + 1
        swm.initialize_ltv(s.A0, s.p0, s.b0, s.Phi(0), None, Gamma, beta, H, gamma, dyn)
    out = []
    for k in range(rec["steps"]):
        xhat, Phat, wx, wP = swm.step(np.asarray(s.zs[k]).reshape(-1), None)
        out.append((xhat, Phat, wx, wP, swm.best_idx, swm.moment_info["fz"][-1], swm.moment_info["err_code"][-1],
                    swm.win_counts.copy()))
    swm.shutdown()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("concurrent", [False, True], ids=["sequential", "concurrent"])
@pytest.mark.parametrize("name", ["window_lti3", "window_ltv4_p2"])
def test_window_bank_matches_reference(name, concurrent):
    rec = _load(name)
    got = _run_bank(rec, concurrent)
    for k, (g, e) in enumerate(zip(got, rec["outputs"])):
        xhat, Phat, wx, wP, best, fz, err, counts = g
        where = f"{name} step {k}"
        assert best == e["best"], f"{where}: best window {best} != {e['best']}"
        assert list(counts) == e["win_counts"], f"{where}: window counts differ"
        assert int(err) == e["err"], f"{where}: error bits differ"
        assert rel_err(np.array([fz]), np.array([e["fz"]])) <= TOL, f"{where}: fz differs"
        for a, b, what in ((xhat, e["xhat"], "xhat"), (Phat, e["Phat"], "Phat"), (wx, e["wavg_xhat"], "avg xhat"),
                           (wP, e["wavg_Phat"], "avg Phat")):
            assert rel_err(np.asarray(a).reshape(-1), np.array(b)) <= TOL, f"{where}: {what} differs"
