"""Deterministic transform of an estimator's characteristic function between updates
(pycauchy_single_step_deterministic_transform, scripts/swig/cauchy/pycauchy.hpp:654-664).

Fixtures (tests/golden/make_transform_golden.py) hold the reference's moments, error codes
and term counts for every update of a run that transforms by (T, bias) after update `at`:
two 3-state LTI runs (with and without a bias) and a 4-state LTV run with two measurements
per step.

The CPU test pins them by what the mathematics fixes: the reference's time propagation maps
each term's hyperplanes A -> A Phi^T and its centre b -> Phi b + B u
(include/cauchy_term.hpp:435-456), and the transform maps A -> A T^T, b -> T b + bias
(include/cauchy_estimator.hpp:1312-1355).  Transforming and then propagating by Phi is
therefore one propagation by Phi T with the control B u = Phi bias, which the CPU
restatement runs as an ordinary controlled update.  The two orders round differently, and
the reduction's approximations (terms merged or g-values dropped at a threshold) turn such
differences into covariance changes far above the rounding: a 1e-15 relative perturbation
of Phi alone moves P by up to 5e-7 on these 3-state systems.  So that check holds term
counts and error codes exactly, fz and xhat within 1e-9 (measured: 1e-10) and P within
1e-4 (measured: 1.2e-5); a transposed T, a dropped bias or a wrong sign moves xhat by O(1).
The GPU test replays the fixtures on this library's estimator, which applies the transform
in the reference's order with unfused multiply-adds: term counts and error codes exactly,
fz, xhat and P within 1e-9."""
import ctypes as ct
import json
import os

import numpy as np
import pytest

from conftest import rel_err
from cauchyfriendly_amd import workloads as W
from oracle import driver as dr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = ["transform_lti3", "transform_lti3_nobias", "transform_ltv4_p2"]
TOL = 1e-9
COMPOSED_P_TOL = 1e-4


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        return json.load(fh)


def _check(rec, k, fz, x, P, err, nt, where, p_tol=TOL):
    n = len(rec["x"][0])
    assert int(err) == rec["err_code"][k], f"{where}: error code"
    if rec["Nt"][k] is not None:
        assert nt == rec["Nt"][k], f"{where}: terms {nt} != {rec['Nt'][k]}"
    assert rel_err(np.array([fz]), np.array([rec["fz"][k]])) <= TOL, f"{where}: fz"
    assert rel_err(np.asarray(x).reshape(-1), np.array(rec["x"][k])) <= TOL, f"{where}: xhat"
    assert rel_err(np.asarray(P).reshape(-1), np.array(rec["P"][k]).reshape(n * n)) <= p_tol, f"{where}: Phat"


@pytest.mark.parametrize("name", CASES)
def test_transform_fixture_is_a_composed_propagation(name):
    rec = _load(name)
    s = W.System.from_dict(rec["system"])
    n, at = s.n, rec["at"]
    T = np.array(rec["Trans"]).reshape(n, n)
    bias = np.array(rec["bias"])
    assert 0 < at < s.num_steps - 1 and abs(np.linalg.det(T)) > 0.1
    est = dr.RefEstimator(s.num_steps, n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=n, seed=s.srand_seed, kind="oracle")
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    zero_B, zero_u = np.zeros(n * n), np.zeros(n)
    for k in range(s.num_steps):
        Phi = s.Phi(k)
        for i in range(s.p):
            if k == at and i == 0:
                r = est.step(s.zs[k, i], (Phi @ T).reshape(-1), Gt, s.beta, s.H[i], s.gamma[i],
                             Phi.reshape(-1), bias)
            else:
                r = est.step(s.zs[k, i], Phi.reshape(-1), Gt, s.beta, s.H[i], s.gamma[i], zero_B, zero_u)
        _check(rec, k, r["fz_public"].real, r["mean"].real, r["cov"].real, r["err"], r["nt_ftr"], f"{name} step {k}",
               p_tol=COMPOSED_P_TOL)
    est.close()


def test_transform_fixtures_move_the_estimate():
    """Each fixture's transform is not a no-op: a run without it ends elsewhere."""
    for name in CASES:
        rec = _load(name)
        s = W.System.from_dict(rec["system"])
        _, recs = dr.run_system(s, kind="oracle")
        last = recs[-1]
        assert rel_err(last["mean"].real, np.array(rec["x"][-1])) > 1e-3, name


def _product_run(rec):
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
    nts = []
    for k in range(s.num_steps):
        if k == rec["at"]:
            est.deterministic_transform(np.array(rec["Trans"]), np.array(rec["bias"]))
        est.step(s.zs[k], None)
        nts.append(est.get_num_CF_terms())
    mi = est.moment_info
    out = [(float(np.real(mi["fz"][k])), mi["x"][k], mi["P"][k], int(np.asarray(mi["err_code"][k]).reshape(-1)[-1]),
            nts[k]) for k in range(s.num_steps)]
    est.shutdown()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_transform_matches_reference(name):
    rec = _load(name)
    for k, (fz, x, P, err, nt) in enumerate(_product_run(rec)):
        _check(rec, k, fz, x, P, err, nt, f"{name} step {k}")
