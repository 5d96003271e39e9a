"""Debug mode (debug_print / print_basic_info): after every term reduction the reference
recomputes its public fz, mean and covariance from the reduced terms, each term's g looked up
in its own g-table at the sign vector of its hyperplanes at the root point
(compute_moments(false), include/cauchy_estimator.hpp:524-602, with eval_g_yei_after_ftr,
include/cauchy_term.hpp:403-433); on the first update these also become the values the
numeric checks fall back to.  The fixtures (tests/golden/debug_*.json, written by
tests/golden/make_debug_golden.py) are the reference's own pycauchy outputs with
debug_print = true: a 3-state LTI window of 6 steps and a 4-state LTV window (two noises, two
measurements per step) of 4 steps."""
import ctypes as ct

import numpy as np
import pytest

from conftest import load_golden, rel_err
from cauchyfriendly_amd import workloads as W
from cauchyfriendly_amd.cauchy_estimator import C_CauchyDynamicsUpdateContainer
from oracle import driver as dr

TOL = 1e-9
CASES = ["debug_lti3", "debug_ltv4_p2"]


def _expected(rec):
    return [(np.asarray(o["fz"]), np.asarray(o["xhat"]), np.asarray(o["Phat"]), o["err"]) for o in rec["outputs"]]


@pytest.mark.parametrize("name", CASES)
def test_restatement_debug_moments_match_reference(name):
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    est = dr.RefEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed, kind="oracle")
    est.set_debug(True)
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    for k, (fz, x, P, err) in enumerate(_expected(rec)):
        # the LTV callback reads the container's step, which a single step leaves at 0
        Phi = s.Phi(0).reshape(-1)
        for i in range(s.p):
            r = est.step(s.zs[k, i], Phi, Gt, s.beta, s.H[i], s.gamma[i])
            f, xh, Ph = est.public_moments()
            where = f"{name} step {k} row {i}"
            assert r["err"] == err[i], f"{where}: error bits"
            assert rel_err(np.array([f.real]), fz[i:i + 1]) <= TOL, f"{where}: fz"
            assert rel_err(xh, x[i * s.n:(i + 1) * s.n]) <= TOL, f"{where}: xhat"
            assert rel_err(Ph.reshape(-1), P[i * s.n * s.n:(i + 1) * s.n * s.n]) <= TOL, f"{where}: Phat"
    est.close()


def test_debug_moments_differ_from_the_default():
    """Without debug mode the public fz after a reduction is 1; with it, the reduced terms'."""
    rec = load_golden("debug_lti3")
    fz_after = [o["fz"][0] for o in rec["outputs"][:-1]]
    assert all(abs(f - 1.0) < 1e-6 for f in fz_after) and any(f != 1.0 for f in fz_after)


def _dynamics(s):
    n = s.n

    def dyn(c_duc):
        duc = c_duc.contents
        k = min(max(int(duc.step), 0), s.num_steps - 1)
        Phi = s.Phi(k).reshape(-1)
        for i in range(n * n):
            duc.Phi[i] = float(Phi[i])
    return ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer))(dyn)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_debug_moments_match_reference(name):
    from cauchyfriendly_amd import pycauchy
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    args = (s.num_steps, s.A0, s.p0, s.b0, s.Phi(0), s.Gamma.reshape(-1), np.zeros(0), np.ravel(s.beta),
            s.H.reshape(-1), np.ravel(s.gamma))
    cb = None
    if rec["mode"] == "lti":
        h = pycauchy.pycauchy_initialize_lti(*args, 0.0, 0, True)
    else:
        cb = _dynamics(s)
        h = pycauchy.pycauchy_initialize_ltv(*args, ct.cast(cb, ct.c_void_p).value, 0.0, 0, True)
    try:
        for k, (fz, x, P, err) in enumerate(_expected(rec)):
            out = pycauchy.pycauchy_single_step_ltiv(h, np.asarray(s.zs[k], dtype=np.float64), None)
            gfz, gx, gP, _, _, _, gerr = out[6:]
            where = f"{name} step {k}"
            assert [int(e) for e in gerr] == err, f"{where}: error bits"
            assert rel_err(np.asarray(gfz), fz) <= TOL, f"{where}: fz"
            assert rel_err(np.asarray(gx), x) <= TOL, f"{where}: xhat"
            assert rel_err(np.asarray(gP), P) <= TOL, f"{where}: Phat"
    finally:
        pycauchy.pycauchy_single_step_shutdown(h)
