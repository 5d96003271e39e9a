"""Extended (nonlinear) mode: pycauchy_initialize_nonlin / pycauchy_single_step_nonlin.

Fixtures (tests/golden/make_nonlin_golden.py) hold the reference's outputs for a
seeded pendulum with a nonlinear measurement (cauchyfriendly_amd/workloads.py):
a single estimator with one and with two measurements per step, and a window
bank.  The CPU tests pin them twice:
* the callback plumbing, bit for bit: zbar = h(x_bar), H = dh/dx(x_bar), the next
  x_bar = f(x_hat) and Phi = the discretised Jacobian at x_hat;
* the estimator, by replaying pycauchy.hpp:545-652 over the CPU restatement (the
  linear update on the residual, then CauchyEstimator::finalize_extended_moments,
  include/cauchy_estimator.hpp:1358-1394) and, where it is built, the reference.
The GPU tests run this library's extended mode through the C ABI against them."""
import json
import os

import numpy as np
import pytest

from conftest import rel_err
from cauchyfriendly_amd import workloads as W
from oracle import driver as dr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TOL = 1e-9
SINGLES = ["nonlin_pend_p1", "nonlin_pend_p2"]


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        return json.load(fh)


def _extended_replay(s, kind):
    """pycauchy_single_step_nonlin over a CPU estimator, step by step."""
    est = dr.RefEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=0, seed=s.srand_seed, kind=kind)
    xbar = s.xbar0.copy()
    Phi, Gamma = np.zeros((2, 2)), np.zeros((2, 1))
    out = []
    for k in range(s.num_steps):
        if k > 0:
            Phi, Gamma = s.discretise(xbar)
            xbar = s.propagate(xbar)
        rec = {"xhat": [], "Phat": [], "xbar": [], "zbar": [], "fz": [], "err": [], "Nt": 0}
        for i in range(s.p):
            zbar = s.h(xbar)[i]
            H = s.H(xbar)[i]
            rec["xbar"].append(xbar.copy())
            rec["zbar"].append(zbar)
            r = est.step(s.zs[k, i] - zbar, Phi.reshape(-1), Gamma.reshape(-1), s.beta, H, s.gamma[i])
            xbar = est.finalize_extended(xbar)
            rec["xhat"].append(xbar.copy())
            rec["Phat"].append(r["cov"].real.reshape(-1))
            rec["fz"].append(r["fz_public"].real)
            rec["err"].append(r["err"])
            rec["Nt"] = r["nt_ftr"]
        out.append(rec)
    est.close()
    return out


@pytest.mark.parametrize("name", SINGLES)
def test_nonlin_fixture_callback_plumbing(name):
    rec = _load(name)
    s = W.Pendulum.from_dict(rec["system"])
    n, p = s.n, s.p
    xhat_prev = None
    for k, o in enumerate(rec["outputs"]):
        xbar = np.array(o["xbar"]).reshape(p, n)
        xhat = np.array(o["xhat"]).reshape(p, n)
        if k == 0:
            assert np.array_equal(xbar[0], s.xbar0)
        else:
            assert np.array_equal(xbar[0], s.propagate(xhat_prev)), f"step {k}: x_bar is not f(x_hat)"
            Phi, Gamma = s.discretise(xhat_prev)
            assert np.array_equal(np.array(o["Phi"]), Phi.reshape(-1)), f"step {k}: Phi"
            assert np.array_equal(np.array(o["Gamma"]), Gamma.reshape(-1)), f"step {k}: Gamma"
        for i in range(p):
            assert o["zbar"][i] == s.h(xbar[i])[i], f"step {k} row {i}: zbar is not h(x_bar)"
            assert np.array_equal(np.array(o["H"]).reshape(p, n)[i], s.H(xbar[i])[i]), f"step {k} row {i}: H"
            if i > 0:
                assert np.array_equal(xbar[i], xhat[i - 1]), f"step {k} row {i}: x_bar is not the last x_hat"
        xhat_prev = xhat[-1]


@pytest.mark.parametrize("kind", ["oracle", "ref"])
@pytest.mark.parametrize("name", SINGLES)
def test_nonlin_fixture_matches_cpu_replay(name, kind):
    if not dr.available(kind):
        pytest.skip(f"{kind} build not present")
    rec = _load(name)
    s = W.Pendulum.from_dict(rec["system"])
    got = _extended_replay(s, kind)
    for k, (g, e) in enumerate(zip(got, rec["outputs"])):
        where = f"{name} [{kind}] step {k}"
        assert g["Nt"] == e["Nt"], f"{where}: terms {g['Nt']} != {e['Nt']}"
        assert g["err"] == e["err"], f"{where}: error bits"
        for f in ("xhat", "Phat", "xbar", "zbar", "fz"):
            assert rel_err(np.ravel(g[f]), np.array(e[f])) <= TOL, f"{where}: {f} differs"


def test_nonlin_fixture_shift_is_not_a_no_op():
    """The finish of an update moves the terms: replaying without it must fail."""
    rec = _load("nonlin_pend_p1")
    s = W.Pendulum.from_dict(rec["system"])
    est = dr.RefEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed, kind="oracle")
    xbar = s.xbar0.copy()
    r = est.step(s.zs[0, 0] - s.h(xbar)[0], np.zeros(4), np.zeros(2), s.beta, s.H(xbar)[0], s.gamma[0])
    xbar = xbar + r["mean"].real          # x_bar moves, the terms do not
    Phi, Gamma = s.discretise(xbar)
    xbar = s.propagate(xbar)
    est.step(s.zs[1, 0] - s.h(xbar)[0], Phi.reshape(-1), Gamma.reshape(-1), s.beta, s.H(xbar)[0], s.gamma[0])
    xhat = est.finalize_extended(xbar)
    est.close()
    assert rel_err(xhat, np.array(rec["outputs"][1]["xhat"])) > 1e-6


def _run_single_gpu(s):
    import ctypes as ct
    from cauchyfriendly_amd.cauchy_estimator import PyCauchyEstimator, Py_CauchyDynamicsUpdateContainer
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    est = PyCauchyEstimator("nonlin", s.num_steps, False)
    dyn, msmt, ext = s.callbacks(Py_CauchyDynamicsUpdateContainer)
    est.initialize_nonlin(s.xbar0, s.A0, s.p0, s.b0, s.beta, s.gamma, dyn, msmt, ext, 0, s.dt, 0)
    out = []
    for k in range(s.num_steps):
        est.step(s.zs[k], None)
        out.append({"fz": est._fz.copy(), "xhat": est._x.copy(), "Phat": est._P.copy(), "xbar": est._xbar.copy(),
                    "zbar": est._zbar.copy(), "err": [int(e) for e in est._err_code], "Nt": est.get_num_CF_terms(),
                    "H": est._H.copy(), "Phi": est._Phi.copy()})
    est.shutdown()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", SINGLES)
def test_nonlin_matches_reference(name):
    rec = _load(name)
    s = W.Pendulum.from_dict(rec["system"])
    got = _run_single_gpu(s)
    for k, (g, e) in enumerate(zip(got, rec["outputs"])):
        where = f"{name} step {k}"
        assert g["Nt"] == e["Nt"], f"{where}: terms {g['Nt']} != {e['Nt']}"
        assert g["err"] == e["err"], f"{where}: error bits"
        for f in ("xhat", "Phat", "xbar", "zbar", "fz", "H"):
            assert rel_err(g[f], np.array(e[f])) <= TOL, f"{where}: {f} differs"
        if k > 0:
            assert rel_err(g["Phi"], np.array(e["Phi"])) <= TOL, f"{where}: Phi differs"


@pytest.mark.gpu
@pytest.mark.parametrize("concurrent", [False, True], ids=["sequential", "concurrent"])
def test_nonlin_window_bank_matches_reference(concurrent):
    import ctypes as ct
    from cauchyfriendly_amd.cauchy_estimator import PySlidingWindowManager, Py_CauchyDynamicsUpdateContainer
    rec = _load("window_nonlin_pend")
    s = W.Pendulum.from_dict(rec["system"])
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    swm = PySlidingWindowManager("nonlin", rec["num_windows"], swm_debug_print=False, concurrent=concurrent)
    dyn, msmt, ext = s.callbacks(Py_CauchyDynamicsUpdateContainer)
    swm.initialize_nonlin(s.xbar0, s.A0, s.p0, s.b0, s.beta, s.gamma, dyn, msmt, ext, 0, s.dt, 0)
    for k, e in enumerate(rec["outputs"]):
        xhat, Phat, wx, wP = swm.step(s.zs[k], None)
        where = f"window_nonlin_pend step {k}"
        assert swm.best_idx == e["best"], f"{where}: best window"
        assert list(swm.win_counts) == e["win_counts"], f"{where}: window counts"
        assert int(swm.moment_info["err_code"][-1]) == e["err"], f"{where}: error bits"
        assert rel_err(np.array([swm.moment_info["fz"][-1]]), np.array([e["fz"]])) <= TOL, f"{where}: fz"
        for a, b, what in ((xhat, e["xhat"], "xhat"), (Phat, e["Phat"], "Phat"), (wx, e["wavg_xhat"], "avg xhat"),
                           (wP, e["wavg_Phat"], "avg Phat")):
            assert rel_err(np.ravel(a), np.array(b)) <= TOL, f"{where}: {what} differs"
    swm.shutdown()
