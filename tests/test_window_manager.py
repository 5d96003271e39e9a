"""The window-manager C functions (pycauchy_initialize_{lti,ltv}_window_manager,
pycauchy_step, pycauchy_shutdown; scripts/swig/cauchy/pycauchy.hpp:80-221) against the
reference's own forked window manager on the same inputs (tests/golden/wm_*.json, written by
tests/golden/make_window_manager_golden.py from the reference build): at every measurement
the reporting window and its error bits are identical, and fz, x_hat, P_hat and their
imaginary-part bounds agree within 1e-9 (relative to the largest magnitude of each output).
Cases: a 3-state LTI bank of 4 windows over 16 measurements; a 3-state bank of 3 windows
with a window-variance boost; a 4-state LTV bank (two process noises, two measurements per
step, Phi_k through the dynamics callback) of 3 windows over 10 steps; the 5-state LTV system
(two process noises, two measurements per step) in a bank of 5 windows over 10 steps.  The bank steps the
windows of a measurement together (merged launches) by default; stepped one window at a time,
or one host thread per window, it must give the same bits."""
import ctypes as ct

import numpy as np
import pytest

from conftest import load_golden
from cauchyfriendly_amd import pycauchy
from cauchyfriendly_amd import workloads as W
from cauchyfriendly_amd.cauchy_estimator import C_CauchyDynamicsUpdateContainer

TOL = 1e-9


def _dynamics(s):
    n = s.n

    def dyn(c_duc):
        duc = c_duc.contents
        k = min(max(int(duc.step), 0), s.num_steps - 1)
        Phi = s.Phi(k).reshape(-1)
        for i in range(n * n):
            duc.Phi[i] = float(Phi[i])
    return ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer))(dyn)


def _run(rec, mode=1):
    s = W.System.from_dict(rec["system"])
    ct.CDLL("libc.so.6").srand(ct.c_uint(rec["srand"]))
    args = (s.A0, s.p0, s.b0, s.Phi(0), s.Gamma.reshape(-1), np.zeros(0), np.ravel(s.beta), s.H.reshape(-1),
            np.ravel(s.gamma))
    boost = np.zeros(0) if rec["win_var_boost"] is None else np.asarray(rec["win_var_boost"], dtype=np.float64)
    nw, steps = rec["num_windows"], rec["steps"]
    cb = None
    if rec["mode"] == "lti":
        pycauchy.pycauchy_initialize_lti_window_manager(nw, steps, *args, False, False, False, None, 0.0, 0, boost)
    else:
        cb = _dynamics(s)
        pycauchy.pycauchy_initialize_ltv_window_manager(nw, steps, *args, cb, False, False, False, None, 0.0, 0,
                                                         boost)
    from cauchyfriendly_amd import _native
    assert _native.lib().mce_window_bank_set_mode(mode) == 0
    out = []
    try:
        for k in range(steps):
            out.append(pycauchy.pycauchy_step(np.asarray(s.zs[k], dtype=np.float64), None))
    finally:
        pycauchy.pycauchy_shutdown()
    return out


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, float(np.max(np.abs(b)))))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["wm_lti3", "wm_lti3_boost", "wm_ltv4_p2", "wm_ltv5_p2"])
def test_window_manager_matches_reference(name):
    rec = load_golden(name)
    print(f"case {name}: {rec['num_windows']} windows, {rec['steps']} measurements", flush=True)
    got = _run(rec)
    worst = 0.0
    for k, (g, e) in enumerate(zip(got, rec["outputs"])):
        fz, x, P, cfz, cx, cP, win, err = g
        assert (win, err) == (e["win_idx"], e["err"]), f"{name} step {k}: window/err {win}/{err} vs reference"
        for a, b in ((fz, e["fz"]), (x, e["xhat"]), (P, e["Phat"])):
            r = _rel(a, b)
            worst = max(worst, r)
            assert r <= TOL, f"{name} step {k}: rel err {r:.2e}"
        for a, b in ((cfz, e["cerr_fz"]), (cx, e["cerr_xhat"]), (cP, e["cerr_Phat"])):
            assert abs(a - b) <= TOL * max(1.0, abs(b)) + 1e-12, f"{name} step {k}: imaginary bound {a} vs {b}"
    print(f"  worst relative error {worst:.2e}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2], ids=["merged", "threaded"])
@pytest.mark.parametrize("name", ["wm_lti3", "wm_lti3_boost", "wm_ltv4_p2", "wm_ltv5_p2"])
def test_window_manager_modes_equal_one_window_at_a_time(name, mode):
    """The windows' updates as merged launches, or on a host thread each, give each window
    the bits of its own updates run one window after another."""
    rec = load_golden(name)
    print(f"case {name}: mode {mode} vs one window at a time", flush=True)
    got, alone = _run(rec, mode=mode), _run(rec, mode=0)
    for k, (g, e) in enumerate(zip(got, alone)):
        fz, x, P, cfz, cx, cP, win, err = g
        fz1, x1, P1, cfz1, cx1, cP1, win1, err1 = e
        assert (win, err) == (win1, err1), f"{name} step {k}: window/err"
        for a, b in ((fz, fz1), (x, x1), (P, P1), (cfz, cfz1), (cx, cx1), (cP, cP1)):
            assert np.array_equal(np.asarray(a), np.asarray(b)), f"{name} step {k}: mode {mode} differs"


def test_window_bank_mode_switch_needs_a_bank():
    from cauchyfriendly_amd import _native
    L = _native.lib()
    assert L.mce_window_bank_set_mode(1) == -1
    assert "no window bank" in L.mce_last_error().decode()


@pytest.mark.gpu
def test_window_manager_leaves_the_rand_state_and_refuses_a_second_bank():
    """Building the bank draws every window's random values from a copy of the rand() state, as
    the reference's forked children do: the caller's next rand() is unchanged.  One bank runs at
    a time, as the reference's single global manager."""
    libc = ct.CDLL("libc.so.6")
    rec = load_golden("wm_lti3")
    s = W.System.from_dict(rec["system"])
    libc.srand(ct.c_uint(77))
    expect = [libc.rand() for _ in range(3)]
    libc.srand(ct.c_uint(77))
    args = (s.A0, s.p0, s.b0, s.Phi(0), s.Gamma.reshape(-1), np.zeros(0), np.ravel(s.beta), s.H.reshape(-1),
            np.ravel(s.gamma))
    pycauchy.pycauchy_initialize_lti_window_manager(3, 4, *args, False, False, False, None, 0.0, 0, np.zeros(0))
    try:
        assert [libc.rand() for _ in range(3)] == expect
        pycauchy.pycauchy_initialize_lti_window_manager(3, 4, *args, False, False, False, None, 0.0, 0, np.zeros(0))
        from cauchyfriendly_amd import _native
        assert "already running" in _native.lib().mce_last_error().decode()
        r = pycauchy.pycauchy_step(np.asarray(s.zs[0], dtype=np.float64), None)
        assert r[6] == 0
    finally:
        pycauchy.pycauchy_shutdown()
