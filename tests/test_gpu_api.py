"""The pycauchy C ABI (the reference's SWIG boundary) on the HIP engine, checked
against the public outputs the reference reports for the same calls."""
import ctypes as ct
import os

import numpy as np
import pytest

from conftest import cplx, load_golden, rel_err
from cauchyfriendly_amd import _native, pycauchy
from cauchyfriendly_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _public(exp):
    """What pycauchy_single_step_ltiv reports per update (pycauchy.hpp:535-541)."""
    mean = cplx(exp["mean"])
    cov = cplx(exp["cov"])
    return (exp["fz_public"][0], mean.real, cov.real, exp["fz_public"][1], np.max(np.abs(mean.imag)),
            np.max(np.abs(cov.imag)), exp["err"])


@pytest.mark.parametrize("name", ["lti3_w8", "controls3", "ref_3state_3msmts_horth", "ref_2state"])
def test_pycauchy_lti_matches_reference(name):
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    _native.libc().srand(ct.c_uint(s.srand_seed))
    B = np.zeros(0) if s.B is None else s.B
    h = pycauchy.pycauchy_initialize_lti(rec["num_steps"], s.A0, s.p0, s.b0, s.Phi(0), s.Gamma, B, s.beta, s.H,
                                         s.gamma, 0.0, 0, False)
    u = 0
    try:
        for k in range(rec["num_steps"]):
            ctrl = np.zeros(0) if s.us is None else s.us[k]
            res = pycauchy.pycauchy_single_step_ltiv(h, s.zs[k], ctrl)
            fz, x, P, cfz, cx, cP, err = res[6:]
            assert fz.size == s.p and x.size == s.p * s.n and P.size == s.p * s.n * s.n
            for i in range(s.p):
                efz, ex, eP, ecfz, ecx, ecP, eerr = _public(rec["updates"][u])
                assert abs(fz[i] - efz) <= TOL * max(1.0, abs(efz))
                assert rel_err(x[i * s.n:(i + 1) * s.n], ex) <= TOL
                assert rel_err(P[i * s.n * s.n:(i + 1) * s.n * s.n], eP) <= TOL
                assert err[i] == eerr
                u += 1
            nt = pycauchy.pycauchy_single_step_get_number_of_terms(h)
            assert nt == rec["updates"][u - 1]["nt_ftr"]
    finally:
        pycauchy.pycauchy_single_step_shutdown(h)


class Duc(ct.Structure):
    _fields_ = [("x", ct.POINTER(ct.c_double)), ("u", ct.POINTER(ct.c_double)), ("dt", ct.c_double),
                ("step", ct.c_int), ("n", ct.c_int), ("cmcc", ct.c_int), ("pncc", ct.c_int), ("p", ct.c_int),
                ("Phi", ct.POINTER(ct.c_double)), ("Gamma", ct.POINTER(ct.c_double)), ("B", ct.POINTER(ct.c_double)),
                ("H", ct.POINTER(ct.c_double)), ("beta", ct.POINTER(ct.c_double)), ("gamma", ct.POINTER(ct.c_double)),
                ("is_xbar_set_for_ece", ct.c_bool), ("other_stuff", ct.c_void_p)]


def test_pycauchy_ltv_callback_matches_reference():
    """LTV mode: a dynamics callback installs Phi_k before every time step
    (pycauchy.hpp:330-370, 512-523)."""
    rec = load_golden("ltv5_p2")
    s = W.System.from_dict(rec["system"])
    state = {"k": 0}

    def cb(duc_ptr):
        duc = ct.cast(duc_ptr, ct.POINTER(Duc)).contents
        Phi = s.Phi(state["k"]).reshape(-1)
        for i in range(s.n * s.n):
            duc.Phi[i] = Phi[i]
        duc.step += 1

    CB = ct.CFUNCTYPE(None, ct.c_void_p)
    fn = CB(cb)
    _native.libc().srand(ct.c_uint(s.srand_seed))
    h = pycauchy.pycauchy_initialize_ltv(rec["num_steps"], s.A0, s.p0, s.b0, s.Phi(0), s.Gamma, np.zeros(0), s.beta,
                                         s.H, s.gamma, ct.cast(fn, ct.c_void_p).value, 0.1, 0, False)
    u = 0
    try:
        for k in range(rec["num_steps"]):
            state["k"] = k
            res = pycauchy.pycauchy_single_step_ltiv(h, s.zs[k], np.zeros(0))
            Phi_out = res[0]
            assert np.array_equal(Phi_out, s.Phi(k).reshape(-1))
            x, P = res[7], res[8]
            for i in range(s.p):
                _, ex, eP, _, _, _, _ = _public(rec["updates"][u])
                assert rel_err(x[i * s.n:(i + 1) * s.n], ex) <= TOL
                assert rel_err(P[i * s.n * s.n:(i + 1) * s.n * s.n], eP) <= TOL
                u += 1
    finally:
        pycauchy.pycauchy_single_step_shutdown(h)


def test_pycauchy_reset_and_search_ordering():
    rec = load_golden("lti3_w8")
    s = W.System.from_dict(rec["system"])
    _native.libc().srand(ct.c_uint(s.srand_seed))
    h = pycauchy.pycauchy_initialize_lti(4, s.A0, s.p0, s.b0, s.Phi(0), s.Gamma, np.zeros(0), s.beta, s.H, s.gamma,
                                         0.0, 0, False)
    try:
        first = [pycauchy.pycauchy_single_step_ltiv(h, s.zs[k], np.zeros(0))[7].copy() for k in range(4)]
        pycauchy.pycauchy_single_step_reset(h, s.A0, s.p0, s.b0, np.zeros(0))
        second = [pycauchy.pycauchy_single_step_ltiv(h, s.zs[k], np.zeros(0))[7].copy() for k in range(4)]
        for a, b in zip(first, second):
            assert np.array_equal(a, b)
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.arange(13)) == 1
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.arange(3)) == 0
    finally:
        pycauchy.pycauchy_single_step_shutdown(h)


@pytest.mark.gpu
def test_first_estimator_of_a_process_draws_like_the_reference():
    """The constructor's rand() draws must come before the process's first HIP call (which
    may draw from rand() itself): a fresh interpreter whose first GPU action is creating
    an estimator reproduces the reference's root point and perturbations."""
    import subprocess
    import sys
    code = (
        "import numpy as np\n"
        "from cauchyfriendly_amd import workloads as W\n"
        "from cauchyfriendly_amd.estimator import GpuEstimator\n"
        "s = W.config0()\n"
        "e = GpuEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed)\n"
        "r, b = e.random_draws()\n"
        "print(' '.join(repr(float(v)) for v in np.concatenate([r, b[0]])))\n"
    )
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], cwd=repo, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    got = np.array([float(v) for v in out.stdout.strip().splitlines()[-1].split()])
    from oracle import driver as dr
    s = W.config0()
    est = dr.RefEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed, kind="oracle")
    exp = np.concatenate([est.root_point(), est.b_pert(0)])
    est.close()
    assert np.array_equal(got, exp)


@pytest.mark.gpu
def test_reused_device_blocks_give_identical_results():
    """Device buffers come from a process-wide block cache, so a new estimator can receive
    memory holding an earlier estimator's data.  In a fresh process, a 5-state LTV window
    run on freshly allocated memory, then again after estimators of other shapes have been
    created, stepped and destroyed, must give bit-identical counts, error bits and moments."""
    import subprocess
    import sys
    code = (
        "import numpy as np\n"
        "from cauchyfriendly_amd import workloads as W\n"
        "from cauchyfriendly_amd.estimator import GpuEstimator\n"
        "def run(s, steps):\n"
        "    e = GpuEstimator(steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed)\n"
        "    out = []\n"
        "    for k in range(steps):\n"
        "        for i in range(s.p):\n"
        "            r = e.step(s.zs[k, i], s.Phi(k).reshape(-1), s.Gamma.reshape(-1), s.beta, s.H[i], s.gamma[i])\n"
        "            out.append((r['nt_mu'], r['nt_ftr'], r['err'], complex(r['fz']), np.asarray(r['mean']).tobytes(),\n"
        "                        np.asarray(r['cov']).tobytes(), np.asarray(r['shape_ftr']).tobytes()))\n"
        "    e.close()\n"
        "    return out\n"
        "s = W.config1(steps=4)\n"
        "first = run(s, 4)\n"
        "for d in (2, 3, 4, 5, 3):\n"
        "    run(W.lti_system(d=d, steps=6, seed_offset=300 + d), 6)\n"
        "run(W.leo7_windows(count=1, window=3)[0], 3)\n"
        "live = [GpuEstimator(3, 4, 1, 1, np.eye(4), np.full(4, 0.1), np.zeros(4), seed=9) for _ in range(3)]\n"
        "again = run(s, 4)\n"
        "print('SAME' if first == again else 'DIFFERENT', len(first), first[-1][0])\n"
    )
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], cwd=repo, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stderr[-2000:]
    last = out.stdout.strip().splitlines()[-1]
    assert last.startswith("SAME"), last
