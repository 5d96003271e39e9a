"""Error paths of the C ABI.  Where the reference prints and calls exit(1) (stepping past
the window, cauchy_estimator.hpp:1220-1225; hyperplane counts beyond 31), this library
returns an error code or NULL and keeps the message for mce_last_error().  Every
condition is forced from the host: no kernel is asked to misbehave."""
import ctypes as ct

import numpy as np
import pytest

from cauchyfriendly_amd import _native, pycauchy
from cauchyfriendly_amd import workloads as W

pytestmark = pytest.mark.gpu
_dp = ct.POINTER(ct.c_double)


def _d(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    return a, a.ctypes.data_as(_dp)


def test_step_past_the_window_returns_an_error_and_keeps_the_estimate():
    L = _native.lib()
    s = W.config0()
    A0, pA = _d(s.A0)
    p0, pp = _d(s.p0)
    b0, pb = _d(s.b0)
    h = L.mce_estimator_create(pA, pp, pb, 2, s.n, 0, s.pncc, s.p, False)
    assert h
    Phi, pPhi = _d(s.Phi(0))
    G, pG = _d(s.Gamma)
    beta, pbeta = _d(s.beta)
    H, pH = _d(s.H[0])
    for k in range(2):
        assert L.mce_estimator_step(h, float(s.zs[k, 0]), pPhi, pG, pbeta, pH, float(s.gamma[0]), None, None) >= 0
    fz0, m0, c0 = np.zeros(2), np.zeros(2 * s.n), np.zeros(2 * s.n * s.n)
    L.mce_estimator_moments(h, fz0.ctypes.data_as(_dp), m0.ctypes.data_as(_dp), c0.ctypes.data_as(_dp))
    assert L.mce_estimator_step(h, float(s.zs[2, 0]), pPhi, pG, pbeta, pH, float(s.gamma[0]), None, None) == -1
    fz1, m1, c1 = np.zeros(2), np.zeros(2 * s.n), np.zeros(2 * s.n * s.n)
    L.mce_estimator_moments(h, fz1.ctypes.data_as(_dp), m1.ctypes.data_as(_dp), c1.ctypes.data_as(_dp))
    assert np.array_equal(fz0, fz1) and np.array_equal(m0, m1) and np.array_equal(c0, c1)
    # after a reset the same estimator steps again
    L.mce_estimator_reset(h)
    assert L.mce_estimator_step(h, float(s.zs[0, 0]), pPhi, pG, pbeta, pH, float(s.gamma[0]), None, None) >= 0
    L.mce_estimator_destroy(h)


def _init(num_steps, n, pncc):
    return pycauchy.pycauchy_initialize_lti(num_steps, np.eye(n), np.full(n, 0.1), np.zeros(n), np.eye(n),
                                            np.ones(n * pncc), np.zeros(0), np.full(pncc, 0.1), np.ones(n),
                                            np.array([0.2]), 0.0, 0, False)


def test_initializer_refuses_windows_beyond_the_hyperplane_limits():
    """A window's largest hyperplane count is (steps - 1) * pncc + d.  Beyond 31 the reference
    refuses it too; this library's LDS cell tables stop at 20.  4 states with two process
    noises: 9 steps reach 20 and build, 10 steps reach 22, 16 steps 34."""
    with pytest.raises(RuntimeError, match="exceeds 31"):
        _init(16, 4, 2)
    with pytest.raises(RuntimeError, match="above 20"):
        _init(10, 4, 2)
    h = _init(9, 4, 2)
    assert h
    pycauchy.pycauchy_single_step_shutdown(h)


def test_window_at_the_hyperplane_limit_steps_like_the_oracle():
    """The 20-hyperplane window the initializer accepts (4 states, two process noises, 9
    steps) steps to its end: row strides, LDS sets and g-table lookups are sized for 20
    hyperplanes.  Every update equals the CPU restatement (counts and error bits exactly,
    moments within 1e-9)."""
    from conftest import rel_err
    from cauchyfriendly_amd.estimator import GpuEstimator
    from oracle import driver
    n, pncc, steps = 4, 2, 9
    zs = np.random.default_rng(5).standard_normal((steps, 1))
    s = W.explicit_system("mm20", np.eye(n), np.ones((n, pncc)), np.ones((1, n)), np.full(pncc, 0.1),
                          np.array([0.2]), np.eye(n), np.full(n, 0.1), np.zeros(n), zs)
    print("case mm20: 4 states, pncc 2, 9 steps", flush=True)
    est = GpuEstimator(steps, n, pncc, 1, s.A0, s.p0, s.b0, seed=s.srand_seed)
    assert est.shape_range == 21
    got = [est.step(s.zs[k, 0], s.Phi(k).reshape(-1), s.Gamma.reshape(-1), s.beta, s.H[0], s.gamma[0])
           for k in range(steps)]
    est.close()
    _, exp = driver.run_system(s, kind="oracle")
    assert len(got) == len(exp) == steps
    for u, (a, b) in enumerate(zip(got, exp)):
        assert (a["nt_mu"], a["nt_ftr"], a["err"]) == (b["nt_mu"], b["nt_ftr"], b["err"]), f"update {u}"
        assert np.array_equal(a["shape_mu"], b["shape_mu"]) and np.array_equal(a["shape_ftr"], b["shape_ftr"])
        assert max(rel_err(a["fz"], b["fz"]), rel_err(a["mean"], b["mean"]), rel_err(a["cov"], b["cov"])) <= 1e-9


def test_density_grid_beyond_1e8_points_returns_no_points():
    from cauchyfriendly_amd.cauchy_estimator import PyCauchyEstimator
    s = W.config0()
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    e = PyCauchyEstimator("lti", s.num_steps, False)
    e.initialize_lti(s.A0, s.p0, s.b0, s.Phi(0), None, s.Gamma, s.beta, s.H, s.gamma)
    e.step(s.zs[0], None)
    e.step(s.zs[1], None)
    X, Y = e.get_marginal_1D_pointwise_cpdf(0, -1.0, 1.0, 1e-9)
    assert X.size == 0 and Y.size == 0
    assert "1e8" in _native.lib().mce_last_error().decode()
    X, Y = e.get_marginal_1D_pointwise_cpdf(0, -1.0, 1.0, 0.5)    # a sane grid afterwards
    assert X.size == 5 and np.all(Y > 0)
    e.shutdown()


def test_search_ordering_beyond_the_state_dimension_is_refused_by_the_step():
    """An ordering accepted by the setter (axes below 12) that names an axis a 3-state
    estimator does not have makes its step fail with a message; nothing reads past a term's
    b vector."""
    from cauchyfriendly_amd.estimator import GpuEstimator
    s = W.config0()
    try:
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.array([5, 0, 1], dtype=np.int32)) == 0
        est = GpuEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed)
        with pytest.raises(RuntimeError, match="search ordering names axis 5"):
            est.step(s.zs[0, 0], s.Phi(0).reshape(-1), s.Gamma.reshape(-1), s.beta, s.H[0], s.gamma[0])
        est.close()
    finally:
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.arange(12, dtype=np.int32)) == 0
