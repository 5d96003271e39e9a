"""The 7-state low-Earth-orbit windows (cauchyfriendly_amd.workloads.leo7_windows).

CPU tests pin the construction from what the mathematics fixes: the closed-form Jacobian
against central differences of the dynamics, Phi against the matrix exponential exp(J dt)
(the order-6 series truncates at (w dt)^7 / 7! ~ 1e-10 here), Gamma against the top-right
block of exp([[J, e7], [0, 0]] dt), which equals the integral of exp(J s) e7 over [0, dt],
and the windows' geometry.  The first window's estimator runs are pinned by the
reference fixture tests/golden/leo7_w0.json (CPU restatement bit-exact, HIP path to 1e-9)."""
import numpy as np
import pytest
import scipy.linalg

from cauchyfriendly_amd import workloads as W


def _x_nominal(k=0, drag_state=0.0):
    x = W.leo_orbit(k + 1)[k].copy()
    x[6] = drag_state
    return x


@pytest.mark.parametrize("drag_state", [0.0, 0.3])
def test_jacobian_matches_central_differences(drag_state):
    x = _x_nominal(0, drag_state)
    J = W.leo_jacobian(x)
    scale = np.array([1e3] * 3 + [1.0] * 3 + [1e-3])
    for j in range(7):
        e = np.zeros(7)
        e[j] = scale[j]
        col = (W.leo_rhs(x + e) - W.leo_rhs(x - e)) / (2 * scale[j])
        np.testing.assert_allclose(J[:, j], col, rtol=1e-6, atol=1e-6 * np.max(np.abs(col)) + 1e-18)


def test_discretisation_matches_matrix_exponentials():
    dt = 60.0
    J = W.leo_jacobian(_x_nominal(0, 0.1))
    Phi, Gamma = W.leo_discretise(J, dt)
    np.testing.assert_allclose(Phi, scipy.linalg.expm(J * dt), rtol=0, atol=1e-9 * np.max(np.abs(Phi)))
    M = np.zeros((8, 8))
    M[:7, :7] = J
    M[6, 7] = 1.0
    G = scipy.linalg.expm(M * dt)[:7, 7:8]
    np.testing.assert_allclose(Gamma, G, rtol=0, atol=1e-9 * np.max(np.abs(G)))
    # the drag state drives the velocities only through the drag term, and itself by ~dt
    assert abs(Gamma[6, 0] - LEO_DECAY(dt)) < 1e-12 * dt


def LEO_DECAY(dt):
    return W.LEO_TAU * (1.0 - np.exp(-dt / W.LEO_TAU))


def test_windows_geometry():
    ws = W.leo7_windows(count=4)
    orbit = W.leo_orbit(4 + 6 - 1, 60.0)
    r = np.linalg.norm(orbit[:, :3], axis=1)
    assert np.all(np.abs(r / W.LEO_R0 - 1.0) < 1e-3)   # circular: drag and the integrator move it little
    for j, s in enumerate(ws):
        assert (s.n, s.pncc, s.p, s.num_steps) == (7, 1, 2, 6)
        np.testing.assert_allclose(np.linalg.norm(s.H, axis=1), 1.0, rtol=1e-14)
        assert np.all(s.H[:, 3:] == 0.0)
        assert s.beta.tolist() == [1e-3] and s.gamma.tolist() == [5e-2, 5e-2]
        np.testing.assert_array_equal(s.Phis[0], W.leo_discretise(W.leo_jacobian(orbit[j]), 60.0)[0])
    # staggered: window 1 starts where window 0's second step is
    np.testing.assert_array_equal(ws[1].Phis[0], ws[0].Phis[1])
    assert len({s.srand_seed for s in ws}) == 4 and not np.array_equal(ws[0].zs, ws[1].zs)
