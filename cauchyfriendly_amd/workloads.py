"""Seeded synthetic linear systems with Cauchy noise.

Every system is produced by numpy's PCG64 generator from a fixed seed, so the
same inputs feed the HIP path, the CPU restatement and the reference build.
The systems follow the reference's LTI/LTV problem setup
(scripts/swig/cauchy/cauchy_estimator.py:658-707: A0, p0, b0, Phi, B, Gamma,
beta, H, gamma) and its measurement convention (one scalar update per row of H,
time propagation only before the first row, include/cauchy_estimator.hpp:614).
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

BASE_SEED = 20261202


@dataclasses.dataclass
class System:
    """One estimation problem: dynamics, noise scales and a measurement record."""

    name: str
    n: int                      # state dimension d
    pncc: int                   # columns of Gamma
    p: int                      # measurements per time step (rows of H)
    A0: np.ndarray              # d x d initial hyperplane directions
    p0: np.ndarray              # d initial Cauchy scales
    b0: np.ndarray              # d initial median
    Phis: List[np.ndarray]      # per time step d x d (index k used before update k)
    Gamma: np.ndarray           # d x pncc
    beta: np.ndarray            # pncc
    H: np.ndarray               # p x d
    gamma: np.ndarray           # p
    zs: np.ndarray              # (num_steps, p) measurements
    xs: np.ndarray              # (num_steps, d) true states
    srand_seed: int             # glibc srand seed used before constructing an estimator
    B: Optional[np.ndarray] = None    # d x cmcc control matrix
    us: Optional[np.ndarray] = None   # (num_steps, cmcc) controls

    @property
    def num_steps(self) -> int:
        return int(self.zs.shape[0])

    @property
    def cmcc(self) -> int:
        return 0 if self.B is None else int(self.B.shape[1])

    def Phi(self, k: int) -> np.ndarray:
        return self.Phis[min(k, len(self.Phis) - 1)]

    def u(self, k: int):
        return None if self.us is None else self.us[k]

    def to_dict(self) -> dict:
        f = lambda a: None if a is None else np.asarray(a, dtype=np.float64).tolist()
        return {"name": self.name, "n": self.n, "pncc": self.pncc, "p": self.p, "A0": f(self.A0), "p0": f(self.p0),
                "b0": f(self.b0), "Phis": [f(P) for P in self.Phis], "Gamma": f(self.Gamma), "beta": f(self.beta),
                "H": f(self.H), "gamma": f(self.gamma), "zs": f(self.zs), "xs": f(self.xs),
                "srand_seed": self.srand_seed, "B": f(self.B), "us": f(self.us)}

    @staticmethod
    def from_dict(o: dict) -> "System":
        a = lambda v: None if v is None else np.asarray(v, dtype=np.float64)
        return System(o["name"], int(o["n"]), int(o["pncc"]), int(o["p"]), a(o["A0"]), a(o["p0"]), a(o["b0"]),
                      [a(P) for P in o["Phis"]], a(o["Gamma"]), a(o["beta"]), a(o["H"]), a(o["gamma"]), a(o["zs"]),
                      a(o["xs"]), int(o["srand_seed"]), a(o.get("B")), a(o.get("us")))


def _orthogonal(rng: np.random.Generator, n: int) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))[None, :]


def _simulate(rng, Phis, Gamma, H, beta, gamma, p0, b0, steps, outlier_prob=0.0, outlier_scale=1.0):
    n = Gamma.shape[0]
    x = b0 + p0 * rng.standard_cauchy(n)
    xs, zs = [], []
    for k in range(steps):
        if k > 0:
            w = beta * rng.standard_cauchy(beta.size)
            x = Phis[min(k, len(Phis) - 1)] @ x + Gamma @ w
        v = gamma * rng.standard_cauchy(gamma.size)
        if outlier_prob > 0.0:
            mask = rng.random(gamma.size) < outlier_prob
            v = np.where(mask, v * outlier_scale, v)
        xs.append(x.copy())
        zs.append(H @ x + v)
    return np.array(xs), np.array(zs)


def lti_system(d: int = 3, steps: int = 8, seed_offset: int = 0, beta: float = 0.1,
               gamma: float = 0.2, p0: float = 0.1, outlier_prob: float = 0.0,
               outlier_scale: float = 50.0, name: str = "lti") -> System:
    """The 3-state LTI recipe: Phi = 0.9 * random orthogonal, Gamma d x 1 and
    H 1 x d standard normal, A0 = I, b0 = 0 (generalised to any d)."""
    rng = np.random.default_rng(BASE_SEED + seed_offset)
    Phi = 0.9 * _orthogonal(rng, d)
    Gamma = rng.standard_normal((d, 1))
    H = rng.standard_normal((1, d))
    beta_v = np.array([beta])
    gamma_v = np.array([gamma])
    A0 = np.eye(d)
    p0_v = np.full(d, p0)
    b0 = np.zeros(d)
    xs, zs = _simulate(rng, [Phi], Gamma, H, beta_v, gamma_v, p0_v, b0, steps,
                       outlier_prob, outlier_scale)
    return System(name, d, 1, 1, A0, p0_v, b0, [Phi], Gamma, beta_v, H, gamma_v,
                  zs, xs, BASE_SEED + seed_offset)


def _expm(M: np.ndarray) -> np.ndarray:
    import scipy.linalg
    return scipy.linalg.expm(M)


def ltv_system(d: int = 5, steps: int = 64, seed_offset: int = 1, dt: float = 0.1,
               pncc: int = 2, p: int = 2, beta: float = 0.05, gamma: float = 0.1,
               p0: float = 0.1, outlier_prob: float = 0.0, outlier_scale: float = 50.0,
               name: str = "ltv") -> System:
    """The 5-state LTV recipe: A_k = (A0 + (k/steps) dA) / max(1, rho), Phi_k = expm(dt A_k),
    Gamma d x pncc and H p x d standard normal."""
    rng = np.random.default_rng(BASE_SEED + seed_offset)
    Ab = rng.standard_normal((d, d))
    dA = 0.25 * rng.standard_normal((d, d))
    Phis = []
    for k in range(steps):
        Ak = Ab + (k / float(steps)) * dA
        rho = max(1.0, float(np.max(np.abs(np.linalg.eigvals(Ak)))))
        Phis.append(_expm(dt * Ak / rho))
    Gamma = rng.standard_normal((d, pncc))
    H = rng.standard_normal((p, d))
    beta_v = np.full(pncc, beta)
    gamma_v = np.full(p, gamma)
    A0 = np.eye(d)
    p0_v = np.full(d, p0)
    b0 = np.zeros(d)
    xs, zs = _simulate(rng, Phis, Gamma, H, beta_v, gamma_v, p0_v, b0, steps,
                       outlier_prob, outlier_scale)
    return System(name, d, pncc, p, A0, p0_v, b0, Phis, Gamma, beta_v, H, gamma_v,
                  zs, xs, BASE_SEED + seed_offset)


def config0() -> System:
    """3-state LTI, beta 0.1, gamma 0.2, p0 0.1, 8 measurement updates."""
    return lti_system(d=3, steps=8, seed_offset=0, name="lti3_w8")


def config1(steps: int = 64) -> System:
    """5-state LTV, two measurements per step, beta 0.05, gamma 0.1."""
    return ltv_system(d=5, steps=steps, seed_offset=1, name="ltv5_p2")


def batch_systems(count: int = 512, seed_offset: int = 2, outlier_prob: float = 0.3,
                  outlier_scale: float = 50.0) -> List[System]:
    """Ragged batch: d ~ U{2..5}, window ~ U{4..8}, 30% of measurements carry a 50x outlier."""
    rng = np.random.default_rng(BASE_SEED + seed_offset)
    out = []
    for j in range(count):
        d = int(rng.integers(2, 6))
        w = int(rng.integers(4, 9))
        s = lti_system(d=d, steps=w, seed_offset=seed_offset + 1000 * (j + 1),
                       outlier_prob=outlier_prob, outlier_scale=outlier_scale,
                       name=f"batch{j}_d{d}_w{w}")
        out.append(s)
    return out


def explicit_system(name, Phi, Gamma, H, beta, gamma, A0, p0, b0, zs, srand_seed=1, B=None, us=None):
    """A system given by explicit matrices and a measurement record (zs: one row per time step)."""
    n = len(b0)
    Gamma = np.asarray(Gamma, dtype=np.float64).reshape(n, -1)
    H = np.asarray(H, dtype=np.float64).reshape(-1, n)
    zs = np.asarray(zs, dtype=np.float64).reshape(-1, H.shape[0])
    return System(name, n, Gamma.shape[1], H.shape[0], np.asarray(A0, dtype=np.float64).reshape(n, n),
                  np.asarray(p0, dtype=np.float64), np.asarray(b0, dtype=np.float64),
                  [np.asarray(Phi, dtype=np.float64).reshape(n, n)], Gamma, np.asarray(beta, dtype=np.float64),
                  H, np.asarray(gamma, dtype=np.float64), zs, np.zeros((zs.shape[0], n)), srand_seed,
                  None if B is None else np.asarray(B, dtype=np.float64).reshape(n, -1),
                  None if us is None else np.asarray(us, dtype=np.float64).reshape(zs.shape[0], -1))


def edge_systems() -> List[System]:
    """Systems that drive the estimator's special paths."""
    rng = np.random.default_rng(BASE_SEED + 7)
    out = []
    # H orthogonal to Gamma and to two initial hyperplanes: non-integrable rows at every update
    Phi = 0.9 * _orthogonal(rng, 3)
    zs = 0.3 * rng.standard_cauchy((7, 1))
    out.append(explicit_system("horthog3", Phi, [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.1], [0.2], np.eye(3),
                               [0.1, 0.08, 0.05], np.zeros(3), zs, srand_seed=11))
    # Gamma parallel to an initial hyperplane: time-propagation coalignment on the first propagation
    zs = 0.3 * rng.standard_cauchy((7, 1))
    out.append(explicit_system("tpcoalign3", np.eye(3), [1.0, 0.0, 0.0], [0.7, -0.4, 0.3], [0.1], [0.2],
                               np.eye(3), [0.1, 0.08, 0.05], np.zeros(3), zs, srand_seed=12))
    # controls B u
    Phi = 0.95 * _orthogonal(rng, 3)
    zs = 0.3 * rng.standard_cauchy((6, 1))
    us = rng.standard_normal((6, 1))
    out.append(explicit_system("controls3", Phi, rng.standard_normal(3), rng.standard_normal(3), [0.1], [0.2],
                               np.eye(3), [0.1, 0.1, 0.1], np.zeros(3), zs, srand_seed=13,
                               B=[[0.5], [-0.2], [0.1]], us=us))
    # 50x impulsive measurement outliers
    s = lti_system(d=3, steps=7, seed_offset=40, outlier_prob=0.3, outlier_scale=50.0, name="outliers3")
    out.append(s)
    # scalar state (every hyperplane coaligns)
    zs = 0.3 * rng.standard_cauchy((10, 1))
    out.append(explicit_system("scalar1", [0.9], [0.4], [2.0], [0.1], [0.2], [1.0], [0.1], [0.0], zs, srand_seed=14))
    # 7 states, two measurements per step
    s7 = ltv_system(d=7, steps=4, seed_offset=41, pncc=1, p=2, beta=1e-3, gamma=5e-2, name="ltv7_p2")
    out.append(s7)
    return out


# ------------------------------------------------------------ nonlinear (extended) mode
@dataclasses.dataclass
class Pendulum:
    """A damped pendulum, state (angle, rate), for the extended estimator.

    Propagation is one classical Runge-Kutta 4 step of
    d(angle)/dt = rate, d(rate)/dt = -(g/L) sin(angle) - damp * rate; Phi and Gamma are
    the second-order Taylor discretisations of its Jacobian (process noise enters the
    rate).  Measurements are the bob's horizontal position L sin(angle) and, when p = 2,
    its height -L cos(angle).  The setup follows the reference's pendulum example
    (scripts/swig/cauchy/test_pycauchy_nonlin.py), with a nonlinear measurement here."""

    name: str
    p: int
    x0: np.ndarray              # true initial state
    xbar0: np.ndarray           # the estimator's initial x_bar
    A0: np.ndarray
    p0: np.ndarray
    b0: np.ndarray
    beta: np.ndarray            # 1 process-noise scale
    gamma: np.ndarray           # p measurement-noise scales
    zs: np.ndarray              # (steps, p) measurements
    L: float = 0.3
    g: float = 9.81
    damp: float = 0.3
    dt: float = 0.05
    srand_seed: int = 1

    n = 2
    pncc = 1

    @property
    def num_steps(self) -> int:
        return int(self.zs.shape[0])

    def ode(self, x):
        return np.array([x[1], -self.g / self.L * np.sin(x[0]) - self.damp * x[1]])

    def propagate(self, x):
        h = self.dt
        k1 = self.ode(x)
        k2 = self.ode(x + 0.5 * h * k1)
        k3 = self.ode(x + 0.5 * h * k2)
        k4 = self.ode(x + h * k3)
        return x + h / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)

    def jacobian(self, x):
        return np.array([[0.0, 1.0], [-self.g / self.L * np.cos(x[0]), -self.damp]])

    def discretise(self, x):
        J = self.jacobian(x)
        h = self.dt
        Phi = np.eye(2) + J * h + J @ J * (h * h / 2.0)
        Gc = np.array([0.0, 1.0])
        Gamma = (np.eye(2) * h + J * (h * h / 2.0) + J @ J * (h ** 3 / 6.0)) @ Gc
        return Phi, Gamma.reshape(2, 1)

    def h(self, x):
        z = [self.L * np.sin(x[0]), -self.L * np.cos(x[0])]
        return np.array(z[:self.p])

    def H(self, x):
        rows = [[self.L * np.cos(x[0]), 0.0], [self.L * np.sin(x[0]), 0.0]]
        return np.array(rows[:self.p])

    # the three callbacks of the extended estimator, over the container view
    def callbacks(self, view_type):
        def dynamics(c_duc):
            v = view_type(c_duc)
            x = v.cget_x()
            Phi, Gamma = self.discretise(x)
            v.cset_x(self.propagate(x))
            v.cset_is_xbar_set_for_ece()
            v.cset_Phi(Phi)
            v.cset_Gamma(Gamma)

        def msmt_model(c_duc, c_zbar):
            v = view_type(c_duc)
            v.cset_zbar(c_zbar, self.h(v.cget_x()))

        def msmt_jacobian(c_duc):
            v = view_type(c_duc)
            v.cset_H(self.H(v.cget_x()))

        return dynamics, msmt_model, msmt_jacobian

    def to_dict(self) -> dict:
        out = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            out[f.name] = v.tolist() if isinstance(v, np.ndarray) else v
        return out

    @staticmethod
    def from_dict(o: dict) -> "Pendulum":
        kw = {}
        for f in dataclasses.fields(Pendulum):
            v = o[f.name]
            kw[f.name] = np.asarray(v, dtype=np.float64) if isinstance(v, list) else v
        return Pendulum(**kw)


def pendulum_system(p: int = 1, steps: int = 8, seed_offset: int = 40, name: str = "pendulum") -> Pendulum:
    """Seeded pendulum run: true trajectory with Cauchy process and measurement noise,
    and the estimator's initial conditions."""
    rng = np.random.default_rng(BASE_SEED + seed_offset)
    beta = np.array([0.08])
    gamma = np.full(p, 0.05)
    # A0 couples angle and rate, so that the rate's deviation is not left at exactly zero
    # by the first angle-only measurements (the moment checks divide by its size)
    A0 = np.array([[1.0, 0.4], [-0.3, 1.0]])
    proto = Pendulum(name=name, p=p, x0=np.array([np.pi / 3, 0.0]), xbar0=np.zeros(2), A0=A0,
                     p0=np.array([0.1, 0.1]), b0=np.zeros(2), beta=beta, gamma=gamma, zs=np.zeros((0, p)))
    x = proto.x0.copy()
    zs = []
    for k in range(steps):
        if k > 0:
            _, G = proto.discretise(x)
            x = proto.propagate(x) + G.reshape(-1) * beta[0] * rng.standard_cauchy()
        zs.append(proto.h(x) + gamma * rng.standard_cauchy(p))
    proto.xbar0 = proto.x0 + proto.p0 * rng.standard_cauchy(2) * 0.5
    proto.zs = np.array(zs)
    return proto


def lti_sequences(count: int = 64, steps: int = 6, seed_offset: int = 3) -> List[System]:
    """One 3-state LTI system (the 8-step recipe's dynamics and scales) with `count` seeded
    measurement sequences of `steps` steps: the likelihood-gradient workload."""
    base = lti_system(d=3, steps=steps, seed_offset=0, name="lti3")
    out = []
    for j in range(count):
        rng = np.random.default_rng(BASE_SEED + seed_offset + 1000 * (j + 1))
        xs, zs = _simulate(rng, base.Phis, base.Gamma, base.H, base.beta, base.gamma, base.p0, base.b0, steps)
        out.append(System(f"lti3_seq{j}", 3, 1, 1, base.A0, base.p0, base.b0, base.Phis, base.Gamma, base.beta,
                          base.H, base.gamma, zs, xs, BASE_SEED + seed_offset + 1000 * (j + 1)))
    return out


# ------------------------------------------------------------ 7-state low-Earth-orbit windows
# Two-body gravity with drag whose air density carries a first-order Gauss-Markov
# perturbation (the seventh state), as in the reference's 7-state satellite model
# (src/leo_satellite_7state.cpp:93-160 constants, 247-262 dynamics, 289-340 discretisation):
# Earth's gravitational parameter, a 200 km circular-speed orbit, the density at 200 km and
# the drag area, coefficient, mass and correlation time of that model.
LEO_MU = 5.9722e24 * 6.674e-11
LEO_R0 = 6378.1e3 + 200e3
LEO_DRAG = 0.5 * 64.0 * 2.0 / 5000.0 * 2.541e-10    # A C_D rho / (2 m)
LEO_TAU = 21600.0


def leo_rhs(x: np.ndarray) -> np.ndarray:
    """Time derivative of (position, velocity, density perturbation)."""
    r, v = x[:3], x[3:6]
    rn, vn = np.linalg.norm(r), np.linalg.norm(v)
    a = -LEO_MU / rn ** 3 * r - LEO_DRAG * (1.0 + x[6]) * vn * v
    return np.concatenate([v, a, [-x[6] / LEO_TAU]])


def leo_jacobian(x: np.ndarray) -> np.ndarray:
    """d leo_rhs / dx, in closed form."""
    r, v = x[:3], x[3:6]
    rn, vn = np.linalg.norm(r), np.linalg.norm(v)
    J = np.zeros((7, 7))
    J[0:3, 3:6] = np.eye(3)
    J[3:6, 0:3] = -LEO_MU * (np.eye(3) / rn ** 3 - 3.0 * np.outer(r, r) / rn ** 5)
    J[3:6, 3:6] = -LEO_DRAG * (1.0 + x[6]) * (vn * np.eye(3) + np.outer(v, v) / vn)
    J[3:6, 6] = -LEO_DRAG * vn * v
    J[6, 6] = -1.0 / LEO_TAU
    return J


def leo_discretise(J: np.ndarray, dt: float, order: int = 6):
    """Phi = sum_{i<=order} (J dt)^i / i! and Gamma = (sum_{i<=order} J^i dt^(i+1) / (i+1)!) e_7,
    the truncated series of the reference's discretisation."""
    n = J.shape[0]
    Phi = np.eye(n)
    Gint = dt * np.eye(n)
    P = np.eye(n)
    fact_i, fact_i1 = 1.0, 1.0
    for i in range(1, order + 1):
        P = P @ J
        fact_i *= i
        fact_i1 *= i + 1
        Phi = Phi + P * dt ** i / fact_i
        Gint = Gint + P * dt ** (i + 1) / fact_i1
    return Phi, Gint[:, 6:7].copy()


def _rk4(x: np.ndarray, dt: float, sub: int = 10) -> np.ndarray:
    h = dt / sub
    for _ in range(sub):
        k1 = leo_rhs(x)
        k2 = leo_rhs(x + 0.5 * h * k1)
        k3 = leo_rhs(x + 0.5 * h * k2)
        k4 = leo_rhs(x + h * k3)
        x = x + h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return x


def leo_orbit(steps: int, dt: float = 60.0) -> np.ndarray:
    """Nominal trajectory (steps x 7): a circular orbit at LEO_R0 whose position and velocity
    directions are those of the reference model's initial condition, (1, sqrt 2, sqrt 3) / sqrt 6
    and (-1, -1, 1) / sqrt 3, with the velocity's radial part removed (left in, it makes the
    orbit eccentric with its perigee below the Earth's surface)."""
    v0 = np.sqrt(LEO_MU / LEO_R0)
    rhat = np.array([1.0, np.sqrt(2), np.sqrt(3)]) / np.sqrt(6)
    u = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    vhat = u - (u @ rhat) * rhat
    vhat /= np.linalg.norm(vhat)
    x = np.concatenate([LEO_R0 * rhat, v0 * vhat, [0.0]])
    out = [x]
    for _ in range(steps - 1):
        x = _rk4(x, dt)
        out.append(x)
    return np.array(out)


def leo7_windows(count: int = 32, window: int = 6, dt: float = 60.0, seed_offset: int = 4,
                 beta: float = 1e-3, gamma: float = 5e-2) -> List[System]:
    """`count` windows of `window` steps staggered one step apart along a low-Earth orbit:
    window j linearises the dynamics about the nominal orbit at steps j .. j + window - 1
    (Phi_k per step, Gamma at its first step), measures ranges to two seeded stations
    (H: their unit line-of-sight vectors at its first step, 2 x 7) and simulates its own
    error-state trajectory and Cauchy noise from the prior (A0 = I, p0 = 0.72 m and m/s for
    position and velocity, 0.0039 for the density perturbation, b0 = 0)."""
    rng = np.random.default_rng(BASE_SEED + seed_offset)
    stations = 3e6 * rng.standard_normal((2, 3))
    orbit = leo_orbit(count + window - 1, dt)
    disc = [leo_discretise(leo_jacobian(x), dt) for x in orbit]
    p0 = np.array([0.7195] * 6 + [0.0039])
    beta_v = np.array([beta])
    gamma_v = np.array([gamma, gamma])
    out = []
    for j in range(count):
        Phis = [disc[j + k][0] for k in range(window)]
        Gamma = disc[j][1]
        los = orbit[j, :3][None, :] - stations
        H = np.zeros((2, 7))
        H[:, :3] = los / np.linalg.norm(los, axis=1, keepdims=True)
        wrng = np.random.default_rng(BASE_SEED + seed_offset + 1000 * (j + 1))
        xs, zs = _simulate(wrng, Phis, Gamma, H, beta_v, gamma_v, p0, np.zeros(7), window)
        out.append(System(f"leo7_w{j}", 7, 1, 2, np.eye(7), p0, np.zeros(7), Phis, Gamma, beta_v, H, gamma_v,
                          zs, xs, BASE_SEED + seed_offset + 1000 * (j + 1)))
    return out
