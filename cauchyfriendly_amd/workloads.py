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
    B: Optional[np.ndarray] = None

    @property
    def num_steps(self) -> int:
        return int(self.zs.shape[0])

    def Phi(self, k: int) -> np.ndarray:
        return self.Phis[min(k, len(self.Phis) - 1)]


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
