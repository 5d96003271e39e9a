"""The derivative with the term structure held fixed, checked on the CPU restatement.

The tangent kernels differentiate the window log-likelihood L = sum_k log fz_k with every
discrete decision of the term tree at its nominal outcome.  On the 7-state orbit at
dt = 1 s, plain central differences change that structure (the reduction's
|p_j - p_i| <= 1e-8 comparison and the g-table negligibility test are decided at rounding
level), so they do not define a derivative there.  The restatement's decision tape
(oracle_tape_*) records the reduction's flag arrays and every negligibility outcome of a
nominal run and replays them in perturbed runs: central differences of the replayed runs
are the frozen-structure derivative, computed independently of the GPU."""
import dataclasses

import numpy as np
import pytest

from cauchyfriendly_amd import workloads as W
from oracle import driver as dr

EPS = np.finfo(float).eps


def _orbit():
    """The first 7-state orbit window at dt = 1 s, 3 steps (6 scalar updates, 17147 terms
    after the last one)."""
    return W.leo7_windows(count=1, window=3, dt=1.0)[0]


def _L(recs):
    return float(sum(np.log(r["fz_raw"].real) for r in recs))


def _structure(recs):
    return [(r["nt_mu"], r["nt_ftr"]) for r in recs]


def _perturbed(s, kind, idx, h):
    beta, gamma = s.beta.copy(), s.gamma.copy()
    (beta if kind == "beta" else gamma)[idx] += h
    return dataclasses.replace(s, beta=beta, gamma=gamma)


def _frozen_cd(s, nominal, kind, idx, rel):
    """Central difference of L along one scale with the nominal run's decisions replayed;
    returns (derivative, step, [replay structures], [replay statuses])."""
    base = (s.beta if kind == "beta" else s.gamma)[idx]
    h = rel * abs(base)
    Ls, structs, stats = [], [], []
    for sgn in (1, -1):
        est, recs = dr.run_system(_perturbed(s, kind, idx, sgn * h), kind="oracle", tape=nominal)
        Ls.append(_L(recs))
        structs.append(_structure(recs))
        stats.append(est.tape_status())
        est.close()
    return (Ls[0] - Ls[1]) / (2 * h), h, structs, stats


def test_replay_at_the_recorded_point_is_bit_identical():
    s = _orbit()
    e0, r0 = dr.run_system(s, kind="oracle", tape="record")
    _, nf, nn = e0.tape_status()
    assert nf > 0 and nn > 0
    e1, r1 = dr.run_system(s, kind="oracle", tape=e0)
    div, uf, un = e1.tape_status()
    assert not div and (uf, un) == (nf, nn), "every recorded decision is replayed once"
    assert [r["fz_raw"] for r in r1] == [r["fz_raw"] for r in r0]
    assert _structure(r1) == _structure(r0)
    e0.close()
    e1.close()


def test_replay_holds_the_structure_where_plain_differences_change_it():
    s = _orbit()
    e0, r0 = dr.run_system(s, kind="oracle", tape="record")
    nominal = _structure(r0)
    plain_changed = 0
    for kind, idx in (("beta", 0), ("gamma", 0), ("gamma", 1)):
        for rel in (1e-4, 1e-5, 1e-6):
            _, _, structs, stats = _frozen_cd(s, e0, kind, idx, rel)
            assert all(st == nominal for st in structs) and not any(d for d, _, _ in stats)
            base = (s.beta if kind == "beta" else s.gamma)[idx]
            for sgn in (1, -1):
                ep, rp = dr.run_system(_perturbed(s, kind, idx, sgn * rel * abs(base)), kind="oracle")
                plain_changed += _structure(rp) != nominal
                ep.close()
    e0.close()
    assert plain_changed > 0, "the orbit window is expected to change structure under plain perturbation"


def test_frozen_differences_equal_plain_ones_where_the_structure_holds():
    """On a 3-state sequence whose perturbed runs keep the nominal structure, replaying the
    decisions changes nothing: the same bits as the plain runs."""
    s = W.lti_sequences(count=1, steps=6)[0]
    e0, r0 = dr.run_system(s, kind="oracle", tape="record")
    for kind, idx in (("beta", 0), ("gamma", 0)):
        base = (s.beta if kind == "beta" else s.gamma)[idx]
        for sgn in (1, -1):
            sp = _perturbed(s, kind, idx, sgn * 1e-6 * abs(base))
            ep, rp = dr.run_system(sp, kind="oracle")
            ef, rf = dr.run_system(sp, kind="oracle", tape=e0)
            if _structure(rp) == _structure(r0):
                assert [r["fz_raw"] for r in rf] == [r["fz_raw"] for r in rp]
            ep.close()
            ef.close()
    e0.close()


def test_frozen_derivative_is_smooth_in_the_step():
    """The replayed function is smooth: the gamma derivatives agree across two decades of h."""
    s = _orbit()
    e0, _ = dr.run_system(s, kind="oracle", tape="record")
    for idx in (0, 1):
        g5 = _frozen_cd(s, e0, "gamma", idx, 1e-5)[0]
        g6 = _frozen_cd(s, e0, "gamma", idx, 1e-6)[0]
        assert abs(g5 - g6) <= 1e-6 * abs(g5), (idx, g5, g6)
    e0.close()


@pytest.mark.gpu
def test_orbit_tangent_gradient_matches_frozen_structure_differences():
    """The GPU tangent gradient of L in (beta, gamma_1, gamma_2) on the orbit window against
    the restatement's frozen-structure central differences: within 1e-5 relative plus the
    differences' own round-off (64 eps sum_k |log fz_k| / h)."""
    from cauchyfriendly_amd.likelihood import scale_directions, window_loglik_tangent
    s = _orbit()
    dirs = scale_directions(s)
    L_gpu, dL, structure = window_loglik_tangent(s, dirs)
    e0, r0 = dr.run_system(s, kind="oracle", tape="record")
    L0 = _L(r0)
    assert abs(L_gpu - L0) <= 1e-9 * abs(L0)
    assert structure == _structure(r0)
    scale = float(sum(abs(np.log(r["fz_raw"].real)) for r in r0))
    for k, (kind, idx) in enumerate(dirs):
        rel = 1e-4 if kind == "beta" else 1e-5   # beta's derivative is 1e5 times smaller than L
        cd, h, structs, stats = _frozen_cd(s, e0, kind, idx, rel)
        assert not any(d for d, _, _ in stats)
        tol = 1e-5 * abs(cd) + 64 * EPS * scale / h
        assert abs(dL[k] - cd) <= tol, f"{kind}[{idx}]: tangent {dL[k]!r} frozen difference {cd!r} (tol {tol:.3g})"
    e0.close()
