"""HIP engine parity: every reference fixture, more seeded systems against the
CPU restatement, and full-size invariants.  Tolerance for fz / x_hat / P_hat is
the 1e-9 relative bound; term counts, error bits and g-table keys are compared
bit-exactly."""
import numpy as np
import pytest

from conftest import compare_terms, compare_update, golden_names, load_golden, rel_err
from cauchyfriendly_amd import workloads as W
from cauchyfriendly_amd.estimator import GpuEstimator

pytestmark = pytest.mark.gpu
TOL = 1e-9


def run_gpu(s, num_steps, term_updates=(), run_steps=None):
    """An estimator for a window of num_steps, stepped for run_steps of them (all by default)."""
    est = GpuEstimator(num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=s.cmcc, seed=s.srand_seed)
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    B = None if s.B is None else np.ascontiguousarray(s.B.reshape(-1))
    out, terms = [], {}
    u = 0
    for k in range(num_steps if run_steps is None else run_steps):
        for i in range(s.p):
            out.append(est.step(s.zs[k, i], s.Phi(k).reshape(-1), Gt, s.beta, s.H[i], s.gamma[i], B, s.u(k)))
            if str(u) in term_updates:
                terms[str(u)] = est.terms()
            u += 1
    return est, out, terms


@pytest.mark.parametrize("name", golden_names())
def test_gpu_matches_reference_fixture(name):
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    print(f"case {name}: {rec['num_steps']} steps", flush=True)
    est, out, terms = run_gpu(s, rec["num_steps"], tuple(rec["terms"].keys()), rec.get("run_steps"))
    worst = 0.0
    for u, (got, exp) in enumerate(zip(out, rec["updates"])):
        worst = max(worst, compare_update(got, exp, TOL, f"{name} update {u}"))
    for u, exp_terms in rec["terms"].items():
        compare_terms(terms[u], exp_terms, TOL, f"{name} after update {u}")
    r, b = est.random_draws()
    np.testing.assert_array_equal(r, np.asarray(rec["root_point"]))
    np.testing.assert_array_equal(b[0], np.asarray(rec["b_pert0"]))
    print(f"  worst relative error {worst:.2e}", flush=True)


def _oracle_run(s, steps):
    from oracle import driver as dr
    return dr.run_system(s, num_steps=steps, kind="oracle")[1]


@pytest.mark.parametrize("d,steps,seed", [(2, 9, 101), (3, 7, 102), (4, 6, 103), (5, 5, 104), (3, 8, 105)])
def test_gpu_matches_oracle_seeded(d, steps, seed):
    s = W.lti_system(d=d, steps=steps, seed_offset=seed, name=f"seeded{seed}")
    print(f"case seeded d={d} steps={steps} seed={seed}", flush=True)
    _, g, _ = run_gpu(s, steps)
    o = _oracle_run(s, steps)
    for u, (a, b) in enumerate(zip(g, o)):
        assert a["nt_mu"] == b["nt_mu"] and a["nt_ftr"] == b["nt_ftr"], f"update {u} counts"
        assert np.array_equal(a["shape_mu"], b["shape_mu"]) and np.array_equal(a["shape_ftr"], b["shape_ftr"])
        assert a["err"] == b["err"]
        assert max(rel_err(a["fz"], b["fz"]), rel_err(a["mean"], b["mean"]), rel_err(a["cov"], b["cov"])) <= TOL


def test_gpu_reset_replays_identically():
    """reset() keeps the construction-time draws, so a second pass repeats the first exactly
    (the reference's demo runs its window twice through reset, src/cauchy_estimator.cpp:111-118)."""
    s = W.config0()
    est = GpuEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, seed=s.srand_seed)
    passes = []
    for _ in range(2):
        recs = []
        for k in range(s.num_steps):
            recs.append(est.step(s.zs[k, 0], s.Phi(k).reshape(-1), s.Gamma.reshape(-1), s.beta, s.H[0], s.gamma[0]))
        passes.append(recs)
        from cauchyfriendly_amd import _native
        _native.lib().mce_estimator_reset(est.h)
    for a, b in zip(*passes):
        assert a["nt_mu"] == b["nt_mu"] and a["nt_ftr"] == b["nt_ftr"]
        assert rel_err(a["mean"], b["mean"]) == 0.0 and rel_err(a["cov"], b["cov"]) == 0.0


def test_gpu_full_window_invariants():
    """Largest fixture window (1.4M terms after the last update): invariants of a valid
    conditional density."""
    rec = load_golden("ref_3state")
    s = W.System.from_dict(rec["system"])
    _, out, _ = run_gpu(s, rec["num_steps"])
    for u, got in enumerate(out):
        assert got["fz"].real > 0
        P = got["cov"].real
        assert np.allclose(P, P.T, rtol=1e-9, atol=1e-12)
        assert sum(got["shape_mu"]) == got["nt_mu"]
        assert sum(got["shape_ftr"]) == got["nt_ftr"]
        if u > 0:
            assert got["nt_ftr"] <= got["nt_mu"]
    assert out[-1]["nt_mu"] == 1403563


def _run_with_big_group_threshold(s, num_steps, threshold, monkeypatch):
    monkeypatch.setenv("MCE_BIG_GROUP_MEMBERS", str(threshold))
    est = GpuEstimator(num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=s.cmcc, seed=s.srand_seed)
    est.profile(enable=True, reset=True)
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    B = None if s.B is None else np.ascontiguousarray(s.B.reshape(-1))
    out = []
    for k in range(num_steps):
        for i in range(s.p):
            out.append(est.step(s.zs[k, i], s.Phi(k).reshape(-1), Gt, s.beta, s.H[i], s.gamma[i], B, s.u(k)))
    prof = est.profile(enable=False)
    est.close()
    return out, prof


def test_large_groups_parallel_member_path_is_identical(monkeypatch):
    """Reduction groups above the member threshold add their members through
    k_group_members / k_group_sum; the result must equal the one-wavefront-per-group
    path bit for bit, on a window whose largest group has hundreds of members.  The
    moments of each update are sums over every g value of the previous update's
    tables, so equal moments pin the tables too."""
    s = W.config1()
    steps = 5
    ref_out, ref_prof = _run_with_big_group_threshold(s, steps, 1 << 30, monkeypatch)
    assert ref_prof[15] > 100  # the window does contain large groups
    for threshold in (16, 0):
        out, _ = _run_with_big_group_threshold(s, steps, threshold, monkeypatch)
        for a, b in zip(out, ref_out):
            assert a["nt_mu"] == b["nt_mu"] and a["nt_ftr"] == b["nt_ftr"]
            assert np.array_equal(a["shape_ftr"], b["shape_ftr"])
            assert a["err"] == b["err"]
            for key in ("fz", "mean", "cov"):
                assert np.array_equal(np.asarray(a[key]), np.asarray(b[key])), (threshold, key)


@pytest.mark.parametrize("mode", [0, 1, 3])
@pytest.mark.parametrize("name", ["lti3_w8", "ltv5_p2", "horthog3"])
def test_gtable_lookup_modes_match_reference(name, mode, monkeypatch):
    """The g-table kernel's fallback parent lookups (binary search over keys in global
    memory or LDS, rank structure with LDS-staged values) give the reference's results."""
    monkeypatch.setenv("MCE_GEVAL_MODE", str(mode))
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    _, out, _ = run_gpu(s, rec["num_steps"])
    for u, (got, exp) in enumerate(zip(out, rec["updates"])):
        compare_update(got, exp, TOL, where=f"{name} mode {mode} update {u}")


@pytest.mark.parametrize("lanes", [128, 256])
@pytest.mark.parametrize("name", ["ltv5_p2", "lti8_w5"])
def test_enumerator_lanes_match_reference(name, lanes, monkeypatch):
    """The TP cell enumerator gives the reference's B-tables at every supported lane count
    per parent (64 is the default the other tests run)."""
    monkeypatch.setenv("MCE_ENUM_THREADS", str(lanes))
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    _, out, _ = run_gpu(s, rec["num_steps"])
    for u, (got, exp) in enumerate(zip(out, rec["updates"])):
        compare_update(got, exp, TOL, where=f"{name} lanes {lanes} update {u}")
