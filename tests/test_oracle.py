"""The CPU restatement (oracle/liboracle.so) against the reference's golden
fixtures, closed forms and brute force.  CPU only."""
import math

import numpy as np
import pytest

from conftest import compare_terms, compare_update, golden_names, load_golden
from cauchyfriendly_amd import workloads as W
from oracle import driver as dr


def run_oracle(rec, term_updates=()):
    s = W.System.from_dict(rec["system"])
    est = dr.RefEstimator(rec["num_steps"], s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=s.cmcc,
                          seed=s.srand_seed, kind="oracle")
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    B = None if s.B is None else np.ascontiguousarray(s.B.reshape(-1))
    out, terms = [], {}
    u = 0
    for k in range(rec.get("run_steps", rec["num_steps"])):
        for i in range(s.p):
            out.append(est.step(s.zs[k, i], s.Phi(k).reshape(-1), Gt, s.beta, s.H[i], s.gamma[i], B, s.u(k)))
            if str(u) in term_updates:
                terms[str(u)] = est.terms()
            u += 1
    return est, out, terms


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    rec = load_golden(name)
    est, out, terms = run_oracle(rec, tuple(rec["terms"].keys()))
    # bit-exact: the restatement keeps the reference's operation and summation order
    for u, (got, exp) in enumerate(zip(out, rec["updates"])):
        compare_update(got, exp, 0.0, f"{name} update {u}")
    for u, exp_terms in rec["terms"].items():
        compare_terms(terms[u], exp_terms, 0.0, f"{name} after update {u}")
    np.testing.assert_array_equal(est.root_point(), np.asarray(rec["root_point"]))
    np.testing.assert_array_equal(est.b_pert(0), np.asarray(rec["b_pert0"]))
    np.testing.assert_array_equal(est.b_pert(7), np.asarray(rec["b_pert7"]))


def test_reference_demo_term_counts():
    """Known answer: the reference's 3-state demo (src/cauchy_estimator.cpp:91-119, default rand
    state) reaches these term counts after each update."""
    rec = load_golden("ref_3state")
    expected = [20, 84, 326, 1196, 4217, 14308, 46823, 150397, 448053, 1403563]
    assert [u["nt_mu"] for u in rec["updates"]][1:] == expected


@pytest.mark.parametrize("name", golden_names())
def test_first_update_closed_form(name):
    """fz after the first update is the Cauchy density of z0 = H x + v with scale
    s = sum_l p0_l |A0_l . H| + gamma and median H . b0."""
    rec = load_golden(name)
    s = W.System.from_dict(rec["system"])
    H = s.H[0]
    scale = float(np.sum(s.p0 * np.abs(s.A0 @ H)) + s.gamma[0])
    r = s.zs[0, 0] - float(H @ s.b0)
    fz = scale / (math.pi * (scale * scale + r * r))
    got = rec["updates"][0]["fz"][0]
    assert abs(got - fz) <= 1e-12 * fz
    # the prior term plus one child per hyperplane that is not orthogonal to H
    assert rec["updates"][0]["nt_mu"] == 1 + int(np.sum(np.abs(s.A0 @ H) >= 1e-10))


def _feasible(A, signs):
    from scipy.optimize import linprog
    m, d = A.shape
    res = linprog(np.zeros(d), A_ub=-(signs[:, None] * A), b_ub=-np.ones(m), bounds=[(None, None)] * d,
                  method="highs")
    return res.status == 0


@pytest.mark.parametrize("name,update", [("lti3_w8", "2"), ("horthog3", "1"), ("lti2_w10", "5")])
def test_gtable_keys_are_arrangement_cells(name, update):
    """Brute force: the keys of every term's g-table are the cells of its central hyperplane
    arrangement whose last hyperplane is positive (half storage), found by LP feasibility of
    all 2^m sign vectors."""
    rec = load_golden(name)
    terms = rec["terms"][update]
    d = rec["system"]["n"]
    checked = 0
    for t in terms[:40]:
        m = t["m"]
        A = np.asarray(t["A"]).reshape(m, d)
        cells = set()
        for key in range(1 << (m - 1)):
            signs = np.array([-1.0 if (key >> l) & 1 else 1.0 for l in range(m)])
            if _feasible(A, signs):
                cells.add(key)
        assert sorted(cells) == sorted(t["keys"]), f"{name} term with m={m}: keys differ from LP cells"
        checked += 1
    assert checked > 0


def test_cell_counts_generic():
    """Buck/Zaslavsky: a central arrangement of m hyperplanes in R^d in general position
    has 2 sum_{i<d} C(m-1, i) cells (fewer otherwise); half of them are stored."""
    import itertools
    rec = load_golden("lti3_w8")
    d = rec["system"]["n"]
    generic = 0
    for t in rec["terms"]["3"]:
        m = t["m"]
        A = np.asarray(t["A"]).reshape(m, d)
        full = (1 << m) if m < d else 2 * sum(math.comb(m - 1, i) for i in range(d))
        assert t["cells"] <= full // 2
        in_general_position = all(abs(np.linalg.det(A[list(c)])) > 1e-9
                                  for c in itertools.combinations(range(m), d)) if m >= d else True
        if in_general_position:
            assert t["cells"] == full // 2
            generic += 1
    assert generic > 0
