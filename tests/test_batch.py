"""Batched stepping: many independent estimators stepped concurrently
(pycauchy_batch_single_step_ltiv / step_estimators).

The batch is the ragged recipe of cauchyfriendly_amd/workloads.batch_systems: state
dimension 2..5, windows of 4..8 steps, 30% of the measurements 50x outliers.  Each
estimator stepped in the batch must give what it gives stepped alone (same kernels,
its own streams), and both must agree with the CPU restatement of the reference
(counts and error bits exact, moments within 1e-9)."""
import ctypes as ct

import numpy as np
import pytest

from conftest import rel_err
from cauchyfriendly_amd import workloads as W
from oracle import driver as dr

TOL = 1e-9
COUNT = 16


def _systems():
    return W.batch_systems(count=COUNT)


def _make(s):
    from cauchyfriendly_amd.cauchy_estimator import PyCauchyEstimator
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    e = PyCauchyEstimator("lti", s.num_steps, False)
    e.initialize_lti(s.A0, s.p0, s.b0, s.Phi(0), None, s.Gamma, s.beta, s.H, s.gamma)
    return e


def _record(e):
    return (e._fz.copy(), e._x.copy(), e._P.copy(), e._err_code.copy(), e.get_num_CF_terms())


def test_batch_recipe_is_ragged():
    ss = _systems()
    assert len({s.n for s in ss}) >= 3 and len({s.num_steps for s in ss}) >= 3


def _check_batch(**step_kw):
    """Steps the recipe's estimators alone and together (step_estimators(**step_kw)); the
    batched records must equal the single ones, and both the oracle's."""
    from cauchyfriendly_amd.cauchy_estimator import step_estimators
    ss = _systems()
    single = []
    for s in ss:
        e = _make(s)
        recs = []
        for k in range(s.num_steps):
            e.step(s.zs[k], None)
            recs.append(_record(e))
        e.shutdown()
        single.append(recs)
    ests = [_make(s) for s in ss]
    batched = [[] for _ in ss]
    for k in range(max(s.num_steps for s in ss)):
        live = [i for i, s in enumerate(ss) if k < s.num_steps]
        step_estimators([ests[i] for i in live], [ss[i].zs[k] for i in live], **step_kw)
        for i in live:
            batched[i].append(_record(ests[i]))
    for e in ests:
        e.shutdown()
    for i, s in enumerate(ss):
        _, oracle_recs = dr.run_system(s, kind="oracle")
        for k in range(s.num_steps):
            where = f"{s.name} step {k}"
            fz, x, P, err, nt = batched[i][k]
            fz1, x1, P1, err1, nt1 = single[i][k]
            assert nt == nt1 and list(err) == list(err1), f"{where}: batched counts/errors differ from single"
            for a, b, what in ((fz, fz1, "fz"), (x, x1, "xhat"), (P, P1, "Phat")):
                # the same kernels with the same arguments: the same bits
                assert np.array_equal(a, b), f"{where}: batched {what} differs from single"
            o = oracle_recs[k]
            assert nt == o["nt_ftr"], f"{where}: terms {nt} != oracle {o['nt_ftr']}"
            assert int(err[-1]) == o["err"], f"{where}: error bits"
            assert rel_err(x, o["mean"].real) <= TOL, f"{where}: xhat vs oracle"
            assert rel_err(P, o["cov"].real.reshape(-1)) <= TOL, f"{where}: Phat vs oracle"
            assert rel_err(np.array([fz[-1]]), np.array([o["fz_public"].real])) <= TOL, f"{where}: fz vs oracle"


@pytest.mark.gpu
def test_batch_step_matches_single_steps_and_oracle():
    _check_batch(num_threads=4)


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [1, 3])
def test_merged_batch_matches_single_steps_and_oracle(groups):
    """Each kernel launched once per group of estimators (merged launches, one-workgroup
    sorts and selects for small inputs) gives each estimator its single-step results."""
    _check_batch(merged_groups=groups)


@pytest.mark.gpu
def test_merged_batch_with_tuned_shapes_matches_single_steps():
    """The autotuner's other launch shapes (enumerator lanes, wavefront-per-member
    threshold, LDS parent values) change no result."""
    from cauchyfriendly_amd import autotune
    from cauchyfriendly_amd.cauchy_estimator import step_estimators
    ss = _systems()
    ref = []
    for s in ss:
        e = _make(s)
        recs = []
        for k in range(s.num_steps):
            e.step(s.zs[k], None)
            recs.append(_record(e))
        e.shutdown()
        ref.append(recs)
    ests = [_make(s) for s in ss]
    autotune.set_shapes(ests, enum_threads=256, big_group_members=8, geval_mode=4)
    for k in range(max(s.num_steps for s in ss)):
        live = [i for i, s in enumerate(ss) if k < s.num_steps]
        step_estimators([ests[i] for i in live], [ss[i].zs[k] for i in live], merged_groups=2)
        for i in live:
            fz, x, P, err, nt = _record(ests[i])
            fz1, x1, P1, err1, nt1 = ref[i][k]
            where = f"{ss[i].name} step {k}"
            assert nt == nt1 and list(err) == list(err1), where
            assert np.array_equal(fz, fz1) and np.array_equal(x, x1) and np.array_equal(P, P1), where
    for e in ests:
        e.shutdown()


def _run_recipe(count, **step_kw):
    from cauchyfriendly_amd.cauchy_estimator import step_estimators
    ss = W.batch_systems(count=count)
    ests = [_make(s) for s in ss]
    recs = [[] for _ in ss]
    for k in range(max(s.num_steps for s in ss)):
        live = [i for i, s in enumerate(ss) if k < s.num_steps]
        step_estimators([ests[i] for i in live], [ss[i].zs[k] for i in live], **step_kw)
        for i in live:
            recs[i].append(_record(ests[i]))
    for e in ests:
        e.shutdown()
    return ss, recs


@pytest.mark.gpu
def test_merged_batch_of_512_matches_threaded_batch():
    """The full 512-estimator ragged batch, merged (4 groups) against the threaded batch,
    every estimator and update: the same counts, error bits and moment bits."""
    ss, merged = _run_recipe(512, merged_groups=4)
    _, threaded = _run_recipe(512, num_threads=16)
    updates = 0
    for i, s in enumerate(ss):
        for k in range(s.num_steps):
            fz, x, P, err, nt = merged[i][k]
            fz1, x1, P1, err1, nt1 = threaded[i][k]
            where = f"{s.name} step {k}"
            assert nt == nt1 and list(err) == list(err1), where
            assert np.array_equal(fz, fz1) and np.array_equal(x, x1) and np.array_equal(P, P1), where
            updates += 1
    assert updates == sum(s.num_steps for s in ss)


@pytest.mark.gpu
def test_batch_rejects_mixed_modes_and_finished_estimators():
    from cauchyfriendly_amd.cauchy_estimator import step_estimators
    s = _systems()[0]
    e = _make(s)
    for k in range(s.num_steps):
        step_estimators([e], [s.zs[k]])
    with pytest.raises(RuntimeError):
        step_estimators([e], [s.zs[0]])
    e.shutdown()


@pytest.mark.gpu
def test_full_batch_of_512_matches_the_restatement_on_every_update():
    """The bench's whole batch (512 estimators) with merged launches at the tuning table's
    settings, against the restatement on 16 host processes: every update's term count after
    reduction and error bits exact, fz, x_hat and P_hat within 1e-9."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from cauchyfriendly_amd import autotune
    from cauchyfriendly_amd.cauchy_estimator import step_estimators
    ss = W.batch_systems(count=512)
    print("full batch: 512 estimators, merged launches", flush=True)
    with ProcessPoolExecutor(max_workers=16, mp_context=mp.get_context("spawn")) as ex:
        futs = [ex.submit(dr.system_records, s) for s in ss]
        ests = [_make(s) for s in ss]
        st = autotune.load_table("batch512") or dict(autotune.DEFAULTS)
        autotune.set_shapes(ests, st["enum_threads"], st["big_group_members"], st["geval_mode"])
        got = [[] for _ in ss]
        for k in range(max(s.num_steps for s in ss)):
            live = [i for i, s in enumerate(ss) if k < s.num_steps]
            step_estimators([ests[i] for i in live], [ss[i].zs[k] for i in live], merged_groups=st["groups"])
            for i in live:
                got[i].append(_record(ests[i]))
        for e in ests:
            e.shutdown()
        want = [f.result(timeout=600) for f in futs]
    worst = 0.0
    for i, s in enumerate(ss):
        assert len(got[i]) == len(want[i]) == s.num_steps, f"{s.name}: update count"
        for k, ((fz, x, P, err, nt), (_, nt_ftr, oerr, ofz, ox, oP)) in enumerate(zip(got[i], want[i])):
            where = f"{s.name} update {k}"
            assert nt == nt_ftr, f"{where}: terms {nt} != restatement {nt_ftr}"
            assert int(err[-1]) == oerr, f"{where}: error bits {int(err[-1])} != {oerr}"
            for a, b, what in ((np.array([fz[-1]]), np.array([ofz]), "fz"), (x, ox, "x_hat"), (P, oP, "P_hat")):
                r = rel_err(a, b)
                worst = max(worst, r)
                assert r <= TOL, f"{where}: {what} rel err {r:.2e}"
    print(f"  512 estimators, worst relative error {worst:.2e}", flush=True)
