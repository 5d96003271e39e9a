"""Measured choice of the merged batch's launch shapes.

The merged batch (pycauchy_batch_single_step_ltiv_merged) has knobs that change its speed
and never its results (every block does the work of its own estimator's launch, and the
parity tests run the non-default values against the reference):

  groups             host threads, one HIP stream each, that the estimators are dealt to
  enum_threads       lanes per parent in the time-propagation cell enumerator (64, 128, 256)
  big_group_members  reduction groups with more members evaluate them one wavefront per
                     member (k_group_members) instead of one wavefront per group
  geval_mode         the g-table kernel's parent lookup: -1 automatic (the rank structure in
                     LDS where it fits), 4 the rank structure with the parent values in LDS

tune_batch() times candidate settings on the caller's own workload and keeps the fastest,
by coordinate descent: one knob at a time, each value timed on whole passes (the best of
`repeats`), starting from the defaults.  A run's choice can be kept as the tuning table
(save_table), which later runs load instead of tuning again (load_table).
"""
from __future__ import annotations

import json
import os

from . import _native

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")

DEFAULTS = {"groups": 2, "enum_threads": 64, "big_group_members": 16, "geval_mode": -1}
CANDIDATES = {
    "groups": (1, 2, 3, 4),
    "enum_threads": (64, 128, 256),
    "big_group_members": (8, 16, 32),
    "geval_mode": (-1, 4),
}


def set_shapes(estimators, enum_threads=0, big_group_members=0, geval_mode=0):
    """Sets the kernel shapes of every estimator (0 keeps a knob as it is)."""
    L = _native.lib()
    for e in estimators:
        if L.mce_estimator_set_shapes(e.py_handle, int(enum_threads), int(big_group_members), int(geval_mode)) != 0:
            raise RuntimeError(L.mce_last_error().decode())


def tune_batch(estimators, run_pass, candidates=None, repeats=2):
    """Chooses the shapes for `estimators`, stepped by run_pass(groups) -> seconds (one
    whole pass of the caller's workload with merged launches from `groups` threads).
    Leaves the chosen kernel shapes set on the estimators and returns (chosen settings,
    [(settings, seconds) for every setting timed])."""
    cand = dict(CANDIDATES, **(candidates or {}))

    def measure(cfg):
        set_shapes(estimators, cfg["enum_threads"], cfg["big_group_members"], cfg["geval_mode"])
        return min(run_pass(cfg["groups"]) for _ in range(max(int(repeats), 1)))

    best = dict(DEFAULTS)
    t_best = measure(best)
    trials = [(dict(best), t_best)]
    for knob, values in cand.items():
        for v in values:
            if v == best[knob]:
                continue
            cfg = dict(best, **{knob: v})
            t = measure(cfg)
            trials.append((cfg, t))
            if t < t_best:
                best, t_best = cfg, t
    set_shapes(estimators, best["enum_threads"], best["big_group_members"], best["geval_mode"])
    return best, trials


def load_table(workload, path=TABLE):
    """The tuned settings recorded for `workload`, or None."""
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    e = t.get(workload)
    return dict(DEFAULTS, **e["settings"]) if e else None


def save_table(workload, settings, trials, path=TABLE):
    """Records the settings chosen for `workload` (with the timings they were chosen by)."""
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        t = {}
    t[workload] = {"settings": dict(settings),
                   "trials_ms": [[dict(c), round(1e3 * sec, 2)] for c, sec in trials]}
    with open(path, "w") as fh:
        json.dump(t, fh, indent=1)
