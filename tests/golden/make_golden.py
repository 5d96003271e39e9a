"""Writes the golden fixtures under tests/golden/ from the reference build.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_golden.py [fixture name ...]

For every system it records the inputs and, per scalar measurement update, the
reference estimator's outputs: fz before term reduction, the public fz, x_hat and
P_hat (complex), the numeric error bits, the term counts after MU/MUC and after
reduction (total and per shape) and the reference thread count.  For a few small
systems it also stores every term (A, p, b and its g-table keys and values) after
selected updates, and the construction-time random draws (root point and the
first b_pert vector).  Fixtures are data only: inputs and expected outputs.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from cauchyfriendly_amd import workloads as W  # noqa: E402
from oracle import driver as dr  # noqa: E402
from tests.golden.demo_systems import reference_demo_systems  # noqa: E402


def cx(z):
    return [float(np.real(z)), float(np.imag(z))]


def cxa(a):
    a = np.asarray(a).reshape(-1)
    return [[float(v.real), float(v.imag)] for v in a]


def record(sysobj, num_steps=None, term_updates=(), kind="ref", run_steps=None):
    """num_steps: the estimator's window (it sets the largest shape and the random draws);
    run_steps: how many of its time steps are stepped and recorded (all by default)."""
    steps = sysobj.num_steps if num_steps is None else num_steps
    est = dr.RefEstimator(steps, sysobj.n, sysobj.pncc, sysobj.p, sysobj.A0, sysobj.p0, sysobj.b0,
                          cmcc=sysobj.cmcc, seed=sysobj.srand_seed, kind=kind)
    out = {"system": sysobj.to_dict(), "num_steps": steps, "root_point": est.root_point().tolist(),
           "b_pert0": est.b_pert(0).tolist(), "b_pert7": est.b_pert(7).tolist(), "updates": [], "terms": {}}
    if run_steps is not None:
        out["run_steps"] = run_steps
    Gt = np.ascontiguousarray(sysobj.Gamma.reshape(-1))
    B = None if sysobj.B is None else np.ascontiguousarray(sysobj.B.reshape(-1))
    u = 0
    for k in range(steps if run_steps is None else run_steps):
        for i in range(sysobj.p):
            r = est.step(sysobj.zs[k, i], sysobj.Phi(k).reshape(-1), Gt, sysobj.beta, sysobj.H[i],
                         sysobj.gamma[i], B, sysobj.u(k))
            out["updates"].append({
                "fz": cx(r["fz"]), "fz_public": cx(r["fz_public"]), "last_fz": cx(r["last_fz"]),
                "mean": cxa(r["mean"]), "cov": cxa(r["cov"]), "err": r["err"], "nt_mu": r["nt_mu"],
                "nt_ftr": r["nt_ftr"], "threads": r["threads"], "shape_mu": r["shape_mu"].tolist(),
                "shape_ftr": r["shape_ftr"].tolist()})
            if u in term_updates:
                terms = []
                for t in est.terms():
                    terms.append({"m": t["m"], "cells": t["cells"], "A": t["A"].reshape(-1).tolist(),
                                  "p": t["p"].tolist(), "b": t["b"].tolist(), "keys": t["keys"].tolist(),
                                  "g": cxa(t["g"])})
                out["terms"][str(u)] = terms
            u += 1
    est.close()
    return out


def _named(sysobj, name):
    sysobj.name = name
    return sysobj


CASES = [
    # (system factory, window in time steps or None, updates whose term tables are stored
    #  [, time steps stepped when fewer than the window])
    (W.config0, None, (0, 1, 2, 3)),
    (lambda: W.config1(), 4, (1,)),
    (lambda: W.lti_system(d=2, steps=10, seed_offset=20, name="lti2_w10"), None, (5,)),
    (lambda: W.lti_system(d=4, steps=7, seed_offset=21, name="lti4_w7"), None, ()),
    (lambda: W.lti_system(d=3, steps=9, seed_offset=10, name="lti3_w9"), None, ()),
    # the first of the staggered 7-state orbit windows, 4 of its 6 steps (8 updates)
    (lambda: W.leo7_windows(count=1)[0], 4, ()),
    # the largest state dimensions the engine takes (6 and its maximum, 8)
    (lambda: W.lti_system(d=6, steps=6, seed_offset=106, name="lti6_w6"), None, (1,)),
    (lambda: W.lti_system(d=8, steps=5, seed_offset=108, name="lti8_w5"), None, (0,)),
    # the orbit window at the 1 s step of the largest configuration (7-state orbit): a window of 6
    # steps (largest shape 12), stepped for its first 4 (8 updates, 750k terms at the last;
    # the reference's own covariance check fails from the 5th update on)
    (lambda: _named(W.leo7_windows(count=1, window=6, dt=1.0)[0], "leo7_dt1_w0"), 6, (), 4),
    # largest shapes the engine takes (20 hyperplanes): 2 states with three process noises
    # over 7 steps, and with two over 10 steps; the cell enumerator then runs up to shape
    # 17 and 18, and the g-tables use the KM = 24 kernels and the binary-search lookups
    (lambda: W.ltv_system(d=2, steps=7, pncc=3, p=1, seed_offset=122, name="ltv2_pn3_w7"), None, ()),
    (lambda: W.ltv_system(d=2, steps=10, pncc=2, p=1, seed_offset=120, name="ltv2_pn2_w10"), None, ()),
    # 16 and 17 hyperplanes: rank-structure g-table lookups with the KM = 16 and KM = 32 rows
    (lambda: W.ltv_system(d=2, steps=8, pncc=2, p=1, seed_offset=124, name="ltv2_pn2_w8"), None, ()),
    (lambda: W.ltv_system(d=3, steps=8, pncc=2, p=1, seed_offset=125, name="ltv3_pn2_w8"), None, ()),
]


def main(only=()):
    """Writes every fixture, or only those named on the command line."""
    if not dr.available("ref"):
        raise SystemExit("build the reference first: make -C oracle ref")
    systems = [(c[0], c[1], c[2], c[3] if len(c) > 3 else None) for c in CASES]
    for sy in W.edge_systems():
        systems.append((lambda sy=sy: sy, None, (1, 3) if sy.name in ("horthog3", "scalar1", "tpcoalign3") else (),
                        None))
    for sy in reference_demo_systems():
        systems.append((lambda sy=sy: sy, None, (2,) if sy.name == "ref_3state_3msmts_horth" else (), None))
    for factory, steps, tu, run_steps in systems:
        sysobj = factory()
        if only and sysobj.name not in only:
            continue
        rec = record(sysobj, steps, tu, run_steps=run_steps)
        path = os.path.join(HERE, f"{sysobj.name}.json")
        with open(path, "w") as fh:
            json.dump(rec, fh, separators=(",", ":"))
        print(f"{sysobj.name}: {len(rec['updates'])} updates, terms {[u['nt_mu'] for u in rec['updates']]} "
              f"-> {os.path.getsize(path) // 1024} KiB")


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
