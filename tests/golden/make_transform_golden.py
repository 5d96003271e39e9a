"""Writes the deterministic-transform fixtures under tests/golden/ from the reference.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_transform_golden.py

The reference transforms an estimator's characteristic function between updates with
pycauchy_single_step_deterministic_transform (scripts/swig/cauchy/pycauchy.hpp:654-664:
deterministic_time_prop(Trans, NULL, NULL) then shift_cf_by_bias(bias),
include/cauchy_estimator.hpp:1312-1355), reached from Python as
PyCauchyEstimator.deterministic_transform (scripts/swig/cauchy/cauchy_estimator.py:809-818).
Each case steps the reference's PyCauchyEstimator (over oracle/_ref/libpycauchy_ref.so,
bound through ctypes; the reference's Python file is loaded from the reference tree in this
container only), applies the transform after update `at`, and records every update's
moments, error code and term count.  What is written is data.
"""
import ctypes as ct
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_window_golden import REF_LIB, lst, reference_module  # noqa: E402
from make_rsys_golden import stepped  # noqa: E402
from cauchyfriendly_amd import workloads as W  # noqa: E402


def run_case(ref, sysobj, mode, at, Trans, bias):
    est = stepped(ref, sysobj, mode, at)
    nts = [None] * at
    est.deterministic_transform(Trans.copy(), bias.copy())
    for k in range(at, sysobj.num_steps):
        est.step(np.asarray(sysobj.zs[k], dtype=np.float64).reshape(-1), None)
        nts.append(int(est.get_num_CF_terms()) if k < sysobj.num_steps - 1 else None)
    mi = est.moment_info
    rec = {"mode": mode, "system": sysobj.to_dict(), "at": at, "Trans": lst(Trans), "bias": lst(bias),
           "fz": [float(np.real(f)) for f in mi["fz"]], "x": [lst(x) for x in mi["x"]],
           "P": [lst(P) for P in mi["P"]], "err_code": [int(e) for e in mi["err_code"]], "Nt": nts}
    est.shutdown()
    return rec


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref = reference_module()
    rng = np.random.default_rng(20261202 + 93)
    T3 = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    T4 = np.eye(4) + 0.2 * rng.standard_normal((4, 4))
    cases = [
        ("transform_lti3", "lti", W.lti_system(d=3, steps=7, seed_offset=60, name="tr_lti3"), 3, T3,
         rng.standard_normal(3)),
        ("transform_lti3_nobias", "lti", W.lti_system(d=3, steps=6, seed_offset=61, name="tr_lti3b"), 2, T3,
         np.zeros(3)),
        ("transform_ltv4_p2", "ltv", W.ltv_system(d=4, steps=4, p=2, seed_offset=62, name="tr_ltv4"), 2, T4,
         rng.standard_normal(4)),
    ]
    for fname, mode, sysobj, at, T, b in cases:
        rec = run_case(ref, sysobj, mode, at, T, b)
        with open(os.path.join(HERE, fname + ".json"), "w") as fh:
            json.dump(rec, fh, separators=(",", ":"))
        print(f"{fname}: {len(rec['fz'])} updates, transform after update {at}, terms {rec['Nt']}")


if __name__ == "__main__":
    main()
