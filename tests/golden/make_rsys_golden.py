"""Writes the relative-system fixtures under tests/golden/ from the reference.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_rsys_golden.py

The reference evaluates the relative system of two estimators through
pycauchy_single_step_eval_2d_rsys_cpdf (scripts/swig/cauchy/pycauchy.hpp:673-750,
include/cauchy_prediction.hpp), called from Python as scripts/swig/leo/gmat_mce.py:950-975
does.  Both estimators are the reference's PyCauchyEstimator over the reference's C++
functions (oracle/_ref/libpycauchy_ref.so, bound through ctypes; the reference's Python
file is loaded from the reference tree in this container only).  What is written is data.
"""
import ctypes as ct
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_window_golden import REF_LIB, lst, reference_module  # noqa: E402
from cauchyfriendly_amd import workloads as W  # noqa: E402


def stepped(ref, sysobj, mode, steps):
    ct.CDLL("libc.so.6").srand(ct.c_uint(sysobj.srand_seed))
    est = ref.PyCauchyEstimator(mode, sysobj.num_steps, False)
    Gamma = sysobj.Gamma if sysobj.pncc > 1 else sysobj.Gamma.reshape(-1)
    H = sysobj.H if sysobj.p > 1 else sysobj.H.reshape(-1)
    beta = np.asarray(sysobj.beta, dtype=np.float64).reshape(-1)
    gamma = np.asarray(sysobj.gamma, dtype=np.float64).reshape(-1)
    n = sysobj.n
    if mode == "lti":
        est.initialize_lti(sysobj.A0.copy(), sysobj.p0.copy(), sysobj.b0.copy(), sysobj.Phi(0).copy(), None,
                           Gamma.copy(), beta.copy(), H.copy(), gamma.copy())
    else:
        def dyn(c_duc):
            duc = c_duc.contents
            k = int(duc.step)
            Phi = sysobj.Phi(k).reshape(-1)
            for i in range(n * n):
                duc.Phi[i] = float(Phi[i])
            duc.step = k + 1
        est.initialize_ltv(sysobj.A0.copy(), sysobj.p0.copy(), sysobj.b0.copy(), sysobj.Phi(0).copy(), None,
                           Gamma.copy(), beta.copy(), H.copy(), gamma.copy(), dyn)
        est._keep_dyn = dyn
    for k in range(steps):
        est.step(np.asarray(sysobj.zs[k], dtype=np.float64).reshape(-1), None)
    return est


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref = reference_module()
    pc = sys.modules["pycauchy"]
    rng = np.random.default_rng(20261202 + 91)
    cases = [
        ("rsys_lti3", "lti", W.lti_system(d=3, steps=8, seed_offset=50, name="rs_p3"),
         W.lti_system(d=3, steps=8, seed_offset=51, name="rs_s3"), 3, 3,
         [np.array([[1.0, 0, 0], [0, 1.0, 0]]), rng.standard_normal((2, 3))], [0.0, 1e-4]),
        ("rsys_ltv5_p2", "ltv", W.ltv_system(d=5, steps=4, seed_offset=52, name="rs_p5"),
         W.ltv_system(d=5, steps=4, seed_offset=53, name="rs_s5"), 2, 2,
         [np.array([[1.0, 0, 0, 0, 0], [0, 0, 0, 0, 1.0]])], [1e-6]),
        ("rsys_lti2", "lti", W.lti_system(d=2, steps=7, seed_offset=54, name="rs_p2"),
         W.lti_system(d=2, steps=7, seed_offset=55, name="rs_s2"), 3, 4, [np.eye(2)], [0.0]),
    ]
    for fname, mode, sp, ss, kp, ks, trels, epss in cases:
        evals = []
        P = stepped(ref, sp, mode, kp)
        S = stepped(ref, ss, mode, ks)
        for Trel in trels:
            for eps in epss:
                # spans that are not multiples of the resolution: the grid's point count
                # (ceil(span / res + 1)) then does not hinge on the last bit of the mean
                grid = [-1.0, 1.0, 0.3, -0.8, 0.8, 0.35]
                out = pc.pycauchy_single_step_eval_2d_rsys_cpdf(Trel.reshape(-1), S.py_handle, P.py_handle, eps,
                                                                *grid)
                fz, xh, Ph, cfz, cxh, cPh, pts, nx, ny = out
                evals.append({"Trel": lst(Trel), "eps": eps, "grid": grid,
                              "fz": lst(fz), "xhat": lst(xh), "Phat": lst(Ph), "cerr_fz": lst(cfz),
                              "cerr_xhat": lst(cxh), "cerr_Phat": lst(cPh), "points": lst(pts), "nx": nx, "ny": ny})
        rec = {"mode": mode, "primary": {"system": sp.to_dict(), "steps": kp, "Nt": int(P.get_num_CF_terms()),
                                         "xhat": lst(P.moment_info["x"][-1]), "Phat": lst(P.moment_info["P"][-1])},
               "secondary": {"system": ss.to_dict(), "steps": ks, "Nt": int(S.get_num_CF_terms()),
                             "xhat": lst(S.moment_info["x"][-1]), "Phat": lst(S.moment_info["P"][-1])},
               "evals": evals}
        P.shutdown()
        S.shutdown()
        with open(os.path.join(HERE, fname + ".json"), "w") as fh:
            json.dump(rec, fh, separators=(",", ":"))
        print(f"{fname}: {rec['primary']['Nt']} x {rec['secondary']['Nt']} pairs, {len(evals)} evaluations, "
              f"grid {evals[0]['nx']} x {evals[0]['ny']}")


if __name__ == "__main__":
    main()
