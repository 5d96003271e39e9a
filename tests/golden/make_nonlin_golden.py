"""Writes the extended-mode (nonlinear) fixtures under tests/golden/ from the reference.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_nonlin_golden.py

The reference's extended estimator is pycauchy_initialize_nonlin /
pycauchy_single_step_nonlin (scripts/swig/cauchy/pycauchy.hpp:365-396, 545-652),
driven by its Python classes PyCauchyEstimator and PySlidingWindowManager in mode
"nonlin" (scripts/swig/cauchy/cauchy_estimator.py:709-751, 1161-1180).  As for the
window-bank fixtures, the SWIG module is the reference's own C++ functions bound
through ctypes (oracle/_ref/libpycauchy_ref.so) and the reference's Python file is
loaded from the reference tree in this container only.  The callbacks are the
seeded pendulum of cauchyfriendly_amd/workloads.py over the reference's own
container view.  What is written is data: the inputs and the reference's outputs.
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


def run_single(ref, sysobj):
    ct.CDLL("libc.so.6").srand(ct.c_uint(sysobj.srand_seed))
    est = ref.PyCauchyEstimator("nonlin", sysobj.num_steps, False)
    dyn, msmt, ext = sysobj.callbacks(ref.Py_CauchyDynamicsUpdateContainer)
    est.initialize_nonlin(sysobj.xbar0.copy(), sysobj.A0.copy(), sysobj.p0.copy(), sysobj.b0.copy(),
                          sysobj.beta.copy(), sysobj.gamma.copy(), dyn, msmt, ext, 0, sysobj.dt, 0)
    out = []
    for k in range(sysobj.num_steps):
        est.step(np.asarray(sysobj.zs[k], dtype=np.float64).reshape(-1), None)
        out.append({"fz": lst(est._fz), "xhat": lst(est._x), "Phat": lst(est._P), "xbar": lst(est._xbar),
                    "zbar": lst(est._zbar), "cerr_fz": lst(est._cerr_fz), "err": [int(e) for e in est._err_code],
                    "Phi": lst(est._Phi), "Gamma": lst(est._Gamma), "H": lst(est._H), "Nt":
                    int(est.get_num_CF_terms())})
    est.shutdown()
    return out


def run_bank(ref, sysobj, num_windows):
    ct.CDLL("libc.so.6").srand(ct.c_uint(sysobj.srand_seed))
    swm = ref.PySlidingWindowManager("nonlin", num_windows, swm_debug_print=False, win_debug_print=False)
    dyn, msmt, ext = sysobj.callbacks(ref.Py_CauchyDynamicsUpdateContainer)
    swm.initialize_nonlin(sysobj.xbar0.copy(), sysobj.A0.copy(), sysobj.p0.copy(), sysobj.b0.copy(),
                          sysobj.beta.copy(), sysobj.gamma.copy(), dyn, msmt, ext, 0, sysobj.dt, 0)
    out = []
    for k in range(sysobj.num_steps):
        xhat, Phat, wx, wP = swm.step(np.asarray(sysobj.zs[k], dtype=np.float64).reshape(-1), None)
        out.append({"xhat": lst(xhat), "Phat": lst(Phat), "wavg_xhat": lst(wx), "wavg_Phat": lst(wP),
                    "best": int(swm.best_idx), "fz": float(swm.moment_info["fz"][-1]),
                    "err": int(swm.moment_info["err_code"][-1]), "win_counts": [int(c) for c in swm.win_counts]})
    swm.shutdown()
    return out


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref = reference_module()
    singles = [("nonlin_pend_p1", W.pendulum_system(p=1, steps=8, seed_offset=40, name="pend_p1")),
               ("nonlin_pend_p2", W.pendulum_system(p=2, steps=6, seed_offset=41, name="pend_p2"))]
    for fname, sysobj in singles:
        out = run_single(ref, sysobj)
        with open(os.path.join(HERE, fname + ".json"), "w") as fh:
            json.dump({"system": sysobj.to_dict(), "outputs": out}, fh, separators=(",", ":"))
        print(f"{fname}: terms per step {[o['Nt'] for o in out]}")
    sysobj = W.pendulum_system(p=1, steps=14, seed_offset=42, name="pend_bank")
    out = run_bank(ref, sysobj, 4)
    with open(os.path.join(HERE, "window_nonlin_pend.json"), "w") as fh:
        json.dump({"system": sysobj.to_dict(), "num_windows": 4, "outputs": out}, fh, separators=(",", ":"))
    print(f"window_nonlin_pend: best windows {[o['best'] for o in out]}")


if __name__ == "__main__":
    main()
