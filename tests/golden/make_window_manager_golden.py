"""Writes the window-manager fixtures under tests/golden/ from the reference build.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_window_manager_golden.py [fixture names]

The reference's C++ window manager (SlidingWindowManager, include/cauchy_windows.hpp,
behind pycauchy_initialize_{lti,ltv}_window_manager / pycauchy_step / pycauchy_shutdown,
scripts/swig/cauchy/pycauchy.hpp:80-221) forks one process per window.  It is driven here
through the reference build's own functions (oracle/_ref/libpycauchy_ref.so) with the
same ctypes bindings as this library's.  What is written is data: the inputs and, per
measurement, the reference's outputs (fz, x_hat, P_hat, their imaginary-part bounds, the
reporting window and its error bits).
"""
import ctypes as ct
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from cauchyfriendly_amd import pycauchy  # noqa: E402
from cauchyfriendly_amd.cauchy_estimator import C_CauchyDynamicsUpdateContainer  # noqa: E402
from cauchyfriendly_amd import workloads as W  # noqa: E402

REF_LIB = os.path.join(REPO, "oracle", "_ref", "libpycauchy_ref.so")


def lst(a):
    return np.asarray(a, dtype=np.float64).reshape(-1).tolist()


def dynamics(sysobj):
    """LTV callback: Phi_k of the system at the container's step (the step is not changed)."""
    n = sysobj.n

    def dyn(c_duc):
        duc = c_duc.contents
        k = min(max(int(duc.step), 0), sysobj.num_steps - 1)
        Phi = sysobj.Phi(k).reshape(-1)
        for i in range(n * n):
            duc.Phi[i] = float(Phi[i])
    return ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer))(dyn)


def run_manager(mod, mode, sysobj, num_windows, steps, seed, boost=None):
    """Steps the window manager of `mod` (a pycauchy-like module) over the system's
    measurements after srand(seed); returns the per-measurement outputs."""
    ct.CDLL("libc.so.6").srand(ct.c_uint(seed))
    s = sysobj
    args = (s.A0, s.p0, s.b0, s.Phi(0), s.Gamma.reshape(-1), np.zeros(0), np.ravel(s.beta), s.H.reshape(-1),
            np.ravel(s.gamma))
    wvb = np.zeros(0) if boost is None else np.asarray(boost, dtype=np.float64)
    cb = None
    if mode == "lti":
        mod.pycauchy_initialize_lti_window_manager(num_windows, steps, *args, False, False, False, None, 0.0, 0, wvb)
    else:
        cb = dynamics(s)
        mod.pycauchy_initialize_ltv_window_manager(num_windows, steps, *args, cb, False, False, False, None, 0.0, 0,
                                                   wvb)
    out = []
    try:
        for k in range(steps):
            fz, x, P, cfz, cx, cP, win, err = mod.pycauchy_step(np.asarray(s.zs[k], dtype=np.float64), None)
            out.append({"fz": fz, "xhat": lst(x), "Phat": lst(P), "cerr_fz": cfz, "cerr_xhat": cx, "cerr_Phat": cP,
                        "win_idx": int(win), "err": int(err)})
    finally:
        mod.pycauchy_shutdown()
    return out


CASES = [
    # (fixture, mode, system, windows, measurements, srand seed, window variance boost)
    ("wm_lti3", "lti", lambda: W.lti_system(d=3, steps=16, seed_offset=30, name="wm_lti3"), 4, 16, 1234, None),
    ("wm_lti3_boost", "lti", lambda: W.lti_system(d=3, steps=12, seed_offset=32, name="wm_lti3_boost"), 3, 12, 99,
     [0.01, 0.02, 0.03]),
    ("wm_ltv4_p2", "ltv", lambda: W.ltv_system(d=4, steps=10, seed_offset=31, p=2, pncc=2, name="wm_ltv4"), 3, 10,
     4321, None),
    # the 5-state LTV system (two process noises, two measurements per step) with 5 windows
    ("wm_ltv5_p2", "ltv", lambda: W.config1(steps=10), 5, 10, W.config1(steps=10).srand_seed, None),
]


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref = pycauchy.bind(ct.CDLL(REF_LIB), sym=lambda name: "refpy_" + name[len("pycauchy_"):], name="pycauchy_ref")
    only = set(sys.argv[1:])  # fixture names to write (all when none are given)
    for fname, mode, make, nw, steps, seed, boost in CASES:
        if only and fname not in only:
            continue
        s = make()
        out = run_manager(ref, mode, s, nw, steps, seed, boost)
        rec = {"system": s.to_dict(), "mode": mode, "num_windows": nw, "steps": steps, "srand": seed,
               "win_var_boost": boost, "outputs": out}
        with open(os.path.join(HERE, fname + ".json"), "w") as fh:
            json.dump(rec, fh, separators=(",", ":"))
        print(f"{fname}: {steps} measurements, reporting windows {[o['win_idx'] for o in out]}", flush=True)


if __name__ == "__main__":
    main()
