"""Writes the debug-mode fixtures under tests/golden/ from the reference build.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_debug_golden.py

With debug_print, the reference's estimator recomputes its public fz, mean and covariance
after every term reduction from the reduced terms (compute_moments(false),
include/cauchy_estimator.hpp:524-602, with eval_g_yei_after_ftr, include/cauchy_term.hpp:403-433)
and prints them.  The fixtures are the outputs of the reference's own
pycauchy_initialize_lti / _ltv and pycauchy_single_step_ltiv with debug_print = true
(scripts/swig/cauchy/pycauchy.hpp, built into oracle/_ref/libpycauchy_ref.so), driven through
the same ctypes bindings as this library's; what the reference prints is discarded.  What is
written is data: the inputs and, per step, the outputs.
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
    n = sysobj.n

    def dyn(c_duc):
        duc = c_duc.contents
        k = min(max(int(duc.step), 0), sysobj.num_steps - 1)
        Phi = sysobj.Phi(k).reshape(-1)
        for i in range(n * n):
            duc.Phi[i] = float(Phi[i])
    return ct.CFUNCTYPE(None, ct.POINTER(C_CauchyDynamicsUpdateContainer))(dyn)


def run(mod, mode, s):
    """Steps the system with debug_print; returns per-step outputs."""
    ct.CDLL("libc.so.6").srand(ct.c_uint(s.srand_seed))
    args = (s.num_steps, s.A0, s.p0, s.b0, s.Phi(0), s.Gamma.reshape(-1), np.zeros(0), np.ravel(s.beta),
            s.H.reshape(-1), np.ravel(s.gamma))
    cb = None
    if mode == "lti":
        h = mod.pycauchy_initialize_lti(*args, 0.0, 0, True)
    else:
        cb = dynamics(s)
        h = mod.pycauchy_initialize_ltv(*args, cb, 0.0, 0, True)
    out = []
    try:
        for k in range(s.num_steps):
            res = mod.pycauchy_single_step_ltiv(h, np.asarray(s.zs[k], dtype=np.float64), None)
            fz, x, P, cfz, cx, cP, err = res[6:]
            out.append({"fz": lst(fz), "xhat": lst(x), "Phat": lst(P), "err": [int(e) for e in err]})
    finally:
        mod.pycauchy_single_step_shutdown(h)
    return out


CASES = [
    ("debug_lti3", "lti", lambda: W.lti_system(d=3, steps=6, seed_offset=40, name="debug_lti3")),
    ("debug_ltv4_p2", "ltv", lambda: W.ltv_system(d=4, steps=4, seed_offset=41, p=2, pncc=2, name="debug_ltv4")),
]


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref = pycauchy.bind(ct.CDLL(REF_LIB), sym=lambda name: "refpy_" + name[len("pycauchy_"):], name="pycauchy_ref")
    for fname, mode, make in CASES:
        s = make()
        sys.stdout.flush()
        saved = os.dup(1)
        null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(null, 1)
        try:
            out = run(ref, mode, s)
        finally:
            ct.CDLL(None).fflush(None)   # the reference's buffered prints go to /dev/null too
            sys.stdout.flush()
            os.dup2(saved, 1)
            os.close(null)
            os.close(saved)
        rec = {"system": s.to_dict(), "mode": mode, "outputs": out}
        with open(os.path.join(HERE, fname + ".json"), "w") as fh:
            json.dump(rec, fh, separators=(",", ":"))
        print(f"{fname}: {len(out)} steps, fz {[o['fz'][-1] for o in out]}", flush=True)


if __name__ == "__main__":
    main()
