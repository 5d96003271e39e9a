"""Writes the window-bank fixtures under tests/golden/ from the reference.

Run in the container that holds the reference tree (after `make -C oracle ref`):

    python tests/golden/make_window_golden.py

The reference's sliding-window bank and window re-initialisation are Python
(scripts/swig/cauchy/cauchy_estimator.py: PySlidingWindowManager,
PyCauchyEstimator.reset_about_estimator, speyers_window_init) over the SWIG
module `pycauchy`.  Here that module is the reference's own C++ functions
(scripts/swig/cauchy/pycauchy.hpp, built by oracle/Makefile into
oracle/_ref/libpycauchy_ref.so) bound through ctypes, and the reference's Python
file is loaded from the reference tree in this container only.  What is written
is data: the inputs and the reference's outputs.
"""
import ctypes as ct
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from cauchyfriendly_amd import pycauchy as product_pycauchy  # noqa: E402
from cauchyfriendly_amd import workloads as W  # noqa: E402

REF_PY = "/root/reference/scripts/swig/cauchy/cauchy_estimator.py"
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libpycauchy_ref.so")


def reference_module():
    lib = ct.CDLL(REF_LIB)
    shim = product_pycauchy.bind(lib, sym=lambda name: "refpy_" + name[len("pycauchy_"):], name="pycauchy")
    sys.modules["pycauchy"] = shim
    spec = importlib.util.spec_from_file_location("reference_cauchy_estimator", REF_PY)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def lst(a):
    return np.asarray(a, dtype=np.float64).reshape(-1).tolist()


def speyer_cases(ref):
    rng = np.random.default_rng(20261202 + 77)
    cases = []
    for n in (2, 3, 4, 5, 7):
        for k in range(3):
            M = rng.standard_normal((n, n))
            Var = M @ M.T / n + 0.05 * np.eye(n)
            if k == 2:  # an eigenvalue below the reference's tolerance: the diagonal is boosted
                w, V = np.linalg.eigh(Var)
                w[0] = -3e-5
                Var = V @ np.diag(w) @ V.T
                Var = 0.5 * (Var + Var.T)
            x = rng.standard_normal(n)
            H = rng.standard_normal(n)
            gamma = float(rng.uniform(0.05, 0.5))
            z = float(H @ x + gamma * rng.standard_cauchy())
            A0, p0, b0 = ref.speyers_window_init(x.copy(), Var.copy(), H.copy(), gamma, z)
            cases.append({"n": n, "xhat": lst(x), "Phat": lst(Var), "H": lst(H), "gamma": gamma, "z": z,
                          "A0": lst(A0), "p0": lst(p0), "b0": lst(b0)})
    return cases


def run_bank(ref, mode, sysobj, num_windows, steps, seed):
    libc = ct.CDLL("libc.so.6")
    libc.srand(ct.c_uint(seed))
    swm = ref.PySlidingWindowManager(mode, num_windows, swm_debug_print=False, win_debug_print=False)
    n = sysobj.n
    Gamma = sysobj.Gamma if sysobj.pncc > 1 else sysobj.Gamma.reshape(-1)
    H = sysobj.H if sysobj.p > 1 else sysobj.H.reshape(-1)
    beta = np.asarray(sysobj.beta, dtype=np.float64).reshape(-1)
    gamma = np.asarray(sysobj.gamma, dtype=np.float64).reshape(-1)
    if mode == "lti":
        swm.initialize_lti(sysobj.A0.copy(), sysobj.p0.copy(), sysobj.b0.copy(), sysobj.Phi(0).copy(), None,
                           Gamma.copy(), beta.copy(), H.copy(), gamma.copy())
    else:
        def dyn(c_duc):
            duc = c_duc.contents
            k = int(duc.step)
            Phi = sysobj.Phi(k)
            for i in range(n * n):
                duc.Phi[i] = float(Phi.reshape(-1)[i])
            duc.step = k + 1
        swm.initialize_ltv(sysobj.A0.copy(), sysobj.p0.copy(), sysobj.b0.copy(), sysobj.Phi(0).copy(), None,
                           Gamma.copy(), beta.copy(), H.copy(), gamma.copy(), dyn)
    out = []
    for k in range(steps):
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
    rec = {"cases": speyer_cases(ref)}
    with open(os.path.join(HERE, "speyer_init.json"), "w") as fh:
        json.dump(rec, fh, separators=(",", ":"))
    print(f"speyer_init: {len(rec['cases'])} cases")
    banks = [
        ("window_lti3", "lti", W.lti_system(d=3, steps=16, seed_offset=30, name="bank_lti3"), 4, 16, 1234),
        ("window_ltv4_p2", "ltv", W.ltv_system(d=4, steps=10, seed_offset=31, p=2, pncc=2, name="bank_ltv4"), 3, 10,
         4321),
    ]
    for fname, mode, sysobj, nw, steps, seed in banks:
        out = run_bank(ref, mode, sysobj, nw, steps, seed)
        rec = {"system": sysobj.to_dict(), "mode": mode, "num_windows": nw, "steps": steps, "srand": seed,
               "outputs": out}
        with open(os.path.join(HERE, fname + ".json"), "w") as fh:
            json.dump(rec, fh, separators=(",", ":"))
        print(f"{fname}: {steps} steps, best windows {[o['best'] for o in out]}")


if __name__ == "__main__":
    main()
