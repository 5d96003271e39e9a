import glob
import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN_DIR = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP engine)")


# fixtures of other record types (window bank, window re-initialisation, extended mode, densities,
# relative systems, deterministic transforms)
NON_ESTIMATOR_FIXTURES = ("speyer_", "window_", "nonlin_", "cpdf_", "rsys_", "transform_", "wm_", "debug_")


def golden_names():
    """Names of the per-update estimator fixtures."""
    names = (os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.json")))
    return sorted(n for n in names if not n.startswith(NON_ESTIMATOR_FIXTURES))


def load_golden(name):
    with open(os.path.join(GOLDEN_DIR, name + ".json")) as fh:
        return json.load(fh)


def cplx(v):
    a = np.asarray(v, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.max(np.abs(b))) if b.size else 0.0
    if den == 0.0:
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / den)


def compare_update(got, exp, tol, where):
    """got: dict from a driver step(); exp: a fixture update record."""
    assert got["nt_mu"] == exp["nt_mu"], f"{where}: terms after MU {got['nt_mu']} != {exp['nt_mu']}"
    assert got["nt_ftr"] == exp["nt_ftr"], f"{where}: terms after reduction {got['nt_ftr']} != {exp['nt_ftr']}"
    assert list(got["shape_mu"]) == exp["shape_mu"], f"{where}: per-shape counts after MU differ"
    assert list(got["shape_ftr"]) == exp["shape_ftr"], f"{where}: per-shape counts after reduction differ"
    assert got["err"] == exp["err"], f"{where}: error bits {got['err']} != {exp['err']}"
    e_fz = rel_err(got["fz"], complex(*exp["fz"]))
    e_x = rel_err(got["mean"], cplx(exp["mean"]))
    e_P = rel_err(got["cov"].reshape(-1), cplx(exp["cov"]))
    assert max(e_fz, e_x, e_P) <= tol, f"{where}: rel err fz {e_fz:.2e} x {e_x:.2e} P {e_P:.2e} > {tol}"
    assert rel_err(got["fz_public"], complex(*exp["fz_public"])) <= tol, f"{where}: public fz differs"
    return max(e_fz, e_x, e_P)


def compare_terms(got_terms, exp_terms, tol, where):
    assert len(got_terms) == len(exp_terms), f"{where}: {len(got_terms)} terms != {len(exp_terms)}"
    for i, (a, b) in enumerate(zip(got_terms, exp_terms)):
        assert a["m"] == b["m"] and a["cells"] == b["cells"], f"{where} term {i}: shape/cells differ"
        assert np.array_equal(np.asarray(a["keys"], dtype=np.int64), np.asarray(b["keys"], dtype=np.int64)), \
            f"{where} term {i}: g-table keys differ"
        for f in ("A", "p", "b"):
            assert rel_err(np.asarray(a[f]).reshape(-1), np.asarray(b[f])) <= tol, f"{where} term {i}: {f} differs"
        if b["cells"]:
            assert rel_err(np.asarray(a["g"]), cplx(b["g"])) <= tol, f"{where} term {i}: g values differ"
