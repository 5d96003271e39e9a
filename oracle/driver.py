"""ctypes drivers for the two CPU checkers, with one interface:

* ``kind="ref"``    -> oracle/_ref/libcauchy_ref.so, the reference estimator built from
  its own sources by oracle/Makefile (present only where the reference tree was);
* ``kind="oracle"`` -> oracle/liboracle.so, this repository's CPU restatement.

Test infrastructure only: used by tests/, tests/golden/make_golden.py,
__graft_entry__.smoke() and bench.py's cpu_baseline leg."""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "ref": (os.path.join(_HERE, "_ref", "libcauchy_ref.so"), "ref_"),
    "oracle": (os.path.join(_HERE, "liboracle.so"), "oracle_"),
}
REF_LIB = LIBS["ref"][0]
_libs = {}

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)


def available(kind: str = "ref") -> bool:
    return os.path.exists(LIBS[kind][0])


class _Prefixed:
    """Attribute access ``L.ref_step`` resolves to ``<prefix>step`` of the loaded library."""

    def __init__(self, cdll, prefix):
        self._cdll = cdll
        self._prefix = prefix

    def __getattr__(self, name):
        if name.startswith("ref_"):
            name = self._prefix + name[4:]
        return getattr(self._cdll, name)


def lib(kind: str = "ref"):
    if kind not in _libs:
        path, pre = LIBS[kind]
        C = ct.CDLL(path)
        L = _Prefixed(C, pre)
        L.ref_create.restype = ct.c_void_p
        L.ref_create.argtypes = [ct.c_int] * 5 + [_dp, _dp, _dp, ct.c_int, ct.c_uint]
        L.ref_destroy.argtypes = [ct.c_void_p]
        for f in ("ref_shape_range", "ref_master_step", "ref_num_terms"):
            getattr(L, f).restype = ct.c_int
            getattr(L, f).argtypes = [ct.c_void_p]
        L.ref_root_point.argtypes = [ct.c_void_p, _dp]
        L.ref_b_pert.argtypes = [ct.c_void_p, ct.c_int, _dp]
        L.ref_step.restype = ct.c_int
        L.ref_step.argtypes = [ct.c_void_p, ct.c_double, _dp, _dp, _dp, _dp, ct.c_double, _dp, _dp, _dp, _ip]
        L.ref_public_moments.argtypes = [ct.c_void_p, _dp, _dp, _dp]
        L.ref_finalize_extended_moments.argtypes = [ct.c_void_p, _dp]
        L.ref_term.restype = ct.c_int
        L.ref_term.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, _ip, _dp, _dp, _dp,
                               ct.POINTER(ct.c_uint), _dp]
        L.ref_set_tr_search_idxs_ordering.restype = ct.c_int
        L.ref_set_tr_search_idxs_ordering.argtypes = [_ip, ct.c_int]
        if kind == "oracle":
            C.oracle_set_random_draws.argtypes = [ct.c_void_p, _dp, _dp]
            C.oracle_path_counters.argtypes = [ct.c_void_p, ct.POINTER(ct.c_longlong)]
            C.oracle_fz_raw.argtypes = [ct.c_void_p, _dp]
            C.oracle_set_debug.argtypes = [ct.c_void_p, ct.c_int]
            C.oracle_tape_mode.argtypes = [ct.c_void_p, ct.c_int]
            C.oracle_tape_copy.argtypes = [ct.c_void_p, ct.c_void_p]
            C.oracle_tape_status.restype = ct.c_int
            C.oracle_tape_status.argtypes = [ct.c_void_p, ct.POINTER(ct.c_longlong)]
        _libs[kind] = L
    return _libs[kind]


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class RefEstimator:
    """One CPU estimator (reference build or restatement); step() returns the
    pre-FTR moments and the term counts after MU/MUC and after FTR."""

    def __init__(self, num_steps, d, pncc, p, A0, p0, b0, cmcc=0, seed=None, kind="ref"):
        if kind == "ref" and not available("ref"):
            raise RuntimeError("reference build oracle/_ref/libcauchy_ref.so is not present")
        self.kind = kind
        L = lib(kind)
        self.d, self.pncc, self.p, self.cmcc = d, pncc, p, cmcc
        self._keep = [_d(A0), _d(p0), _d(b0)]
        self.h = L.ref_create(num_steps, d, cmcc, pncc, p, self._keep[0][1], self._keep[1][1],
                              self._keep[2][1], 1 if seed is not None else 0,
                              ct.c_uint(seed if seed is not None else 0))
        self.shape_range = L.ref_shape_range(self.h)

    def close(self):
        if self.h:
            lib(self.kind).ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def root_point(self):
        out = np.zeros(self.d)
        lib(self.kind).ref_root_point(self.h, out.ctypes.data_as(_dp))
        return out

    def b_pert(self, tid):
        out = np.zeros(self.shape_range - 1)
        lib(self.kind).ref_b_pert(self.h, tid, out.ctypes.data_as(_dp))
        return out

    def step(self, z, Phi, Gamma, beta, H_row, gamma, B=None, u=None):
        d = self.d
        sr = self.shape_range
        mom = np.zeros(6 + 2 * d + 2 * d * d)
        cnt = np.zeros(4 + 2 * sr, dtype=np.int32)
        Phi_, Gam_, beta_, H_ = _d(Phi), _d(Gamma), _d(beta), _d(H_row)
        B_ = _d(B if B is not None else np.zeros(1))
        u_ = _d(u if u is not None else np.zeros(1))
        rc = lib(self.kind).ref_step(self.h, float(z), Phi_[1], Gam_[1], beta_[1], H_[1], float(gamma),
                            B_[1] if B is not None else None, u_[1] if u is not None else None,
                            mom.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip))
        c = lambda off: complex(mom[off], mom[off + 1])
        mean = mom[6:6 + 2 * d].reshape(d, 2)
        cov = mom[6 + 2 * d:].reshape(d * d, 2)
        extra = {}
        if self.kind == "oracle":
            raw = np.zeros(2)
            lib("oracle")._cdll.oracle_fz_raw(self.h, raw.ctypes.data_as(_dp))
            extra["fz_raw"] = complex(raw[0], raw[1])
        return {**extra,
            "rc": int(rc),
            "fz_public": c(0),
            "fz": c(2),
            "last_fz": c(4),
            "mean": mean[:, 0] + 1j * mean[:, 1],
            "cov": (cov[:, 0] + 1j * cov[:, 1]).reshape(d, d),
            "err": int(cnt[0]),
            "nt_mu": int(cnt[1]),
            "nt_ftr": int(cnt[2]),
            "threads": int(cnt[3]),
            "shape_mu": cnt[4:4 + sr].copy(),
            "shape_ftr": cnt[4 + sr:4 + 2 * sr].copy(),
        }

    def set_debug(self, on=True):
        """Restatement only: the reference's debug mode (public moments after reduction)."""
        lib("oracle")._cdll.oracle_set_debug(self.h, 1 if on else 0)

    def tape_record(self):
        """Restatement only: record the reduction and negligibility decisions from now on."""
        lib("oracle")._cdll.oracle_tape_mode(self.h, 1)

    def tape_replay(self, recorded):
        """Restatement only: replay `recorded`'s decisions instead of deciding."""
        lib("oracle")._cdll.oracle_tape_copy(self.h, recorded.h)

    def tape_status(self):
        """(diverged, reduction decisions used, negligibility decisions used)."""
        used = (ct.c_longlong * 2)()
        div = lib("oracle")._cdll.oracle_tape_status(self.h, used)
        return bool(div), int(used[0]), int(used[1])

    def finalize_extended(self, x_bar):
        """Extended-mode finish of the last update; returns the new x_bar."""
        xb = np.ascontiguousarray(np.asarray(x_bar, dtype=np.float64).reshape(-1).copy())
        lib(self.kind).ref_finalize_extended_moments(self.h, xb.ctypes.data_as(_dp))
        return xb

    def public_moments(self):
        d = self.d
        fz = np.zeros(2)
        x = np.zeros(d)
        P = np.zeros(d * d)
        lib(self.kind).ref_public_moments(self.h, fz.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                                 P.ctypes.data_as(_dp))
        return complex(fz[0], fz[1]), x, P.reshape(d, d)

    def terms(self, with_tables=True):
        """All terms (after an FTR step) in storage order: shape ascending, index ascending."""
        L = lib(self.kind)
        d = self.d
        out = []
        info = np.zeros(3, dtype=np.int32)
        for m in range(self.shape_range):
            i = 0
            while True:
                if L.ref_term(self.h, m, i, info.ctypes.data_as(_ip), None, None, None, None, None) != 0:
                    break
                mm, cells = int(info[0]), int(info[1])
                A = np.zeros(mm * d)
                pp = np.zeros(mm)
                b = np.zeros(d)
                keys = np.zeros(max(cells, 1), dtype=np.uint32)
                g = np.zeros(2 * max(cells, 1))
                L.ref_term(self.h, m, i, info.ctypes.data_as(_ip), A.ctypes.data_as(_dp),
                           pp.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                           keys.ctypes.data_as(ct.POINTER(ct.c_uint)) if with_tables else None,
                           g.ctypes.data_as(_dp) if with_tables else None)
                t = {"m": mm, "cells": cells, "A": A.reshape(mm, d), "p": pp, "b": b}
                if with_tables:
                    t["keys"] = keys[:cells].copy()
                    t["g"] = (g[0::2] + 1j * g[1::2])[:cells].copy()
                out.append(t)
                i += 1
        return out


def run_system(sysobj, num_steps=None, seed=None, collect_terms=False, kind="ref", tape=None):
    """Runs one estimator over a System's measurement record (num_steps time steps).
    tape (restatement only): "record", or an estimator whose recorded decisions to replay."""
    steps = sysobj.num_steps if num_steps is None else num_steps
    est = RefEstimator(steps, sysobj.n, sysobj.pncc, sysobj.p, sysobj.A0, sysobj.p0, sysobj.b0,
                       cmcc=sysobj.cmcc, seed=sysobj.srand_seed if seed is None else seed, kind=kind)
    if tape == "record":
        est.tape_record()
    elif tape is not None:
        est.tape_replay(tape)
    recs = []
    Gt = np.ascontiguousarray(sysobj.Gamma.reshape(-1))
    B = None if sysobj.B is None else np.ascontiguousarray(sysobj.B.reshape(-1))
    for k in range(steps):
        Phi = sysobj.Phi(k)
        for i in range(sysobj.p):
            r = est.step(sysobj.zs[k, i], Phi.reshape(-1), Gt, sysobj.beta, sysobj.H[i], sysobj.gamma[i],
                         B, sysobj.u(k))
            if collect_terms and r["nt_ftr"] > 0 and not (k == steps - 1 and i == sysobj.p - 1):
                r["terms"] = est.terms()
            recs.append(r)
    return est, recs


def system_records(sysobj):
    """The restatement over a System's whole record, one tuple per update: (terms after
    MU+MUC, terms after reduction, error bits, public fz, public x_hat, public P_hat).  A
    module-level function, so that process pools can run it."""
    est, recs = run_system(sysobj, kind="oracle")
    est.close()
    return [(r["nt_mu"], r["nt_ftr"], r["err"], float(r["fz_public"].real), r["mean"].real.copy(),
             r["cov"].real.reshape(-1).copy()) for r in recs]
