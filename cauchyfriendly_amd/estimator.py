"""`GpuEstimator`: one MCE on the HIP engine, driven one scalar update at a time
through the CauchyEstimator-level C ABI (mce_estimator_*), with the per-update
diagnostics the parity tests compare (pre-reduction moments, term counts per
shape after MU/MUC and after reduction)."""
from __future__ import annotations

import ctypes as ct
import threading

import numpy as np

from . import _native

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)

# The engine draws its construction-time random values (the root point and the first
# perturbation vector) from glibc's rand(), whose state is shared by every thread of the
# process.  Seeding and creating under one lock keeps each estimator's draws its own seed's
# when estimators are created from several threads at once.
_CREATE_LOCK = threading.Lock()


def _d(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    return a, a.ctypes.data_as(_dp)


class GpuEstimator:
    def __init__(self, num_steps, d, pncc, p, A0, p0, b0, cmcc=0, seed=None):
        L = _native.lib()
        self.d, self.pncc, self.p, self.cmcc = d, pncc, p, cmcc
        a0, p0_, b0_ = _d(A0), _d(p0), _d(b0)
        with _CREATE_LOCK:
            if seed is not None:
                _native.libc().srand(ct.c_uint(seed))
            self.h = L.mce_estimator_create(a0[1], p0_[1], b0_[1], int(num_steps), d, cmcc, pncc, p, False)
        if not self.h:
            raise RuntimeError("mce_estimator_create failed: " + L.mce_last_error().decode())
        self.shape_range = L.mce_shape_range(self.h)

    def close(self):
        if self.h:
            _native.lib().mce_estimator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_random_draws(self, root_point, b_pert):
        r, b = _d(root_point), _d(b_pert)
        _native.lib().mce_set_random_draws(self.h, r[1], b[1])

    def random_draws(self):
        r = np.zeros(self.d)
        b = np.zeros(8 * (self.shape_range - 1))
        _native.lib().mce_get_random_draws(self.h, r.ctypes.data_as(_dp), b.ctypes.data_as(_dp))
        return r, b.reshape(8, -1)

    def enable_tangents(self, num_directions):
        """Tangent mode: every later step also carries num_directions derivative directions
        (before the window's first update)."""
        L = _native.lib()
        if L.mce_estimator_enable_tangents(self.h, int(num_directions)) != 0:
            raise RuntimeError("mce_estimator_enable_tangents failed: " + L.mce_last_error().decode())
        self.ntan = int(num_directions)

    def step(self, z, Phi, Gamma, beta, H_row, gamma, B=None, u=None, tangents=None):
        """One scalar update.  In tangent mode, tangents = (dPhi (P, d, d), dbeta (P, pncc),
        dH (P, d), dgamma (P,)), the update's seed directions, and the record gains "dfz":
        d fz / d theta (complex, P) of the pre-reduction fz.  "fz" is the update's fz after the
        numeric checks (the previous one when both its mean and covariance fail them), "fz_raw"
        the update's own."""
        L = _native.lib()
        Phi_, Gam_, beta_, H_ = _d(Phi), _d(Gamma), _d(beta), _d(H_row)
        B_ = _d(B) if B is not None else (None, None)
        u_ = _d(u) if u is not None else (None, None)
        dfz = None
        if tangents is None:
            rc = L.mce_estimator_step(self.h, float(z), Phi_[1], Gam_[1], beta_[1], H_[1], float(gamma), B_[1], u_[1])
        else:
            dP, db, dH, dg = (_d(a) for a in tangents)
            out = np.zeros(2 * self.ntan)
            rc = L.mce_estimator_step_tangent(self.h, float(z), Phi_[1], Gam_[1], beta_[1], H_[1], float(gamma), B_[1],
                                              u_[1], dP[1], db[1], dH[1], dg[1], out.ctypes.data_as(_dp))
            dfz = out[0::2] + 1j * out[1::2]
        if rc == -2:
            raise RuntimeError("mce_estimator_step failed: " + L.mce_last_error().decode())
        d, sr = self.d, self.shape_range
        mom = np.zeros(6 + 2 * d + 2 * d * d)
        cnt = np.zeros(4 + 2 * sr, dtype=np.int32)
        L.mce_last_update(self.h, mom.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip))
        c = lambda off: complex(mom[off], mom[off + 1])
        mean = mom[6:6 + 2 * d].reshape(d, 2)
        cov = mom[6 + 2 * d:].reshape(d * d, 2)
        raw = np.zeros(2)
        L.mce_last_fz_raw(self.h, raw.ctypes.data_as(_dp))
        return {
            "rc": int(rc),
            "fz_public": c(0),
            "fz": c(2),
            "fz_raw": complex(raw[0], raw[1]),
            "last_fz": c(4),
            "mean": mean[:, 0] + 1j * mean[:, 1],
            "cov": (cov[:, 0] + 1j * cov[:, 1]).reshape(d, d),
            "err": int(cnt[0]),
            "nt_mu": int(cnt[1]),
            "nt_ftr": int(cnt[2]),
            "threads": int(cnt[3]),
            "shape_mu": cnt[4:4 + sr].copy(),
            "shape_ftr": cnt[4 + sr:4 + 2 * sr].copy(),
            "dfz": dfz,
        }

    def terms(self, with_tables=True):
        L = _native.lib()
        d = self.d
        out = []
        info = np.zeros(3, dtype=np.int32)
        for m in range(self.shape_range):
            i = 0
            while L.mce_term(self.h, m, i, info.ctypes.data_as(_ip), None, None, None, None, None) == 0:
                mm, cells = int(info[0]), int(info[1])
                A = np.zeros(mm * d)
                pp = np.zeros(mm)
                b = np.zeros(d)
                keys = np.zeros(max(cells, 1), dtype=np.uint32)
                g = np.zeros(2 * max(cells, 1))
                L.mce_term(self.h, m, i, info.ctypes.data_as(_ip), A.ctypes.data_as(_dp), pp.ctypes.data_as(_dp),
                           b.ctypes.data_as(_dp), keys.ctypes.data_as(ct.POINTER(ct.c_uint)) if with_tables else None,
                           g.ctypes.data_as(_dp) if with_tables else None)
                t = {"m": mm, "cells": cells, "A": A.reshape(mm, d), "p": pp, "b": b}
                if with_tables:
                    t["keys"] = keys[:cells].copy()
                    t["g"] = (g[0::2] + 1j * g[1::2])[:cells].copy()
                out.append(t)
                i += 1
        return out

    def profile(self, enable=None, reset=False):
        L = _native.lib()
        if enable is not None:
            L.mce_profile_enable(self.h, int(bool(enable)))
        if reset:
            L.mce_profile_reset(self.h)
        out = np.zeros(18)
        L.mce_profile_read(self.h, out.ctypes.data_as(_dp))
        return out

    def reset(self):
        _native.lib().mce_estimator_reset(self.h)


def run_system(sysobj, num_steps=None, seed=None, collect_terms=False):
    """Runs one GPU estimator over a System's measurement record (num_steps time steps)."""
    steps = sysobj.num_steps if num_steps is None else num_steps
    est = GpuEstimator(steps, sysobj.n, sysobj.pncc, sysobj.p, sysobj.A0, sysobj.p0, sysobj.b0,
                       cmcc=sysobj.cmcc, seed=sysobj.srand_seed if seed is None else seed)
    recs = []
    Gt = np.ascontiguousarray(sysobj.Gamma.reshape(-1))
    B = None if sysobj.B is None else np.ascontiguousarray(sysobj.B.reshape(-1))
    for k in range(steps):
        Phi = sysobj.Phi(k)
        for i in range(sysobj.p):
            r = est.step(sysobj.zs[k, i], Phi.reshape(-1), Gt, sysobj.beta, sysobj.H[i], sysobj.gamma[i],
                         B, sysobj.u(k))
            recs.append(r)
    return est, recs
