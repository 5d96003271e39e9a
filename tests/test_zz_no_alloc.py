"""mce_cpdf1d_noalloc, the 1D marginal density written into caller-provided device arrays
with all working memory in the library's static per-GPU array (no device allocation),
against pycauchy_get_marginal_1D_pointwise_cpdf on the same estimator: identical grids and
densities, bit for bit, on two calls in a row (the second with fewer terms and fewer
points), and clear errors beyond the output capacity or the working memory."""
import ctypes as ct

import numpy as np
import pytest

from cauchyfriendly_amd import _native, pycauchy
from cauchyfriendly_amd import workloads as W
from cauchyfriendly_amd.estimator import GpuEstimator

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the engine itself is linked against (already loaded with it)."""
    _native.lib()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln and ln.split()[-1].startswith("/")}
    eng = [p for p in paths if "/torch/" not in p]
    H = ct.CDLL(sorted(eng)[0])
    H.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    H.hipFree.argtypes = [ct.c_void_p]
    H.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    return H


class DevArray:
    def __init__(self, H, n):
        self.H, self.n, self.p = H, n, ct.c_void_p()
        assert H.hipMalloc(ct.byref(self.p), 8 * n) == 0

    def host(self, n):
        out = np.zeros(n)
        assert self.H.hipDeviceSynchronize() == 0
        assert self.H.hipMemcpy(out.ctypes.data_as(ct.c_void_p), self.p, 8 * n, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = ct.c_void_p()


def _stepped(s, updates):
    est = GpuEstimator(s.num_steps, s.n, s.pncc, s.p, s.A0, s.p0, s.b0, cmcc=s.cmcc, seed=s.srand_seed)
    Gt = np.ascontiguousarray(s.Gamma.reshape(-1))
    u = 0
    for k in range(s.num_steps):
        for i in range(s.p):
            if u == updates:
                return est
            est.step(s.zs[k, i], s.Phi(k).reshape(-1), Gt, s.beta, s.H[i], s.gamma[i])
            u += 1
    return est


def _noalloc(est, marg, lo, hi, res, xs, fs, capacity):
    n = ct.c_int(-7)
    rc = _native.lib().mce_cpdf1d_noalloc(est.h, marg, lo, hi, res, xs.p, fs.p, capacity, None, ct.byref(n))
    return rc, n.value


def test_noalloc_density_equals_the_allocating_entry_point():
    H = _hip()
    big = _stepped(W.config1(steps=6), 6)       # 5 states, 6 updates: several 1024-term chunks
    small = _stepped(W.config0(), 3)            # 3 states, 3 updates: one chunk
    xs, fs = DevArray(H, 4096), DevArray(H, 4096)
    try:
        for est, marg, lo, hi, res in ((big, 2, -3.0, 3.0, 0.0025), (small, 1, -1.5, 1.0, 0.02)):
            d = est.d
            fz, mean, cov = np.zeros(2), np.zeros(2 * d), np.zeros(2 * d * d)
            dp = ct.POINTER(ct.c_double)
            nt = _native.lib().mce_estimator_moments(est.h, fz.ctypes.data_as(dp), mean.ctypes.data_as(dp),
                                                     cov.ctypes.data_as(dp))
            print(f"case d={d} terms={nt} marg={marg} grid=[{lo}, {hi}] step {res}", flush=True)
            ref, num = pycauchy.pycauchy_get_marginal_1D_pointwise_cpdf(est.h, marg, lo, hi, res)
            ref = ref.reshape(-1, 2)
            assert num == ref.shape[0] > 100
            rc, n = _noalloc(est, marg, lo, hi, res, xs, fs, xs.n)
            assert rc == 0, _native.lib().mce_last_error().decode()
            assert n == num
            np.testing.assert_array_equal(xs.host(n), ref[:, 0])
            np.testing.assert_array_equal(fs.host(n), ref[:, 1])
            assert np.all(ref[:, 1] > 0)
    finally:
        xs.free()
        fs.free()
        big.close()
        small.close()


def test_noalloc_density_refuses_what_does_not_fit():
    H = _hip()
    est = _stepped(W.config1(steps=6), 6)
    xs, fs = DevArray(H, 100), DevArray(H, 100)
    try:
        # 2401 points into arrays of 100
        rc, n = _noalloc(est, 0, -3.0, 3.0, 0.0025, xs, fs, xs.n)
        assert rc == -1 and n == 0
        assert "output arrays hold 100" in _native.lib().mce_last_error().decode()
    finally:
        xs.free()
        fs.free()
    # a grid whose partial sums (8 B per 1024-term chunk and point) exceed the 128 MiB array
    d = est.d
    dp = ct.POINTER(ct.c_double)
    nt = _native.lib().mce_estimator_moments(est.h, np.zeros(2).ctypes.data_as(dp), np.zeros(2 * d).ctypes.data_as(dp),
                                             np.zeros(2 * d * d).ctypes.data_as(dp))
    chunks = (nt + 1023) // 1024
    lo, res = -1000.0, 0.001
    hi = lo + res * ((1 << 27) // (8 * chunks) + 1000)
    big_n = int((hi - lo + res - 1e-15) / res) + 1     # the grid's point count
    print(f"case {nt} terms, {big_n} points", flush=True)
    xs, fs = DevArray(H, big_n), DevArray(H, big_n)
    try:
        rc, n = _noalloc(est, 0, lo, hi, res, xs, fs, big_n)
        assert rc == -1 and n == 0
        assert "no-allocation form has 134217728" in _native.lib().mce_last_error().decode()
    finally:
        xs.free()
        fs.free()
        est.close()
