"""gcm_mlp_predict_llh_noalloc (the allocation-free surrogate likelihood)
through the C ABI against gcm_mlp_predict_llh on the same rows: equal bit for
bit on two calls in a row, the second with fewer rows, on the 5M-event
model's network (300 -> 512 -> 512 -> 6000, no split-K) and on a network
whose first layer takes the split-K path (1100 inputs); oversized calls and
calls before the likelihood head is set are refused.

The device buffers are allocated here through the HIP runtime the library
itself links (libamdhip64.so.7), not through torch, which bundles its own."""
import ctypes as C

import numpy as np
import pytest

from gundam_amd import _native as N
from gundam_amd.surrogate import MLP

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DeviceArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, out):
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _net(n_in, h, n_out, seed):
    rng = np.random.default_rng(seed)
    m = MLP(n_in, h, n_out, seed=seed)
    p = m.params()
    m.set_params((p + 0.05 * rng.standard_normal(p.size)).astype(np.float32))
    mean = np.log(50.0) + 0.3 * rng.standard_normal(n_out)
    scale = 0.1 + 0.02 * rng.random(n_out)
    data = np.round(np.exp(mean))
    ratio = 1.0 + 0.1 * rng.random(n_out)
    N.check(N.lib().gcm_mlp_set_head(m._h, mean.ctypes.data, scale.ctypes.data,
                                     data.ctypes.data, ratio.ctypes.data))
    return m, rng


@pytest.mark.parametrize("shape,rows", [((300, 512, 6000), (64, 17)),
                                        ((1100, 64, 96), (40, 3))])
def test_noalloc_matches_predict_llh(shape, rows):
    print(f"noalloc predict_llh shape={shape} rows={rows}", flush=True)
    lib, hip = N.lib(), _hip()
    n_in, h, n_out = shape
    m, rng = _net(n_in, h, n_out, seed=11)
    dx = DeviceArray(hip, max(rows) * n_in * 4)
    dc = DeviceArray(hip, max(rows) * 8)
    try:
        for r in rows:  # the second call has fewer rows
            x = rng.standard_normal((r, n_in)).astype(np.float32)
            want = np.empty(r)
            N.check(lib.gcm_mlp_predict_llh(m._h, x.ctypes.data, r, want.ctypes.data))
            dx.upload(x)
            N.check(lib.gcm_mlp_predict_llh_noalloc(m._h, dx.ptr, r, dc.ptr, None))
            assert hip.hipDeviceSynchronize() == 0
            got = dc.download(np.empty(r))
            assert np.all(np.isfinite(want))
            np.testing.assert_array_equal(got, want)
    finally:
        dx.free()
        dc.free()
        m.close()


def test_noalloc_refuses_oversized_and_unset_head():
    print("noalloc refusals", flush=True)
    lib, hip = N.lib(), _hip()
    m = MLP(300, 512, 6000, seed=3)
    dx = DeviceArray(hip, 4 * 300 * 4)
    dc = DeviceArray(hip, 4 * 8)
    try:
        # no head yet: refused before any launch
        assert lib.gcm_mlp_predict_llh_noalloc(m._h, dx.ptr, 4, dc.ptr, None) == N.GCM_ERR_STATE
        m2, _ = _net(300, 512, 6000, seed=5)
        # 1000 rows need ~33 MB of scratch: refused by the size check on the
        # host, with no kernel launched (the buffers passed are never read)
        rc = lib.gcm_mlp_predict_llh_noalloc(m2._h, dx.ptr, 1000, dc.ptr, None)
        assert rc == N.GCM_ERR_ARG
        assert b"scratch" in lib.gcm_last_error()
        assert lib.gcm_mlp_predict_llh_noalloc(m2._h, None, 4, dc.ptr, None) == N.GCM_ERR_ARG
        # a valid call afterwards still works
        x = np.random.default_rng(1).standard_normal((4, 300)).astype(np.float32)
        dx.upload(x)
        N.check(lib.gcm_mlp_predict_llh_noalloc(m2._h, dx.ptr, 4, dc.ptr, None))
        assert hip.hipDeviceSynchronize() == 0
        want = np.empty(4)
        N.check(lib.gcm_mlp_predict_llh(m2._h, x.ctypes.data, 4, want.ctypes.data))
        np.testing.assert_array_equal(dc.download(np.empty(4)), want)
        m2.close()
    finally:
        dx.free()
        dc.free()
        m.close()
