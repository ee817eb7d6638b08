"""GPU: the allocation-free gradient entry point (hdcx_rc_gradient_noalloc)
against hdcx_rc_gradient on the same network and samples: two calls in a
row, the second with fewer samples (the shared scratch keeps the first
call's contents, which must not leak into the second), and a call beyond
the scratch's 4096 samples is rejected."""

import ctypes

import numpy as np
import pytest

import hdcodex_amd as hx
from oracle import mlp

pytestmark = pytest.mark.gpu


def _target(q):
    p = q["position"].astype(np.float64)
    return (1.0 + 0.5 * np.sin(2 * np.pi * p[:, 0]) + 0.2 * q["lane"]).astype(np.float32)


class _Hip:
    """The test's own device buffers, from the HIP runtime the library uses
    (the entry point under test allocates nothing; its caller does)."""

    def __init__(self):
        # The HIP runtime libhdcx.so is linked against (one runtime per process).
        hx.library()
        path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
        self.lib = ctypes.CDLL(path)
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.buffers = []

    def upload(self, array: np.ndarray) -> int:
        a = np.ascontiguousarray(array)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        self.buffers.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.buffers:
            self.lib.hipFree(p)
        self.buffers.clear()


def test_noalloc_gradient_equals_the_allocating_entry_point():
    t = hx.PathTracer(0)
    hip = _Hip()
    try:
        t.set_radiance_cache(True, seed=11)
        params = np.random.default_rng(1).normal(scale=0.12, size=hx.RC_PARAMS).astype(np.float32)
        t.rc_set_params(params)
        for count, seed in ((4096, 2), (1000, 3)):
            q = mlp.random_queries(count, seed, _target)
            ref_loss, ref_grad = t.rc_gradient(q)
            dev_q = hip.upload(q)
            dev_g = hip.upload(np.full(hx.RC_PARAMS, np.nan, dtype=np.float32))
            dev_l = hip.upload(np.zeros(1, dtype=np.float32))
            t.rc_gradient_noalloc(dev_q, count, dev_g, dev_l)
            grad = hip.download(dev_g, np.zeros(hx.RC_PARAMS, dtype=np.float32))
            loss = float(hip.download(dev_l, np.zeros(1, dtype=np.float32))[0])
            print(count, "max |diff|", float(np.abs(grad - ref_grad).max()), "loss", loss, ref_loss)
            # The same kernels and reduction order: equal to the last bit.
            assert np.array_equal(grad, ref_grad)
            assert loss == pytest.approx(ref_loss, rel=1e-6)
        q = mlp.random_queries(4097, 4, _target)
        dev_q = hip.upload(q)
        dev_g = hip.upload(np.zeros(hx.RC_PARAMS, dtype=np.float32))
        with pytest.raises(hx.HdcxError) as e:
            t.rc_gradient_noalloc(dev_q, 4097, dev_g, 0)
        assert e.value.code == hx.ERR_INVALID_ARGUMENT
    finally:
        hip.free()
        t.close()


def test_noalloc_gradient_needs_a_network():
    t = hx.PathTracer(0)
    try:
        assert hx.library().hdcx_rc_gradient_noalloc(t._h, ctypes.c_void_p(16), 8, ctypes.c_void_p(16), None) == \
            hx.ERR_INVALID_ARGUMENT
    finally:
        t.close()
