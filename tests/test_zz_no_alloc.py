"""phafd_hip_gradphi_noalloc (the gradient without device allocation) against
phafd_hip_gradphi, through the C ABI.

The test owns its device memory and stream, allocated through the HIP
runtime the library is bound to (hip* symbols looked up through the
library's own handle: in a process that imported torch first, that is the
runtime torch loaded), so that pointers and the stream handle belong to one
runtime.  The entry point runs
the same kernels as phafd_hip_gradphi, so the outputs must agree bit for bit.
"""
import ctypes

import numpy as np
import pytest

import phafd_amd.api as A

pytestmark = pytest.mark.gpu

_V = ctypes.c_void_p
_H2D, _D2H = 1, 2
E_ARG = -1   # PHAFD_E_ARG


def _hip():
    h = ctypes.CDLL(A.LIB_PATH)   # dlsym through it searches the library's dependencies
    h.hipMalloc.argtypes = [ctypes.POINTER(_V), ctypes.c_size_t]
    h.hipFree.argtypes = [_V]
    h.hipMemcpy.argtypes = [_V, _V, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(_V)]
    h.hipStreamDestroy.argtypes = [_V]
    h.hipStreamSynchronize.argtypes = [_V]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    return h


def _ok(rc, what):
    assert rc == 0, f"{what} returned {rc}"


class _Dev:
    """Device arrays of one shape and a stream of the library's runtime."""

    def __init__(self, hip, shape):
        self.hip, self.shape = hip, shape
        self.nbytes = int(np.prod(shape)) * 8
        self.ptrs = []
        for _ in range(4):
            p = _V()
            _ok(hip.hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
            self.ptrs.append(p)
        self.stream = _V()
        _ok(hip.hipStreamCreate(ctypes.byref(self.stream)), "hipStreamCreate")

    def put(self, i, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        _ok(self.hip.hipMemcpy(self.ptrs[i], a.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def get(self, i):
        out = np.empty(self.shape)
        _ok(self.hip.hipMemcpy(out.ctypes.data, self.ptrs[i], self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        _ok(self.hip.hipStreamDestroy(self.stream), "hipStreamDestroy")
        for p in self.ptrs:
            _ok(self.hip.hipFree(p), "hipFree")


def _noalloc(dev, period):
    nz, ny, nx = dev.shape
    pe = np.ascontiguousarray(period, dtype=np.float64)
    p = dev.ptrs
    print(f"gradphi_noalloc nz,ny,nx={dev.shape}", flush=True)
    rc = A.lib().phafd_hip_gradphi_noalloc(nx, ny, nz, pe.ctypes.data_as(A._dp), p[0], p[1], p[2], p[3],
                                           dev.stream)
    assert rc == 0, (rc, A.lib().phafd_hip_last_error(None))
    _ok(dev.hip.hipStreamSynchronize(dev.stream), "hipStreamSynchronize")
    return [dev.get(i) for i in (1, 2, 3)]


def test_noalloc_gradphi_matches_allocating_entry_point_twice():
    hip = _hip()
    _ok(hip.hipSetDevice(0), "hipSetDevice")
    # first call: 128^3, the largest cubic grid within the scratch limit;
    # second call: a smaller anisotropic grid, so that a stale spectrum or
    # twiddle view left by the first call would show
    cases = [((128, 128, 128), [128.0, 128.0, 128.0], 11), ((16, 8, 32), [24.0, 10.0, 40.0], 12)]
    for shape, period, seed in cases:
        phi = np.random.default_rng(seed).standard_normal(shape)
        dev = _Dev(hip, shape)
        try:
            dev.put(0, phi)
            nan = np.full(shape, np.nan)
            for i in (1, 2, 3):
                dev.put(i, nan)       # every output element must be written
            got = _noalloc(dev, period)
        finally:
            dev.free()
        want = A.gradphi(phi, period)
        for c, (g, w) in enumerate(zip(got, want)):
            assert np.isfinite(g).all(), (shape, c)
            assert np.array_equal(g, w), (shape, c, float(np.max(np.abs(g - w))))


def test_noalloc_gradphi_rejects_oversize_and_bad_shapes():
    hip = _hip()
    _ok(hip.hipSetDevice(0), "hipSetDevice")
    dev = _Dev(hip, (4, 4, 4))
    try:
        ones = np.ones(3)
        pe = ones.ctypes.data_as(A._dp)
        p = dev.ptrs
        f = A.lib().phafd_hip_gradphi_noalloc
        # (Nx/2+1)*Ny*Nz = 129*128*128 > 1308672: rejected before any launch
        assert f(256, 128, 128, pe, p[0], p[1], p[2], p[3], dev.stream) == E_ARG
        msg = A.lib().phafd_hip_last_error(None).decode()
        assert "1308672" in msg and "40 MiB" in msg, msg
        assert f(6, 8, 8, pe, p[0], p[1], p[2], p[3], dev.stream) == E_ARG
        assert f(8, 8, 8, pe, None, p[1], p[2], p[3], dev.stream) == E_ARG
    finally:
        dev.free()
