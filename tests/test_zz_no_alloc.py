"""sisi_gemm_device_noalloc (the GEMM entry point that allocates nothing)
against sisi_gemm_host on the same inputs: bit-identical results for two
calls in a row, the second smaller, and a clear error when a requested
split-K plan does not fit the scratch array.

Device buffers come from the HIP runtime the engine library itself links
(hipMalloc / hipMemcpy through ctypes), so the test shares one runtime with
the engine in whatever order it runs."""

import ctypes

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _hip():
    from sisi4s_amd import engine
    engine.load()  # the engine library first: its rpath names the runtime below
    h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemcpy.restype = ctypes.c_int
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    h.hipDeviceSynchronize.argtypes = []
    h.hipDeviceSynchronize.restype = ctypes.c_int
    return h


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Dev:
    """A device copy of a float64 array (column-major), freed on close."""

    def __init__(self, hip, x):
        self.hip = hip
        self.host = np.ascontiguousarray(np.asfortranarray(x).ravel(order="F"), dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def close(self):
        self.hip.hipFree(self.ptr)


def test_noalloc_gemm_matches_gemm_host_on_two_calls():
    from sisi4s_amd import engine
    hip = _hip()
    lib = engine.load()
    f = lib.sisi_gemm_device_noalloc
    rng = np.random.default_rng(11)
    # first call: few output tiles and a long k (split-K partial sums in the
    # scratch array); then smaller products
    for prec, ta, tb, M, N, K in [(0, 1, 0, 300, 200, 40000), (0, 0, 1, 64, 48, 3000),
                                  (1, 0, 0, 130, 129, 700)]:
        A = rng.standard_normal((K, M) if ta else (M, K))
        B = rng.standard_normal((N, K) if tb else (K, N))
        C0 = rng.standard_normal((M, N))
        alpha, beta = 0.75, -0.5
        announce("noalloc gemm prec=%d ta=%d tb=%d M=%d N=%d K=%d" % (prec, ta, tb, M, N, K))
        ref = engine.gemm(A, B, C0, alpha=alpha, beta=beta, transA=ta, transB=tb, precision=prec)
        a, b, c = Dev(hip, A), Dev(hip, B), Dev(hip, C0)
        try:
            rc = f(prec, ta, tb, M, N, K, alpha, a.ptr.value, A.shape[0], b.ptr.value, B.shape[0],
                   beta, c.ptr.value, M, 0, None)
            assert rc == 0, lib.sisi_last_error()
            assert hip.hipDeviceSynchronize() == 0
            got = c.read().reshape((M, N), order="F")
        finally:
            for d in (a, b, c):
                d.close()
        assert np.array_equal(got, ref), float(np.abs(got - ref).max())


def test_noalloc_gemm_rejects_a_split_beyond_the_scratch():
    from sisi4s_amd import engine
    hip = _hip()
    lib = engine.load()
    M, N, K = 2048, 2048, 4096
    a, b, c = Dev(hip, np.zeros((M, K))), Dev(hip, np.zeros((K, N))), Dev(hip, np.zeros((M, N)))
    announce("noalloc gemm split beyond the scratch")
    try:
        rc = lib.sisi_gemm_device_noalloc(0, 0, 0, M, N, K, 1.0, a.ptr.value, M, b.ptr.value, K,
                                          0.0, c.ptr.value, M, 16, None)
        assert rc != 0
        assert b"exceed the workspace" in lib.sisi_last_error()
        assert hip.hipDeviceSynchronize() == 0
    finally:
        for d in (a, b, c):
            d.close()
