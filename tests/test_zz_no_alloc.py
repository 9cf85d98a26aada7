"""ptn_nesterov_mnls_noalloc (the entry point that allocates no device memory) against
ptn_nesterov_mnls on the same inputs: bitwise-equal factors, step counts and Lipschitz
constants, on two calls in a row (the second smaller, so it runs on the first call's
leftover scratch contents), and the size limit turned into an error by the host."""
import ctypes

import numpy as np
import pytest

import ptn_hip as P

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class DeviceArray:
    """A device buffer from the HIP runtime the library itself loaded (test plumbing)."""

    def __init__(self, host):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.n = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.n)) == 0
        src = np.ascontiguousarray(host)
        assert self.rt.hipMemcpy(self.ptr, src.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.n), _H2D) == 0

    def get(self, shape):
        out = np.empty(self.n // 8, dtype=np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.n), _D2H) == 0
        return out.reshape(shape, order="F")

    def free(self):
        self.rt.hipFree(self.ptr)


def _problem(I, R, seed):
    rng = np.random.default_rng(seed)
    B = rng.random((3 * R, R))
    Z = B.T @ B
    W = rng.standard_normal((I, R)) * 4 + 2
    A0 = rng.random((I, R))
    return Z, W, A0


def test_noalloc_matches_allocating_entry_point():
    with P.Handle(0) as h:
        for I, R, seed in [(1500, 32, 1), (129, 17, 2)]:          # second call smaller than the first
            print("[gpu-test] no-alloc nesterov I=%d R=%d" % (I, R), flush=True)
            Z, W, A0 = _problem(I, R, seed)
            Aref, itr, Lr, mur = h.nesterov(Z, W, A0, 1e-2, 1e-2, 50)
            bufs = [DeviceArray(np.asfortranarray(x).ravel(order="F")) for x in (Z, W, A0)]
            try:
                it, L, mu = h.nesterov_noalloc(I, R, bufs[0].ptr.value, bufs[1].ptr.value, bufs[2].ptr.value,
                                               1e-2, 1e-2, 50)
                got = bufs[2].get((I, R))
            finally:
                for b in bufs:
                    b.free()
            assert (it, L, mu) == (itr, Lr, mur)
            assert np.array_equal(got, Aref)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    with P.Handle(0) as h:
        Z, W, A0 = _problem(4, 64, 3)
        bufs = [DeviceArray(np.asfortranarray(x).ravel(order="F")) for x in (Z, W, A0)]
        try:
            # 20000 rows of rank 64 need ~140 MB of scratch; the host check fails before any launch
            with pytest.raises(P.PtnError) as e:
                h.nesterov_noalloc(20000, 64, bufs[0].ptr.value, bufs[1].ptr.value, bufs[2].ptr.value)
            assert "PTN_NOALLOC_SCRATCH_BYTES" in str(e.value)
        finally:
            for b in bufs:
                b.free()
