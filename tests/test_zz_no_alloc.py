"""gf_btdsystem_noalloc (the block-tridiagonal solve on the caller's device
buffers, its workspace in the library's fixed device scratch) against
gf_btdsystem on the same systems: two calls in a row, the second smaller, then
a singular system, then a request past the scratch capacity.

The caller's device buffers come from the HIP runtime the library itself links
(hipMalloc through ctypes): the caller allocates, the entry point does not."""
import ctypes as C
import os

import numpy as np
import pytest

from giraffe_amd import abi, runtime

pytestmark = pytest.mark.gpu


def _hip():
    runtime.load()  # the library's own libamdhip64 is then already in the process
    h = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, host):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.host.nbytes) == 0
        assert hip.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _system(m, nblk, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (nblk, m, m))
    Cb = rng.uniform(-1, 1, (nblk, m, m))
    B = rng.uniform(-1, 1, (nblk, m, m)) + 4.0 * m * np.eye(m)
    for i in range(0, nblk, 3):  # some blocks that need row interchanges
        B[i] = B[i][rng.permutation(m)]
    b = rng.uniform(-1, 1, nblk * m)
    return A, B, Cb, b


def _noalloc(A, B, Cb, b):
    L = runtime.load()
    hip = _hip()
    L.gf_btdsystem_noalloc.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 7
    nblk, m = B.shape[0], B.shape[1]
    bufs = [DevBuf(hip, v) for v in (A, B, Cb, b, np.full(nblk * m, np.nan), np.full(1, -1, np.int32))]
    st = L.gf_btdsystem_noalloc(m, nblk, *[q.p for q in bufs], None)  # default stream
    assert hip.hipDeviceSynchronize() == 0
    x, f = bufs[4].get(), int(bufs[5].get()[0])
    for q in bufs:
        q.free()
    return st, x, f


def test_noalloc_matches_btdsystem_twice():
    for m, nblk, seed in [(16, 400, 1), (6, 57, 2)]:
        print(f"noalloc m={m} nblk={nblk}", flush=True)
        A, B, Cb, b = _system(m, nblk, seed)
        x_ref, f_ref = runtime.btdsystem(A, B, Cb, b)
        st, x, f = _noalloc(A, B, Cb, b)
        assert st == abi.GF_OK
        assert f == f_ref == 0
        # same kernels in the same order: identical results
        assert np.array_equal(x, x_ref)


def test_noalloc_singular_and_capacity():
    A, B, Cb, b = _system(16, 200, 3)
    A[:] = 0.0
    Cb[:] = 0.0
    B[37] = 0.0
    x_ref, f_ref = runtime.btdsystem(A, B, Cb, b)
    st, x, f = _noalloc(A, B, Cb, b)
    assert st == abi.GF_OK and f == f_ref == 1
    # past the scratch capacity (nblk m^2 > 2^17): refused before any device work
    A, B, Cb, b = _system(32, 129, 4)
    st, x, f = _noalloc(A, B, Cb, b)
    assert st == abi.GF_E_CAPACITY and f == -1 and np.isnan(x).all()
