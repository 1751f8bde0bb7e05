"""cs_ordered_sum_noalloc: the exact ordered sum on device pointers with its
scratch in a static device array (no allocation), against the existing
entry point cs_selftest_ordered_sum (which allocates its scratch) and the
left-to-right loop, on two calls in a row, the second with a smaller input;
and the capacity error beyond the array's maximum.  Device buffers and a
stream come from the HIP runtime itself (ctypes), as a caller of the C-ABI
would hold them."""
import ctypes as C
import os

import numpy as np
import pytest

HIP_LIB = "/opt/rocm/lib/libamdhip64.so"


@pytest.fixture(scope="module")
def lib():
    from cloth_sim_self_collision_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("needs a HIP device")
    return _lib


@pytest.fixture(scope="module")
def hip():
    if not os.path.exists(HIP_LIB):
        pytest.skip("HIP runtime library not found")
    h = C.CDLL(HIP_LIB)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipStreamCreate, h.hipStreamDestroy,
              h.hipStreamSynchronize, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.size:
            assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def get(self, n):
        out = np.empty(n, np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _sequential(t):
    s = 0.0
    for x in t:
        s += float(x)
    return s


def _noalloc(lib, hip, t, stream=None):
    src = DevBuf(hip, 8 * len(t))
    out = DevBuf(hip, 8)
    try:
        src.put(t)
        out.put(np.array([np.nan]))
        st = lib.load().cs_ordered_sum_noalloc(len(t), src.p, out.p, stream)
        lib.check(st)
        assert (hip.hipStreamSynchronize(stream) if stream else hip.hipDeviceSynchronize()) == 0
        return float(out.get(1)[0])
    finally:
        src.free()
        out.free()


@pytest.mark.gpu
def test_noalloc_sum_equals_the_allocating_entry_point(lib, hip):
    from cloth_sim_self_collision_amd.sim import ordered_sum
    rng = np.random.default_rng(17)
    big = rng.standard_normal(300_000) * np.exp2(rng.integers(-40, 40, 300_000))
    small = np.round(rng.standard_normal(1500) * 8) / 4  # exact halves: rounding ties
    for t in (big, small):
        want = ordered_sum(t)  # cs_selftest_ordered_sum
        got = _noalloc(lib, hip, t)
        assert got == want == _sequential(t), (len(t), got, want)
    # on a stream of the caller's, at the largest size the array holds
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    try:
        t = rng.standard_normal(4096 * 1024)
        got = _noalloc(lib, hip, t, stream=s)
        assert got == ordered_sum(t) == _sequential(t)
    finally:
        hip.hipStreamDestroy(s)
    assert _noalloc(lib, hip, np.zeros(0)) == 0.0


@pytest.mark.gpu
def test_noalloc_rejects_sizes_beyond_its_scratch(lib, hip):
    out = DevBuf(hip, 8)
    try:
        h = lib.load()
        st = h.cs_ordered_sum_noalloc(4096 * 1024 + 1, out.p, out.p, None)
        assert st == 7  # CS_ERR_CAPACITY, nothing launched
        assert b"exceeds the maximum" in h.cs_last_error()
    finally:
        out.free()
