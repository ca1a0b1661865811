"""sbdg_sigma_device, the sigma entry point that allocates nothing (its scratch is a static
device array per device), through the C-ABI against sbdg_sigma on the same inputs.

Both run the same kernels on the same values, so the results must be bit-identical, in
FP64 and FP32.  Two calls in a row share the device's scratch array, the second on a
smaller problem (which must not see the first call's contents); a problem beyond the
array's size is rejected with SBDG_ERR_LIMIT and an unbuilt context with SBDG_ERR_STATE.
"""
import ctypes as C

import numpy as np
import pytest

import sbdg
from helpers import workload_small
from sbdg import synthetic as syn

pytestmark = pytest.mark.gpu


def _solver(case):
    return sbdg.Solver(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"])


class DeviceBuffer:
    """n doubles of device memory from the HIP runtime the library itself loaded (torch
    ships another runtime of the same soname, which must not be initialised beside it)."""

    H2D, D2H = 1, 2

    def __init__(self, n):
        sbdg.lib()  # the library first: its runtime is the one dlopen finds by soname
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.n = int(n)
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), 8 * self.n) == 0

    @property
    def ptr(self):
        return self.p.value

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, self.H2D) == 0

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, 8 * self.n, self.D2H) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _device_sigma(s, x, precision):
    xd, yd = DeviceBuffer(s.ndet), DeviceBuffer(s.ndet)
    try:
        xd.put(x)
        yd.put(np.full(s.ndet, np.nan))  # nothing of y survives but the result
        s.sigma_device(xd.ptr, yd.ptr, precision)
        return yd.get()
    finally:
        xd.free()
        yd.free()


def test_sigma_device_equals_sigma_two_calls():
    """A 120 x 8,000 problem at norb=36 (960,000 determinants; the alpha-beta kernel takes its
    ket columns in two chunks), then a 400 x 300 one at norb=26 (120,000): each call
    bit-identical to sbdg_sigma, FP64 and FP32."""
    big = workload_small(36, 9, 9, 120, 8000, 29)
    small = workload_small(26, 5, 5, 400, 300, 31)
    rng = np.random.default_rng(17)
    for case in (big, small):
        s = _solver(case)
        s.build()
        x = rng.normal(size=s.ndet)
        for prec in (0, 1):
            want = s.sigma(x, precision=prec)
            got = _device_sigma(s, x, prec)
            print(f"{s.na} x {s.nb}, precision {prec}: max |diff| {np.abs(got - want).max():.3e}", flush=True)
            assert np.array_equal(got, want), (s.na, s.nb, prec)
        s.close()


def test_sigma_device_limits():
    """norb=16, 5+5 full space (19.1M determinants) is beyond the scratch in both precisions;
    an unbuilt context is refused; the error leaves the context usable."""
    wl = syn.config_workload("c3_L16_5e5e_full")
    s = sbdg.Solver(wl.norb, wl.core, wl.h1, wl.eri, wl.astr, wl.bstr)
    x, y = DeviceBuffer(s.ndet), DeviceBuffer(s.ndet)
    try:
        with pytest.raises(sbdg.SbdgError) as e:
            s.sigma_device(x.ptr, y.ptr, 0)
        assert e.value.code == sbdg.ERR_STATE
        s.build()
        for prec in (0, 1):
            with pytest.raises(sbdg.SbdgError) as e:
                s.sigma_device(x.ptr, y.ptr, prec)
            assert e.value.code == sbdg.ERR_LIMIT and "scratch" in str(e.value)
        with pytest.raises(sbdg.SbdgError) as e:
            s.sigma_device(x.ptr, x.ptr, 0)
        assert e.value.code == sbdg.ERR_ARG
    finally:
        x.free()
        y.free()
    s.close()
