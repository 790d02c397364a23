"""Device-memory helpers over the library's own C entry points (used by the GPU tests)."""
import ctypes as C

import numpy as np

from gausslectronitegrate_amd import _abi, _lib


class DevArray:
    def __init__(self, nbytes):
        self.L = _lib.require_gpu()
        self.ptr = C.c_void_p()
        _lib.check(self.L.gli_device_malloc(C.byref(self.ptr), max(int(nbytes), 8)), "gli_device_malloc")
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            _lib.check(d.L.gli_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0), "gli_memcpy")
        return d

    def to_numpy(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _lib.check(self.L.gli_memcpy(out.ctypes.data, self.ptr, out.nbytes, 1), "gli_memcpy")
        return out

    def free(self):
        if self.ptr:
            self.L.gli_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


def grid6(spec):
    return np.array([v for ax in spec for v in ax], dtype=np.int64)


def eri_values_device(coeffs, spec, indices, tol, method=0):
    """values of host 'indices' through gli_eri_values_device (device-resident path)."""
    L = _lib.require_gpu()
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    g = grid6(spec)
    di = DevArray.from_numpy(np.ascontiguousarray(indices, dtype=np.int64))
    dv = DevArray(8 * len(indices))
    _lib.check(L.gli_eri_values_device(c.ctypes.data_as(_abi.P_double), len(c), g.ctypes.data_as(_abi.P_glong),
                                       di.ptr, len(indices), float(tol), method, dv.ptr), "gli_eri_values_device")
    return dv.to_numpy(np.float64, len(indices))


def eri_enumerate_device(n):
    """(device array, count) of the unique four-indices of a symmetric n^3 grid."""
    L = _lib.require_gpu()
    p = C.c_void_p()
    num = C.c_int64()
    _lib.check(L.gli_eri_enumerate_device(n, C.byref(p), C.byref(num)), "gli_eri_enumerate_device")
    d = DevArray.__new__(DevArray)
    d.L, d.ptr, d.nbytes = L, p, 8 * num.value
    return d, num.value
