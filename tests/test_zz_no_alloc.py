"""gli_eri_values_device_noalloc (the allocation-free variant of gli_eri_values_device, whose workspace
is one static scratch array per GPU) against gli_eri_values_device through the C ABI: bit-identical
values on two calls in a row, the second with a smaller input, and a third on another grid; sizes
beyond the scratch and other methods are refused with a message, and the library stays usable."""
import ctypes as C
import sys

import numpy as np
import pytest

from gausslectronitegrate_amd import _abi, _lib
from tests.conftest import golden
from tests.gpu_helpers import DevArray, eri_enumerate_device, eri_values_device

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sym6(n):
    h = (n - 1) // 2
    return np.array([-h, n, -h, n, -h, n], dtype=np.int64)


def noalloc(L, c, n, d_idx, num, tol, method=0):
    """values through the allocation-free entry point; (status, values or None)"""
    c = np.ascontiguousarray(c, dtype=np.float64)
    g = sym6(n)
    dv = DevArray(8 * max(num, 1))
    st = L.gli_eri_values_device_noalloc(c.ctypes.data_as(_abi.P_double), len(c), g.ctypes.data_as(_abi.P_glong),
                                         d_idx, num, tol, method, dv.ptr)
    return st, (dv.to_numpy(np.float64, num) if st == 0 else None)


def test_noalloc_matches_device_entry_two_calls():
    L = _lib.require_gpu()
    # call 1: the full 7^3 list (3,571,977 elements), K = 57 (82 MB of the 96 MB scratch)
    say("noalloc call 1: 7^3 K57 full list")
    c57 = golden("coeffs_K57_s0.npy")
    d7, n7 = eri_enumerate_device(7)
    st, got7 = noalloc(L, c57, 7, d7.ptr, n7, 1e-16)
    assert st == 0, L.gli_last_error()
    idx7 = d7.to_numpy(np.int64, n7)
    ref7 = eri_values_device(c57, ((-3, 7),) * 3, idx7, 1e-16, 0)
    assert np.array_equal(got7, ref7)
    # call 2, smaller: the reference build's 4096-element 5^3 sample, K = 29; the scratch still holds
    # call 1's contents
    say("noalloc call 2: 5^3 K29 sample")
    g5 = golden("golden_sample_n5_K29s0.npz")
    di5 = DevArray.from_numpy(g5["eri_indices"])
    st, got5 = noalloc(L, g5["coeffs"], 5, di5.ptr, len(g5["eri_indices"]), float(g5["tol"]))
    assert st == 0, L.gli_last_error()
    ref5 = eri_values_device(g5["coeffs"], ((-2, 5),) * 3, g5["eri_indices"], float(g5["tol"]), 0)
    assert np.array_equal(got5, ref5)
    assert np.max(np.abs(got5 - g5["eri_values"])) <= 1e-13
    # call 3: the full 3^3 set (K = 57, 1027 elements) against the reference build's values
    say("noalloc call 3: 3^3 K57")
    g3 = golden("golden_n3_K57s0.npz")
    di3 = DevArray.from_numpy(g3["eri_indices"])
    st, got3 = noalloc(L, g3["coeffs"], 3, di3.ptr, len(g3["eri_indices"]), float(g3["tol"]))
    assert st == 0, L.gli_last_error()
    assert np.array_equal(got3, eri_values_device(g3["coeffs"], ((-1, 3),) * 3, g3["eri_indices"], float(g3["tol"]), 0))
    assert np.max(np.abs(got3 - g3["eri_values"])) <= 1e-13
    # an empty list is a no-op
    st, _ = noalloc(L, c57, 5, None, 0, 1e-16)
    assert st == 0
    # refused: workspaces beyond the scratch (9^3 with K = 57 needs 214 MB; 3^3 with the 135
    # coefficients of G10 needs 234 MB: the kernel tensor grows with K), and methods other than FP64
    say("noalloc refusals")
    st, _ = noalloc(L, c57, 9, di5.ptr, 16, 1e-16)
    assert st == -1 and b"scratch" in L.gli_last_error()
    L.gli_clear_error()
    st, _ = noalloc(L, golden("ref_eri_3x3x3.npz")["coeffs"], 3, di3.ptr, 16, 1e-16)
    assert st == -1 and b"scratch" in L.gli_last_error()
    L.gli_clear_error()
    st, _ = noalloc(L, c57, 5, di5.ptr, 16, 1e-16, method=1)
    assert st == -1 and b"FP64" in L.gli_last_error()
    L.gli_clear_error()
    # still usable, and still bit-identical
    st, again = noalloc(L, g5["coeffs"], 5, di5.ptr, len(g5["eri_indices"]), float(g5["tol"]))
    assert st == 0 and np.array_equal(again, ref5)
