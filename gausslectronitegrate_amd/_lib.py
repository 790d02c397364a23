"""Loader of the C-ABI library lib/libgli_amd.so.

The library is built in-tree by gausslectronitegrate_amd.build; nothing here
falls back to another implementation: a missing library or a missing GPU is
an error.
"""
import ctypes as C
import os

from . import _abi

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libgli_amd.so")

_lib = None

# additive entry points of include/gli_amd.h: name -> (restype, argtypes)
EXTRA = {
    "gli_set_eri_method": (C.c_int, [C.c_int]),
    "gli_get_eri_method": (C.c_int, []),
    "gli_last_error": (C.c_char_p, []),
    "gli_last_status": (C.c_int, []),
    "gli_clear_error": (None, []),
    "gli_stream": (C.c_void_p, []),
    "gli_synchronize": (C.c_int, []),
    "gli_enable_phase_timing": (C.c_int, [C.c_int]),
    "gli_set_tuning": (C.c_int, [C.c_char_p, C.c_int]),
    "gli_get_tuning": (C.c_int, [C.c_char_p]),
    "gli_phase_times": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_char_p), C.c_int]),
    "gli_enable_direct_flop_count": (C.c_int, [C.c_int]),
    "gli_eri_integrals_points": (C.c_int, [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64), C.c_int64,
                                           C.c_double, C.c_int, C.POINTER(C.c_double)]),
    "gli_erida_integrals_points": (C.c_int, [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64), C.c_int64,
                                             C.c_double, C.POINTER(C.c_double)]),
    "gli_direct_flops": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gli_sparse_eri_compute": (C.c_int, [C.POINTER(_abi.GaussletData), C.POINTER(_abi.Grid), C.c_double,
                                         C.POINTER(_abi.SparseEri)]),
    "gli_kinetic_matrix": (C.c_int, [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "gli_nuclear_matrix": (C.c_int, [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int), C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "gli_device_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64]),
    "gli_device_free": (None, [C.c_void_p]),
    "gli_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
    "gli_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "gli_eri_enumerate_device": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "gli_erida_enumerate_device": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "gli_eri_values_device": (C.c_int, [_abi.P_double, C.c_int64, _abi.P_glong, C.c_void_p, C.c_int64,
                                        C.c_double, C.c_int, C.c_void_p]),
    "gli_erida_values_device": (C.c_int, [_abi.P_double, C.c_int64, _abi.P_glong, C.c_void_p, C.c_int64,
                                          C.c_double, C.c_void_p]),
    "gli_eri_values_device_noalloc": (C.c_int, [_abi.P_double, C.c_int64, _abi.P_glong, C.c_void_p, C.c_int64,
                                                C.c_double, C.c_int, C.c_void_p]),
    "gli_eri_values_batch_device": (C.c_int, [C.c_int, C.POINTER(_abi.P_double), _abi.P_glong, _abi.P_glong,
                                              C.POINTER(C.c_void_p), _abi.P_glong, C.c_double, C.c_int,
                                              C.POINTER(C.c_void_p)]),
    "gli_eri_batch_create": (C.c_int, [C.c_int, C.POINTER(_abi.P_double), _abi.P_glong, _abi.P_glong,
                                        C.POINTER(C.c_void_p), _abi.P_glong, C.c_double, C.c_int,
                                        C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "gli_eri_batch_run": (C.c_int, [C.c_void_p, C.POINTER(_abi.P_double), C.POINTER(C.c_double)]),
    "gli_eri_batch_describe": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "gli_eri_batch_destroy": (None, [C.c_void_p]),
    "gli_eri_key_count": (C.c_int, [C.c_int64, C.c_int]),
    "gli_eri_separable_rule": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
    "gli_eri_sparse_key": (C.c_int64, [C.POINTER(_abi.Grid), _abi.P_glong * 48, C.POINTER(_abi.GridPoint)]),
    "gli_erida_sparse_key": (C.c_int64, [C.POINTER(_abi.Grid), _abi.P_glong * 48, C.POINTER(_abi.GridPoint)]),
    "gli_release_workspace": (None, []),
    "gli_mode_product_split_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                                C.c_int64, C.c_void_p, C.c_int64]),
    "gli_mode_product_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                          C.c_int64, C.c_void_p, C.c_int64]),
}

ERI_TENSOR_FP64 = 0
ERI_TENSOR_MIXED = 1
ERI_DIRECT_FP64 = 2
ERI_SEPARABLE_FP64 = 6


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgli_amd.so is not built: run `python -m gausslectronitegrate_amd.build`")
        L = C.CDLL(LIB_PATH)
        _abi.bind(L)
        for name, (res, args) in EXTRA.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def require_gpu():
    """Initialise the device context; raise if no GPU is usable."""
    L = lib()
    if not L.gli_stream():
        raise RuntimeError("no usable HIP device: %s" % L.gli_last_error().decode())
    return L


def check(status, what):
    if status != 0:
        raise RuntimeError("%s failed: %s" % (what, lib().gli_last_error().decode()))
