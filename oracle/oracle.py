"""ctypes binding of the oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libgli_ref.so")

_i64 = C.c_int64
_pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_pl = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_pi = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_octahedral_matrices.argtypes = [_pl]
        L.orc_octahedral_perms.argtypes = [_pl, _pl]
        L.orc_factors.argtypes = [_pd, _i64, _i64, _i64, C.c_double, _pd, _pl, _pl]
        L.orc_eri_products.argtypes = [_pd, _i64, _i64, _i64, C.c_double, _pd, _pl, _pl]
        L.orc_eri_coulomb_table.argtypes = [_i64, _pd]
        L.orc_erida_coulomb_table.argtypes = [_i64, _pd]
        L.orc_eri_table_len.argtypes = [_i64, _i64]
        L.orc_eri_table_len.restype = _i64
        L.orc_erida_table_len.argtypes = [_i64, _i64]
        L.orc_erida_table_len.restype = _i64
        L.orc_eri_sparse.argtypes = [_pd, _i64, _pl, _pl, _i64, C.c_double, _pd]
        L.orc_eri_flops.argtypes = [_pd, _i64, _pl, _pl, _i64, C.c_double, _pd]
        L.orc_eri_single.argtypes = [_pd, _i64, _pl, C.c_double]
        L.orc_eri_single.restype = C.c_double
        L.orc_min_orbit_eri.argtypes = [_i64, _pl, _i64, _i64, _i64, _i64]
        L.orc_min_orbit_eri.restype = _i64
        L.orc_eri_enumerate.argtypes = [_pl, C.c_void_p]
        L.orc_eri_enumerate.restype = _i64
        L.orc_erida_products.argtypes = [_pd, _i64, C.c_double, _pd, C.POINTER(_i64), C.POINTER(_i64)]
        L.orc_erida_sparse.argtypes = [_pd, _i64, _pl, _pl, _i64, C.c_double, _pd]
        L.orc_erida_single.argtypes = [_pd, _i64, _pl, C.c_double]
        L.orc_erida_single.restype = C.c_double
        L.orc_min_orbit_erida.argtypes = [_i64, _pl, _i64, _i64]
        L.orc_min_orbit_erida.restype = _i64
        L.orc_erida_enumerate.argtypes = [_pl, C.c_void_p]
        L.orc_erida_enumerate.restype = _i64
        L.orc_nuclear_matrix.argtypes = [_pd, _i64, _pl, _pd, _pi, C.c_int, C.c_double, _pd]
        L.orc_nuclear_single.argtypes = [_pd, _i64, _pl, _pd, _pi, C.c_int, C.c_double]
        L.orc_nuclear_single.restype = C.c_double
        L.orc_kinetic_1d.argtypes = [_pd, _i64, _i64, _pd]
        L.orc_kinetic_matrix.argtypes = [_pd, _i64, _pl, _pd]
        L.orc_lookup.argtypes = [_pl, _pd, _i64, _i64]
        L.orc_lookup.restype = C.c_double
        L.orc_fill_dense_eri.argtypes = [_pl, _pl, _pd, _i64, _pl, _pd]
        L.orc_fill_dense_erida.argtypes = [_pl, _pl, _pd, _i64, _pl, _pd]
        L.orc_num_threads.restype = C.c_int
        L.orc_eri_canonical_keys.argtypes = [_pl, _pl, _i64, _pl]
        _lib = L
    return _lib


def _g(grid):
    return np.array([v for ax in grid for v in ax], dtype=np.int64)


def _c(c):
    return np.ascontiguousarray(c, dtype=np.float64)


def sym_grid(n):
    h = (n - 1) // 2
    return ((-h, n), (-h, n), (-h, n))


def num_points(grid):
    return int(grid[0][1] * grid[1][1] * grid[2][1])


def octahedral_matrices():
    out = np.zeros(48 * 9, dtype=np.int64)
    lib().orc_octahedral_matrices(out)
    return out.reshape(48, 3, 3)


def octahedral_perms(grid):
    N = num_points(grid)
    out = np.zeros(48 * N, dtype=np.int64)
    lib().orc_octahedral_perms(_g(grid), out)
    return out.reshape(48, N)


def factors(c, s0, ns, tol):
    c = _c(c)
    K = len(c)
    fac = np.zeros(ns * (2 * K - 1))
    st = np.zeros(ns, dtype=np.int64)
    ln = np.zeros(ns, dtype=np.int64)
    lib().orc_factors(c, K, s0, ns, tol, fac, st, ln)
    return fac.reshape(ns, 2 * K - 1), st, ln


def eri_products(c, s0, ns, tol):
    c = _c(c)
    K = len(c)
    LP = 4 * K - 3
    prod = np.zeros(ns * ns * LP)
    st = np.zeros(ns * ns, dtype=np.int64)
    ln = np.zeros(ns * ns, dtype=np.int64)
    lib().orc_eri_products(c, K, s0, ns, tol, prod, st, ln)
    return prod.reshape(ns * ns, LP), st, ln


def eri_table(n_entries):
    T = np.zeros(n_entries)
    lib().orc_eri_coulomb_table(n_entries, T)
    return T


def eri_table_len(max_range, K):
    return int(lib().orc_eri_table_len(max_range, K))


def eri_sparse(c, grid, indices, tol):
    c = _c(c)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    out = np.zeros(len(idx))
    lib().orc_eri_sparse(c, len(c), _g(grid), idx, len(idx), tol, out)
    return out


def eri_flops(c, grid, indices, tol):
    """per-element flops of the reference's binned loops, 2 n_x n_y + 2 R_xy (n_z + 1)"""
    c = _c(c)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    out = np.zeros(len(idx))
    lib().orc_eri_flops(c, len(c), _g(grid), idx, len(idx), tol, out)
    return out


def eri_single(c, points, tol):
    c = _c(c)
    return lib().orc_eri_single(c, len(c), np.ascontiguousarray(points, dtype=np.int64).reshape(12), tol)


def min_orbit_eri(grid, perm, i, j, k, l):
    return int(lib().orc_min_orbit_eri(num_points(grid), np.ascontiguousarray(perm).reshape(-1), i, j, k, l))


def eri_enumerate(grid):
    g = _g(grid)
    n = int(lib().orc_eri_enumerate(g, None))
    out = np.zeros(n, dtype=np.int64)
    lib().orc_eri_enumerate(g, out.ctypes.data)
    return out


def erida_sparse(c, grid, indices, tol):
    c = _c(c)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    out = np.zeros(len(idx))
    lib().orc_erida_sparse(c, len(c), _g(grid), idx, len(idx), tol, out)
    return out


def erida_single(c, points, tol):
    c = _c(c)
    return lib().orc_erida_single(c, len(c), np.ascontiguousarray(points, dtype=np.int64).reshape(6), tol)


def erida_enumerate(grid):
    g = _g(grid)
    n = int(lib().orc_erida_enumerate(g, None))
    out = np.zeros(n, dtype=np.int64)
    lib().orc_erida_enumerate(g, out.ctypes.data)
    return out


def min_orbit_erida(grid, perm, i, j):
    return int(lib().orc_min_orbit_erida(num_points(grid), np.ascontiguousarray(perm).reshape(-1), i, j))


def nuclear_matrix(c, grid, pos, charges, tol):
    c = _c(c)
    N = num_points(grid)
    out = np.zeros(N * N)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    chg = np.ascontiguousarray(charges, dtype=np.int32)
    lib().orc_nuclear_matrix(c, len(c), _g(grid), pos, chg, len(chg), tol, out)
    return out.reshape(N, N)


def nuclear_single(c, points, pos, charges, tol):
    c = _c(c)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    chg = np.ascontiguousarray(charges, dtype=np.int32)
    return lib().orc_nuclear_single(c, len(c), np.ascontiguousarray(points, dtype=np.int64).reshape(6), pos, chg, len(chg), tol)


def kinetic_1d(c, nshift):
    c = _c(c)
    out = np.zeros(nshift)
    lib().orc_kinetic_1d(c, len(c), nshift, out)
    return out


def kinetic_matrix(c, grid):
    c = _c(c)
    N = num_points(grid)
    out = np.zeros(N * N)
    lib().orc_kinetic_matrix(c, len(c), _g(grid), out)
    return out.reshape(N, N)


def lookup(keys, vals, index):
    return lib().orc_lookup(np.ascontiguousarray(keys, dtype=np.int64), np.ascontiguousarray(vals, dtype=np.float64), len(keys), index)


def fill_dense_eri(grid_sparse, keys, vals, grid_dense):
    Nd = num_points(grid_dense)
    out = np.zeros(Nd ** 4)
    lib().orc_fill_dense_eri(_g(grid_sparse), np.ascontiguousarray(keys, dtype=np.int64),
                             np.ascontiguousarray(vals, dtype=np.float64), len(keys), _g(grid_dense), out)
    return out.reshape(Nd, Nd, Nd, Nd)


def fill_dense_erida(grid_sparse, keys, vals, grid_dense):
    Nd = num_points(grid_dense)
    out = np.zeros(Nd ** 2)
    lib().orc_fill_dense_erida(_g(grid_sparse), np.ascontiguousarray(keys, dtype=np.int64),
                               np.ascontiguousarray(vals, dtype=np.float64), len(keys), _g(grid_dense), out)
    return out.reshape(Nd, Nd)


def num_threads():
    return int(lib().orc_num_threads())


def eri_canonical_keys(grid, indices):
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    out = np.zeros(len(idx), dtype=np.int64)
    lib().orc_eri_canonical_keys(_g(grid), idx, len(idx), out)
    return out
