"""Python surface of the integral engine, mirroring the reference module
`gausslectronitegrate` (python/pymodule.c:1839-1921): the same function and
method names, argument meaning and error types, evaluated through the C ABI of
lib/libgli_amd.so (GPU path).

Grids are sequences ((istart, num), (istart, num), (istart, num)); points are
integer triples; coefficients a 1D array of odd length K.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from . import _lib
from ._lib import lib, require_gpu


# ---------------------------------------------------------------------------
# argument parsing (pymodule.c:17-297)

def _coeffs(gausslet_coeffs, syntax):
    try:
        c = np.ascontiguousarray(gausslet_coeffs, dtype=np.float64)
    except (TypeError, ValueError):
        raise ValueError("converting input argument 'gausslet_coeffs' to a NumPy array with degree 1 failed; syntax: %s" % syntax)
    if c.ndim != 1:
        raise ValueError("'gausslet_coeffs' must have degree 1, received %d; syntax: %s" % (c.ndim, syntax))
    if c.shape[0] % 2 != 1:
        raise ValueError("'gausslet_coeffs' must have an odd number of entries; syntax: %s" % syntax)
    gd = _abi.GaussletData()
    gd.coefficients = c.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(c.shape[0] - 1) // 2
    gd.indices.num = c.shape[0]
    return gd, c


def _grid(grid, syntax):
    try:
        seq = [tuple(ax) for ax in grid]
    except TypeError:
        raise ValueError("cannot interpret 'grid' as a sequence; syntax: %s" % syntax)
    if len(seq) != 3:
        raise ValueError("'grid' must be a sequence of length three (corresponding to the Cartesian axis directions); syntax: %s" % syntax)
    for i, ax in enumerate(seq):
        if len(ax) != 2:
            raise ValueError("'grid[%d]' must be a sequence of length two (start coordinate and number of grid points along direction %d); syntax: %s" % (i, i, syntax))
        if int(ax[1]) <= 0:
            raise ValueError("number of grid points along direction %d (second entry in 'grid[%d]') must be a positive integer; syntax: %s" % (i, i, syntax))
    return _abi.Grid.of(seq)


def _points(points, count, syntax):
    pts = np.asarray(points, dtype=np.int64)
    if pts.shape != (count, 3):
        raise ValueError("'grid_points' must be a sequence of %d points with three integer coordinates each; syntax: %s" % (count, syntax))
    arr = (_abi.GridPoint * count)()
    for p in range(count):
        for a in range(3):
            arr[p].c[a] = int(pts[p, a])
    return arr


def _tol(tol, syntax):
    tol = float(tol)
    if tol < 0:
        raise ValueError("'tol' cannot be negative, received %g; syntax: %s" % (tol, syntax))
    return tol


def _nuclei(positions, charges, syntax):
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    chg = np.asarray(charges).reshape(-1)
    if len(chg) != len(pos):
        raise ValueError("number of nuclear positions and charges must agree; syntax: %s" % syntax)
    arr = (_abi.Nucleus * max(len(chg), 1))()
    for k in range(len(chg)):
        for a in range(3):
            arr[k].pos[a] = float(pos[k, a])
        arr[k].charge = int(chg[k])
    return arr, len(chg)


def _grid6(g):
    """the grid as {istart_x, num_x, istart_y, num_y, istart_z, num_z}"""
    return np.array([v for a in range(3) for v in (g.coord_range[a].istart, g.coord_range[a].num)], dtype=np.int64)


def _symmetric(g, syntax):
    spec = g.spec()
    extent = max(max(abs(s), abs(s + n - 1)) for s, n in spec)
    for s, n in spec:
        if s != -extent or n != 2 * extent + 1:
            raise ValueError("specified grid is not rotationally symmetric; syntax: %s" % syntax)


# ---------------------------------------------------------------------------
# grid helpers

def grid_num_points(grid):
    g = _grid(grid, "grid_num_points(grid)")
    return int(g.coord_range[0].num * g.coord_range[1].num * g.coord_range[2].num)


def grid_point_to_linear_index(grid, point):
    syntax = "grid_point_to_linear_index(grid, point)"
    g = _grid(grid, syntax)
    p = [int(v) for v in point]
    off = []
    for a in range(3):
        o = p[a] - g.coord_range[a].istart
        if not (0 <= o < g.coord_range[a].num):
            raise ValueError("point coordinate %d out of range; syntax: %s" % (a, syntax))
        off.append(o)
    return (off[0] * g.coord_range[1].num + off[1]) * g.coord_range[2].num + off[2]


def linear_index_to_grid_point(grid, index):
    syntax = "linear_index_to_grid_point(grid, index)"
    g = _grid(grid, syntax)
    n = [g.coord_range[a].num for a in range(3)]
    if not (0 <= index < n[0] * n[1] * n[2]):
        raise ValueError("linear index out of range; syntax: %s" % syntax)
    return (g.coord_range[0].istart + index // (n[1] * n[2]),
            g.coord_range[1].istart + (index // n[2]) % n[1],
            g.coord_range[2].istart + index % n[2])


# ---------------------------------------------------------------------------
# one-body

def compute_kinetic_gausslet_integral(gausslet_coeffs, grid_points):
    syntax = "compute_kinetic_gausslet_integral(gausslet_coeffs, grid_points)"
    gd, keep = _coeffs(gausslet_coeffs, syntax)
    pts = _points(grid_points, 2, syntax)
    return float(require_gpu().compute_kinetic_gausslet_integral(C.byref(gd), pts))


def compute_kinetic_gausslet_integrals(gausslet_coeffs, grid):
    syntax = "compute_kinetic_gausslet_integrals(gausslet_coeffs, grid)"
    gd, keep = _coeffs(gausslet_coeffs, syntax)
    g = _grid(grid, syntax)
    L = require_gpu()
    N = grid_num_points(grid)
    m = np.empty((N, N))
    g6 = _grid6(g)
    _lib.check(L.gli_kinetic_matrix(gd.coefficients, gd.indices.num, g6.ctypes.data_as(_abi.P_glong),
                                    m.ctypes.data_as(_abi.P_double)), "compute_kinetic_gausslet_integrals")
    return m


def compute_nuclear_gausslet_integral(gausslet_coeffs, grid_points, nuclear_positions, nuclear_charges, tol):
    syntax = "compute_nuclear_gausslet_integral(gausslet_coeffs, grid_points, nuclear_positions, nuclear_charges, tol)"
    gd, keep = _coeffs(gausslet_coeffs, syntax)
    pts = _points(grid_points, 2, syntax)
    nuc, nn = _nuclei(nuclear_positions, nuclear_charges, syntax)
    tol = _tol(tol, syntax)
    return float(require_gpu().compute_nuclear_gausslet_integral(C.byref(gd), pts, nuc, nn, tol))


def compute_nuclear_gausslet_integrals(gausslet_coeffs, grid, nuclear_positions, nuclear_charges, tol):
    syntax = "compute_nuclear_gausslet_integrals(gausslet_coeffs, grid, nuclear_positions, nuclear_charges, tol)"
    gd, keep = _coeffs(gausslet_coeffs, syntax)
    g = _grid(grid, syntax)
    nuc, nn = _nuclei(nuclear_positions, nuclear_charges, syntax)
    tol = _tol(tol, syntax)
    L = require_gpu()
    N = grid_num_points(grid)
    m = np.empty((N, N))
    g6 = _grid6(g)
    pos = np.ascontiguousarray([[nuc[k].pos[a] for a in range(3)] for k in range(nn)], dtype=np.float64).reshape(-1)
    chg = np.ascontiguousarray([nuc[k].charge for k in range(nn)], dtype=np.int32)
    _lib.check(L.gli_nuclear_matrix(gd.coefficients, gd.indices.num, g6.ctypes.data_as(_abi.P_glong),
                                    pos.ctypes.data_as(_abi.P_double), chg.ctypes.data_as(C.POINTER(C.c_int)), nn, tol,
                                    m.ctypes.data_as(_abi.P_double)), "compute_nuclear_gausslet_integrals")
    return m


# ---------------------------------------------------------------------------
# ERI

def compute_eri_gausslet_integral(gausslet_coeffs, grid_points, tol):
    syntax = "compute_eri_gausslet_integral(gausslet_coeffs, grid_points, tol)"
    gd, keep = _coeffs(gausslet_coeffs, syntax)
    pts = _points(grid_points, 4, syntax)
    tol = _tol(tol, syntax)
    return float(require_gpu().compute_eri_gausslet_integral(C.byref(gd), pts, tol))


def _point_tuples(points, count, syntax):
    pts = np.ascontiguousarray(points, dtype=np.int64)
    if pts.ndim != 3 or pts.shape[1:] != (count, 3):
        raise ValueError("'grid_points' must have shape (n, %d, 3): n tuples of %d points with three integer "
                         "coordinates each; syntax: %s" % (count, count, syntax))
    return pts


def compute_eri_gausslet_integral_batch(gausslet_coeffs, grid_points, tol, method=None):
    """(ij|kl) for n tuples of four grid points (shape (n, 4, 3)) in one GPU evaluation: the batched
    form of compute_eri_gausslet_integral (pymodule.c:414-640).  method: one of the ERI_* constants,
    default the current library method (gli_get_eri_method).  The GIL is released during the call."""
    syntax = "compute_eri_gausslet_integral_batch(gausslet_coeffs, grid_points, tol, method=None)"
    _, c = _coeffs(gausslet_coeffs, syntax)
    pts = _point_tuples(grid_points, 4, syntax)
    tol = _tol(tol, syntax)
    L = require_gpu()
    m = L.gli_get_eri_method() if method is None else int(method)
    out = np.empty(pts.shape[0])
    _lib.check(L.gli_eri_integrals_points(c.ctypes.data_as(_abi.P_double), len(c), pts.ctypes.data_as(_abi.P_glong),
                                          pts.shape[0], tol, m, out.ctypes.data_as(_abi.P_double)),
               "compute_eri_gausslet_integral_batch")
    return out


def compute_erida_gausslet_integral_batch(gausslet_coeffs, grid_points, tol):
    """(ii|jj) for n pairs of grid points (shape (n, 2, 3)) in one GPU evaluation: the batched form
    of compute_erida_gausslet_integral.  The GIL is released during the call."""
    syntax = "compute_erida_gausslet_integral_batch(gausslet_coeffs, grid_points, tol)"
    _, c = _coeffs(gausslet_coeffs, syntax)
    pts = _point_tuples(grid_points, 2, syntax)
    tol = _tol(tol, syntax)
    L = require_gpu()
    out = np.empty(pts.shape[0])
    _lib.check(L.gli_erida_integrals_points(c.ctypes.data_as(_abi.P_double), len(c), pts.ctypes.data_as(_abi.P_glong),
                                            pts.shape[0], tol, out.ctypes.data_as(_abi.P_double)),
               "compute_erida_gausslet_integral_batch")
    return out


class _Perms:
    """O_h grid permutations of a sparse grid, kept for lookups (pymodule.c:1169)."""

    def __init__(self, g):
        L = lib()
        self.arr = (_abi.P_glong * 48)()
        L.evaluate_octahedral_grid_permutations(C.byref(g), self.arr)
        self._free = C.CDLL(None).free

    def __del__(self):
        for e in range(48):
            if self.arr[e]:
                self._free(C.cast(self.arr[e], C.c_void_p))


class SparseERI:
    """Symmetry-reduced ERI set of a centred cubic grid (pymodule.c:647-1108)."""

    def __init__(self, gausslet_coeffs, grid, tol):
        syntax = "compute_sparse_eri_gausslet_integrals(gausslet_coeffs, grid, tol)"
        self._tol = _tol(tol, syntax)
        self._gd, self._coeffs = _coeffs(gausslet_coeffs, syntax)
        g = _grid(grid, syntax)
        _symmetric(g, syntax)
        L = require_gpu()
        self._perm = _Perms(g)
        # enumerate_symmetry_reduced_eri_indices + compute_sparse_eri_gausslet_integrals in one call, the
        # index list staying on the GPU in between
        eri = _abi.SparseEri()
        _lib.check(L.gli_sparse_eri_compute(C.byref(self._gd), C.byref(g), self._tol, C.byref(eri)),
                   "compute_sparse_eri_gausslet_integrals")
        self._eri = eri

    @classmethod
    def _from_arrays(cls, eri_struct, coeffs, tol):
        """a set read from a file: the struct owns its arrays (released by delete_*)"""
        self = cls.__new__(cls)
        self._tol = float(tol)
        self._gd, self._coeffs = _coeffs(coeffs, "load_sparse_eri_gausslet_integrals(filename)")
        self._perm = _Perms(eri_struct.grid)
        self._eri = eri_struct
        return self

    def __del__(self):
        if getattr(self, "_eri", None) is not None and self._eri.integral_values:
            lib().delete_sparse_eri_gausslet_integrals(C.byref(self._eri))

    @property
    def grid(self):
        return self._eri.grid.spec()

    @property
    def tol(self):
        return self._tol

    @property
    def num_entries(self):
        return int(self._eri.num_entries)

    @property
    def gausslet_coeffs(self):
        return self._coeffs.copy()

    @property
    def four_indices(self):
        return np.ctypeslib.as_array(self._eri.four_indices, shape=(self.num_entries,)).copy()

    @property
    def integral_values(self):
        return np.ctypeslib.as_array(self._eri.integral_values, shape=(self.num_entries,)).copy()

    def integral_value(self, grid_points):
        syntax = "integral_value(grid_points)"
        pts = _points(grid_points, 4, syntax)
        key = lib().gli_eri_sparse_key(C.byref(self._eri.grid), self._perm.arr, pts)
        if key < 0:
            raise SyntaxError("grid point out of range (after box center translation); syntax: %s" % syntax)
        return float(lib().sparse_eri_gausslet_integrals_get_value(C.byref(self._eri), key))

    def fill_tensor(self, grid):
        g = _grid(grid, "fill_tensor(grid)")
        N = grid_num_points(grid)
        out = np.zeros((N, N, N, N))
        lib().fill_dense_eri_tensor(C.byref(self._eri), C.byref(g), out.ctypes.data_as(_abi.P_double))
        return out

    def project(self, basis):
        b = np.ascontiguousarray(basis, dtype=np.float64)
        N = grid_num_points(self.grid)
        if b.ndim != 2 or b.shape[0] != N:
            raise ValueError("'basis' must be a matrix with %d rows; syntax: project(basis)" % N)
        S = b.shape[1]
        out = np.zeros((S, S, S, S))
        lib().project_sparse_eri_gausslet_integrals(C.byref(self._eri), b.ctypes.data_as(_abi.P_double), S,
                                                    out.ctypes.data_as(_abi.P_double))
        return out

    def apply(self, states):
        x = np.ascontiguousarray(states, dtype=np.float64)
        N = grid_num_points(self.grid)
        if x.ndim != 2 or x.shape[0] != N * N:
            raise ValueError("'states' must be a matrix with %d rows; syntax: apply(states)" % (N * N))
        out = np.zeros_like(x)
        lib().apply_sparse_eri_tensor(C.byref(self._eri), x.ctypes.data_as(_abi.P_double), x.shape[1],
                                      out.ctypes.data_as(_abi.P_double))
        return out


def compute_sparse_eri_gausslet_integrals(gausslet_coeffs, grid, tol):
    return SparseERI(gausslet_coeffs, grid, tol)


_HDF5 = None


def _hdf5_lib():
    """lib/libgli_amd_hdf5.so: the reference's sparse ERI file layout (built when HDF5 is present)."""
    global _HDF5
    if _HDF5 is None:
        path = os.path.join(os.path.dirname(_lib.LIB_PATH), "libgli_amd_hdf5.so")
        if not os.path.exists(path):
            raise RuntimeError("libgli_amd_hdf5.so is not built (no HDF5 installation was found at build time)")
        L = C.CDLL(path)
        L.gli_eri_load_hdf5.restype = C.c_int
        L.gli_eri_load_hdf5.argtypes = [C.c_char_p, C.POINTER(_abi.SparseEri), C.POINTER(_abi.P_double),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.gli_eri_save_hdf5.restype = C.c_int
        L.gli_eri_save_hdf5.argtypes = [C.c_char_p, C.c_int64, C.c_double, _abi.P_double, C.c_int64,
                                        C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
        _HDF5 = L
    return _HDF5


def load_sparse_eri_gausslet_integrals(filename):
    """Load a sparse ERI set from an HDF5 file in the reference's layout (pymodule.c:1181-1326)."""
    syntax = "load_sparse_eri_gausslet_integrals(filename)"
    if not isinstance(filename, str):
        raise SyntaxError("error parsing input; syntax: %s" % syntax)
    L = _hdf5_lib()
    eri = _abi.SparseEri()
    pc = _abi.P_double()
    K = C.c_int64()
    tol = C.c_double()
    if L.gli_eri_load_hdf5(filename.encode(), C.byref(eri), C.byref(pc), C.byref(K), C.byref(tol)) != 0:
        raise RuntimeError("reading the sparse ERI set from HDF5 file '%s' failed; syntax: %s" % (filename, syntax))
    coeffs = np.ctypeslib.as_array(pc, shape=(K.value,)).copy()
    C.CDLL(None).free(C.cast(pc, C.c_void_p))
    return SparseERI._from_arrays(eri, coeffs, tol.value)


def save_sparse_eri_gausslet_integrals(eri, filename):
    """Write a sparse ERI set in the reference's HDF5 layout (the layout cli/gli_eri_integrals.c writes)."""
    if not isinstance(eri, SparseERI) or not isinstance(filename, str):
        raise SyntaxError("error parsing input; syntax: save_sparse_eri_gausslet_integrals(eri, filename)")
    g = eri._eri.grid.spec()
    if not (g[0] == g[1] == g[2] and g[0][0] == -(g[0][1] - 1) // 2):
        raise ValueError("the HDF5 layout stores a centred cubic grid")
    c = np.ascontiguousarray(eri._coeffs, dtype=np.float64)
    if _hdf5_lib().gli_eri_save_hdf5(filename.encode(), g[0][1], eri.tol, c.ctypes.data_as(_abi.P_double), len(c),
                                     C.cast(eri._eri.four_indices, C.c_void_p), C.cast(eri._eri.integral_values, C.c_void_p),
                                     eri.num_entries, 0) != 0:
        raise RuntimeError("writing HDF5 file '%s' failed" % filename)


# ---------------------------------------------------------------------------
# ERIDA

def compute_erida_gausslet_integral(gausslet_coeffs, grid_points, tol):
    syntax = "compute_erida_gausslet_integral(gausslet_coeffs, grid_points, tol)"
    gd, keep = _coeffs(gausslet_coeffs, syntax)
    pts = _points(grid_points, 2, syntax)
    tol = _tol(tol, syntax)
    return float(require_gpu().compute_erida_gausslet_integral(C.byref(gd), pts, tol))


class SparseERIDA:
    """Symmetry-reduced ERIDA set of a centred cubic grid (pymodule.c:1375-1752)."""

    def __init__(self, gausslet_coeffs, grid, tol):
        syntax = "compute_sparse_erida_gausslet_integrals(gausslet_coeffs, grid, tol)"
        self._tol = _tol(tol, syntax)
        self._gd, self._coeffs = _coeffs(gausslet_coeffs, syntax)
        g = _grid(grid, syntax)
        _symmetric(g, syntax)
        L = require_gpu()
        self._perm = _Perms(g)
        idx = _abi.SparseEridaIndices()
        L.enumerate_symmetry_reduced_erida_indices(C.byref(g), C.byref(idx))
        self._erida = _abi.SparseErida()
        L.compute_sparse_erida_gausslet_integrals(C.byref(self._gd), C.byref(g), idx.two_indices, idx.num,
                                                  self._tol, C.byref(self._erida))
        L.delete_sparse_erida_indices(C.byref(idx))

    def __del__(self):
        if getattr(self, "_erida", None) is not None and self._erida.integral_values:
            lib().delete_sparse_erida_gausslet_integrals(C.byref(self._erida))

    @property
    def grid(self):
        return self._erida.grid.spec()

    @property
    def tol(self):
        return self._tol

    @property
    def num_entries(self):
        return int(self._erida.num_entries)

    @property
    def gausslet_coeffs(self):
        return self._coeffs.copy()

    @property
    def two_indices(self):
        return np.ctypeslib.as_array(self._erida.two_indices, shape=(self.num_entries,)).copy()

    @property
    def integral_values(self):
        return np.ctypeslib.as_array(self._erida.integral_values, shape=(self.num_entries,)).copy()

    def integral_value(self, grid_points):
        syntax = "integral_value(grid_points)"
        pts = _points(grid_points, 2, syntax)
        key = lib().gli_erida_sparse_key(C.byref(self._erida.grid), self._perm.arr, pts)
        if key < 0:
            raise SyntaxError("grid point out of range (after box center translation); syntax: %s" % syntax)
        return float(lib().sparse_erida_gausslet_integrals_get_value(C.byref(self._erida), key))

    def fill_matrix(self, grid):
        g = _grid(grid, "fill_matrix(grid)")
        N = grid_num_points(grid)
        out = np.zeros((N, N))
        lib().fill_dense_erida_matrix(C.byref(self._erida), C.byref(g), out.ctypes.data_as(_abi.P_double))
        return out

    def project(self, basis):
        b = np.ascontiguousarray(basis, dtype=np.float64)
        N = grid_num_points(self.grid)
        if b.ndim != 2 or b.shape[0] != N:
            raise ValueError("'basis' must be a matrix with %d rows; syntax: project(basis)" % N)
        S = b.shape[1]
        out = np.zeros((S, S, S, S))
        lib().project_sparse_erida_gausslet_integrals(C.byref(self._erida), b.ctypes.data_as(_abi.P_double), S,
                                                      out.ctypes.data_as(_abi.P_double))
        return out


def compute_sparse_erida_gausslet_integrals(gausslet_coeffs, grid, tol):
    return SparseERIDA(gausslet_coeffs, grid, tol)


def get_max_openmp_threads():
    """Host threads of the library's host-side loops (the GPU path does not use them)."""
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
