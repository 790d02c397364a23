"""ctypes mirror of the public C ABI in include/*.h.

The struct layouts are those of the reference headers (range.h:10-14,
grid.h:10-26, gausslet.h:10-14, nuclear_integrals.h:11-37,
eri_integrals.h:23-47, erida_integrals.h:21-45, kinetic_integrals.h:14-18) on
x86-64 SysV, so the same declarations drive both the product library and the
reference build used by the tests.
"""
import ctypes as C

glong = C.c_int64
P_double = C.POINTER(C.c_double)
P_glong = C.POINTER(glong)


class Range(C.Structure):
    _fields_ = [("istart", glong), ("num", glong)]


class Grid(C.Structure):
    _fields_ = [("coord_range", Range * 3)]

    @classmethod
    def of(cls, spec):
        """spec: ((istart, num), (istart, num), (istart, num))"""
        g = cls()
        for a in range(3):
            g.coord_range[a].istart = int(spec[a][0])
            g.coord_range[a].num = int(spec[a][1])
        return g

    def spec(self):
        return tuple((int(r.istart), int(r.num)) for r in self.coord_range)


class GridPoint(C.Union):
    _fields_ = [("c", glong * 3)]

    @classmethod
    def of(cls, p):
        q = cls()
        for a in range(3):
            q.c[a] = int(p[a])
        return q


class GaussletData(C.Structure):
    _fields_ = [("coefficients", P_double), ("indices", Range)]


class Nucleus(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("charge", C.c_int)]


class SparseEriIndices(C.Structure):
    _fields_ = [("four_indices", P_glong), ("num", glong), ("grid", Grid)]


class SparseEri(C.Structure):
    _fields_ = [("integral_values", P_double), ("four_indices", P_glong),
                ("num_entries", glong), ("grid", Grid)]


class SparseEridaIndices(C.Structure):
    _fields_ = [("two_indices", P_glong), ("num", glong), ("grid", Grid)]


class SparseErida(C.Structure):
    _fields_ = [("integral_values", P_double), ("two_indices", P_glong),
                ("num_entries", glong), ("grid", Grid)]


class KineticIntegrals(C.Structure):
    _fields_ = [("integral_values", P_double), ("grid", Grid)]


class NuclearIntegrals(C.Structure):
    _fields_ = [("integral_values", P_double), ("grid", Grid),
                ("nuclei", C.POINTER(Nucleus)), ("num_nuclei", C.c_int)]


class IMat3x3(C.Structure):
    _fields_ = [("data", (glong * 3) * 3)]


# sizes fixed by the reference ABI (x86-64)
ABI_SIZES = {
    Range: 16, Grid: 48, GridPoint: 24, GaussletData: 24, Nucleus: 32,
    SparseEriIndices: 64, SparseEri: 72, SparseEridaIndices: 64, SparseErida: 72,
    KineticIntegrals: 56, NuclearIntegrals: 72, IMat3x3: 72,
}

# name -> (restype, argtypes) of every C entry point of the reference API
SIGNATURES = {
    "compute_eri_gausslet_integral": (C.c_double, [C.POINTER(GaussletData), C.POINTER(GridPoint), C.c_double]),
    "minimum_octahedral_orbit_eri_tensor_index": (glong, [glong, C.POINTER(P_glong), glong, glong, glong, glong]),
    "enumerate_symmetry_reduced_eri_indices": (None, [C.POINTER(Grid), C.POINTER(SparseEriIndices)]),
    "delete_sparse_eri_indices": (None, [C.POINTER(SparseEriIndices)]),
    "compute_sparse_eri_gausslet_integrals": (None, [C.POINTER(GaussletData), C.POINTER(Grid), P_glong, glong, C.c_double, C.POINTER(SparseEri)]),
    "sparse_eri_gausslet_integrals_get_value": (C.c_double, [C.POINTER(SparseEri), glong]),
    "project_sparse_eri_gausslet_integrals": (None, [C.POINTER(SparseEri), P_double, glong, P_double]),
    "fill_dense_eri_tensor": (None, [C.POINTER(SparseEri), C.POINTER(Grid), P_double]),
    "apply_sparse_eri_tensor": (None, [C.POINTER(SparseEri), P_double, glong, P_double]),
    "delete_sparse_eri_gausslet_integrals": (None, [C.POINTER(SparseEri)]),
    "compute_erida_gausslet_integral": (C.c_double, [C.POINTER(GaussletData), C.POINTER(GridPoint), C.c_double]),
    "minimum_octahedral_orbit_erida_tensor_index": (glong, [glong, C.POINTER(P_glong), glong, glong]),
    "enumerate_symmetry_reduced_erida_indices": (None, [C.POINTER(Grid), C.POINTER(SparseEridaIndices)]),
    "delete_sparse_erida_indices": (None, [C.POINTER(SparseEridaIndices)]),
    "compute_sparse_erida_gausslet_integrals": (None, [C.POINTER(GaussletData), C.POINTER(Grid), P_glong, glong, C.c_double, C.POINTER(SparseErida)]),
    "sparse_erida_gausslet_integrals_get_value": (C.c_double, [C.POINTER(SparseErida), glong]),
    "project_sparse_erida_gausslet_integrals": (None, [C.POINTER(SparseErida), P_double, glong, P_double]),
    "fill_dense_erida_matrix": (None, [C.POINTER(SparseErida), C.POINTER(Grid), P_double]),
    "delete_sparse_erida_gausslet_integrals": (None, [C.POINTER(SparseErida)]),
    "compute_nuclear_gausslet_integral": (C.c_double, [C.POINTER(GaussletData), C.POINTER(GridPoint), C.POINTER(Nucleus), C.c_int, C.c_double]),
    "compute_nuclear_gausslet_integrals": (None, [C.POINTER(GaussletData), C.POINTER(Grid), C.POINTER(Nucleus), C.c_int, C.c_double, C.POINTER(NuclearIntegrals)]),
    "delete_nuclear_gausslet_integrals": (None, [C.POINTER(NuclearIntegrals)]),
    "compute_kinetic_gausslet_integral": (C.c_double, [C.POINTER(GaussletData), C.POINTER(GridPoint)]),
    "compute_kinetic_gausslet_integrals": (None, [C.POINTER(GaussletData), C.POINTER(Grid), C.POINTER(KineticIntegrals)]),
    "delete_kinetic_gausslet_integrals": (None, [C.POINTER(KineticIntegrals)]),
    "compute_cartesian_grid_3d_permutation": (None, [C.POINTER(Grid), C.POINTER(IMat3x3), P_glong]),
    "evaluate_octahedral_grid_permutations": (None, [C.POINTER(Grid), P_glong * 48]),
    "imat3x3_multiply": (None, [C.POINTER(IMat3x3), C.POINTER(IMat3x3), C.POINTER(IMat3x3)]),
    "imat3x3_equal": (C.c_bool, [C.POINTER(IMat3x3), C.POINTER(IMat3x3)]),
    "is_permutation": (C.c_bool, [P_glong, glong]),
    "is_identity_permutation": (C.c_bool, [P_glong, glong]),
    "compose_permutations": (None, [glong, P_glong, P_glong, P_glong]),
}

# exported data symbols
DATA_SYMBOLS = ("octahedral_matrep", "octahedral_multiplication_table")


def bind(lib, names=None):
    """Attach restype/argtypes for every known entry point present in 'lib'."""
    for name, (res, args) in SIGNATURES.items():
        if names is not None and name not in names:
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
