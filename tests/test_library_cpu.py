"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared in include/*.h, keeps the reference struct layouts and its group
tables agree with the O_h axioms.  No GPU computation is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from gausslectronitegrate_amd import _abi, _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HDF5_HEADERS = ("hdf5_util.h", "gli_hdf5.h")  # exported by lib/libgli_amd_hdf5.so


def declared_functions(headers=None):
    names = set()
    for h in os.listdir(os.path.join(REPO, "include")):
        if (h in HDF5_HEADERS) != (headers == "hdf5"):
            continue
        text = open(os.path.join(REPO, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "for", "while", "return", "sizeof"):
                names.add(name)
        for m in re.finditer(r"^extern const [^;]*?\b(\w+)\s*\[", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert "compute_sparse_eri_gausslet_integrals" in names and "gli_eri_values_device" in names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert missing == []


def test_library_exports_nothing_else():
    """kernels, C++ helpers and the private host/device interface stay local to the library"""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert exported - declared_functions() == set()


def test_hdf5_library_exports_every_declared_symbol():
    path = os.path.join(os.path.dirname(_lib.LIB_PATH), "libgli_amd_hdf5.so")
    L = ctypes.CDLL(path)
    names = declared_functions("hdf5")
    assert {"read_hdf5_dataset", "write_hdf5_vector_attribute", "gli_eri_load_hdf5", "gli_eri_save_hdf5"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert missing == []


def test_abi_struct_sizes():
    for cls, size in _abi.ABI_SIZES.items():
        assert ctypes.sizeof(cls) == size, cls.__name__


def test_exported_group_tables():
    L = _lib.lib()
    M = (_abi.IMat3x3 * 48).in_dll(L, "octahedral_matrep")
    T = ((ctypes.c_int * 48) * 48).in_dll(L, "octahedral_multiplication_table")
    mats = np.array([[[M[e].data[r][c] for c in range(3)] for r in range(3)] for e in range(48)])
    index = {m.tobytes(): e for e, m in enumerate(mats)}
    assert len(index) == 48 and np.array_equal(mats[0], np.eye(3, dtype=np.int64))
    for i in range(48):
        for j in range(48):
            assert index[(mats[i] @ mats[j]).tobytes()] == T[i][j]


def test_group_tables_match_reference_build(ref_lib):
    L = _lib.lib()
    for name, typ in (("octahedral_matrep", _abi.IMat3x3 * 48), ("octahedral_multiplication_table", (ctypes.c_int * 48) * 48)):
        a = bytes(typ.in_dll(L, name))
        b = bytes(typ.in_dll(ref_lib, name))
        assert a == b, name


def test_host_helpers_match_reference_semantics(orc):
    """grid permutation, orbit minimum and permutation helpers run on the host; compare with the oracle."""
    L = _lib.lib()
    g = _abi.Grid.of(orc.sym_grid(5))
    arr = (_abi.P_glong * 48)()
    L.evaluate_octahedral_grid_permutations(ctypes.byref(g), arr)
    P = np.array([np.ctypeslib.as_array(arr[e], shape=(125,)) for e in range(48)])
    assert np.array_equal(P, orc.octahedral_perms(orc.sym_grid(5)))
    assert L.is_permutation(arr[17], 125)
    rng = np.random.default_rng(5)
    for _ in range(50):
        i, j, k, l = (int(x) for x in rng.integers(0, 125, size=4))
        assert L.minimum_octahedral_orbit_eri_tensor_index(125, arr, i, j, k, l) == orc.min_orbit_eri(orc.sym_grid(5), P, i, j, k, l)
        assert L.minimum_octahedral_orbit_erida_tensor_index(125, arr, i, j) == orc.min_orbit_erida(orc.sym_grid(5), P, i, j)


def test_no_gpu_raises_cleanly():
    import subprocess, sys
    code = ("import os; os.environ['HIP_VISIBLE_DEVICES']='-1'\n"
            "from gausslectronitegrate_amd import _lib\n"
            "try:\n    _lib.require_gpu()\nexcept RuntimeError as e:\n    print('raised', e)\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=REPO, timeout=120)
    assert "raised" in out.stdout, out.stdout + out.stderr


def _separable_rule(L, M):
    Q = L.gli_eri_separable_rule(M, None, None, 0)
    s = np.zeros(Q)
    w = np.zeros(Q)
    dp = ctypes.POINTER(ctypes.c_double)
    assert L.gli_eri_separable_rule(M, s.ctypes.data_as(dp), w.ctypes.data_as(dp), Q) == Q
    return s, w


@pytest.mark.parametrize("n,K", [(3, 57), (5, 29), (7, 43), (7, 57), (9, 57), (11, 57), (13, 57), (15, 57)])
def test_separable_rule_reproduces_coulomb_table(n, K):
    """the separable method's quadrature rule against the closed form of the Coulomb table,
    T[i] = erf(d / (sqrt(2) w)) / d with d = sqrt(i) / 6, w = 1/3 (eri_integrals.c:18-27, 569-574), i.e.
    6 erf(sqrt(i / 8)) / sqrt(i), for every table index i <= 3 M^2 of the grid's folded range M"""
    from scipy.special import erf
    L = _lib.lib()
    M = 6 * (n - 1) + 2 * (K - 1)
    s, w = _separable_rule(L, M)
    assert 0 < len(s) <= 128 and np.all((s > 0) & (s < 1)) and np.all(w > 0)
    assert abs(w.sum() - 1.0) <= 1e-15   # the rule integrates 1 over [0, 1]: T at i = 0
    i = np.arange(3 * M * M + 1, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        table = np.where(i > 0, 6.0 * erf(np.sqrt(i / 8.0)) / np.sqrt(i), 6.0 / np.sqrt(2.0 * np.pi))
    approx = np.zeros_like(i)
    for q in range(len(s)):   # one node at a time keeps the memory small
        approx += w[q] * np.exp(-i * (s[q] * s[q] / 8.0))
    approx *= 6.0 / np.sqrt(2.0 * np.pi)
    assert np.max(np.abs(approx - table) / table) <= 3e-15


def test_separable_rule_edges():
    L = _lib.lib()
    assert L.gli_eri_separable_rule(-1, None, None, 0) == -1
    s, w = _separable_rule(L, 0)
    assert len(s) > 0 and abs(w.sum() - 1.0) <= 1e-15
