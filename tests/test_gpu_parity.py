"""GPU parity of the HIP path against the reference fixtures, the golden
vectors of the reference build, the oracle, and size-independent invariants
at the full workload sizes.  Every call goes through the C ABI of
lib/libgli_amd.so.

Tolerances: integer results (index lists) are bit-exact; FP64 values are
compared with the reference tests' own bounds (1e-13 ERI/nuclear/kinetic,
1e-12 ERIDA dense, test/test_*.c).
"""
import ctypes as C
import itertools
import math
import sys

import numpy as np
import pytest

import gausslectronitegrate_amd as gli
from gausslectronitegrate_amd import _abi, _lib
from tests import burnside
from tests.conftest import golden
from tests.gpu_helpers import DevArray, eri_enumerate_device, eri_values_device

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    return _lib.require_gpu()


def sym(n):
    h = (n - 1) // 2
    return ((-h, n), (-h, n), (-h, n))


# ---------------------------------------------------------------------------
# the MFMA mode product (pins the FP64 16x16x4 accumulator layout)

@pytest.mark.parametrize("Kd,P,Cc", [(4, 16, 16), (17, 70, 33), (176, 300, 130)])
def test_mode_product_vs_numpy(L, Kd, P, Cc):
    say("mode product", Kd, P, Cc)
    rng = np.random.default_rng(Kd + P)
    X = rng.standard_normal((Kd, P))
    B = rng.standard_normal((Kd, Cc))
    dX, dB, dY = DevArray.from_numpy(X), DevArray.from_numpy(B), DevArray(8 * P * Cc)
    _lib.check(L.gli_mode_product_device(dX.ptr, P, P, dB.ptr, Cc, Cc, Kd, dY.ptr, Cc), "mode product")
    Y = dY.to_numpy(np.float64, P * Cc).reshape(P, Cc)
    assert np.max(np.abs(Y - X.T @ B)) <= 1e-12 * max(1.0, np.abs(X.T @ B).max())


# ---------------------------------------------------------------------------
# enumeration

def test_eri_enumeration_bit_exact_small(L):
    say("enumerate 1, 3, 5")
    d, n = eri_enumerate_device(1)   # one point: the single four-index 0 (Burnside count 1)
    assert n == burnside.orbit_count(1) == 1 and d.to_numpy(np.int64, n).tolist() == [0]
    g = gli.SparseERI  # noqa: F841 (surface import check)
    ref3 = golden("ref_eri_3x3x3.npz")["four_indices"]
    d, n = eri_enumerate_device(3)
    assert np.array_equal(d.to_numpy(np.int64, n), ref3)
    ref5 = golden("golden_enum_n5.npz")["eri_indices"]
    d, n = eri_enumerate_device(5)
    assert np.array_equal(d.to_numpy(np.int64, n), ref5)


def test_eri_enumeration_7_complete_and_canonical(L, orc):
    say("enumerate 7")
    d, n = eri_enumerate_device(7)
    got = d.to_numpy(np.int64, n)
    assert n == burnside.orbit_count(7)
    assert np.all(np.diff(got) > 0)
    assert np.array_equal(orc.eri_canonical_keys(sym(7), got), got)


@pytest.mark.parametrize("n", [9, 11])
def test_eri_enumeration_large_counts(L, orc, n):
    """Full-size lists: Burnside count, strictly ascending, every sampled entry canonical."""
    say("enumerate", n)
    d, num = eri_enumerate_device(n)
    assert num == burnside.orbit_count(n)
    got = d.to_numpy(np.int64, num)
    d.free()
    assert got[0] >= 0 and got[-1] < (n ** 3) ** 4
    assert np.all(np.diff(got) > 0)
    rng = np.random.default_rng(n)
    sample = np.sort(rng.choice(got, size=20000, replace=False))
    assert np.array_equal(orc.eri_canonical_keys(sym(n), sample), sample)


@pytest.mark.parametrize("method", [0, 1])
def test_host_api_streamed_matches_device(L, method):
    """The reference entry points with host arrays at 9^3 (39,277,113 entries: ten transfer chunks, the
    last one ragged): enumerate_symmetry_reduced_eri_indices equals the device-resident list, and
    compute_sparse_eri_gausslet_integrals returns bit-identical values to the device-resident path
    together with an exact copy of its input indices."""
    say("host API streamed 9, method", method)
    n = 9
    g = _abi.Grid.of(sym(n))
    ri = _abi.SparseEriIndices()
    L.enumerate_symmetry_reduced_eri_indices(C.byref(g), C.byref(ri))
    host = np.ctypeslib.as_array(ri.four_indices, shape=(ri.num,)).copy()
    L.delete_sparse_eri_indices(C.byref(ri))
    d, num = eri_enumerate_device(n)
    dev = d.to_numpy(np.int64, num)
    d.free()
    assert num == burnside.orbit_count(n) and np.array_equal(host, dev)
    c = golden("coeffs_K57_s0.npy")
    gd = _abi.GaussletData()
    cc = np.ascontiguousarray(c)
    gd.coefficients = cc.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(len(c) - 1) // 2
    gd.indices.num = len(c)
    L.gli_set_eri_method(method)
    try:
        res = _abi.SparseEri()
        L.compute_sparse_eri_gausslet_integrals(C.byref(gd), C.byref(g), host.ctypes.data_as(_abi.P_glong), num, 1e-16,
                                                C.byref(res))
        vals = np.ctypeslib.as_array(res.integral_values, shape=(num,)).copy()
        keys = np.ctypeslib.as_array(res.four_indices, shape=(num,)).copy()
        L.delete_sparse_eri_gausslet_integrals(C.byref(res))
    finally:
        L.gli_set_eri_method(0)
    assert np.array_equal(keys, host)
    ref = eri_values_device(c, sym(n), host, 1e-16, method)
    assert np.array_equal(vals, ref)
    # the fused entry (index list kept on the GPU) fills the same struct
    L.gli_set_eri_method(method)
    try:
        fused = _abi.SparseEri()
        _lib.check(L.gli_sparse_eri_compute(C.byref(gd), C.byref(g), 1e-16, C.byref(fused)), "gli_sparse_eri_compute")
        assert fused.num_entries == num and fused.grid.spec() == g.spec()
        assert np.array_equal(np.ctypeslib.as_array(fused.four_indices, shape=(num,)), host)
        assert np.array_equal(np.ctypeslib.as_array(fused.integral_values, shape=(num,)), ref)
        L.delete_sparse_eri_gausslet_integrals(C.byref(fused))
    finally:
        L.gli_set_eri_method(0)
    bad = _abi.Grid.of(((-4, 9), (-4, 9), (-3, 9)))   # not centred: an error status, no abort
    assert L.gli_sparse_eri_compute(C.byref(gd), C.byref(bad), 1e-16, C.byref(_abi.SparseEri())) == -1


def test_eri_enumeration_13_count(L):
    say("enumerate 13")
    d, num = eri_enumerate_device(13)
    assert num == burnside.orbit_count(13) == 1250296187
    # spot-check order on a strided slice without copying 10 GB to the host
    head = d.to_numpy(np.int64, 1 << 20)
    assert np.all(np.diff(head) > 0)
    d.free()


def test_erida_enumeration(L):
    for n in (3, 5, 7, 9, 11, 13):
        e = gli.SparseERIDA(golden("coeffs_K29_s0.npy"), sym(n), 1e-3) if n == 3 else None
        out = _abi.SparseEridaIndices()
        g = _abi.Grid.of(sym(n))
        L.enumerate_symmetry_reduced_erida_indices(C.byref(g), C.byref(out))
        got = np.ctypeslib.as_array(out.two_indices, shape=(out.num,)).copy()
        L.delete_sparse_erida_indices(C.byref(out))
        assert len(got) == math.comb(n + 2, 3) == burnside.orbit_count(n, 2)
        assert np.all(np.diff(got) > 0)
        if n == 7:
            assert np.array_equal(got, golden("golden_n7_K57s0.npz")["erida_indices"])


# ---------------------------------------------------------------------------
# ERI values

def test_eri_3x3x3_fixture_through_python_surface(L):
    """examples/eri_3x3x3.hdf5: G10, tol 1e-16, 1027 values; notebook values (eri_integrals.ipynb:59-192)."""
    say("eri 3x3x3")
    ref = golden("ref_eri_3x3x3.npz")
    eri = gli.compute_sparse_eri_gausslet_integrals(ref["coeffs"], sym(3), 1e-16)
    assert eri.num_entries == 1027
    assert np.array_equal(eri.four_indices, ref["four_indices"])
    assert np.max(np.abs(eri.integral_values - ref["values"])) <= 1e-13
    assert eri.integral_value(([0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0])) == pytest.approx(1.6252009000829197, abs=1e-13)
    pts = ([1, 0, 0], [0, 0, -1], [0, 0, 1], [0, 0, 0])
    assert eri.integral_value(pts) == pytest.approx(-0.0005864045861050727, abs=1e-13)
    assert gli.compute_eri_gausslet_integral(ref["coeffs"], pts, 1e-16) == pytest.approx(-0.0005864045861050708, abs=1e-13)


@pytest.mark.parametrize("method", [0, 2])
def test_eri_dense_fixture(L, orc, method):
    """test_eri_integrals.hdf5: sparse 3^3 (tol 1e-3) -> dense 12^4, <= 1e-13 (test_eri_integrals.c:12-85)."""
    say("eri dense fixture, method", method)
    ref = golden("ref_eri.npz")
    tol = float(ref["tol"])
    idx = golden("ref_eri_3x3x3.npz")["four_indices"]
    vals = eri_values_device(ref["coeffs"], sym(3), idx, tol, method)
    dense = orc.fill_dense_eri(sym(3), idx, vals, ((-1, 2), (0, 2), (-1, 3)))
    assert np.max(np.abs(dense - ref["eri"])) <= 1e-13


def test_eri_fill_tensor_and_single(L):
    """fill_tensor of the GPU set vs the fixture, and single-element evaluation vs dense (test_eri_integrals.c:87-112)."""
    ref = golden("ref_eri.npz")
    eri = gli.compute_sparse_eri_gausslet_integrals(ref["coeffs"], sym(3), float(ref["tol"]))
    dense = eri.fill_tensor(((-1, 2), (0, 2), (-1, 3)))
    assert np.max(np.abs(dense - ref["eri"])) <= 1e-13
    pts = [(x, y, z) for x in range(-1, 1) for y in range(0, 2) for z in range(-1, 2)]
    rng = np.random.default_rng(2)
    for _ in range(12):
        q = rng.integers(0, 12, size=4)
        v = gli.compute_eri_gausslet_integral(ref["coeffs"], [pts[i] for i in q], float(ref["tol"]))
        assert abs(v - dense[tuple(q)]) <= 1e-13


def test_eri_synthetic_3_full(L):
    g = golden("golden_n3_K57s0.npz")
    for method in (0, 2, 6):
        got = eri_values_device(g["coeffs"], sym(3), g["eri_indices"], float(g["tol"]), method)
        assert np.max(np.abs(got - g["eri_values"])) <= 1e-13


@pytest.mark.parametrize("name,n", [("golden_sample_n5_K29s0.npz", 5), ("golden_sample_n7_K43s0.npz", 7),
                                    ("golden_sample_n9_K57s1.npz", 9), ("golden_sample_n11_K57s0.npz", 11),
                                    ("golden_sample_n13_K57s0.npz", 13)])
@pytest.mark.parametrize("method", [0, 2, 6])
def test_eri_golden_samples(L, name, n, method):
    """seeded canonical samples with values of the reference build (tol 1e-16); method 6 is the separable
    (Gauss-Legendre) expansion of the Coulomb kernel, held to the same bound as the exact methods"""
    say("eri golden", name, "method", method)
    g = golden(name)
    got = eri_values_device(g["coeffs"], sym(n), g["eri_indices"], float(g["tol"]), method)
    assert np.max(np.abs(got - g["eri_values"])) <= 1e-13


def test_eri_nonsymmetric_coefficients(L, orc):
    """coefficients without mirror symmetry use the general key set; compare with the oracle on a 5^3 sample."""
    rng = np.random.default_rng(9)
    c = rng.standard_normal(21) * np.exp(-(np.arange(21) - 10) ** 2 / 60.0)
    idx = golden("golden_enum_n5.npz")["eri_indices"][::97]
    ref = orc.eri_sparse(c, sym(5), idx, 1e-12)
    for method in (0, 2, 6):
        got = eri_values_device(c, sym(5), idx, 1e-12, method)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.abs(ref).max())


def test_eri_arbitrary_grid_and_indices(L, orc):
    """compute_sparse_* accepts any index list on any grid (eri_integrals.h:50-52)."""
    c = golden("coeffs_K29_s0.npy")
    grid = ((-2, 4), (0, 3), (1, 5))
    N = 4 * 3 * 5
    rng = np.random.default_rng(4)
    idx = np.unique(rng.integers(0, N ** 4, size=300))
    ref = orc.eri_sparse(c, grid, idx, 1e-10)
    for method in (0, 2, 6):
        got = eri_values_device(c, grid, idx, 1e-10, method)
        assert np.max(np.abs(got - ref)) <= 1e-13
        bad = eri_values_device(c, grid, np.array([-1, N ** 4], dtype=np.int64), 1e-10, method)
        assert np.all(np.isnan(bad))


@pytest.mark.parametrize("n,K,tol", [(5, 29, 1e-16), (7, 43, 1e-12)])
def test_direct_flop_count_matches_oracle(L, orc, n, K, tol):
    """the per-element path's device counter of the reference's loop flops equals the oracle's count
    (eri_integrals.c:624-669) on a seeded sample; its values match the oracle's"""
    say("direct flops", n, K)
    c = golden("coeffs_K%d_s0.npy" % K)
    idx = golden("golden_sample_n%d_K%ds0.npz" % (n, K))["eri_indices"]
    assert L.gli_enable_direct_flop_count(1) == 0
    try:
        got = eri_values_device(c, sym(n), idx, tol, 2)
        ref_form = C.c_double()
        executed = C.c_double()
        assert L.gli_direct_flops(C.byref(ref_form), C.byref(executed)) == 0
    finally:
        L.gli_enable_direct_flop_count(0)
    want = orc.eri_flops(c, sym(n), idx, tol)
    assert ref_form.value == pytest.approx(want.sum(), rel=1e-9)
    assert 0 < executed.value
    assert np.max(np.abs(got - orc.eri_sparse(c, sym(n), idx, tol))) <= 1e-13
    assert L.gli_direct_flops(C.byref(ref_form), C.byref(executed)) == -1   # counting is off again


@pytest.mark.parametrize("n,K,seed", [(7, 57, 0), (9, 57, 1)])
def test_separable_full_list_vs_tensor(L, n, K, seed):
    """the separable method (Gauss-Legendre expansion of the Coulomb kernel, one rank-Q product) against
    the exact FP64 key-tensor method over every unique element of the grid: within 1e-14 absolute"""
    say("separable vs tensor", n)
    c = golden("coeffs_K%d_s%d.npy" % (K, seed))
    d, num = eri_enumerate_device(n)
    idx = d.to_numpy(np.int64, num)
    exact = eri_values_device(c, sym(n), idx, 1e-16, 0)
    sep = eri_values_device(c, sym(n), idx, 1e-16, 6)
    err = float(np.max(np.abs(sep - exact)))
    say("  max abs diff %.3g over %d elements" % (err, num))
    assert err <= 1e-14
    assert not np.array_equal(sep, exact)   # a different evaluation, not a copy of the exact one


def test_eri_7_full_list_tensor_vs_direct_sample(L, orc):
    """full 7^3 list on the tensor path; a seeded sample checked against the oracle."""
    say("eri 7 full")
    c = golden("coeffs_K57_s0.npy")
    d, n = eri_enumerate_device(7)
    idx = d.to_numpy(np.int64, n)
    vals = eri_values_device(c, sym(7), idx, 1e-16, 0)
    assert np.all(np.isfinite(vals))
    rng = np.random.default_rng(77)
    s = np.sort(rng.choice(n, size=300, replace=False))
    ref = orc.eri_sparse(c, sym(7), idx[s], 1e-16)
    assert np.max(np.abs(vals[s] - ref)) <= 1e-13


# ---------------------------------------------------------------------------
# ERIDA

def test_erida_fixture_and_notebook(L, orc):
    say("erida")
    ref = golden("ref_erida.npz")
    e = gli.compute_sparse_erida_gausslet_integrals(ref["coeffs"], sym(3), float(ref["tol"]))
    dense = e.fill_matrix(((0, 2), (-1, 3), (-1, 2)))
    assert np.max(np.abs(dense - ref["erida"])) <= 1e-12
    g10 = golden("ref_eri.npz")["coeffs"]
    assert gli.compute_erida_gausslet_integral(g10, ([0, 0, 0], [0, 0, 0]), 1e-16) == pytest.approx(2.4104836730129793, abs=1e-13)
    assert gli.compute_erida_gausslet_integral(g10, ([-1, 0, 2], [0, 1, -1]), 1e-16) == pytest.approx(0.3024908490984703, abs=1e-13)
    e7 = gli.compute_sparse_erida_gausslet_integrals(g10, sym(7), 1e-16)
    assert e7.num_entries == 84
    assert e7.integral_value(([-1, 0, 2], [0, 1, -1])) == pytest.approx(0.3024908490984723, abs=1e-13)


def test_erida_golden_7(L):
    g = golden("golden_n7_K57s0.npz")
    e = gli.compute_sparse_erida_gausslet_integrals(g["coeffs"], sym(7), float(g["tol"]))
    assert np.array_equal(e.two_indices, g["erida_indices"])
    assert np.max(np.abs(e.integral_values - g["erida_values"])) <= 1e-13


def _reference_test_basis(num_points, num_states):
    """The deterministic pseudo-random basis of the reference's projection tests
    (test/test_erida_integrals.c:164-168, test/test_eri_integrals.c)."""
    i = np.arange(num_points * num_states, dtype=np.int64)
    b = -0.12 * np.cos(math.sqrt(7) * math.pi * ((409 * i) % 349) - 6) / (1.7 + np.sin(math.sqrt(3) * math.pi * ((557 * i) % 281)))
    return b.reshape(num_points, num_states)


def test_erida_project_5(L):
    """test_project_erida_integrals (test/test_erida_integrals.c:112-228): 5^3 grid, tol 1e-5, five states;
    the projection equals sum_ij b_ip b_iq b_jr b_js M_ij over the dense matrix M, to 1e-12."""
    say("erida project 5")
    c = golden("ref_erida.npz")["coeffs"]
    e = gli.compute_sparse_erida_gausslet_integrals(c, sym(5), 1e-5)
    basis = _reference_test_basis(125, 5)
    proj = e.project(basis)
    M = e.fill_matrix(sym(5))
    pq = np.einsum("ip,iq->ipq", basis, basis).reshape(125, 25)
    ref = (pq.T @ M @ pq).reshape(5, 5, 5, 5)
    assert np.max(np.abs(proj - ref)) <= 1e-12
    assert np.max(np.abs(proj - proj.transpose(2, 3, 0, 1))) <= 1e-14


def test_erida_sparse_completeness_7(L, orc):
    """test_erida_integral_sparse_completeness (test/test_erida_integrals.c:230-299): 7^3 grid, tol 1e-6;
    the dense matrix filled from the symmetry-reduced set has no zero entry.  A seeded sample of entries
    is also checked against the oracle's single-pair sum (erida_integrals.c:119-201)."""
    say("erida completeness 7")
    c = golden("ref_erida.npz")["coeffs"]
    e = gli.compute_sparse_erida_gausslet_integrals(c, sym(7), 1e-6)
    M = e.fill_matrix(sym(7))
    assert M.shape == (343, 343)
    assert np.count_nonzero(M == 0) == 0
    assert np.array_equal(M, M.T)
    pts = np.array(list(itertools.product(range(-3, 4), repeat=3)))
    rng = np.random.default_rng(31)
    for i, j in rng.integers(0, 343, size=(40, 2)):
        assert abs(M[i, j] - orc.erida_single(c, (pts[i], pts[j]), 1e-6)) <= 1e-12


# ---------------------------------------------------------------------------
# one-body

def test_kinetic(L):
    say("kinetic")
    ref = golden("ref_kinetic.npz")
    k = gli.compute_kinetic_gausslet_integrals(ref["coeffs"], ((0, 2), (1, 3), (-1, 3)))
    assert np.max(np.abs(k - ref["kgi"])) <= 1e-13
    g10 = golden("ref_eri.npz")["coeffs"]
    kg = gli.compute_kinetic_gausslet_integrals(g10, ((-1, 3), (-2, 5), (-2, 4)))
    assert kg[7, 7] == pytest.approx(5.218509050588766, abs=1e-13)
    assert gli.compute_kinetic_gausslet_integral(g10, ([0, -2, -1], [0, -2, 1])) == pytest.approx(0.3010258456920169, abs=1e-13)
    assert np.array_equal(kg, kg.T)
    g = golden("golden_n3_K57s0.npz")
    assert np.max(np.abs(gli.compute_kinetic_gausslet_integrals(g["coeffs"], sym(3)) - g["kinetic"])) <= 1e-13
    # the reference's struct-returning entry point gives the same matrix bit for bit
    gd = _abi.GaussletData()
    cc = np.ascontiguousarray(g10)
    gd.coefficients = cc.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(len(g10) - 1) // 2
    gd.indices.num = len(g10)
    grid = _abi.Grid.of(((-1, 3), (-2, 5), (-2, 4)))
    out = _abi.KineticIntegrals()
    L.compute_kinetic_gausslet_integrals(C.byref(gd), C.byref(grid), C.byref(out))
    assert np.array_equal(np.ctypeslib.as_array(out.integral_values, shape=kg.shape), kg)
    L.delete_kinetic_gausslet_integrals(C.byref(out))


def test_nuclear(L):
    say("nuclear")
    ref = golden("ref_nuclear.npz")
    m = gli.compute_nuclear_gausslet_integrals(ref["coeffs"], ((0, 2), (2, 2), (-1, 3)), ref["pos"], ref["charges"], float(ref["tol"]))
    assert np.max(np.abs(m - ref["ngi"])) <= 1e-13
    g10 = golden("ref_eri.npz")["coeffs"]
    pos = [(-0.2, 1.3, 0.5), (0.7, 0, -2.1)]
    ngi = gli.compute_nuclear_gausslet_integrals(g10, ((-1, 3), (0, 2), (-2, 4)), pos, [2, 1], 1e-16)
    assert ngi[10, 10] == pytest.approx(1.8699450271662796, abs=1e-13)
    assert np.array_equal(ngi, ngi.T)
    v = gli.compute_nuclear_gausslet_integral(g10, ([-1, 0, -2], [0, 1, 1]), pos, [2, 1], 1e-16)
    assert v == pytest.approx(-0.0003044157658974286, abs=1e-13)
    g = golden("golden_n3_K57s0.npz")
    m3 = gli.compute_nuclear_gausslet_integrals(g["coeffs"], sym(3), [[0, 0, 0]], [1], float(g["tol"]))
    assert np.max(np.abs(m3 - g["nuclear"])) <= 1e-13
    # the reference's struct-returning entry point gives the same matrix bit for bit
    gd = _abi.GaussletData()
    cc = np.ascontiguousarray(g10)
    gd.coefficients = cc.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(len(g10) - 1) // 2
    gd.indices.num = len(g10)
    grid = _abi.Grid.of(((-1, 3), (0, 2), (-2, 4)))
    nuc = (_abi.Nucleus * 2)()
    for q in range(2):
        for a in range(3):
            nuc[q].pos[a] = pos[q][a]
        nuc[q].charge = [2, 1][q]
    out = _abi.NuclearIntegrals()
    L.compute_nuclear_gausslet_integrals(C.byref(gd), C.byref(grid), nuc, 2, 1e-16, C.byref(out))
    assert np.array_equal(np.ctypeslib.as_array(out.integral_values, shape=ngi.shape), ngi)
    L.delete_nuclear_gausslet_integrals(C.byref(out))


def test_nuclear_golden_7(L):
    """7^3, two unit charges at (0, 0, +-0.7): sampled entries, row sums and trace of the reference build.

    The reference sums each entry sequentially over about 2.9 million signed
    terms (113^3 primitive centres x 2 nuclei, random-sign coefficients), which
    carries a rounding error of order 1e-13 by itself; the bound is 3e-13."""
    say("nuclear 7")
    g = golden("golden_n7_K57s0.npz")
    m = gli.compute_nuclear_gausslet_integrals(g["coeffs"], sym(7), g["nuc_pos"], g["nuc_charges"], float(g["tol"]))
    assert np.max(np.abs(m[g["nuc_rows"], g["nuc_cols"]] - g["nuc_sample"])) <= 3e-13
    assert np.max(np.abs(m.sum(axis=1) - g["nuc_rowsum"])) <= 1e-11
    assert abs(np.trace(m) - g["nuc_trace"]) <= 1e-11
    assert np.array_equal(m, m.T)


# ---------------------------------------------------------------------------
# post-processing through the surface (notebook checks, eri_integrals.ipynb:17-21)

def test_project_and_apply(L):
    ref = golden("ref_eri_3x3x3.npz")
    eri = gli.compute_sparse_eri_gausslet_integrals(ref["coeffs"], sym(3), 1e-16)
    rng = np.random.default_rng(293)
    basis = np.linalg.qr(rng.standard_normal((27, 27)))[0][:, :4]
    dense = eri.fill_tensor(sym(3))
    proj = eri.project(basis)
    ref_proj = np.einsum("ip,jq,kr,ls,ijkl->pqrs", basis, basis, basis, basis, dense, optimize=True)
    assert np.linalg.norm(proj - ref_proj) <= 1e-12
    states = rng.normal(size=(27 * 27, 3))
    out = eri.apply(states)
    vint = dense.transpose((0, 2, 1, 3)).reshape(27 * 27, 27 * 27)
    assert np.linalg.norm(out - vint @ states) <= 1e-12


# ---------------------------------------------------------------------------
# split-precision (mixed) path: error <= 1e-10 absolute against FP64

@pytest.mark.parametrize("Kd,P,Cc", [(32, 64, 64), (176, 300, 130), (149, 1000, 77), (300, 200, 100), (520, 70, 90)])
def test_split_mode_product_vs_numpy(L, Kd, P, Cc):
    """BF16 digit slices with exact FP32 sums: error bounded by ~2^-40 of the column scales."""
    say("split mode product", Kd, P, Cc)
    rng = np.random.default_rng(Kd * 7 + P)
    X = rng.standard_normal((Kd, P)) * np.exp(rng.uniform(-8, 2, size=(1, P)))
    B = rng.standard_normal((Kd, Cc))
    dX, dB, dY = DevArray.from_numpy(X), DevArray.from_numpy(B), DevArray(8 * P * Cc)
    _lib.check(L.gli_mode_product_split_device(dX.ptr, P, P, dB.ptr, Cc, Cc, Kd, dY.ptr, Cc), "split mode product")
    Y = dY.to_numpy(np.float64, P * Cc).reshape(P, Cc)
    ref = X.T @ B
    bound = Kd * 2.0 ** -38 * np.abs(X).max(axis=0)[:, None] * np.abs(B).max(axis=0)[None, :]
    assert np.all(np.abs(Y - ref) <= bound)


@pytest.mark.parametrize("name,n", [("golden_sample_n9_K57s1.npz", 9), ("golden_sample_n11_K57s0.npz", 11),
                                    ("golden_sample_n13_K57s0.npz", 13)])
def test_eri_mixed_golden(L, name, n):
    say("eri mixed golden", name)
    g = golden(name)
    got = eri_values_device(g["coeffs"], sym(n), g["eri_indices"], float(g["tol"]), 1)
    assert np.max(np.abs(got - g["eri_values"])) <= 1e-10


def test_eri_mixed_full_7(L):
    """every unique 7^3 element: mixed vs the FP64 tensor path, max abs error <= 1e-10"""
    say("eri mixed full 7")
    c = golden("coeffs_K57_s1.npy")
    d, n = eri_enumerate_device(7)
    idx = d.to_numpy(np.int64, n)
    v64 = eri_values_device(c, sym(7), idx, 1e-16, 0)
    vmx = eri_values_device(c, sym(7), idx, 1e-16, 1)
    assert np.max(np.abs(vmx - v64)) <= 1e-10


# ---------------------------------------------------------------------------
# tuning knobs: every variant computes the same terms in the same order per element

def test_tuning_variants_bit_identical(L):
    """tile order of the mode products, gather order, elements per lane and workgroup size change the schedule only:
    the full 7^3 list (and an unsorted, ragged copy of it) must come out bit-identical"""
    say("tuning variants")
    c = golden("coeffs_K57_s0.npy")
    d, n = eri_enumerate_device(7)
    idx = d.to_numpy(np.int64, n)
    rng = np.random.default_rng(5)
    shuffled = idx[rng.permutation(n)][: n - 77]
    knobs = [(b"tile_order", (0, 1)), (b"gather_order", (0, 1, 2)), (b"gather_ept", (2, 4, 8, 16)), (b"gather_lds", (0, 1)),
             (b"gemm_shape", (0, 3, 4, 5)), (b"mode1_half", (0, 1)), (b"gather_threads", (256, 512, 1024)),
             (b"gather_bucketed", (0, 1)), (b"gather_classed", (0, 1))]
    saved = {k: L.gli_get_tuning(k) for k, _ in knobs}
    try:
        ref, ref_s = {}, {}
        for to, go, ept, lds, shape in itertools.product((0, 1), (0, 1, 2), (2, 4, 8, 16), (0, 1), (0, 3, 4, 5)):
            if (to + go + ept + lds + shape) % 2 and ept not in (4, 16):
                continue   # a spread of the combinations keeps the test short; every value of every knob is used
            half = (to + lds + shape + ept // 4) % 2   # the half-stored first mode: both values with every shape
            threads = (256, 512, 1024)[(to + go + shape + ept) % 3]
            bucketed = (go + lds + ept // 4) % 2   # window-bucketed key-tensor reads: both with every gather order
            classed = (to + go + lds) % 2   # line-class grouped reads: both with every gather order and table place
            for k, v in zip((b"tile_order", b"gather_order", b"gather_ept", b"gather_lds", b"gemm_shape", b"mode1_half",
                             b"gather_threads", b"gather_bucketed", b"gather_classed"),
                            (to, go, ept, lds, shape, half, threads, bucketed, classed)):
                assert L.gli_set_tuning(k, v) == 0
            for method in (0, 1, 6):
                got = eri_values_device(c, sym(7), idx, 1e-16, method)
                got_s = eri_values_device(c, sym(7), shuffled, 1e-16, method)
                if method not in ref:
                    ref[method], ref_s[method] = got, got_s
                assert np.array_equal(got, ref[method]), (to, go, ept, lds, shape, half, method)
                assert np.array_equal(got_s, ref_s[method]), (to, go, ept, lds, shape, half, method)
        assert np.array_equal(ref_s[0], ref[0][np.searchsorted(idx, shuffled)])
        assert L.gli_set_tuning(b"gather_ept", 3) == -1
        assert L.gli_set_tuning(b"gather_lds", 2) == -1
        assert L.gli_set_tuning(b"mode1_half", 2) == -1
        assert L.gli_set_tuning(b"batch_sorted", 2) == -1
        assert L.gli_set_tuning(b"gather_threads", 128) == -1
        assert L.gli_set_tuning(b"gather_order", 3) == -1
        assert L.gli_set_tuning(b"gather_bucketed", 3) == -1
        assert L.gli_set_tuning(b"gather_classed", 2) == -1
        assert L.gli_set_tuning(b"batch_positions", -1) == -1
        assert L.gli_get_tuning(b"batch_sorted") in (0, 1)
        assert L.gli_set_tuning(b"no_such_knob", 1) == -1
        assert L.gli_get_tuning(b"no_such_knob") == -1
    finally:
        L.gli_clear_error()
        for k, v in saved.items():
            L.gli_set_tuning(k, v)


def _coeffs_sym(K, seed, mirror=True):
    """seeded coefficients of K entries (K odd), mirror-symmetric unless asked otherwise"""
    z = np.random.default_rng(seed).standard_normal(K) * np.exp(-(4 * np.arange(-(K // 2), K // 2 + 1) / 57.0) ** 2)
    return 0.5 * (z + z[::-1]) if mirror else z


@pytest.mark.parametrize("n,K,mirror", [(3, 3, True), (5, 15, True), (7, 57, True), (7, 89, True), (5, 57, False)])
def test_rankq_tet_bit_identical_to_tile_kernel(L, n, K, mirror):
    """the separable key tensor from the tetrahedral rank-Q kernel (16 x 16 key blocks per wave over chunks of
    16 k3) equals the 64 x 64 tile kernel's bit for bit on every element of the full n^3 list: both add the
    same products in the same order.  The cases take S = ceil(Q / 4) = 13, 17, 21, 25 MFMA steps
    (folded ranges M = 16, 52, 148, 212) and a coefficient set without mirror symmetry (n^3 keys per axis)"""
    say("rankq tet", n, K, mirror)
    c = _coeffs_sym(K, 11 + n + K, mirror)
    d, num = eri_enumerate_device(n)
    idx = d.to_numpy(np.int64, num)
    saved = L.gli_get_tuning(b"rankq_tet")
    try:
        assert L.gli_set_tuning(b"rankq_tet", 0) == 0
        tile = eri_values_device(c, sym(n), idx, 1e-16, 6)
        assert L.gli_set_tuning(b"rankq_tet", 1) == 0
        tet = eri_values_device(c, sym(n), idx, 1e-16, 6)
        assert np.all(np.isfinite(tet))
        assert np.array_equal(tet.view(np.uint64), tile.view(np.uint64)), float(np.max(np.abs(tet - tile)))
        exact = eri_values_device(c, sym(n), idx, 1e-16, 0)
        assert np.max(np.abs(tet - exact)) <= 4e-15 * max(1.0, float(np.max(np.abs(exact))))
        assert L.gli_set_tuning(b"rankq_tet", 2) == -1
    finally:
        L.gli_clear_error()
        L.gli_set_tuning(b"rankq_tet", saved)


@pytest.mark.parametrize("n,K", [(1, 3), (3, 301)])
def test_rankq_tet_smallest_and_largest_rules(L, orc, n, K):
    """the shortest and the longest node rules the tetrahedral rank-Q kernel is built for: the one-point grid with
    K = 3 (folded range 4, Q = 34, S = 9) and 3^3 with K = 301 (range 612, Q = 114, S = 29, coefficients from device
    memory).  Bit-identical to the tile kernel, and within 1e-13 relative of the oracle's reference sums"""
    say("rankq tet rules", n, K)
    c = _coeffs_sym(K, 5 + K)
    d, num = eri_enumerate_device(n)
    idx = d.to_numpy(np.int64, num)
    sample = idx[:: max(1, num // 64)]
    saved = L.gli_get_tuning(b"rankq_tet")
    try:
        assert L.gli_set_tuning(b"rankq_tet", 0) == 0
        tile = eri_values_device(c, sym(n), idx, 1e-16, 6)
        assert L.gli_set_tuning(b"rankq_tet", 1) == 0
        tet = eri_values_device(c, sym(n), idx, 1e-16, 6)
        assert np.array_equal(tet.view(np.uint64), tile.view(np.uint64)), float(np.max(np.abs(tet - tile)))
        ref = orc.eri_sparse(c, sym(n), sample, 1e-16)
        got = eri_values_device(c, sym(n), sample, 1e-16, 6)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, float(np.max(np.abs(ref))))
    finally:
        L.gli_clear_error()
        L.gli_set_tuning(b"rankq_tet", saved)


# ---------------------------------------------------------------------------
# batch plans: several grids, grouped launches, autotuner, one replayed hipGraph

def _batch_setup(specs):
    """specs: list of (grid spec, coefficients, indices or None for the full enumerated list)"""
    entries = []
    for spec, c, idx in specs:
        if idx is None:
            n = spec[0][1]
            d, num = eri_enumerate_device(n)
            idx = d.to_numpy(np.int64, num)
        di = DevArray.from_numpy(np.ascontiguousarray(idx, dtype=np.int64))
        dv = DevArray(8 * max(len(idx), 1))
        entries.append((spec, np.ascontiguousarray(c, dtype=np.float64), np.ascontiguousarray(idx), di, dv))
    return entries


def _batch_arrays(entries, coeffs=None):
    n = len(entries)
    cs = coeffs if coeffs is not None else [e[1] for e in entries]
    cptr = (_abi.P_double * n)(*[c.ctypes.data_as(_abi.P_double) for c in cs])
    Ks = np.array([len(c) for c in cs], dtype=np.int64)
    grids = np.concatenate([np.array([v for ax in e[0] for v in ax], dtype=np.int64) for e in entries])
    dptr = (C.c_void_p * n)(*[e[3].ptr for e in entries])
    nums = np.array([len(e[2]) for e in entries], dtype=np.int64)
    vptr = (C.c_void_p * n)(*[e[4].ptr for e in entries])
    return cptr, Ks, grids, dptr, nums, vptr


@pytest.mark.parametrize("method", [0, 1, 6])
@pytest.mark.parametrize("branches,positions,sorted_windows", [(1, 1, 1), (2, 0, 1), (1, 1, 0)])
def test_batch_plan_matches_single_grids(L, method, branches, positions, sorted_windows):
    """a 4-grid plan (uneven K, one non-cubic grid with an unsorted sample): every replay equals the
    single-grid path bit for bit (FP64 and separable; the same kernels in the same summation order) or
    within 1e-10 of FP64 (mixed); output
    scales, new coefficients and a symmetry change behave as documented.  branches = 2 forces the
    two-branch graph (largest grid beside the others on a second stream) wherever it applies;
    positions = 1 caches every element's key-tensor position at plan creation, 0 decodes per replay;
    sorted_windows = 1 also sorts the cached positions inside windows of 8192 elements (the 3^3 list and
    the last window of every list are partial)."""
    say("batch plan method", method, "branches", branches, "positions", positions, "sorted", sorted_windows)
    saved = L.gli_get_tuning(b"batch_branches"), L.gli_get_tuning(b"batch_positions"), L.gli_get_tuning(b"batch_sorted")
    assert L.gli_set_tuning(b"batch_branches", branches) == 0
    assert L.gli_set_tuning(b"batch_positions", positions) == 0
    assert L.gli_set_tuning(b"batch_sorted", sorted_windows) == 0
    rng = np.random.default_rng(11)
    ragged = rng.integers(0, (5 * 3 * 4) ** 4, size=5000)
    specs = [(sym(3), golden("coeffs_K29_s0.npy"), None),
             (sym(5), golden("coeffs_K43_s0.npy"), None),
             (sym(7), golden("coeffs_K57_s1.npy"), None),
             (((-2, 5), (-1, 3), (-2, 4)), golden("coeffs_K57_s0.npy"), ragged)]
    entries = _batch_setup(specs)
    cptr, Ks, grids, dptr, nums, vptr = _batch_arrays(entries)
    plan = C.c_void_p()
    try:
        _lib.check(L.gli_eri_batch_create(len(entries), cptr, Ks.ctypes.data_as(_abi.P_glong),
                                          grids.ctypes.data_as(_abi.P_glong), dptr, nums.ctypes.data_as(_abi.P_glong), 1e-16,
                                          method, vptr, 1, C.byref(plan)), "gli_eri_batch_create")
    finally:
        L.gli_set_tuning(b"batch_branches", saved[0])
        L.gli_set_tuning(b"batch_positions", saved[1])
        L.gli_set_tuning(b"batch_sorted", saved[2])
    try:
        import json
        buf = C.create_string_buffer(1 << 16)
        assert L.gli_eri_batch_describe(plan, buf, len(buf)) > 0
        report = json.loads(buf.value.decode())
        assert report["gather"]
        assert report["gather_chosen"]["cached_positions"] == bool(positions)
        assert report["gather_chosen"]["sorted_windows"] == bool(positions and sorted_windows)
        if method == 6:
            # one grouped rank-Q product for all grids; no two-branch graph
            assert len(report["modes"]) == 1 and len(report["modes"][0]["rank_q"]) == 2 and "branches" not in report
        else:
            assert len(report["modes"]) == 3
            split_any = any(m["chosen"] != "fp64" for m in report["modes"])
            if not split_any:
                assert report["branches"]["chosen"] == ("two" if branches == 2 else report["branches"]["chosen"])
                assert report["branches"]["lead_grid"] == 2   # the full 7^3 list has the most elements
        single_method = 6 if method == 6 else 0
        singles = [eri_values_device(e[1], e[0], e[2], 1e-16, single_method) for e in entries]
        if method == 6:
            for e, ref in zip(entries, singles):
                exact = eri_values_device(e[1], e[0], e[2], 1e-16, 0)
                assert np.max(np.abs(ref - exact)) <= 1e-14
        for rep in range(3):
            _lib.check(L.gli_eri_batch_run(plan, None, None), "gli_eri_batch_run")
            assert L.gli_synchronize() == 0
            for e, ref in zip(entries, singles):
                got = e[4].to_numpy(np.float64, len(e[2]))
                if method in (0, 6):
                    assert np.array_equal(got, ref), (rep, e[0])
                else:
                    assert np.max(np.abs(got - ref)) <= 1e-10, (rep, e[0])
            # eager work between replays
            eri_values_device(entries[0][1], entries[0][0], entries[0][2][:100], 1e-16, 0)
        scales = np.array([0.5, 1.0, 2.0, 1.0 / 1.5])
        _lib.check(L.gli_eri_batch_run(plan, None, scales.ctypes.data_as(C.POINTER(C.c_double))), "run scaled")
        assert L.gli_synchronize() == 0
        for e, ref, sc in zip(entries, singles, scales):
            got = e[4].to_numpy(np.float64, len(e[2]))
            if method in (0, 6):
                assert np.array_equal(got, ref * sc)
            else:
                assert np.max(np.abs(got - ref * sc)) <= 2e-10
        # new coefficients with the same mirror symmetry
        new = [golden("coeffs_K29_s0.npy")[::-1].copy(), golden("coeffs_K43_s0.npy") * 0.5,
               golden("coeffs_K57_s0.npy"), golden("coeffs_K57_s1.npy")]
        assert all(np.array_equal(c, c[::-1]) == np.array_equal(e[1], e[1][::-1]) for c, e in zip(new, entries))
        ncptr = (_abi.P_double * len(new))(*[c.ctypes.data_as(_abi.P_double) for c in new])
        _lib.check(L.gli_eri_batch_run(plan, ncptr, None), "run new coefficients")
        assert L.gli_synchronize() == 0
        for e, c in zip(entries, new):
            ref = eri_values_device(c, e[0], e[2], 1e-16, single_method)
            got = e[4].to_numpy(np.float64, len(e[2]))
            if method in (0, 6):
                assert np.array_equal(got, ref)
            else:
                assert np.max(np.abs(got - ref)) <= 1e-10
        # a coefficient set without mirror symmetry does not fit a plan built for symmetric ones
        bad = [c.copy() for c in new]
        bad[1][0] += 1e-3
        bcptr = (_abi.P_double * len(bad))(*[c.ctypes.data_as(_abi.P_double) for c in bad])
        assert L.gli_eri_batch_run(plan, bcptr, None) == -1
        assert b"symmetry" in L.gli_last_error()
        L.gli_clear_error()
    finally:
        L.gli_eri_batch_destroy(plan)


@pytest.mark.parametrize("method", [0, 6])
def test_batch_plan_indices_outside_grid_in_whole_windows(L, method):
    """indices outside the grid inside whole windows of 8192 (and in the partial last one) of a plan with
    cached, window-sorted positions give NaN there and the single-call values everywhere else"""
    say("batch plan, indices outside the grid, method", method)
    saved = L.gli_get_tuning(b"batch_positions"), L.gli_get_tuning(b"batch_sorted")
    assert L.gli_set_tuning(b"batch_positions", 1) == 0
    assert L.gli_set_tuning(b"batch_sorted", 1) == 0
    N4 = 5 ** 12
    idx = golden("golden_enum_n5.npz")["eri_indices"][:3 * 8192 + 777].astype(np.int64).copy()
    bad_at = np.array([0, 1, 4000, 8191, 8192, 12345, 16383, 20000, len(idx) - 1])
    idx[bad_at] = np.array([-1, N4, -5, N4 + 7, 2 ** 40, -2 ** 40, N4, -1, N4 + 1])
    entries = _batch_setup([(sym(5), golden("coeffs_K29_s0.npy"), idx)])
    cptr, Ks, grids, dptr, nums, vptr = _batch_arrays(entries)
    plan = C.c_void_p()
    try:
        _lib.check(L.gli_eri_batch_create(1, cptr, Ks.ctypes.data_as(_abi.P_glong), grids.ctypes.data_as(_abi.P_glong),
                                          dptr, nums.ctypes.data_as(_abi.P_glong), 1e-16, method, vptr, 1,
                                          C.byref(plan)), "gli_eri_batch_create")
    finally:
        L.gli_set_tuning(b"batch_positions", saved[0])
        L.gli_set_tuning(b"batch_sorted", saved[1])
    try:
        ref = eri_values_device(entries[0][1], entries[0][0], idx, 1e-16, method)
        assert np.all(np.isnan(ref[bad_at])) and np.isnan(ref).sum() == len(bad_at)
        for rep in range(2):
            _lib.check(L.gli_eri_batch_run(plan, None, None), "gli_eri_batch_run")
            assert L.gli_synchronize() == 0
            got = entries[0][4].to_numpy(np.float64, len(idx))
            assert np.array_equal(got, ref, equal_nan=True), rep
    finally:
        L.gli_eri_batch_destroy(plan)


def test_batch_one_shot_entry(L):
    """gli_eri_values_batch_device (a plan without autotuning, run once) equals the single-grid path"""
    say("batch one shot")
    specs = [(sym(5), golden("coeffs_K29_s0.npy"), None), (sym(3), golden("coeffs_K57_s0.npy"), None)]
    entries = _batch_setup(specs)
    cptr, Ks, grids, dptr, nums, vptr = _batch_arrays(entries)
    _lib.check(L.gli_eri_values_batch_device(len(entries), cptr, Ks.ctypes.data_as(_abi.P_glong), grids.ctypes.data_as(_abi.P_glong),
                                             dptr, nums.ctypes.data_as(_abi.P_glong), 1e-16, 0, vptr), "batch")
    for e in entries:
        ref = eri_values_device(e[1], e[0], e[2], 1e-16, 0)
        assert np.array_equal(e[4].to_numpy(np.float64, len(e[2])), ref)


def test_precision_ladder_7(L):
    """the comparison methods on every unique 7^3 element against FP64: the 5-digit split stays within
    1e-10, the BF16 hi/lo split, plain FP32 and the FP32 hi/lo split land near their unit roundoffs
    (bounds 1e-3 and 1e-4 catch a wrong accumulator layout, which gives O(1) errors), and the ladder is
    ordered"""
    say("precision ladder 7")
    c = golden("coeffs_K57_s1.npy")
    d, n = eri_enumerate_device(7)
    idx = d.to_numpy(np.int64, n)
    v64 = eri_values_device(c, sym(7), idx, 1e-16, 0)
    err = {}
    for method, bound in ((1, 1e-10), (3, 1e-3), (4, 1e-4), (5, 1e-4)):
        got = eri_values_device(c, sym(7), idx, 1e-16, method)
        err[method] = float(np.max(np.abs(got - v64)))
        say("method", method, "max abs err", err[method])
        assert err[method] <= bound, (method, err[method])
    assert err[1] < err[4] and err[1] < err[3] and err[1] < err[5]
    assert err[3] > 0 and err[4] > 0 and err[5] > 0


# ---------------------------------------------------------------------------
# dense views on the GPU: canonicalisation of arbitrary tuples + lookup

def test_dense_fill_7_sub_box_bit_exact(L, orc):
    """fill_tensor of the full 7^3 sparse set on an off-centre 3x2x4 dense box equals the oracle's
    fill (reference canonicalisation with bias placements, eri_integrals.c:837-967) bit for bit"""
    say("dense fill 7 sub box")
    c = golden("coeffs_K57_s0.npy")
    eri = gli.compute_sparse_eri_gausslet_integrals(c, sym(7), 1e-16)
    box = ((-3, 3), (-1, 2), (0, 4))
    got = eri.fill_tensor(box)
    ref = orc.fill_dense_eri(sym(7), eri.four_indices, eri.integral_values, box)
    assert got.shape == ref.shape
    assert np.array_equal(got.ravel(), np.asarray(ref).ravel())


def test_apply_5_columns_and_linearity(L, orc):
    """apply_sparse_eri_tensor on 5^3 (2.4e8 tuples): a unit state at (j0, l0) returns the column of
    values (i j0 | k l0), checked through the oracle's canonical keys, and apply is linear"""
    say("apply 5")
    c = golden("coeffs_K43_s0.npy")
    eri = gli.compute_sparse_eri_gausslet_integrals(c, sym(5), 1e-16)
    keys, vals = eri.four_indices, eri.integral_values
    N = 125
    rng = np.random.default_rng(41)
    jl = rng.integers(0, N * N, size=2)
    X = np.zeros((N * N, 4))
    X[jl[0], 0] = 1.0
    X[jl[1], 1] = 1.0
    X[:, 2] = rng.standard_normal(N * N)
    X[:, 3] = 2.0 * X[:, 0] - 3.0 * X[:, 2]
    out = eri.apply(X)
    for u in range(2):
        j0, l0 = divmod(int(jl[u]), N)
        ik = np.arange(N * N)
        i, k = ik // N, ik % N
        t = ((i * N + j0) * N + k) * N + l0
        ck = orc.eri_canonical_keys(sym(5), t)
        pos = np.clip(np.searchsorted(keys, ck), 0, len(keys) - 1)
        col = np.where(keys[pos] == ck, vals[pos], 0.0)
        assert np.array_equal(out[:, u], col), u
    scale = np.max(np.abs(out[:, 2])) + 1.0
    assert np.max(np.abs(out[:, 3] - (2.0 * out[:, 0] - 3.0 * out[:, 2]))) <= 1e-13 * scale


def test_hdf5_streamed_save_from_device(L, tmp_path):
    """the 9^3 set (39M entries, two 2^25-entry slices) written straight from device memory loads
    back bit-identical in the reference's layout"""
    say("hdf5 streamed save 9")
    from gausslectronitegrate_amd import pymodule
    c = golden("coeffs_K57_s1.npy")
    d, n = eri_enumerate_device(9)
    dv = DevArray(8 * n)
    g = np.array([-4, 9, -4, 9, -4, 9], dtype=np.int64)
    _lib.check(L.gli_eri_values_device(c.ctypes.data_as(_abi.P_double), len(c), g.ctypes.data_as(_abi.P_glong), d.ptr, n,
                                       1e-16, 0, dv.ptr), "values")
    out = str(tmp_path / "eri9.hdf5")
    H = pymodule._hdf5_lib()
    assert H.gli_eri_save_hdf5(out.encode(), 9, 1e-16, c.ctypes.data_as(_abi.P_double), len(c), d.ptr, dv.ptr, n, 1) == 0
    back = gli.load_sparse_eri_gausslet_integrals(out)
    assert back.num_entries == n and back.grid == sym(9)
    assert np.array_equal(back.four_indices, d.to_numpy(np.int64, n))
    assert np.array_equal(back.integral_values, dv.to_numpy(np.float64, n))


def test_batch_split_products_in_graph(L):
    """a mixed-method plan without autotuning keeps the split-precision products inside its captured
    graph (per-grid digit extraction in a plan-owned workspace): every replay equals the single-grid
    mixed path bit for bit"""
    say("batch split in graph")
    specs = [(sym(5), golden("coeffs_K29_s0.npy"), None), (sym(7), golden("coeffs_K57_s1.npy"), None)]
    entries = _batch_setup(specs)
    cptr, Ks, grids, dptr, nums, vptr = _batch_arrays(entries)
    plan = C.c_void_p()
    _lib.check(L.gli_eri_batch_create(len(entries), cptr, Ks.ctypes.data_as(_abi.P_glong), grids.ctypes.data_as(_abi.P_glong),
                                      dptr, nums.ctypes.data_as(_abi.P_glong), 1e-16, 1, vptr, 0, C.byref(plan)),
               "gli_eri_batch_create")
    try:
        refs = [eri_values_device(e[1], e[0], e[2], 1e-16, 1) for e in entries]
        for rep in range(3):
            _lib.check(L.gli_eri_batch_run(plan, None, None), "run")
            assert L.gli_synchronize() == 0
            for e, ref in zip(entries, refs):
                assert np.array_equal(e[4].to_numpy(np.float64, len(e[2])), ref), rep
            eri_values_device(entries[1][1], entries[1][0], entries[1][2][:1000], 1e-16, 1)   # eager work between replays
    finally:
        L.gli_eri_batch_destroy(plan)


def test_batched_single_elements(L, orc):
    """compute_eri/erida_gausslet_integral_batch on seeded random point tuples (negative coordinates,
    repeated points, tuples of different spans) against the oracle's single-element sums
    (eri_integrals.c:133-232, erida_integrals.c:119-201), with both ERI methods; the notebook value of
    four orbitals at the origin (eri_integrals.ipynb:59); shape and span errors"""
    say("batched single elements")
    c = golden("coeffs_K29_s0.npy")
    rng = np.random.default_rng(31)
    pts = rng.integers(-3, 4, size=(120, 4, 3))
    pts[:10, 1] = pts[:10, 0]                      # repeated points
    pts[10:20] = pts[10:20] // 3                   # small spans
    ref = np.array([orc.eri_single(c, pts[e], 1e-12) for e in range(len(pts))])
    for method in (None, 0, 2):
        got = gli.compute_eri_gausslet_integral_batch(c, pts, 1e-12, method)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.abs(ref).max()), method
    pairs = rng.integers(-4, 5, size=(60, 2, 3))
    ref2 = np.array([orc.erida_single(c, pairs[e], 1e-12) for e in range(len(pairs))])
    got2 = gli.compute_erida_gausslet_integral_batch(c, pairs, 1e-12)
    assert np.max(np.abs(got2 - ref2)) <= 1e-13 * max(1.0, np.abs(ref2).max())
    g10 = golden("ref_eri.npz")["coeffs"]
    v = gli.compute_eri_gausslet_integral_batch(g10, np.zeros((2, 4, 3), dtype=np.int64), 1e-16)
    assert v[0] == pytest.approx(1.6252009000829197, abs=1e-13) and v[1] == v[0]
    assert len(gli.compute_eri_gausslet_integral_batch(c, np.zeros((0, 4, 3), dtype=np.int64), 1e-12)) == 0
    with pytest.raises(ValueError):
        gli.compute_eri_gausslet_integral_batch(c, np.zeros((3, 2, 3), dtype=np.int64), 1e-12)
    far = np.zeros((1, 4, 3), dtype=np.int64)
    far[0, 3] = (40, 40, 40)
    with pytest.raises(RuntimeError):
        gli.compute_eri_gausslet_integral_batch(c, far, 1e-12)


def test_concurrent_host_threads(L):
    """entry points called from several Python threads at once (ctypes releases the GIL; the library
    serialises its stream work) give the same values as sequential calls"""
    say("concurrent host threads")
    from concurrent.futures import ThreadPoolExecutor
    c = golden("coeffs_K29_s0.npy")
    rng = np.random.default_rng(8)
    jobs = [rng.integers(-2, 3, size=(64, 4, 3)) for _ in range(8)]

    def work(q):
        if q % 2 == 0:
            return gli.compute_eri_gausslet_integral_batch(c, jobs[q], 1e-12)
        return gli.compute_sparse_eri_gausslet_integrals(c, sym(5), 1e-12).integral_values[::97]

    seq = [work(q) for q in range(8)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        par = list(ex.map(work, range(8)))
    for a, b in zip(seq, par):
        assert np.array_equal(a, b)


def test_empty_and_ragged_inputs(L, orc):
    """empty index lists through the reference entry points (every ERI method) and the device calls,
    a batch whose first grid has no elements, and list lengths around the 64-lane and 512-element
    boundaries against the oracle"""
    say("empty and ragged")
    c = golden("coeffs_K29_s0.npy")
    gd = _abi.GaussletData()
    gd.coefficients = c.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(len(c) - 1) // 2
    gd.indices.num = len(c)
    g5 = _abi.Grid.of(sym(5))
    empty = np.zeros(1, dtype=np.int64)
    for method in range(6):
        assert L.gli_set_eri_method(method) == 0
        res = _abi.SparseEri()
        L.compute_sparse_eri_gausslet_integrals(C.byref(gd), C.byref(g5), empty.ctypes.data_as(_abi.P_glong), 0, 1e-12,
                                                C.byref(res))
        assert res.num_entries == 0
        L.delete_sparse_eri_gausslet_integrals(C.byref(res))
    assert L.gli_set_eri_method(0) == 0
    resd = _abi.SparseErida()
    L.compute_sparse_erida_gausslet_integrals(C.byref(gd), C.byref(g5), empty.ctypes.data_as(_abi.P_glong), 0, 1e-12,
                                              C.byref(resd))
    assert resd.num_entries == 0
    L.delete_sparse_erida_gausslet_integrals(C.byref(resd))
    P = c.ctypes.data_as(_abi.P_double)
    g = np.array([v for ax in sym(5) for v in ax], dtype=np.int64)
    dv = DevArray(8)
    for method in (0, 1, 2):
        assert L.gli_eri_values_device(P, len(c), g.ctypes.data_as(_abi.P_glong), None, 0, 1e-12, method, dv.ptr) == 0
    # ragged lengths
    full = golden("golden_enum_n5.npz")["eri_indices"]
    for m in (1, 63, 64, 65, 511, 513, 4097):
        idx = full[:: max(1, len(full) // m)][:m]
        ref = orc.eri_sparse(c, sym(5), idx, 1e-12)
        for method in (0, 2):
            got = eri_values_device(c, sym(5), idx, 1e-12, method)
            assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.abs(ref).max()), (m, method)
    # a batch whose first grid is empty
    d3, n3 = eri_enumerate_device(3)
    out3 = DevArray(8 * n3)
    outE = DevArray(8)
    cptr = (_abi.P_double * 2)(P, P)
    Ks = np.array([len(c), len(c)], dtype=np.int64)
    grids = np.array([v for ax in sym(5) for v in ax] + [v for ax in sym(3) for v in ax], dtype=np.int64)
    dptr = (C.c_void_p * 2)(d3.ptr, d3.ptr)
    nums = np.array([0, n3], dtype=np.int64)
    vptr = (C.c_void_p * 2)(outE.ptr, out3.ptr)
    _lib.check(L.gli_eri_values_batch_device(2, cptr, Ks.ctypes.data_as(_abi.P_glong), grids.ctypes.data_as(_abi.P_glong),
                                             dptr, nums.ctypes.data_as(_abi.P_glong), 1e-12, 0, vptr), "batch")
    want = eri_values_device(c, sym(3), d3.to_numpy(np.int64, n3), 1e-12, 0)
    assert np.array_equal(out3.to_numpy(np.float64, n3), want)


def test_additive_api_errors(L):
    """bad arguments of the additive entry points return -1 with a message and leave the library usable"""
    say("api errors")
    c = golden("coeffs_K57_s0.npy")
    d, n = eri_enumerate_device(3)
    dv = DevArray(8 * n)
    P = c.ctypes.data_as(_abi.P_double)
    bad_grid = np.array([-1, 0, -1, 3, -1, 3], dtype=np.int64)
    assert L.gli_eri_values_device(P, len(c), bad_grid.ctypes.data_as(_abi.P_glong), d.ptr, n, 1e-16, 0, dv.ptr) == -1
    assert L.gli_last_error()
    L.gli_clear_error()
    g = np.array([-1, 3, -1, 3, -1, 3], dtype=np.int64)
    assert L.gli_eri_values_device(P, len(c), g.ctypes.data_as(_abi.P_glong), d.ptr, n, 1e-16, 7, dv.ptr) == -1
    assert L.gli_eri_values_device(P, len(c) - 1, g.ctypes.data_as(_abi.P_glong), d.ptr, n, 1e-16, 0, dv.ptr) == -1
    assert L.gli_eri_values_device(P, len(c), g.ctypes.data_as(_abi.P_glong), d.ptr, n, -1.0, 0, dv.ptr) == -1
    p = C.c_void_p()
    num = C.c_int64()
    assert L.gli_eri_enumerate_device(4, C.byref(p), C.byref(num)) == -1   # even grids have no centred enumeration
    plan = C.c_void_p()
    cptr = (_abi.P_double * 1)(P)
    Ks = np.array([len(c)], dtype=np.int64)
    dptr = (C.c_void_p * 1)(d.ptr)
    nums = np.array([n], dtype=np.int64)
    vptr = (C.c_void_p * 1)(dv.ptr)
    assert L.gli_eri_batch_create(1, cptr, Ks.ctypes.data_as(_abi.P_glong), g.ctypes.data_as(_abi.P_glong), dptr,
                                  nums.ctypes.data_as(_abi.P_glong), 1e-16, 2, vptr, 0, C.byref(plan)) == -1
    assert b"batch" in L.gli_last_error()
    L.gli_clear_error()
    # still usable
    _lib.check(L.gli_eri_values_device(P, len(c), g.ctypes.data_as(_abi.P_glong), d.ptr, n, 1e-16, 0, dv.ptr), "values")
    assert np.array_equal(dv.to_numpy(np.float64, n), eri_values_device(c, sym(3), d.to_numpy(np.int64, n), 1e-16, 0))
