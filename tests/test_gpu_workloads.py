"""GPU parity at the workload sizes of the benchmark configurations, against vectors of the
reference build (tests/golden/, written by tests/golden/make_fixtures.py in the development
container):

* the mixed-size batch (5^3 K29, 7^3 K43, 9^3 K57 seed 1, 11^3 K57) as one plan over the full
  enumerated lists, replayed several times, against the reference's values of seeded canonical
  samples of every grid (eri_integrals.c:544-678 per grid);
* the reference's sparse-completeness, projection and apply tests on the GPU path
  (test/test_eri_integrals.c:125-419);
* every ERI-DA pair of the 13^3 grid (erida_integrals.c:429-542);
* the precision study at 13^3 on the reference's sample values;
* the complete 9^3 and 11^3 enumeration lists against the count and SHA-256 of the reference
  build's lists, and the 13^3 list against reference-canonicalised members (eri_integrals.c:382-526);
* the nuclear matrix for more nuclei than fit one LDS chunk (nuclear_integrals.c:133-266);
* an unchanged C caller of the reference API, compiled against include/ and linked with -lgli_amd.

Every call goes through the C ABI of lib/libgli_amd.so."""
import ctypes as C
import hashlib
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import gausslectronitegrate_amd as gli
from gausslectronitegrate_amd import _abi, _lib
from tests import burnside
from tests.conftest import REPO, golden
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


def g6(n):
    return np.array([v for ax in sym(n) for v in ax], dtype=np.int64)


# ---------------------------------------------------------------------------
# the mixed-size batch over the full lists

# (n, K, coefficient seed, spacing a): the benchmark's batch; values are scaled by 1/a
BATCH = [(5, 29, 0, 0.5), (7, 43, 0, 0.75), (9, 57, 1, 1.0), (11, 57, 0, 1.5)]


@pytest.mark.parametrize("method,bound", [(0, 1e-13), (1, 1e-10), (6, 1e-13)])
def test_batch_workload_vs_reference(L, method, bound):
    """One plan over the full enumerated lists of all four grids (304,231,528 elements), its graph
    replayed three times with eager work in between; every replay's values at the reference build's
    seeded canonical samples equal the reference values times 1/a within bound/a (FP64 and separable 1e-13,
    mixed 1e-10).  Every sampled canonical index must also be present in the enumerated list."""
    say("batch workload, method", method)
    n_g = len(BATCH)
    coeffs, lists, vals, samples = [], [], [], []
    for n, K, seed, a in BATCH:
        g = golden("golden_sample_n%d_K%ds%d.npz" % (n, K, seed))
        c = golden("coeffs_K%d_s%d.npy" % (K, seed))
        assert np.array_equal(g["coeffs"], c)
        d, num = eri_enumerate_device(n)
        assert num == burnside.orbit_count(n)
        host = d.to_numpy(np.int64, num)
        pos = np.searchsorted(host, g["eri_indices"])
        assert np.all(pos < num) and np.array_equal(host[pos], g["eri_indices"]), n
        del host
        coeffs.append(np.ascontiguousarray(c))
        lists.append((d, num))
        vals.append(DevArray(8 * num))
        samples.append((pos, g["eri_values"] / a, bound / a))
    cptr = (_abi.P_double * n_g)(*[c.ctypes.data_as(_abi.P_double) for c in coeffs])
    Ks = np.array([len(c) for c in coeffs], dtype=np.int64)
    grids = np.concatenate([g6(b[0]) for b in BATCH])
    dptr = (C.c_void_p * n_g)(*[d.ptr for d, _ in lists])
    nums = np.array([num for _, num in lists], dtype=np.int64)
    assert int(nums.sum()) == 304231528
    vptr = (C.c_void_p * n_g)(*[v.ptr for v in vals])
    scales = np.array([1.0 / b[3] for b in BATCH])
    plan = C.c_void_p()
    _lib.check(L.gli_eri_batch_create(n_g, cptr, Ks.ctypes.data_as(_abi.P_glong), grids.ctypes.data_as(_abi.P_glong), dptr,
                                      nums.ctypes.data_as(_abi.P_glong), 1e-16, method, vptr, 1, C.byref(plan)),
               "gli_eri_batch_create")
    try:
        for rep in range(3):
            for v, (_, num) in zip(vals, lists):   # clear, so that each replay must rewrite every value
                _lib.check(L.gli_memset(v.ptr, 0xff, 8 * num), "memset")
            _lib.check(L.gli_eri_batch_run(plan, None, scales.ctypes.data_as(C.POINTER(C.c_double))), "gli_eri_batch_run")
            assert L.gli_synchronize() == 0
            for (n, K, seed, a), v, (_, num), (pos, ref, tol) in zip(BATCH, vals, lists, samples):
                got = v.to_numpy(np.float64, num)
                assert np.all(np.isfinite(got)), (rep, n)
                err = float(np.max(np.abs(got[pos] - ref)))
                say("  replay", rep, "grid", n, "max abs err vs reference x 1/a: %.3g" % err)
                assert err <= tol, (rep, n, err)
                del got
            # eager work on the library stream between replays
            eri_values_device(coeffs[0], sym(5), golden("golden_sample_n5_K29s0.npz")["eri_indices"][:100], 1e-16, 0)
    finally:
        L.gli_eri_batch_destroy(plan)


# ---------------------------------------------------------------------------
# the reference's sparse-completeness test, on the GPU path

def test_eri_sparse_completeness_5(L, orc):
    """test_eri_integral_sparse_completeness (test/test_eri_integrals.c:342-419): G10 coefficients of
    test_eri_integrals.hdf5, a 5^3 grid, tol 1e-5; the dense 125^4 tensor filled from the
    symmetry-reduced set has no zero entry.  A seeded sample of the sparse values is also checked
    against the oracle at the same tolerance (the tolerance skips of eri_integrals.c:629-658)."""
    say("eri completeness 5")
    c = golden("ref_eri.npz")["coeffs"]
    eri = gli.compute_sparse_eri_gausslet_integrals(c, sym(5), 1e-5)
    assert eri.num_entries == burnside.orbit_count(5)
    dense = eri.fill_tensor(sym(5))
    assert dense.shape == (125,) * 4
    assert np.count_nonzero(dense == 0) == 0
    del dense
    rng = np.random.default_rng(342)
    s = np.sort(rng.choice(eri.num_entries, size=2000, replace=False))
    ref = orc.eri_sparse(c, sym(5), eri.four_indices[s], 1e-5)
    assert np.max(np.abs(eri.integral_values[s] - ref)) <= 1e-13


# ---------------------------------------------------------------------------
# the reference's ERI projection and apply tests, on the GPU path

def _g10_3_tol5():
    """the sparse set both reference tests build: G10 coefficients of test_eri_integrals.hdf5, the
    3^3 grid, tol 1e-5"""
    c = golden("ref_eri.npz")["coeffs"]
    return gli.compute_sparse_eri_gausslet_integrals(c, sym(3), 1e-5)


def test_project_eri_reference(L):
    """test_project_eri_integrals (test/test_eri_integrals.c:125-231): four states with the test's
    pseudo-random entries 0.1 sin(sqrt(3) pi ((557 i) mod 311) - 8) / (1.5 + cos(sqrt(5) pi ((541 i)
    mod 229))); the projection equals the dense tensor contracted with the basis on all four axes,
    uniform distance <= 1e-12"""
    say("project eri reference")
    eri = _g10_3_tol5()
    i = np.arange(27 * 4)
    basis = (0.1 * np.sin(np.sqrt(3) * np.pi * ((557 * i) % 311) - 8) /
             (1.5 + np.cos(np.sqrt(5) * np.pi * ((541 * i) % 229)))).reshape(27, 4)
    proj = eri.project(basis)
    dense = eri.fill_tensor(sym(3))
    ref = np.einsum("ip,jq,kr,ls,ijkl->pqrs", basis, basis, basis, basis, dense, optimize=True)
    assert np.max(np.abs(proj - ref)) <= 1e-12


def test_apply_eri_reference(L):
    """test_apply_eri_integrals (test/test_eri_integrals.c:233-340): three two-particle states with the
    test's pseudo-random entries 0.12 sin(sqrt(7) pi ((523 i) mod 379) - 4) / (1.8 + cos(sqrt(11) pi
    ((563 i) mod 281))); ret[(i,k), s] = sum_jl (ij|kl) states[(j,l), s] over the dense tensor, uniform
    distance <= 1e-12"""
    say("apply eri reference")
    eri = _g10_3_tol5()
    i = np.arange(27 * 27 * 3)
    states = (0.12 * np.sin(np.sqrt(7) * np.pi * ((523 * i) % 379) - 4) /
              (1.8 + np.cos(np.sqrt(11) * np.pi * ((563 * i) % 281)))).reshape(27 * 27, 3)
    out = eri.apply(states)
    dense = eri.fill_tensor(sym(3))
    ref = np.einsum("ijkl,jls->iks", dense, states.reshape(27, 27, 3)).reshape(27 * 27, 3)
    assert np.max(np.abs(out - ref)) <= 1e-12


# ---------------------------------------------------------------------------
# ERI-DA at 13^3

def test_erida_13_all_pairs_vs_reference(L):
    """all 455 unique ERI-DA pairs of the 13^3 grid (K = 57, seed 0, tol 1e-16): the list equals the
    reference build's and every value is within 1e-13 of it"""
    say("erida 13 all pairs")
    g = golden("golden_erida_n13_K57s0.npz")
    e = gli.compute_sparse_erida_gausslet_integrals(g["coeffs"], sym(13), float(g["tol"]))
    assert e.num_entries == 455 == math.comb(15, 3)
    assert np.array_equal(e.two_indices, g["erida_indices"])
    err = float(np.max(np.abs(e.integral_values - g["erida_values"])))
    say("  max abs err %.3g" % err)
    assert err <= 1e-13


# ---------------------------------------------------------------------------
# the precision study at 13^3

def test_precision_ladder_13_vs_reference(L):
    """FP64, the separable expansion, the 5-digit split (mixed), BF16 hi/lo with FP32 accumulation (bf16x2), FP32
    and FP32 hi/lo on the 2048 seeded canonical 13^3 indices with the reference build's values: FP64 and
    separable within 1e-13,
    mixed within 1e-10, the comparison methods near their unit roundoffs (bounds 1e-3 and 1e-4 catch
    a wrong accumulator layout, which gives O(1) errors), and the ladder ordered"""
    say("precision ladder 13")
    g = golden("golden_sample_n13_K57s0.npz")
    err = {}
    for method, bound in ((0, 1e-13), (6, 1e-13), (1, 1e-10), (3, 1e-3), (4, 1e-4), (5, 1e-4)):
        got = eri_values_device(g["coeffs"], sym(13), g["eri_indices"], float(g["tol"]), method)
        err[method] = float(np.max(np.abs(got - g["eri_values"])))
        say("  method", method, "max abs err vs reference %.3g" % err[method])
        assert err[method] <= bound, (method, err[method])
    assert err[0] <= err[1] < min(err[3], err[4], err[5])
    assert err[3] > 0 and err[4] > 0 and err[5] > 0


# ---------------------------------------------------------------------------
# enumeration against the reference build's complete lists

@pytest.mark.parametrize("n", [9, 11])
def test_enumeration_digest_vs_reference(L, n):
    """the complete GPU list of the n^3 grid has the reference build's count and SHA-256; on a
    mismatch, the strided subsample, head and tail locate the difference"""
    say("enumeration digest", n)
    ref = golden("golden_enum_digests.npz")
    d, num = eri_enumerate_device(n)
    got = d.to_numpy(np.int64, num)
    d.free()
    stride = int(ref["stride"])
    assert num == int(ref["n%d_count" % n])
    assert np.array_equal(got[:1024], ref["n%d_head" % n])
    assert np.array_equal(got[-1024:], ref["n%d_tail" % n])
    assert np.array_equal(got[::stride], ref["n%d_strided" % n])
    assert hashlib.sha256(got.astype("<i8").tobytes()).hexdigest() == str(ref["n%d_sha256" % n])


def test_enumeration_13_sorted_canonical_and_complete(L, orc):
    """13^3 (1,250,296,187 entries, 10 GB): the Burnside count, strictly ascending over the whole list,
    every one of the reference build's 2048 seeded canonical indices present, and a seeded sample of
    20,000 entries canonical under the oracle's orbit minimum"""
    say("enumeration 13 full")
    d, num = eri_enumerate_device(13)
    assert num == burnside.orbit_count(13) == 1250296187
    got = d.to_numpy(np.int64, num)
    d.free()
    assert got[0] >= 0 and got[-1] < 2197 ** 4
    chunk = 1 << 26
    for s in range(0, num - 1, chunk):
        seg = got[s:min(num, s + chunk + 1)]
        assert np.all(seg[1:] > seg[:-1]), s
    members = golden("golden_sample_n13_K57s0.npz")["eri_indices"]
    pos = np.searchsorted(got, members)
    assert np.all(pos < num) and np.array_equal(got[pos], members)
    rng = np.random.default_rng(13)
    sample = np.sort(got[rng.choice(num, size=20000, replace=False)])
    del got
    assert np.array_equal(orc.eri_canonical_keys(sym(13), sample), sample)


# ---------------------------------------------------------------------------
# nuclear matrix with many nuclei

def nuclear_bound(spread):
    """With many nuclei an entry of the nuclear matrix sums ~10^5 signed terms per nucleus, and the
    reference's own result moves when the same terms are associated differently: the fixture holds
    that spread of the reference build (its matrix over all nuclei against the sum of its per-nucleus
    matrices).  Another summation order is compared within four times the spread, and never within
    less than the reference tests' 1e-13."""
    return max(4.0 * float(spread), 1e-13)


def test_nuclear_many_nuclei(L):
    """100 nuclei (two LDS chunks of the potential kernel; the reference loops over any num_nuclei,
    nuclear_integrals.c:234-244) on a 3x4x5 grid against the reference build's matrix, through the
    Python surface and the struct-returning C entry point; the first 64 of them alone (one chunk) as
    well; and the matrix of all 100 equals the sum of those of the first 64 and the last 36 (a dropped
    or repeated chunk would be off by a whole nucleus's contribution, O(1))"""
    say("nuclear 100 nuclei")
    g = golden("golden_nuclear_many.npz")
    c, tol, pos, chg = g["many_coeffs"], float(g["many_tol"]), g["many_pos"], g["many_charges"]
    grid = tuple(tuple(int(v) for v in ax) for ax in g["many_grid"])
    m = gli.compute_nuclear_gausslet_integrals(c, grid, pos, chg, tol)
    err, bound = float(np.max(np.abs(m - g["many_m100"]))), nuclear_bound(g["many_spread100"])
    say("  100 nuclei: max |ref| %.3g, max abs err %.3g, bound %.3g" % (np.abs(g["many_m100"]).max(), err, bound))
    assert err <= bound
    assert np.array_equal(m, m.T)
    m64 = gli.compute_nuclear_gausslet_integrals(c, grid, pos[:64], chg[:64], tol)
    err64, bound64 = float(np.max(np.abs(m64 - g["many_m64"]))), nuclear_bound(g["many_spread64"])
    say("  64 nuclei: max abs err %.3g, bound %.3g" % (err64, bound64))
    assert err64 <= bound64
    m36 = gli.compute_nuclear_gausslet_integrals(c, grid, pos[64:], chg[64:], tol)
    assert np.max(np.abs(m - (m64 + m36))) <= bound
    gd = _abi.GaussletData()
    cc = np.ascontiguousarray(c)
    gd.coefficients = cc.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(len(c) - 1) // 2
    gd.indices.num = len(c)
    nuc = (_abi.Nucleus * 100)()
    for q in range(100):
        for a in range(3):
            nuc[q].pos[a] = pos[q][a]
        nuc[q].charge = int(chg[q])
    out = _abi.NuclearIntegrals()
    L.compute_nuclear_gausslet_integrals(C.byref(gd), C.byref(_abi.Grid.of(grid)), nuc, 100, tol, C.byref(out))
    assert out.num_nuclei == 100
    assert np.array_equal(np.ctypeslib.as_array(out.integral_values, shape=m.shape), m)
    L.delete_nuclear_gausslet_integrals(C.byref(out))


# ---------------------------------------------------------------------------
# projection in chunks of product states

def test_project_chunked_bit_identical(L):
    """project_sparse_eri_gausslet_integrals with the product states sent in many small chunks (a
    16 KiB chunk holds 2 of the 36 state columns of a 3^3 grid) equals the one-chunk result bit for
    bit and the dense contraction within 1e-12"""
    say("project chunked")
    c = golden("coeffs_K57_s0.npy")
    eri = gli.compute_sparse_eri_gausslet_integrals(c, sym(3), 1e-16)
    rng = np.random.default_rng(6)
    basis = rng.standard_normal((27, 6))
    saved = L.gli_get_tuning(b"project_chunk_kib")
    try:
        one = eri.project(basis)
        assert L.gli_set_tuning(b"project_chunk_kib", 16) == 0
        many = eri.project(basis)
        assert L.gli_set_tuning(b"project_chunk_kib", 0) == -1
    finally:
        L.gli_clear_error()
        L.gli_set_tuning(b"project_chunk_kib", saved)
    assert np.array_equal(one, many)
    dense = eri.fill_tensor(sym(3))
    ref = np.einsum("ip,jq,kr,ls,ijkl->pqrs", basis, basis, basis, basis, dense, optimize=True)
    assert np.max(np.abs(many - ref)) <= 1e-12 * max(1.0, np.abs(ref).max())


# ---------------------------------------------------------------------------
# a C caller of the reference API

def test_c_caller(L, tmp_path):
    """tests/c/gli_caller (plain C11, compiled against include/ and linked with -lgli_amd) runs the
    reference entry points on the 3^3 grid: its ERI list and values equal the reference build's
    golden set, lookups return the values, single-element ERIs agree with the sparse ones, the
    ERI-DA set and kinetic matrix match the golden set, and the nuclear matrix of 70 nuclei matches
    the reference build's within the bound of nuclear_bound"""
    say("C caller")
    exe = os.path.join(REPO, "tests", "c", "gli_caller")
    if not os.path.exists(exe):
        pytest.fail("tests/c/gli_caller is not built (python -c 'import __graft_entry__ as g; g.build()')")
    g = golden("golden_n3_K57s0.npz")
    gn = golden("golden_nuclear_many.npz")
    c = np.ascontiguousarray(g["coeffs"], dtype="<f8")
    assert np.array_equal(c, gn["caller_coeffs"])
    nuc = gn["caller_nuclei"]
    (tmp_path / "c.f64").write_bytes(c.tobytes())
    (tmp_path / "n.f64").write_bytes(np.ascontiguousarray(nuc, dtype="<f8").tobytes())
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(tmp_path / "c.f64"), str(tmp_path / "n.f64"), "3", str(out)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stdout

    def rd(name, dt):
        return np.fromfile(str(out / name), dtype=dt)

    idx = rd("eri_indices.i64", "<i8")
    vals = rd("eri_values.f64", "<f8")
    assert np.array_equal(idx, g["eri_indices"])
    assert np.max(np.abs(vals - g["eri_values"])) <= 1e-13
    assert np.array_equal(rd("eri_lookup.f64", "<f8"), vals)
    single = rd("eri_single.f64", "<f8")
    assert len(single) == (len(idx) + 6) // 7
    assert np.max(np.abs(single - vals[::7])) <= 1e-13
    assert np.array_equal(rd("erida_indices.i64", "<i8"), g["erida_indices"])
    assert np.max(np.abs(rd("erida_values.f64", "<f8") - g["erida_values"])) <= 1e-13
    assert np.max(np.abs(rd("kinetic.f64", "<f8").reshape(27, 27) - g["kinetic"])) <= 1e-13
    nm = rd("nuclear.f64", "<f8").reshape(27, 27)
    bound = nuclear_bound(gn["caller_spread70"])
    assert np.max(np.abs(nm - gn["caller_m70"])) <= bound
    assert abs(rd("nuclear_single.f64", "<f8")[0] - nm[0, 26]) <= bound
