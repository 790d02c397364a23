"""Pins of the oracle (CPU restatement) before it is trusted as the checker.

Sources of truth, none of them the oracle's own formulas:
  * the reference's HDF5 test data (test/data/*.hdf5, tolerances of
    test/test_*.c) and examples/eri_3x3x3.hdf5 (bit-exact index list);
  * values stored in the reference notebooks (examples/*.ipynb);
  * the reference build oracle/_ref (where the reference tree is present);
  * mathematics: O_h group axioms, Burnside orbit counts, brute-force
    canonicalisation on 3^3, closed forms of the coincident-point integrals.
"""
import itertools
import math

import numpy as np
import pytest

from tests.conftest import golden
from tests import burnside

G10 = None


def g10():
    global G10
    if G10 is None:
        G10 = golden("ref_eri.npz")["coeffs"]
    return G10


# ---------------------------------------------------------------------------
# symmetry tables

def test_octahedral_group_closure(orc):
    """48 distinct signed permutation matrices closed under products (test_symmetry.c:6-34)."""
    M = orc.octahedral_matrices()
    keys = {m.tobytes(): e for e, m in enumerate(M)}
    assert len(keys) == 48
    for m in M:
        assert abs(round(np.linalg.det(m))) == 1
        assert sorted(np.abs(m).sum(axis=0)) == [1, 1, 1]
    for a in M:
        for b in M:
            assert (a @ b).tobytes() in keys


def test_octahedral_matrices_match_reference_order(orc, ref_lib):
    import ctypes
    from gausslectronitegrate_amd import _abi
    arr = (_abi.IMat3x3 * 48).in_dll(ref_lib, "octahedral_matrep")
    ref = np.array([[[arr[e].data[r][c] for c in range(3)] for r in range(3)] for e in range(48)])
    assert np.array_equal(ref, orc.octahedral_matrices())


def test_grid_permutation_inverse(orc):
    """perm of element 17 composed with its inverse is the identity on 5^3 (test_symmetry.c:37-92)."""
    grid = orc.sym_grid(5)
    P = orc.octahedral_perms(grid)
    M = orc.octahedral_matrices()
    inv = [e for e in range(48) if np.array_equal(M[e] @ M[17], np.eye(3, dtype=np.int64))][0]
    assert sorted(P[17]) == list(range(125))
    assert np.array_equal(P[inv][P[17]], np.arange(125))


# ---------------------------------------------------------------------------
# enumeration

def test_eri_enumeration_3_bit_exact(orc):
    ref = golden("ref_eri_3x3x3.npz")
    got = orc.eri_enumerate(orc.sym_grid(3))
    assert got.dtype == np.int64
    assert np.array_equal(got, ref["four_indices"])


def test_eri_enumeration_5_bit_exact(orc):
    ref = golden("golden_enum_n5.npz")["eri_indices"]
    got = orc.eri_enumerate(orc.sym_grid(5))
    assert len(got) == burnside.orbit_count(5) == 138187
    assert np.array_equal(got, ref)


def test_erida_enumeration_counts(orc):
    for n in (3, 5, 7, 9):
        got = orc.erida_enumerate(orc.sym_grid(n))
        assert len(got) == burnside.orbit_count(n, 2) == math.comb(n + 2, 3)
        assert np.all(np.diff(got) > 0)
    ref7 = golden("golden_n7_K57s0.npz")["erida_indices"]
    assert np.array_equal(orc.erida_enumerate(orc.sym_grid(7)), ref7)


def test_burnside_counts_match_stored_workload_sizes():
    """Stored reference counts: 1027 (eri_integrals.ipynb:170), 84 ERIDA at 7^3 (erida_integrals.ipynb:171)."""
    assert burnside.orbit_count(3) == 1027
    assert burnside.orbit_count(7, 2) == 84
    assert burnside.axis_class_count(11) == 4641


def _brute_canonical_set(n):
    """Canonical four-indices of a 3^3-type grid by brute force over all
    configurations, all centred placements and all 384 symmetry images."""
    h = (n - 1) // 2
    M = np.array(list(itertools.product(range(-h, h + 1), repeat=3)))
    pts_lin = {tuple(p): i for i, p in enumerate(M)}
    N = len(M)
    Ms = []
    for perm in itertools.permutations(range(3)):
        for s in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), dtype=np.int64)
            for r in range(3):
                m[r, perm[r]] = s[r]
            Ms.append(m)
    labels = burnside._label_group(4)
    seen = set()
    canon = set()
    for tup in itertools.product(range(N), repeat=4):
        P = M[list(tup)]
        lo, hi = P.min(axis=0), P.max(axis=0)
        c = lo + hi
        if np.any(np.abs(c) > 1):
            continue
        if tup in seen:
            continue
        # orbit: all images, all centred placements
        best = None
        orbit = set()
        for m in Ms:
            Q = P @ m.T
            qlo, qhi = Q.min(axis=0), Q.max(axis=0)
            base = Q - (qlo + qhi) // 2
            cc = base.min(axis=0) + base.max(axis=0)
            for b in itertools.product(*[range(abs(int(x)) + 1) for x in cc]):
                R = base - np.array(b) * cc
                lin = [pts_lin[tuple(r)] for r in R]
                for lab in labels:
                    t = tuple(lin[lab[a]] for a in range(4))
                    orbit.add(t)
                    v = ((t[0] * N + t[1]) * N + t[2]) * N + t[3]
                    best = v if best is None else min(best, v)
        seen |= orbit
        canon.add(best)
    return np.array(sorted(canon), dtype=np.int64)


@pytest.mark.slow
def test_eri_enumeration_brute_force_3(orc):
    assert np.array_equal(_brute_canonical_set(3), golden("ref_eri_3x3x3.npz")["four_indices"])


def test_min_orbit_is_idempotent_on_enumeration(orc):
    grid = orc.sym_grid(3)
    P = orc.octahedral_perms(grid)
    idx = orc.eri_enumerate(grid)
    N = 27
    for t in idx[::7]:
        l, k, j, i = t % N, (t // N) % N, (t // N ** 2) % N, t // N ** 3
        assert orc.min_orbit_eri(grid, P, i, j, k, l) == t


# ---------------------------------------------------------------------------
# ERI values

def test_eri_values_3x3x3_fixture(orc):
    ref = golden("ref_eri_3x3x3.npz")
    got = orc.eri_sparse(ref["coeffs"], orc.sym_grid(3), ref["four_indices"], float(ref["tol"]))
    assert np.max(np.abs(got - ref["values"])) <= 1e-13


def test_eri_dense_fixture(orc):
    """Sparse 3^3 -> dense 12^4 on grid ((-1,2),(0,2),(-1,3)) vs test_eri_integrals.hdf5, <= 1e-13
    (test_eri_integrals.c:12-85)."""
    ref = golden("ref_eri.npz")
    gs = orc.sym_grid(3)
    tol = float(ref["tol"])
    idx = orc.eri_enumerate(gs)
    vals = orc.eri_sparse(ref["coeffs"], gs, idx, tol)
    dense = orc.fill_dense_eri(gs, idx, vals, ((-1, 2), (0, 2), (-1, 3)))
    assert np.max(np.abs(dense - ref["eri"])) <= 1e-13
    # completeness: no zero entry (test_eri_integrals.c:342-419, on the 3^3 grid)
    full = orc.fill_dense_eri(gs, idx, vals, gs)
    assert np.all(full != 0)


def test_eri_single_vs_dense(orc):
    ref = golden("ref_eri.npz")
    tol = float(ref["tol"])
    gd = ((-1, 2), (0, 2), (-1, 3))
    pts = np.array(list(itertools.product(range(-1, 1), range(0, 2), range(-1, 2))))
    rng = np.random.default_rng(3)
    for _ in range(40):
        q = rng.integers(0, 12, size=4)
        v = orc.eri_single(ref["coeffs"], pts[q], tol)
        assert abs(v - ref["eri"][tuple(q)]) <= 1e-13


def test_eri_notebook_values(orc):
    c = g10()
    tol = 1e-16
    assert orc.eri_single(c, [[0, 0, 0]] * 4, tol) == pytest.approx(1.6252009000829197, abs=1e-14)
    v = orc.eri_single(c, [[1, 0, 0], [0, 0, -1], [0, 0, 1], [0, 0, 0]], tol)
    assert v == pytest.approx(-0.0005864045861050708, abs=1e-15)


def test_eri_coulomb_table_closed_forms(orc):
    T = orc.eri_table(200)
    assert T[0] == pytest.approx(3 * math.sqrt(2 / math.pi), rel=1e-15)
    r = np.sqrt(np.arange(1, 200)) / 6
    assert np.allclose(T[1:], np.array([math.erf(x / (math.sqrt(2) / 3)) for x in r]) / r, rtol=1e-14, atol=0)


def test_eri_matches_reference_build(orc, ref_lib):
    """Oracle vs the reference build on synthetic K=57 coefficients (3^3 full list)."""
    g = golden("golden_n3_K57s0.npz")
    got = orc.eri_sparse(g["coeffs"], orc.sym_grid(3), g["eri_indices"], float(g["tol"]))
    assert np.array_equal(got, g["eri_values"])


@pytest.mark.parametrize("name,n", [("golden_sample_n5_K29s0.npz", 5), ("golden_sample_n7_K43s0.npz", 7),
                                    ("golden_sample_n9_K57s1.npz", 9), ("golden_sample_n11_K57s0.npz", 11)])
def test_eri_golden_samples(orc, name, n):
    g = golden(name)
    got = orc.eri_sparse(g["coeffs"], orc.sym_grid(n), g["eri_indices"], float(g["tol"]))
    assert np.max(np.abs(got - g["eri_values"])) <= 1e-15


def test_eri_brute_force_tiny_K(orc):
    """Direct sum over primitive quadruples with the closed-form Coulomb integral
    of two normalised Gaussians (K = 3, random coefficients, tol = 0).

    g_a(x - p) g_b(x - q) = (sqrt(pi)/3) e^{-(a - b + 3(p - q))^2/4} * n(x; ctr, w = 1/3)
    per axis, ctr = (3(p + q) + a + b)/6, with n a normalised Gaussian; two
    normalised Gaussians at distance d interact with erf(d / (sqrt(2) w)) / d."""
    from scipy.special import erf
    rng = np.random.default_rng(11)
    K = 3
    c = rng.standard_normal(K)
    js = np.arange(K) - (K - 1) // 2
    pts = np.array([[1, 0, -1], [0, 0, 0], [0, 1, 0], [-1, 0, 1]])
    prims = list(itertools.product(range(K), repeat=3))

    def weights_centres(p, q):
        W, Cn = [], []
        for ia in prims:
            for ib in prims:
                w = 1.0
                ctr = np.zeros(3)
                for ax in range(3):
                    a, b = js[ia[ax]], js[ib[ax]]
                    w *= c[ia[ax]] * c[ib[ax]] * (math.sqrt(math.pi) / 3) * math.exp(-((a - b + 3 * (p[ax] - q[ax])) ** 2) / 4)
                    ctr[ax] = (3 * (p[ax] + q[ax]) + a + b) / 6
                W.append(w)
                Cn.append(ctr)
        return np.array(W), np.array(Cn)

    W1, C1 = weights_centres(pts[0], pts[1])
    W2, C2 = weights_centres(pts[2], pts[3])
    d = np.linalg.norm(C1[:, None, :] - C2[None, :, :], axis=2)
    wv = 1.0 / 3
    with np.errstate(divide="ignore", invalid="ignore"):
        kern = np.where(d > 0, erf(d / (math.sqrt(2) * wv)) / d, math.sqrt(2 / math.pi) / wv)
    brute = float(W1 @ kern @ W2)
    v = orc.eri_single(c, pts, 0.0)
    assert v == pytest.approx(brute, rel=1e-12, abs=1e-14)


def test_eri_single_equals_sparse_for_distant_pairs(orc):
    """orc.eri_single on tuples whose two pairs lie far apart (a table sized from the pair spreads
    alone, as in eri_integrals.c:134-159, would be read past its end) equals the sparse path on the
    box holding the points (eri_integrals.c:551-574, table sized from the grid)"""
    rng = np.random.default_rng(31)
    c = golden("coeffs_K29_s0.npy")
    pts = rng.integers(-3, 4, size=(40, 4, 3))
    pts[:, 1] = pts[:, 0]          # i = j and k = l: zero pair spreads, centres up to 6 apart
    pts[:, 3] = pts[:, 2]
    grid = ((-3, 7), (-3, 7), (-3, 7))
    N = 343
    lin = ((pts[..., 0] + 3) * 7 + (pts[..., 1] + 3)) * 7 + (pts[..., 2] + 3)
    idx = ((lin[:, 0] * N + lin[:, 1]) * N + lin[:, 2]) * N + lin[:, 3]
    want = orc.eri_sparse(c, grid, idx, 1e-12)
    got = np.array([orc.eri_single(c, pts[e], 1e-12) for e in range(len(pts))])
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - want)) <= 1e-15 * max(1.0, np.abs(want).max())


def test_eri_flop_count_matches_instrumented_loops(orc):
    """orc.eri_flops (the closed count 2 n_x n_y + 2 R_xy (n_z + 1)) against the reference's loops
    (eri_integrals.c:608-669) run in Python with one count per executed multiply and add; the loop's
    values also reproduce orc.eri_sparse.  3^3 grid, K = 7, tol 1e-12, a seeded sample."""
    rng = np.random.default_rng(5)
    c = rng.standard_normal(7)
    n = 3
    grid = orc.sym_grid(n)
    R = n
    s0, ns = -R + 1, 2 * R - 1
    prod, st, ln = orc.eri_products(c, s0, ns, 1e-12)
    T = orc.eri_table(orc.eri_table_len(R, len(c)))
    N = n ** 3
    idx = np.sort(rng.choice(N ** 4, size=40, replace=False)).astype(np.int64)
    want_w = orc.eri_flops(c, grid, idx, 1e-12)
    want_v = orc.eri_sparse(c, grid, idx, 1e-12)
    for e, t in enumerate(idx):
        t = int(t)
        ps = []
        for _ in range(4):
            ps.append(t % N)
            t //= N
        l, k, j, i = ps
        pts = [((p // (n * n)) - 1, ((p // n) % n) - 1, (p % n) - 1) for p in (i, j, k, l)]
        arrs, cds = [], []
        for a in range(3):
            cds.append(3 * ((pts[0][a] + pts[1][a]) - (pts[2][a] + pts[3][a])))
            kl = ((pts[0][a] - pts[1][a]) - s0) * ns + ((pts[2][a] - pts[3][a]) - s0)
            arrs.append([(int(st[kl]) + q, prod[kl][q]) for q in range(int(ln[kl]))])
        ops = 0
        bins = {}
        for tx, px in arrs[0]:
            if abs(px) <= 1e-12:
                continue
            for ty, py in arrs[1]:
                if abs(py) <= 1e-12:
                    continue
                r = (cds[0] + tx) ** 2 + (cds[1] + ty) ** 2
                bins[r] = bins.get(r, 0.0) + px * py
                ops += 2
        val = 0.0
        for r in sorted(bins):
            if bins[r] == 0:
                continue
            acc = 0.0
            for tz, pz in arrs[2]:
                if abs(pz) <= 1e-12:
                    continue
                acc += pz * T[r + (cds[2] + tz) ** 2]
                ops += 2
            val += acc * bins[r]
            ops += 2
        assert ops == want_w[e]
        assert val == pytest.approx(want_v[e], rel=1e-12, abs=1e-15)


# ---------------------------------------------------------------------------
# ERIDA

def test_erida_dense_fixture(orc):
    """Sparse 3^3 -> dense 12x12 on ((0,2),(-1,3),(-1,2)) vs test_erida_integrals.hdf5, <= 1e-12."""
    ref = golden("ref_erida.npz")
    gs = orc.sym_grid(3)
    tol = float(ref["tol"])
    idx = orc.erida_enumerate(gs)
    vals = orc.erida_sparse(ref["coeffs"], gs, idx, tol)
    dense = orc.fill_dense_erida(gs, idx, vals, ((0, 2), (-1, 3), (-1, 2)))
    assert np.max(np.abs(dense - ref["erida"])) <= 1e-12


def test_erida_notebook_values(orc):
    c = g10()
    assert orc.erida_single(c, [[0, 0, 0], [0, 0, 0]], 1e-16) == pytest.approx(2.4104836730129793, abs=1e-14)
    assert orc.erida_single(c, [[-1, 0, 2], [0, 1, -1]], 1e-16) == pytest.approx(0.3024908490984703, abs=1e-14)
    gs = orc.sym_grid(7)
    idx = orc.erida_enumerate(gs)
    vals = orc.erida_sparse(c, gs, idx, 1e-16)
    dense = orc.fill_dense_erida(gs, idx, vals, gs)
    assert np.allclose(dense[:3, :3], [[2.41048367, 1.05555191, 0.46842533],
                                       [1.05555191, 2.41048367, 1.05555191],
                                       [0.46842533, 1.05555191, 2.41048367]], atol=5e-9)
    assert np.array_equal(dense, dense.T)


def test_erida_golden(orc):
    g = golden("golden_n7_K57s0.npz")
    got = orc.erida_sparse(g["coeffs"], orc.sym_grid(7), g["erida_indices"], float(g["tol"]))
    assert np.array_equal(got, g["erida_values"])


# ---------------------------------------------------------------------------
# one-body

def test_kinetic_fixture_and_notebook(orc):
    ref = golden("ref_kinetic.npz")
    got = orc.kinetic_matrix(ref["coeffs"], ((0, 2), (1, 3), (-1, 3)))
    assert np.max(np.abs(got - ref["kgi"])) <= 1e-13
    c = g10()
    kg = orc.kinetic_matrix(c, ((-1, 3), (-2, 5), (-2, 4)))
    assert kg[7, 7] == pytest.approx(5.218509050588766, abs=1e-14)
    t = orc.kinetic_1d(c, 3)
    assert t[2] == pytest.approx(0.3010258456920169, abs=1e-14)


def test_kinetic_closed_form_diagonal(orc):
    """t(0) = 1/2 sum c_a c_b (3 sqrt(pi)/2)(1 - d^2/2) e^{-d^2/4}, d = a - b."""
    c = golden("coeffs_K29_s0.npy")
    j = np.arange(29)
    d = (j[:, None] - j[None, :]).astype(float)
    t0 = 0.5 * c @ (1.5 * math.sqrt(math.pi) * (1 - 0.5 * d * d) * np.exp(-0.25 * d * d)) @ c
    assert orc.kinetic_1d(c, 1)[0] == pytest.approx(t0, rel=1e-13)


def test_nuclear_fixture_and_notebook(orc):
    ref = golden("ref_nuclear.npz")
    got = orc.nuclear_matrix(ref["coeffs"], ((0, 2), (2, 2), (-1, 3)), ref["pos"], ref["charges"], float(ref["tol"]))
    assert np.max(np.abs(got - ref["ngi"])) <= 1e-13
    c = g10()
    grid = ((-1, 3), (0, 2), (-2, 4))
    pos = [(-0.2, 1.3, 0.5), (0.7, 0, -2.1)]
    chg = [2, 1]
    ngi = orc.nuclear_matrix(c, grid, pos, chg, 1e-16)
    assert ngi[10, 10] == pytest.approx(1.8699450271662796, abs=1e-13)
    v = orc.nuclear_single(c, [[-1, 0, -2], [0, 1, 1]], pos, chg, 1e-16)
    assert v == pytest.approx(-0.0003044157658974286, abs=1e-15)
    assert np.array_equal(ngi, ngi.T)


def test_one_body_golden(orc):
    g = golden("golden_n3_K57s0.npz")
    gs = orc.sym_grid(3)
    assert np.array_equal(orc.kinetic_matrix(g["coeffs"], gs), g["kinetic"])
    assert np.array_equal(orc.nuclear_matrix(g["coeffs"], gs, [[0, 0, 0]], [1], float(g["tol"])), g["nuclear"])


def test_nuclear_many_nuclei_golden(orc):
    """the oracle's single-entry nuclear integral (nuclear_integrals.c:31-126) for 100 seeded nuclei on a
    3x4x5 grid at seeded entries, against the reference build's matrix within four times the reference's
    own spread under re-association (stored with the fixture)"""
    g = golden("golden_nuclear_many.npz")
    grid = [tuple(int(v) for v in ax) for ax in g["many_grid"]]
    n = [ax[1] for ax in grid]
    bound = max(4.0 * float(g["many_spread100"]), 1e-13)
    rng = np.random.default_rng(5)
    for I, J in rng.integers(0, n[0] * n[1] * n[2], size=(6, 2)):
        pts = []
        for p in (I, J):
            x, y, z = p // (n[1] * n[2]), (p // n[2]) % n[1], p % n[2]
            pts.append([grid[0][0] + x, grid[1][0] + y, grid[2][0] + z])
        v = orc.nuclear_single(g["many_coeffs"], pts, g["many_pos"], g["many_charges"], float(g["many_tol"]))
        assert abs(v - g["many_m100"][I, J]) <= bound, (I, J)
