#!/usr/bin/env python3
"""Regenerate the committed fixtures under tests/golden/.

Runs only in the development container (it reads the reference tree and
needs h5dump and the reference build oracle/_ref/libgli_ref.so):

  1. the reference's own HDF5 test data (test/data/*.hdf5,
     examples/eri_3x3x3.hdf5) -> .npz via `h5dump -b LE` (data only);
  2. the seeded synthetic gausslet coefficient sets of the workloads;
  3. golden vectors of the reference build for the synthetic workloads:
     full results on the small grids, seeded samples of canonical indices
     on the large ones, all ERI-DA pairs of the 13^3 grid, nuclear matrices
     for many nuclei with the reference's own re-association spread;
  4. count, SHA-256 and a strided subsample of the reference build's
     complete 9^3 and 11^3 enumeration lists.

Usage: python tests/golden/make_fixtures.py [--skip-ref-runs] [--parts a,b,...]
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from gausslectronitegrate_amd import _abi  # noqa: E402
from gausslectronitegrate_amd.coeffs import synthetic_coefficients  # noqa: E402

REF = "/root/reference"
H5DUMP = "/opt/conda/bin/h5dump"


def h5_read(path, obj, dtype, shape, attribute=False):
    """Read one dataset/attribute as raw little-endian bytes."""
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "x.bin")
        flag = "-a" if attribute else "-d"
        subprocess.run([H5DUMP, "-b", "LE", flag, obj, "-o", out, path],
                       check=True, stdout=subprocess.DEVNULL, cwd=td)
        a = np.fromfile(out, dtype=dtype)
    return a.reshape(shape)


def convert_reference_fixtures():
    d = os.path.join(REF, "test", "data")
    f = os.path.join(d, "test_eri_integrals.hdf5")
    np.savez_compressed(os.path.join(HERE, "ref_eri.npz"),
                        coeffs=h5_read(f, "/gausslet_coeffs", "<f8", (135,), True),
                        tol=h5_read(f, "/tol", "<f8", (), True),
                        eri=h5_read(f, "/eri", "<f8", (12, 12, 12, 12)))
    f = os.path.join(d, "test_erida_integrals.hdf5")
    np.savez_compressed(os.path.join(HERE, "ref_erida.npz"),
                        coeffs=h5_read(f, "/gausslet_coeffs", "<f8", (135,), True),
                        tol=h5_read(f, "/tol", "<f8", (), True),
                        erida=h5_read(f, "/erida", "<f8", (12, 12)))
    f = os.path.join(d, "test_kinetic_integrals.hdf5")
    np.savez_compressed(os.path.join(HERE, "ref_kinetic.npz"),
                        coeffs=h5_read(f, "/gausslet_coeffs", "<f8", (135,), True),
                        kgi=h5_read(f, "/kgi", "<f8", (18, 18)))
    f = os.path.join(d, "test_nuclear_integrals.hdf5")
    nn = int(h5_read(f, "/num_nuclei", "<i8", (), True))
    pos = np.stack([h5_read(f, "/nuclear_pos_%d" % k, "<f8", (3,), True) for k in range(nn)])
    chg = np.array([int(h5_read(f, "/nuclear_charge_%d" % k, "<i8", (), True)) for k in range(nn)])
    np.savez_compressed(os.path.join(HERE, "ref_nuclear.npz"),
                        coeffs=h5_read(f, "/gausslet_coeffs", "<f8", (135,), True),
                        tol=h5_read(f, "/tol", "<f8", (), True),
                        pos=pos, charges=chg,
                        ngi=h5_read(f, "/ngi", "<f8", (12, 12)))
    f = os.path.join(REF, "examples", "eri_3x3x3.hdf5")
    np.savez_compressed(os.path.join(HERE, "ref_eri_3x3x3.npz"),
                        grid_length=h5_read(f, "/grid_length", "<i8", (), True),
                        tol=h5_read(f, "/tol", "<f8", (), True),
                        coeffs=h5_read(f, "/gausslet_coeffs", "<f8", (135,)),
                        four_indices=h5_read(f, "/four_indices", "<i8", (1027,)),
                        values=h5_read(f, "/integral_values", "<f8", (1027,)))


# ---------------------------------------------------------------------------
# the reference build, driven through the C ABI

def load_ref():
    lib = C.CDLL(os.path.join(REPO, "oracle", "_ref", "libgli_ref.so"))
    return _abi.bind(lib)


def gdata_of(c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    g = _abi.GaussletData()
    g.coefficients = c.ctypes.data_as(_abi.P_double)
    g.indices.istart = -(len(c) - 1) // 2
    g.indices.num = len(c)
    return g, c


def sym_grid(n):
    h = (n - 1) // 2
    return ((-h, n), (-h, n), (-h, n))


def ref_enumerate(lib, n):
    g = _abi.Grid.of(sym_grid(n))
    out = _abi.SparseEriIndices()
    lib.enumerate_symmetry_reduced_eri_indices(C.byref(g), C.byref(out))
    a = np.ctypeslib.as_array(out.four_indices, shape=(out.num,)).copy()
    lib.delete_sparse_eri_indices(C.byref(out))
    return a


def ref_enumerate_erida(lib, n):
    g = _abi.Grid.of(sym_grid(n))
    out = _abi.SparseEridaIndices()
    lib.enumerate_symmetry_reduced_erida_indices(C.byref(g), C.byref(out))
    a = np.ctypeslib.as_array(out.two_indices, shape=(out.num,)).copy()
    lib.delete_sparse_erida_indices(C.byref(out))
    return a


def ref_eri_values(lib, c, n, idx, tol):
    gd, keep = gdata_of(c)
    g = _abi.Grid.of(sym_grid(n))
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = _abi.SparseEri()
    lib.compute_sparse_eri_gausslet_integrals(C.byref(gd), C.byref(g), idx.ctypes.data_as(_abi.P_glong),
                                              len(idx), tol, C.byref(out))
    v = np.ctypeslib.as_array(out.integral_values, shape=(out.num_entries,)).copy()
    lib.delete_sparse_eri_gausslet_integrals(C.byref(out))
    return v


def ref_erida_values(lib, c, n, idx, tol):
    gd, keep = gdata_of(c)
    g = _abi.Grid.of(sym_grid(n))
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = _abi.SparseErida()
    lib.compute_sparse_erida_gausslet_integrals(C.byref(gd), C.byref(g), idx.ctypes.data_as(_abi.P_glong),
                                                len(idx), tol, C.byref(out))
    v = np.ctypeslib.as_array(out.integral_values, shape=(out.num_entries,)).copy()
    lib.delete_sparse_erida_gausslet_integrals(C.byref(out))
    return v


def ref_nuclear(lib, c, grid, pos, chg, tol):
    gd, keep = gdata_of(c)
    g = _abi.Grid.of(grid)
    nuc = (_abi.Nucleus * len(chg))()
    for k in range(len(chg)):
        for a in range(3):
            nuc[k].pos[a] = float(pos[k][a])
        nuc[k].charge = int(chg[k])
    out = _abi.NuclearIntegrals()
    lib.compute_nuclear_gausslet_integrals(C.byref(gd), C.byref(g), nuc, len(chg), tol, C.byref(out))
    N = grid[0][1] * grid[1][1] * grid[2][1]
    v = np.ctypeslib.as_array(out.integral_values, shape=(N, N)).copy()
    lib.delete_nuclear_gausslet_integrals(C.byref(out))
    return v


def ref_kinetic(lib, c, grid):
    gd, keep = gdata_of(c)
    g = _abi.Grid.of(grid)
    out = _abi.KineticIntegrals()
    lib.compute_kinetic_gausslet_integrals(C.byref(gd), C.byref(g), C.byref(out))
    N = grid[0][1] * grid[1][1] * grid[2][1]
    v = np.ctypeslib.as_array(out.integral_values, shape=(N, N)).copy()
    lib.delete_kinetic_gausslet_integrals(C.byref(out))
    return v


def ref_perms(lib, n):
    g = _abi.Grid.of(sym_grid(n))
    arr = (_abi.P_glong * 48)()
    lib.evaluate_octahedral_grid_permutations(C.byref(g), arr)
    return g, arr


def sample_canonical(lib, n, count, seed):
    """Seeded sample of canonical ERI four-indices: random 4-point tuples are
    centred and canonicalised with the reference's own orbit minimum, the
    same route its lookups take (pymodule.c:710-780)."""
    N = n ** 3
    h = (n - 1) // 2
    g, perm = ref_perms(lib, n)
    rng = np.random.default_rng(seed)
    keys = set()
    while len(keys) < count:
        pts = rng.integers(-h, h + 1, size=(4, 3))
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        pts = pts - (lo + hi) // 2          # centre: lo + hi in {0, 1}
        c = pts.min(axis=0) + pts.max(axis=0)
        best = None
        for bx in range(abs(int(c[0])) + 1):
            for by in range(abs(int(c[1])) + 1):
                for bz in range(abs(int(c[2])) + 1):
                    q = pts - np.array([bx, by, bz]) * c
                    lin = [((int(p[0]) + h) * n + int(p[1]) + h) * n + int(p[2]) + h for p in q]
                    v = lib.minimum_octahedral_orbit_eri_tensor_index(N, perm, *lin)
                    best = v if best is None else min(best, v)
        keys.add(int(best))
    return np.array(sorted(keys), dtype=np.int64)


def golden_synthetic(lib):
    tol = 1e-16
    c57s0 = synthetic_coefficients(57, 0)
    # small grid: everything in full (3^3, K=57, seed 0)
    n = 3
    idx = ref_enumerate(lib, n)
    vals = ref_eri_values(lib, c57s0, n, idx, tol)
    kin = ref_kinetic(lib, c57s0, sym_grid(n))
    nuc = ref_nuclear(lib, c57s0, sym_grid(n), [[0.0, 0.0, 0.0]], [1], tol)
    eda_idx = ref_enumerate_erida(lib, n)
    eda = ref_erida_values(lib, c57s0, n, eda_idx, tol)
    np.savez_compressed(os.path.join(HERE, "golden_n3_K57s0.npz"), coeffs=c57s0, tol=tol,
                        eri_indices=idx, eri_values=vals, kinetic=kin, nuclear=nuc,
                        erida_indices=eda_idx, erida_values=eda)
    # ERI enumeration lists of the reference at 5^3 (bit-exact pin for the GPU enumeration)
    idx5 = ref_enumerate(lib, 5)
    np.savez_compressed(os.path.join(HERE, "golden_enum_n5.npz"), eri_indices=idx5)
    # 7^3: ERIDA list + values, nuclear matrix with two unit charges at (0, 0, +-0.7)
    n = 7
    eda_idx = ref_enumerate_erida(lib, n)
    eda = ref_erida_values(lib, c57s0, n, eda_idx, tol)
    pos = [[0.0, 0.0, 0.7], [0.0, 0.0, -0.7]]
    nuc = ref_nuclear(lib, c57s0, sym_grid(n), pos, [1, 1], tol)
    rng = np.random.default_rng(7)
    N = n ** 3
    rows = rng.integers(0, N, size=2048)
    cols = rng.integers(0, N, size=2048)
    np.savez_compressed(os.path.join(HERE, "golden_n7_K57s0.npz"), coeffs=c57s0, tol=tol,
                        erida_indices=eda_idx, erida_values=eda, nuc_pos=np.array(pos),
                        nuc_charges=np.array([1, 1]), nuc_rows=rows, nuc_cols=cols,
                        nuc_sample=nuc[rows, cols], nuc_rowsum=nuc.sum(axis=1), nuc_trace=np.trace(nuc))


def golden_samples(lib):
    """Seeded samples of canonical ERI indices with reference values at the large grids: the
    mixed-size batch (5^3 K29, 7^3 K43, 9^3 K57 seed 1, 11^3 K57) and 13^3.  A larger count
    with the same seed extends the smaller sample (the draws are the same sequence)."""
    tol = 1e-16
    cases = [(5, 29, synthetic_coefficients(29, 0), "K29s0", 4096),
             (7, 43, synthetic_coefficients(43, 0), "K43s0", 4096),
             (9, 57, synthetic_coefficients(57, 1), "K57s1", 4096),
             (11, 57, synthetic_coefficients(57, 0), "K57s0", 4096),
             (13, 57, synthetic_coefficients(57, 0), "K57s0", 2048)]
    for n, K, c, tag, count in cases:
        idx = sample_canonical(lib, n, count, seed=1000 + n)
        vals = ref_eri_values(lib, c, n, idx, tol)
        np.savez_compressed(os.path.join(HERE, "golden_sample_n%d_%s.npz" % (n, tag)), coeffs=c, tol=tol,
                            eri_indices=idx, eri_values=vals)
        print("sample n=%d K=%d: %d values, max |v| %.3g" % (n, K, len(vals), np.abs(vals).max()))


def golden_erida13(lib):
    """every one of the 455 unique ERI-DA pairs of the 13^3 grid (K = 57, seed 0) with the
    reference build's values"""
    tol = 1e-16
    c = synthetic_coefficients(57, 0)
    idx = ref_enumerate_erida(lib, 13)
    vals = ref_erida_values(lib, c, 13, idx, tol)
    np.savez_compressed(os.path.join(HERE, "golden_erida_n13_K57s0.npz"), coeffs=c, tol=tol,
                        erida_indices=idx, erida_values=vals)
    print("erida 13^3: %d pairs" % len(idx))


def golden_nuclear_many(lib):
    """Nuclear matrices of the reference build for many nuclei: 100 seeded nuclei (charges 1-3) on a
    3x4x5 grid with K = 29 (tol 1e-14), the first 64 of them alone, and 70 seeded nuclei (charges 1-2)
    on the 3^3 grid with K = 57 (tol 1e-16).  With this many nuclei an entry sums ~10^5 signed terms
    per nucleus, so each case also stores the reference's own spread under re-association: its matrix
    over all nuclei against the sum of its per-nucleus matrices."""
    def spread(c, grid, pos, chg, tol, full):
        per = np.zeros_like(full)
        for k in range(len(chg)):
            per += ref_nuclear(lib, c, grid, pos[k:k + 1], chg[k:k + 1], tol)
        return float(np.max(np.abs(full - per)))
    out = {}
    c29 = synthetic_coefficients(29, 0)
    rng = np.random.default_rng(100)
    pos = rng.uniform(-2.5, 2.5, size=(100, 3))
    chg = rng.integers(1, 4, size=100).astype(np.int32)
    grid = ((-1, 3), (-2, 4), (-2, 5))
    m100 = ref_nuclear(lib, c29, grid, pos, chg, 1e-14)
    m64 = ref_nuclear(lib, c29, grid, pos[:64], chg[:64], 1e-14)
    out.update(many_coeffs=c29, many_tol=1e-14, many_grid=np.array(grid, dtype=np.int64), many_pos=pos, many_charges=chg,
               many_m100=m100, many_spread100=spread(c29, grid, pos, chg, 1e-14, m100),
               many_m64=m64, many_spread64=spread(c29, grid, pos[:64], chg[:64], 1e-14, m64))
    c57 = synthetic_coefficients(57, 0)
    rng = np.random.default_rng(70)
    nuc = np.zeros((70, 4))
    nuc[:, :3] = rng.uniform(-1.5, 1.5, size=(70, 3))
    nuc[:, 3] = rng.integers(1, 3, size=70)
    g3 = sym_grid(3)
    m70 = ref_nuclear(lib, c57, g3, nuc[:, :3], nuc[:, 3].astype(np.int32), 1e-16)
    out.update(caller_coeffs=c57, caller_tol=1e-16, caller_nuclei=nuc, caller_m70=m70,
               caller_spread70=spread(c57, g3, nuc[:, :3], nuc[:, 3].astype(np.int32), 1e-16, m70))
    np.savez_compressed(os.path.join(HERE, "golden_nuclear_many.npz"), **out)
    print("nuclear many: spreads %.3g (100), %.3g (64), %.3g (70)" % (out["many_spread100"], out["many_spread64"],
                                                                       out["caller_spread70"]))


ENUM_STRIDE = 1 << 16


def enum_digests(lib, grids=(9, 11)):
    """Count, SHA-256 of the little-endian int64 bytes and a strided subsample (every 2^16-th entry,
    plus the first and last 1024) of the reference build's complete enumeration lists.  The 11^3
    run holds about 17 GB of list nodes."""
    import hashlib
    import time
    out = {}
    for n in grids:
        t0 = time.time()
        idx = ref_enumerate(lib, n)
        dt = time.time() - t0
        out["n%d_count" % n] = np.int64(len(idx))
        out["n%d_sha256" % n] = np.array(hashlib.sha256(idx.astype("<i8").tobytes()).hexdigest())
        out["n%d_strided" % n] = idx[::ENUM_STRIDE].copy()
        out["n%d_head" % n] = idx[:1024].copy()
        out["n%d_tail" % n] = idx[-1024:].copy()
        print("enumeration %d^3: %d entries in %.1f s, sha256 %s" % (n, len(idx), dt, out["n%d_sha256" % n]))
        del idx
    out["stride"] = np.int64(ENUM_STRIDE)
    np.savez_compressed(os.path.join(HERE, "golden_enum_digests.npz"), **out)


def main():
    parts_all = ("convert", "coeffs", "synthetic", "samples", "erida13", "nuclear-many", "enum-digests")
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-ref-runs", action="store_true", help="only the parts that need no reference build")
    ap.add_argument("--parts", default=",".join(parts_all),
                    help="comma-separated subset of %s" % ", ".join(parts_all))
    args = ap.parse_args()
    parts = set(args.parts.split(","))
    unknown = parts - set(parts_all)
    if unknown:
        ap.error("unknown parts: %s" % ", ".join(sorted(unknown)))
    if "convert" in parts:
        convert_reference_fixtures()
    if "coeffs" in parts:
        for K, seed in ((29, 0), (43, 0), (57, 0), (57, 1)):
            np.save(os.path.join(HERE, "coeffs_K%d_s%d.npy" % (K, seed)), synthetic_coefficients(K, seed))
    if not args.skip_ref_runs:
        lib = load_ref()
        if "synthetic" in parts:
            golden_synthetic(lib)
        if "samples" in parts:
            golden_samples(lib)
        if "erida13" in parts:
            golden_erida13(lib)
        if "nuclear-many" in parts:
            golden_nuclear_many(lib)
        if "enum-digests" in parts:
            enum_digests(lib)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
