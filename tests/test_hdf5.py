"""HDF5 layout of sparse ERI sets (libgli_amd_hdf5.so): the reference's own example file
(examples/eri_3x3x3.hdf5, copied as a data fixture) loads to the same arrays as its h5dump
conversion, and save -> load and save -> h5dump round trips keep the layout the reference's CLI
writes (cli/gli_eri_integrals.c:119-156).  Host only: no GPU computation."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import gausslectronitegrate_amd as gli
from tests.conftest import golden

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "eri_3x3x3.hdf5")


def test_load_reference_example_file():
    eri = gli.load_sparse_eri_gausslet_integrals(FIXTURE)
    ref = golden("ref_eri_3x3x3.npz")
    assert eri.grid == ((-1, 3), (-1, 3), (-1, 3))
    assert eri.num_entries == 1027
    assert np.array_equal(eri.four_indices, ref["four_indices"])
    assert np.array_equal(eri.integral_values, ref["values"])
    assert np.array_equal(eri.gausslet_coeffs, ref["coeffs"])
    assert eri.tol == float(ref["tol"])
    # lookups of the loaded set work on the host (binary search, eri_integrals.c:685-707)
    assert eri.integral_value([(0, 0, 0)] * 4) == ref["values"][-1]


def test_save_load_round_trip(tmp_path):
    src = gli.load_sparse_eri_gausslet_integrals(FIXTURE)
    out = str(tmp_path / "copy.hdf5")
    gli.save_sparse_eri_gausslet_integrals(src, out)
    back = gli.load_sparse_eri_gausslet_integrals(out)
    assert back.grid == src.grid and back.tol == src.tol
    assert np.array_equal(back.four_indices, src.four_indices)
    assert np.array_equal(back.integral_values, src.integral_values)
    assert np.array_equal(back.gausslet_coeffs, src.gausslet_coeffs)


def test_saved_layout_matches_reference_cli(tmp_path):
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        pytest.skip("h5dump not available")
    src = gli.load_sparse_eri_gausslet_integrals(FIXTURE)
    out = str(tmp_path / "layout.hdf5")
    gli.save_sparse_eri_gausslet_integrals(src, out)
    mine = subprocess.run([h5dump, "-H", out], capture_output=True, text=True).stdout
    theirs = subprocess.run([h5dump, "-H", FIXTURE], capture_output=True, text=True).stdout
    strip = lambda t: [l.strip() for l in t.splitlines() if not l.startswith("HDF5")]
    assert strip(mine) == strip(theirs)


def test_load_errors(tmp_path):
    with pytest.raises(RuntimeError):
        gli.load_sparse_eri_gausslet_integrals(str(tmp_path / "missing.hdf5"))
    with pytest.raises(SyntaxError):
        gli.load_sparse_eri_gausslet_integrals(3)
