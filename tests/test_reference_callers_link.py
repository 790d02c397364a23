"""The reference's own C callers build unchanged against this library's headers and link against
libgli_amd.so (plus libgli_amd_hdf5.so for the HDF5 users) with every symbol resolved: the CLI tools
(cli/*.c), the perf harness (perf/perf_eri_integrals.c), the C test driver with every test file
(test/run_tests.c, test/test_*.c) and the CPython extension (python/pymodule.c).

Runs only where the reference tree is present (the development container); it compiles into a
temporary directory and keeps nothing.  Nothing here runs the programs."""
import os
import subprocess
import sys
import sysconfig

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
INC = os.path.join(REPO, "include")
LIB = os.path.join(REPO, "gausslectronitegrate_amd", "lib")
HDF5 = os.environ.get("HDF5_ROOT", "/opt/conda")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree absent")


def _need_libs():
    for name in ("libgli_amd.so", "libgli_amd_hdf5.so"):
        if not os.path.exists(os.path.join(LIB, name)):
            pytest.skip("library not built: %s" % name)
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        pytest.skip("HDF5 headers absent")


def _link(tmp_path, sources, out, extra=(), libs=()):
    # no HDF5 directory in the rpath: the linker would then resolve the HIP runtime's C++ library
    # needs against HDF5's older copy of libstdc++
    cmd = (["gcc", "-std=gnu11", "-O1", "-fopenmp", "-I" + INC, "-I" + os.path.join(HDF5, "include")] + list(extra) +
           sources + ["-L" + LIB, "-lgli_amd", "-lgli_amd_hdf5", "-L" + os.path.join(HDF5, "lib"), "-lhdf5",
                      "-Wl,-rpath," + LIB, "-Wl,--no-undefined"] + list(libs) + ["-lm", "-o", str(tmp_path / out)])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    assert os.path.getsize(tmp_path / out) > 0


@pytest.mark.parametrize("src", ["cli/gli_eri_integrals.c", "cli/gli_eri_symmetry_reduction.c",
                                 "cli/gli_erida_symmetry_reduction.c", "perf/perf_eri_integrals.c"])
def test_cli_and_perf_link(tmp_path, src):
    _need_libs()
    _link(tmp_path, [os.path.join(REF, src)], "prog")


def test_run_tests_links(tmp_path):
    _need_libs()
    tdir = os.path.join(REF, "test")
    srcs = [os.path.join(tdir, f) for f in sorted(os.listdir(tdir)) if f.endswith(".c")]
    assert "run_tests.c" in [os.path.basename(s) for s in srcs]
    _link(tmp_path, srcs, "run_tests")


def test_pymodule_links(tmp_path):
    _need_libs()
    pyinc = sysconfig.get_paths()["include"]
    libdir = sysconfig.get_config_var("LIBDIR")
    ver = sysconfig.get_config_var("LDVERSION") or "%d.%d" % sys.version_info[:2]
    if not os.path.exists(os.path.join(pyinc, "Python.h")):
        pytest.skip("Python headers absent")
    _link(tmp_path, [os.path.join(REF, "python", "pymodule.c")], "pymodule.so",
          extra=["-fPIC", "-shared", "-I" + pyinc, "-I" + np.get_include()], libs=["-L" + libdir, "-lpython" + ver])
