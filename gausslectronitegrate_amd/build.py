"""Build the C-ABI library lib/libgli_amd.so (HIP kernels for gfx950 + host C).

    python -m gausslectronitegrate_amd.build      (or build() from Python)

Every HIP source is compiled with hipcc for gfx950 only; the host C sources
(the reference-compatible API) with gcc; everything links into one shared
library whose exported symbols are the declarations of include/*.h.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libgli_amd.so")
LIB_HDF5 = os.path.join(LIB_DIR, "libgli_amd_hdf5.so")
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")
INCLUDE = os.path.join(REPO, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

# -amdgpu-mfma-vgpr-form: MFMA accumulators live in VGPRs.  Without it the compiler keeps the
# loop-carried accumulators of the mode-product tiles in VGPRs anyway and copies them to and from
# AGPRs around every slice (64 extra moves and a wait for the last MFMA per slice), and the
# combined register count limits the tiles to 3 waves per SIMD instead of 5.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-mfma-vgpr-form",
             "-Wno-unused-result"]
C_FLAGS = ["-O3", "-fPIC", "-std=gnu11", "-fopenmp", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def sources():
    hip = sorted(os.path.join(CSRC, "hip", f) for f in os.listdir(os.path.join(CSRC, "hip")) if f.endswith(".hip"))
    c = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".c"))
    return hip, c


HDF5_HEADERS = ("hdf5_util.h", "gli_hdf5.h")   # exported by libgli_amd_hdf5.so instead


def exported_symbols():
    """every function and table declared in include/*.h (but the HDF5 headers): the library's
    whole dynamic interface; everything else (kernels, C++ helpers) stays local"""
    import re
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if h in HDF5_HEADERS or not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "for", "while", "return", "sizeof"):
                names.add(m.group(1))
        for m in re.finditer(r"^extern const [^;]*?\b(\w+)\s*\[", text, flags=re.M):
            names.add(m.group(1))
    return names


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    _run([sys.executable, os.path.join(CSRC, "host", "gen_symmetry.py"), os.path.join(BUILD, "oh_tables.inc")])
    incs = ["-I" + INCLUDE, "-I" + CSRC, "-I" + BUILD]
    hip, c = sources()
    jobs = []
    for s in hip:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + incs + ["-c", s, "-o", o], o))
    for s in c:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        jobs.append(([CC] + C_FLAGS + incs + ["-c", s, "-o", o], o))
    workers = min(8, max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [o for _, o in jobs]
    vmap = os.path.join(BUILD, "exports.map")
    with open(vmap, "w") as f:
        f.write("{\n  global:\n%s  local: *;\n};\n" % "".join("    %s;\n" % n for n in sorted(exported_symbols())))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-Wl,-Bsymbolic", "-Wl,--no-undefined", "-Wl,--version-script=" + vmap, "-lgomp", "-lm"])
    os.replace(tmp, LIB)
    build_hdf5(incs)
    build_c_caller()
    return LIB


C_CALLER_SRC = os.path.join(REPO, "tests", "c", "gli_caller.c")
C_CALLER = os.path.join(REPO, "tests", "c", "gli_caller")


def build_c_caller():
    """tests/c/gli_caller: a plain C11 program built with gcc against include/ and linked with
    -lgli_amd, the way an unchanged C caller of the reference links (run by the GPU tests)."""
    if not os.path.exists(C_CALLER_SRC):
        return None
    tmp = C_CALLER + ".tmp"
    _run([CC, "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + INCLUDE, C_CALLER_SRC, "-o", tmp,
          "-L" + LIB_DIR, "-lgli_amd", "-Wl,-rpath,$ORIGIN/../../gausslectronitegrate_amd/lib", "-Wl,--no-undefined"])
    os.replace(tmp, C_CALLER)
    return C_CALLER


def build_hdf5(incs):
    """libgli_amd_hdf5.so: the reference's HDF5 helpers and sparse ERI file layout, kept out of the
    core library so that libgli_amd.so needs no HDF5 installation; skipped when HDF5 is absent."""
    h5inc = os.path.join(HDF5_ROOT, "include")
    h5lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(h5inc, "hdf5.h")):
        print("HDF5 not found under %s: libgli_amd_hdf5.so not built" % HDF5_ROOT, file=sys.stderr)
        return None
    src = os.path.join(CSRC, "hdf5", "hdf5_io.c")
    obj = os.path.join(BUILD, "hdf5_io.c.o")
    _run([CC] + C_FLAGS + incs + ["-I" + h5inc, "-c", src, "-o", obj])
    tmp = LIB_HDF5 + ".tmp"
    _run([CC, "-shared", "-fPIC", "-o", tmp, obj, "-L" + LIB_DIR, "-lgli_amd", "-Wl,-rpath,$ORIGIN",
          "-L" + h5lib, "-lhdf5", "-Wl,-rpath," + h5lib, "-Wl,--no-undefined"])
    os.replace(tmp, LIB_HDF5)
    return LIB_HDF5


if __name__ == "__main__":
    print(build(verbose=True))
