#!/usr/bin/env python3
"""Benchmark: unique ERI elements/s on an 11^3 gausslet grid (K = 57, seed 0, tol 1e-16).

One step = the ERI values of every symmetry-unique four-index of the grid
(261,244,251 elements), from the coefficients to the values in HBM: the
factor/product/fold tables, the key-tensor contraction and the per-element
gather.  The index list is enumerated on the GPU beforehand (timed and
reported separately) and stays resident in HBM, as do the values.

Printed: one JSON line with the throughput, the roofline of the dominant
kernel (HIP events on the library stream), the per-element FP64 path on a
seeded sample, the error of the evaluated path against the per-element FP64
path and the reference C build, and a CPU baseline timed on the host cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
# OpenMP threads of the host-side baseline stay close to each other (libgomp reads this when the
# first library linked with it is loaded, which happens below, after this line)
os.environ.setdefault("OMP_PROC_BIND", "close")

from gausslectronitegrate_amd import _abi, _lib  # noqa: E402
from gausslectronitegrate_amd.coeffs import synthetic_coefficients  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
METHODS = {"tensor": 0, "mixed": 1, "direct": 2, "bf16x2": 3, "fp32": 4, "fp32x2": 5, "separable": 6}
DTYPES = {"tensor": "f64", "mixed": "f64+split-bf16", "direct": "f64", "bf16x2": "bf16x2+f32acc", "fp32": "f32",
          "fp32x2": "f32x2+f32acc", "separable": "f64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sym_grid6(n):
    h = (n - 1) // 2
    return np.array([-h, n, -h, n, -h, n], dtype=np.int64)


class Dev:
    def __init__(self, L, nbytes):
        self.L = L
        self.ptr = C.c_void_p()
        _lib.check(L.gli_device_malloc(C.byref(self.ptr), max(8, int(nbytes))), "device malloc")

    def __del__(self):
        if self.ptr:
            self.L.gli_device_free(self.ptr)


def eri_step(L, c, g, d_idx, num, tol, method, d_val):
    _lib.check(L.gli_eri_values_device(c.ctypes.data_as(_abi.P_double), len(c), g.ctypes.data_as(_abi.P_glong),
                                       d_idx, num, tol, method, d_val), "gli_eri_values_device")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 counter summary, else None.

    The summary line is `traffic_bytes_per_launch: <bytes>`: FETCH_SIZE + WRITE_SIZE of one dispatch,
    each scaled by the ratio the same counter shows for an elementwise kernel of known size."""
    path = os.path.join(REPO, "profiles", "%s_%s_traffic.md" % (kernel, workload))
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].strip().split()[0])
    except OSError:
        return None
    return None


def calibrate_copy(nbytes):
    """One elementwise kernel of known traffic (read nbytes, write nbytes), for counter calibration."""
    import torch
    n = int(nbytes) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "torch.neg float64", "bytes_read": 8 * n, "bytes_written": 8 * n, "launches": 3}),
          flush=True)


def phase_times(L):
    ms = (C.c_double * 16)()
    names = (C.c_char_p * 16)()
    n = L.gli_phase_times(ms, names, 16)
    return {names[i].decode(): ms[i] for i in range(max(n, 0))}


def host_cpu():
    """(model name, set of ISA flags) of the host CPU from /proc/cpuinfo"""
    model, flags = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if model is None and line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                elif not flags and line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                if model is not None and flags:
                    break
    except OSError:
        pass
    return model, flags


# x86-64 micro-architecture levels of the prebuilt reference variants (oracle/Makefile) and the CPU
# flags each one needs
ISA_LEVELS = [("x86-64-v4", "_v4", {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl", "avx2", "fma", "bmi2"}),
              ("x86-64-v3", "_v3", {"avx2", "fma", "bmi1", "bmi2", "movbe", "f16c", "abm"})]


def ref_runner(path, c, n, tol):
    """values of an index sample through a build of the reference library at `path`"""
    ref = _abi.bind(C.CDLL(path))
    cc = np.ascontiguousarray(c)
    h = (n - 1) // 2

    def run(idx):
        gd = _abi.GaussletData()
        gd.coefficients = cc.ctypes.data_as(_abi.P_double)
        gd.indices.istart = -(len(cc) - 1) // 2
        gd.indices.num = len(cc)
        grid = _abi.Grid.of(((-h, n), (-h, n), (-h, n)))
        ii = np.ascontiguousarray(idx)
        out = _abi.SparseEri()
        ref.compute_sparse_eri_gausslet_integrals(C.byref(gd), C.byref(grid), ii.ctypes.data_as(_abi.P_glong),
                                                  len(ii), tol, C.byref(out))
        v = np.ctypeslib.as_array(out.integral_values, shape=(len(ii),)).copy()
        ref.delete_sparse_eri_gausslet_integrals(C.byref(out))
        return v
    return run


def timed_sample(run, all_idx, pos, rng):
    """run the sample all_idx[pos] as 8 seeded sub-batches: (values, seconds, per-batch rates)"""
    count = len(pos)
    part = rng.permutation(count) % 8
    vals = np.empty(count)
    dt = 0.0
    rates = []
    for b in range(8):
        sel = np.nonzero(part == b)[0]
        sub = np.ascontiguousarray(all_idx[pos[sel]])
        t0 = time.perf_counter()
        vals[sel] = run(sub)
        d = time.perf_counter() - t0
        dt += d
        rates.append(len(sel) / d)
    return vals, dt, np.array(rates)


def ci95(rates):
    half = 2.365 * rates.std(ddof=1) / np.sqrt(len(rates))   # Student t, 7 degrees of freedom, 95%
    return [float(rates.mean() - half), float(rates.mean() + half)]


def cpu_baseline(c, n, all_idx, tol, target_s, seed, label):
    """Reference C build (oracle/_ref, the reference's Release flags) if present, else the oracle
    restatement, on all host threads.  Where a build for a wider ISA level of the host CPU exists too,
    it runs the same sample as `isa_variant`."""
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    model, flags = host_cpu()
    if os.path.exists(orc.REF_LIB_PATH):
        run = ref_runner(orc.REF_LIB_PATH, c, n, tol)
        kind = "reference"
    else:
        h = (n - 1) // 2

        def run(idx):
            return orc.eri_sparse(c, ((-h, n), (-h, n), (-h, n)), idx, tol)
        kind = "port"
    # calibrate on a small seeded sample, then time a seeded sample sized to about target_s
    rng = np.random.default_rng(2024)
    probe_pos = np.sort(rng.choice(len(all_idx), size=min(len(all_idx), max(threads * 64, 1024)), replace=False))
    t0 = time.perf_counter()
    run(all_idx[probe_pos])
    per = (time.perf_counter() - t0) / len(probe_pos)
    count = int(min(len(all_idx), max(len(probe_pos), target_s / max(per, 1e-9))))
    count = max(count, 8)
    pos = np.sort(rng.choice(len(all_idx), size=count, replace=False))
    vals, dt, rates = timed_sample(run, all_idx, pos, np.random.default_rng(7))
    rate = count / dt
    sample = ("%d seeded uniformly drawn symmetry-unique %s four-indices (K=%d, seed %d, tol %g) in 8 sub-batches, "
              "%.1f s" % (count, label, len(c), seed, tol, dt))
    out = {"value": rate, "unit": "elements/s", "cores": threads, "kind": kind, "ci95": ci95(rates),
           "full_set_s": len(all_idx) / rate, "sample": sample, "cpu_model": model,
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "build": "reference sources, gcc -O3 -DNDEBUG -fopenmp (the reference's Release flags)" if kind == "reference"
                    else "oracle restatement, gcc -O3 -fopenmp"}
    if kind == "reference":
        for isa, suffix, need in ISA_LEVELS:
            path = orc.REF_LIB_PATH.replace(".so", suffix + ".so")
            if os.path.exists(path) and need <= flags:
                run_isa = ref_runner(path, c, n, tol)
                run_isa(all_idx[probe_pos])
                v2, dt2, rates2 = timed_sample(run_isa, all_idx, pos, np.random.default_rng(7))
                out["isa_variant"] = {"value": count / dt2, "unit": "elements/s", "isa": isa, "ci95": ci95(rates2),
                                      "build": "reference sources, gcc -O3 -DNDEBUG -fopenmp -march=%s" % isa,
                                      "seconds": dt2, "max_abs_diff_vs_release_build": float(np.max(np.abs(v2 - vals)))}
                break
    return out, pos, vals


# grids of the mixed-size batch: (n, K, coefficient seed, spacing a in bohr); outputs are scaled by 1/a
BATCH = [(5, 29, 0, 0.5), (7, 43, 0, 0.75), (9, 57, 1, 1.0), (11, 57, 0, 1.5)]


def mixed_vs_fp64(compare, method, value):
    """the fast method's rate against the exact FP64 key-tensor method's (tensor) on the same step, and the
    split-precision (mixed) method's, when they were timed in this run"""
    by = {c["method"]: c for c in compare}
    if method == "tensor":
        fp64, fast, fast_value = None, None, None
        fp64_value = value
    elif "tensor" in by:
        fp64 = by["tensor"]
        fp64_value = fp64["value"]
        fast, fast_value = method, value
    else:
        return None
    out = {"fp64_exact_elements_per_s": fp64_value}
    if fast is not None:
        out.update({"fast_method": fast, "ratio": fast_value / fp64_value, "fast_elements_per_s": fast_value,
                    "max_abs_err_vs_fp64": fp64["max_abs_err_vs_%s_all_elements" % method],
                    "fast_note": "FP64 throughout: the Coulomb kernel T[i] = 6 erf(sqrt(i/8))/sqrt(i) written as a "
                                 "Gauss-Legendre sum of separable Gaussians (within 2e-15 relative of the table), so the "
                                 "key tensor is one rank-Q product instead of three mode products"})
    if "mixed" in by:
        m = by["mixed"]
        out.update({"mixed_ratio": m["value"] / fp64_value, "mixed_elements_per_s": m["value"],
                    "mixed_max_abs_err_vs_%s" % method: m["max_abs_err_vs_%s_all_elements" % method],
                    "mixed_note": "the split BF16 products with their digit passes are slower than FP64 MFMA "
                                  "(78.6 TF/s) on MI355X; the step is bound by the per-element gather"})
    return out


def batch_measure(L, args, stream, method_name=None):
    """One step = every unique four-index of every grid of the batch, from the coefficients to the
    scaled values in HBM, as one replay of the plan's captured graph.  The error is taken over every
    element against the single-grid FP64 path (times 1/a).  method_name overrides the bench's method.
    Returns the measurement as a dict."""
    import torch
    name = method_name or args.method
    method = METHODS[name] if name in ("tensor", "mixed", "separable") else 0
    n_g = len(BATCH)
    coeffs, grids, idxs, nums, vals = [], [], [], [], []
    enum_ms = 0.0
    for n, K, seed, a in BATCH:
        coeffs.append(synthetic_coefficients(K, seed))
        grids.append(sym_grid6(n))
        p = C.c_void_p()
        num = C.c_int64()
        t0 = time.perf_counter()
        _lib.check(L.gli_eri_enumerate_device(n, C.byref(p), C.byref(num)), "enumeration")
        enum_ms += (time.perf_counter() - t0) * 1e3
        idxs.append(p)
        nums.append(num.value)
        vals.append(Dev(L, 8 * num.value))
    total = int(sum(nums))
    log("batch: %s, %d unique four-indices" % ([b[:2] for b in BATCH], total))
    cptr = (_abi.P_double * n_g)(*[c.ctypes.data_as(_abi.P_double) for c in coeffs])
    Ks = np.array([len(c) for c in coeffs], dtype=np.int64)
    g6 = np.concatenate(grids)
    dptr = (C.c_void_p * n_g)(*[p.value for p in idxs])
    nn = np.array(nums, dtype=np.int64)
    vptr = (C.c_void_p * n_g)(*[v.ptr.value for v in vals])
    scales = np.array([1.0 / b[3] for b in BATCH])
    plan = C.c_void_p()
    t0 = time.perf_counter()
    _lib.check(L.gli_eri_batch_create(n_g, cptr, Ks.ctypes.data_as(_abi.P_glong), g6.ctypes.data_as(_abi.P_glong), dptr,
                                      nn.ctypes.data_as(_abi.P_glong), args.tol, method, vptr, 1, C.byref(plan)),
               "gli_eri_batch_create")
    create_ms = (time.perf_counter() - t0) * 1e3
    buf = C.create_string_buffer(1 << 16)
    L.gli_eri_batch_describe(plan, buf, len(buf))
    tuned = json.loads(buf.value.decode())
    sc = scales.ctypes.data_as(C.POINTER(C.c_double))
    for _ in range(args.warmup):
        _lib.check(L.gli_eri_batch_run(plan, cptr, sc), "batch run")
    L.gli_synchronize()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        _lib.check(L.gli_eri_batch_run(plan, cptr, sc), "batch run")
    ev1.record(stream)
    L.gli_synchronize()
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1) / args.steps
    log("batch step: %.3f ms" % step_ms)
    # errors over every element against the single-grid FP64 path
    errs = []
    for g, (n, K, seed, a) in enumerate(BATCH):
        got = np.empty(nums[g])
        _lib.check(L.gli_memcpy(got.ctypes.data, vals[g].ptr, 8 * nums[g], 1), "copy batch values")
        ref_d = Dev(L, 8 * nums[g])
        eri_step(L, coeffs[g], grids[g], idxs[g], nums[g], args.tol, 0, ref_d.ptr)
        L.gli_synchronize()
        ref = np.empty(nums[g])
        _lib.check(L.gli_memcpy(ref.ctypes.data, ref_d.ptr, 8 * nums[g], 1), "copy reference values")
        ref *= scales[g]
        d = np.abs(got - ref)
        big = np.abs(ref) >= 1e-10
        errs.append({"grid": n, "K": K, "seed": seed, "spacing": a, "elements": nums[g],
                     "max_abs_err_vs_fp64": float(d.max()),
                     "max_rel_err_vs_fp64_where_abs_ge_1e-10": float((d[big] / np.abs(ref[big])).max()) if big.any() else None})
        del got, ref, d, big, ref_d
    L.gli_eri_batch_destroy(plan)
    for p in idxs:
        L.gli_device_free(p)
    return {"workload": "eri_batch_5-7-9-11", "grids": BATCH, "method": {0: "tensor", 1: "mixed", 6: "separable"}[method],
            "unique_elements": total, "ms_per_step": step_ms, "value": total / (step_ms / 1e3), "unit": "elements/s",
            "steps": args.steps, "plan_create_ms": create_ms, "enumeration_ms": enum_ms, "autotune": tuned,
            "errors": errs}


def large_measure(L, args, stream):
    """The largest single-GPU case: every unique four-index of the n^3 grid (n = --large-n, K = 57,
    seed 0), timed with the bench's method, the exact FP64 key-tensor method and the BF16 / FP32 key-tensor
    variants, each with its error against the exact FP64 method over every element (computed on the
    device), and the ERI-DA set of the same grid.  Returns the measurement as a dict."""
    import torch
    import gausslectronitegrate_amd as gli
    n, K, seed = args.large_n, 57, 0
    c = synthetic_coefficients(K, seed)
    g = sym_grid6(n)
    p = C.c_void_p()
    num = C.c_int64()
    t0 = time.perf_counter()
    _lib.check(L.gli_eri_enumerate_device(n, C.byref(p), C.byref(num)), "enumeration")
    L.gli_synchronize()
    enum_ms = (time.perf_counter() - t0) * 1e3
    num = num.value
    log("large: %d^3, %d unique four-indices (%.1f ms)" % (n, num, enum_ms))
    steps = max(1, min(args.steps, 5))
    names = [args.method] + [m for m in ("tensor", "bf16x2", "fp32") if m != args.method]
    out = {}
    ref = torch.empty(num, dtype=torch.float64, device="cuda")
    cur = torch.empty(num, dtype=torch.float64, device="cuda")
    for name in ["tensor"] + [m for m in names if m != "tensor"]:
        dst = ref if name == "tensor" else cur
        eri_step(L, c, g, p, num, args.tol, METHODS[name], C.c_void_p(dst.data_ptr()))
        L.gli_synchronize()
        torch.cuda.synchronize()
        m0 = torch.cuda.Event(enable_timing=True)
        m1 = torch.cuda.Event(enable_timing=True)
        m0.record(stream)
        for _ in range(steps):
            eri_step(L, c, g, p, num, args.tol, METHODS[name], C.c_void_p(dst.data_ptr()))
        m1.record(stream)
        L.gli_synchronize()
        torch.cuda.synchronize()
        ms = m0.elapsed_time(m1) / steps
        rec = {"dtype": DTYPES[name], "ms_per_step": ms, "value": num / (ms / 1e3)}
        if name != "tensor":
            d = (cur - ref).abs_()
            big = ref.abs() >= 1e-10
            rec["max_abs_err_vs_fp64_all_elements"] = float(d.max())
            rec["max_rel_err_vs_fp64_where_abs_ge_1e-10"] = float((d[big] / ref[big].abs()).max()) if bool(big.any()) else None
            del d, big
        out[name] = rec
        log("large %s: %.3f ms/step" % (name, ms))
    del ref, cur
    L.gli_device_free(p)
    h = (n - 1) // 2
    t0 = time.perf_counter()
    erida = gli.compute_sparse_erida_gausslet_integrals(c, ((-h, n), (-h, n), (-h, n)), args.tol)
    erida_ms = (time.perf_counter() - t0) * 1e3
    return {"workload": "eri_%dx%dx%d_K%d_seed%d" % (n, n, n, K, seed), "unique_elements": num, "steps": steps,
            "enumeration_ms": enum_ms, "methods": out,
            "erida": {"unique_elements": int(erida.num_entries), "ms_incl_host_copies": erida_ms}}


def batch_bench(args):
    """the mixed-size batch alone, as its own JSON line"""
    import torch
    L = _lib.require_gpu()
    for kv in args.tune:
        knob, value = kv.split("=", 1)
        _lib.check(L.gli_set_tuning(knob.encode(), int(value)), "gli_set_tuning " + kv)
    stream = torch.cuda.ExternalStream(L.gli_stream())
    b = batch_measure(L, args, stream)
    line = {
        "metric": "unique ERI elements/s, mixed-size batch (5^3 K29, 7^3 K43, 9^3 K57, 11^3 K57), one graph replay per step",
        "value": b["value"],
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[b["method"]],
        "data": "synthetic (seeded gausslet coefficients)",
        "config": {"workload": b["workload"], "grids": BATCH, "tol": args.tol, "method": b["method"],
                   "unique_elements": b["unique_elements"]},
        "plan_create_ms": b["plan_create_ms"],
        "enumeration_ms": b["enumeration_ms"],
        "autotune": b["autotune"],
        "errors": b["errors"],
    }
    print(json.dumps(line), flush=True)


def aux_bench(args):
    """ERI-DA set, nuclear matrix (unit charges at (0, 0, +-0.7 / a) grid units, a = 1) and kinetic matrix on
    an n^3 grid: GPU time and the max abs error against the reference C build on the same inputs (the
    full ERI-DA set and kinetic matrix; a seeded sample of nuclear entries through the reference's
    single-entry function, whose full matrix takes minutes on the host)."""
    import gausslectronitegrate_amd as gli
    from oracle import oracle as orc
    _lib.require_gpu()
    n, K = args.n, args.K
    c = synthetic_coefficients(K, args.seed)
    h = (n - 1) // 2
    grid = ((-h, n), (-h, n), (-h, n))
    N = n ** 3
    pos = np.array([[0.0, 0.0, 0.7], [0.0, 0.0, -0.7]])
    chg = np.array([1, 1], dtype=np.int32)
    have_ref = os.path.exists(orc.REF_LIB_PATH)
    ref = _abi.bind(C.CDLL(orc.REF_LIB_PATH)) if have_ref else None
    gd = _abi.GaussletData()
    cc = np.ascontiguousarray(c)
    gd.coefficients = cc.ctypes.data_as(_abi.P_double)
    gd.indices.istart = -(K - 1) // 2
    gd.indices.num = K
    g = _abi.Grid.of(grid)

    def timed(f, reps=3):
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = f()
        return r, (time.perf_counter() - t0) / reps * 1e3

    out = {"metric": "enumeration, one-body and ERI-DA assembly on an %d^3 grid (K=%d, seed %d), GPU vs reference C build"
                     % (n, K, args.seed),
           "grid": n, "K": K, "nuclei": pos.tolist(), "charges": chg.tolist(), "tol": args.tol}
    # ERI enumeration (symmetry-unique four-index list): GPU against the reference on all host threads
    L = _lib.require_gpu()
    gpu_idx = {}

    def enum_gpu():
        p = C.c_void_p()
        cnt = C.c_int64()
        _lib.check(L.gli_eri_enumerate_device(n, C.byref(p), C.byref(cnt)), "enumeration")
        L.gli_synchronize()
        gpu_idx["p"], gpu_idx["n"] = p, cnt.value
        return cnt.value

    enum_gpu()
    L.gli_device_free(gpu_idx["p"])
    t0 = time.perf_counter()
    num = enum_gpu()
    rec = {"unique_elements": num, "gpu_ms": (time.perf_counter() - t0) * 1e3}
    if ref is not None and n <= args.enum_ref_max:
        host = np.empty(num, dtype=np.int64)
        _lib.check(L.gli_memcpy(host.ctypes.data, gpu_idx["p"], 8 * num, 1), "copy indices")
        ri = _abi.SparseEriIndices()
        t0 = time.perf_counter()
        ref.enumerate_symmetry_reduced_eri_indices(C.byref(g), C.byref(ri))
        rec["reference_ms"] = (time.perf_counter() - t0) * 1e3
        rec["reference_threads"] = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        rl = np.ctypeslib.as_array(ri.four_indices, shape=(ri.num,))
        rec["bit_exact"] = bool(ri.num == num and np.array_equal(rl, host))
        ref.delete_sparse_eri_indices(C.byref(ri))
    L.gli_device_free(gpu_idx["p"])
    out["eri_enumeration"] = rec
    # ERI-DA: all unique pairs
    erida, t_g = timed(lambda: gli.compute_sparse_erida_gausslet_integrals(c, grid, args.tol))
    rec = {"unique_elements": erida.num_entries, "gpu_ms_incl_host_copies": t_g}
    if ref is not None:
        idx = np.ascontiguousarray(erida.two_indices)
        res = _abi.SparseErida()
        t0 = time.perf_counter()
        ref.compute_sparse_erida_gausslet_integrals(C.byref(gd), C.byref(g), idx.ctypes.data_as(_abi.P_glong), len(idx),
                                                    args.tol, C.byref(res))
        rec["reference_ms"] = (time.perf_counter() - t0) * 1e3
        rv = np.ctypeslib.as_array(res.integral_values, shape=(len(idx),)).copy()
        ref.delete_sparse_erida_gausslet_integrals(C.byref(res))
        rec["max_abs_err_vs_reference"] = float(np.max(np.abs(erida.integral_values - rv)))
    out["erida"] = rec
    # kinetic: full matrix
    kin, t_g = timed(lambda: gli.compute_kinetic_gausslet_integrals(c, grid))
    rec = {"matrix": [N, N], "gpu_ms_incl_host_copies": t_g}
    if ref is not None:
        res = _abi.KineticIntegrals()
        t0 = time.perf_counter()
        ref.compute_kinetic_gausslet_integrals(C.byref(gd), C.byref(g), C.byref(res))
        rec["reference_ms"] = (time.perf_counter() - t0) * 1e3
        rk = np.ctypeslib.as_array(res.integral_values, shape=(N * N,)).reshape(N, N).copy()
        ref.delete_kinetic_gausslet_integrals(C.byref(res))
        rec["max_abs_err_vs_reference"] = float(np.max(np.abs(np.asarray(kin).reshape(N, N) - rk)))
    out["kinetic"] = rec
    # nuclear: full matrix on the GPU, seeded sample of entries through the reference
    nuc, t_g = timed(lambda: gli.compute_nuclear_gausslet_integrals(c, grid, pos, chg, args.tol))
    nuc = np.asarray(nuc).reshape(N, N)
    rec = {"matrix": [N, N], "gpu_ms_incl_host_copies": t_g}
    if ref is not None:
        rng = np.random.default_rng(7)
        pairs = rng.integers(0, N, size=(48, 2))
        nucl = (_abi.Nucleus * 2)()
        for q in range(2):
            for a in range(3):
                nucl[q].pos[a] = pos[q][a]
            nucl[q].charge = int(chg[q])
        pts = (_abi.GridPoint * 2)()
        t0 = time.perf_counter()
        errs = []
        for i, j in pairs:
            for slot, p in enumerate((i, j)):
                x, y, z = p // (n * n) - h, (p // n) % n - h, p % n - h
                pts[slot].c[0], pts[slot].c[1], pts[slot].c[2] = x, y, z
            v = ref.compute_nuclear_gausslet_integral(C.byref(gd), pts, nucl, 2, args.tol)
            errs.append(abs(nuc[i, j] - v))
        rec["reference_ms_per_entry"] = (time.perf_counter() - t0) * 1e3 / len(pairs)
        rec["reference_sample"] = len(pairs)
        rec["max_abs_err_vs_reference_sample"] = float(max(errs))
        rec["symmetric"] = bool(np.array_equal(nuc, nuc.T))
    out["nuclear"] = rec
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=11)
    ap.add_argument("--K", type=int, default=57)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-16)
    ap.add_argument("--method", choices=sorted(METHODS), default="separable",
                    help="method of the timed step: separable (default; FP64, the Coulomb kernel as a Gauss-Legendre "
                         "sum of separable Gaussians, within ~1e-15 of tensor), tensor (the exact FP64 re-association), "
                         "mixed (split BF16), direct (per element) or a comparison method")
    ap.add_argument("--direct-sample", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--asymmetric", action="store_true",
                    help="coefficients without mirror symmetry (general axis-key set, GEMM-dominated)")
    ap.add_argument("--compare", default="tensor,mixed",
                    help="comma-separated methods timed beside the main one, each with its error against the main "
                         "method's values over every element (e.g. tensor,mixed,bf16x2,fp32,fp32x2)")
    ap.add_argument("--no-compare", action="store_true", help="time no other method")
    ap.add_argument("--no-host-api", action="store_true",
                    help="skip timing the reference entry points with host arrays (PCIe included)")
    ap.add_argument("--no-plan", action="store_true",
                    help="skip the one-grid batch-plan replay reported beside the single-call rate")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the mixed-size batch (5^3/7^3/9^3/11^3 plan replay) reported beside the single-call rate")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the largest single-GPU case (--large-n grid, its methods and ERI-DA) in the default line")
    ap.add_argument("--large-n", type=int, default=13)
    ap.add_argument("--tune", action="append", default=[], metavar="KNOB=VALUE",
                    help="set a kernel tuning knob (gli_set_tuning) before timing; repeatable")
    ap.add_argument("--batch", action="store_true",
                    help="mixed-size batch: 5^3 (K=29), 7^3 (K=43), 9^3 and 11^3 (K=57) in one autotuned plan, "
                         "replayed as one hipGraph per step")
    ap.add_argument("--aux", action="store_true",
                    help="ERI-DA set, nuclear and kinetic matrices on the --n grid against the reference C build")
    ap.add_argument("--enum-ref-max", type=int, default=9,
                    help="with --aux: largest grid whose enumeration is also timed on the reference build (9^3 takes "
                         "~40 s on 8 host threads; the reference's list needs 64 B of host memory per element)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="BYTES",
                    help="only run an elementwise kernel of known traffic (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus == 1 and args.batch:
        batch_bench(args)
        return
    if args.gpus == 1 and args.aux:
        aux_bench(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would take a contiguous, "
              "cost-balanced slice of the sorted index list); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch  # device memory and events are plumbing here; the work runs in libgli_amd.so
    L = _lib.require_gpu()
    stream = torch.cuda.ExternalStream(L.gli_stream())
    for kv in args.tune:
        knob, value = kv.split("=", 1)
        _lib.check(L.gli_set_tuning(knob.encode(), int(value)), "gli_set_tuning " + kv)
    tuning = {k: L.gli_get_tuning(k.encode()) for k in ("tile_order", "gemm_shape", "gather_ept", "gather_order", "gather_lds",
                                                      "mode1_half", "gather_threads", "gather_bucketed", "gather_classed")}
    method = METHODS[args.method]
    c = synthetic_coefficients(args.K, args.seed, symmetrise=not args.asymmetric)
    g = sym_grid6(args.n)

    # enumeration (grid only), resident in HBM
    log("enumerating %d^3 ..." % args.n)
    p = C.c_void_p()
    num = C.c_int64()
    t0 = time.perf_counter()
    _lib.check(L.gli_eri_enumerate_device(args.n, C.byref(p), C.byref(num)), "enumeration")
    enum_ms = (time.perf_counter() - t0) * 1e3
    num = num.value
    d_idx = p
    log("unique four-indices: %d (%.1f ms)" % (num, enum_ms))
    d_val = Dev(L, 8 * num)

    # warmup, then K timed steps bracketed by synchronisation
    for _ in range(args.warmup):
        eri_step(L, c, g, d_idx, num, args.tol, method, d_val.ptr)
    L.gli_synchronize()
    torch.cuda.synchronize()
    # the library records HIP events at its phase boundaries on its own stream during the timed
    # steps (no synchronisation between steps); phases are averaged over the K steps
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    L.gli_enable_phase_timing(1)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        eri_step(L, c, g, d_idx, num, args.tol, method, d_val.ptr)
    ev1.record(stream)
    L.gli_synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    phases = phase_times(L)
    L.gli_enable_phase_timing(0)
    step_ms = ev0.elapsed_time(ev1) / args.steps
    log("step: %.3f ms (events), wall %.3f ms/step" % (step_ms, wall * 1e3 / args.steps))
    log("phases (ms):", {k: round(v, 4) for k, v in phases.items()})

    # per-element FP64 path on a seeded sample, and errors against it
    rng = np.random.default_rng(12345)
    host_idx = np.empty(num, dtype=np.int64)
    _lib.check(L.gli_memcpy(host_idx.ctypes.data, d_idx, 8 * num, 1), "copy indices")
    vals = np.empty(num)
    _lib.check(L.gli_memcpy(vals.ctypes.data, d_val.ptr, 8 * num, 1), "copy values")
    samp = np.sort(rng.choice(num, size=min(args.direct_sample, num), replace=False))
    s_idx = np.ascontiguousarray(host_idx[samp])
    ds_idx = Dev(L, 8 * len(s_idx))
    ds_val = Dev(L, 8 * len(s_idx))
    _lib.check(L.gli_memcpy(ds_idx.ptr, s_idx.ctypes.data, 8 * len(s_idx), 0), "copy sample")
    eri_step(L, c, g, ds_idx.ptr, len(s_idx), args.tol, 2, ds_val.ptr)
    L.gli_synchronize()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    eri_step(L, c, g, ds_idx.ptr, len(s_idx), args.tol, 2, ds_val.ptr)
    e1.record(stream)
    L.gli_synchronize()
    torch.cuda.synchronize()
    direct_ms = e0.elapsed_time(e1)
    # flops of the same run, counted on the device in a separate pass
    _lib.check(L.gli_enable_direct_flop_count(1), "direct flop count")
    eri_step(L, c, g, ds_idx.ptr, len(s_idx), args.tol, 2, ds_val.ptr)
    w_ref, w_exe = C.c_double(), C.c_double()
    _lib.check(L.gli_direct_flops(C.byref(w_ref), C.byref(w_exe)), "direct flops")
    L.gli_enable_direct_flop_count(0)
    direct_vals = np.empty(len(s_idx))
    _lib.check(L.gli_memcpy(direct_vals.ctypes.data, ds_val.ptr, 8 * len(s_idx), 1), "copy direct")
    err_abs = float(np.max(np.abs(vals[samp] - direct_vals)))
    err_rel = float(np.max(np.abs(vals[samp] - direct_vals) / np.maximum(np.abs(direct_vals), 1e-300)))
    direct_rate = len(s_idx) / (direct_ms / 1e3)
    log("direct FP64 sample: %d elements in %.2f ms (%.3g reference-form TF/s, %.3g executed TF/s); max abs err %.3g"
        % (len(s_idx), direct_ms, w_ref.value / direct_ms / 1e9, w_exe.value / direct_ms / 1e9, err_abs))

    # other methods on the same workload: time and error over every element
    compare = []
    for name in ([] if args.no_compare else [m for m in args.compare.split(",") if m and m != args.method]):
        other = METHODS[name]
        d_oth = Dev(L, 8 * num)
        eri_step(L, c, g, d_idx, num, args.tol, other, d_oth.ptr)
        L.gli_synchronize()
        torch.cuda.synchronize()
        m0 = torch.cuda.Event(enable_timing=True)
        m1 = torch.cuda.Event(enable_timing=True)
        m0.record(stream)
        for _ in range(args.steps):
            eri_step(L, c, g, d_idx, num, args.tol, other, d_oth.ptr)
        m1.record(stream)
        L.gli_synchronize()
        torch.cuda.synchronize()
        oth_ms = m0.elapsed_time(m1) / args.steps
        oth_vals = np.empty(num)
        _lib.check(L.gli_memcpy(oth_vals.ctypes.data, d_oth.ptr, 8 * num, 1), "copy values")
        del d_oth
        diff = np.abs(oth_vals - vals)
        big = np.abs(vals) >= 1e-10
        compare.append({"method": name, "dtype": DTYPES[name], "ms_per_step": oth_ms, "value": num / (oth_ms / 1e3),
                        "max_abs_err_vs_%s_all_elements" % args.method: float(diff.max()),
                        "max_rel_err_vs_%s_where_abs_ge_1e-10" % args.method:
                            float((diff[big] / np.abs(vals[big])).max()) if big.any() else None,
                        "max_abs_err_vs_fp64_direct_sample": float(np.max(np.abs(oth_vals[samp] - direct_vals)))})
        del oth_vals, diff, big
        log("%s: %.3f ms/step, max abs err vs %s over all elements %.3g"
            % (name, oth_ms, args.method, compare[-1]["max_abs_err_vs_%s_all_elements" % args.method]))

    # the same workload as a one-grid batch plan (repeated evaluation): every element's key-tensor
    # position is computed once at plan creation, then each replay of the captured graph runs the
    # tables, the three products and a gather from those positions; reported beside `value`
    plan_line = None
    if not args.no_plan:
        plan = C.c_void_p()
        cptr = (_abi.P_double * 1)(c.ctypes.data_as(_abi.P_double))
        Ks = np.array([len(c)], dtype=np.int64)
        dptr = (C.c_void_p * 1)(d_idx.value)
        nn = np.array([num], dtype=np.int64)
        d_pl = Dev(L, 8 * num)
        vptr = (C.c_void_p * 1)(d_pl.ptr.value)
        t0 = time.perf_counter()
        _lib.check(L.gli_eri_batch_create(1, cptr, Ks.ctypes.data_as(_abi.P_glong), g.ctypes.data_as(_abi.P_glong), dptr,
                                          nn.ctypes.data_as(_abi.P_glong), args.tol, method if method in (0, 1, 6) else 0,
                                          vptr, 1, C.byref(plan)), "gli_eri_batch_create")
        create_ms = (time.perf_counter() - t0) * 1e3
        buf = C.create_string_buffer(1 << 16)
        L.gli_eri_batch_describe(plan, buf, len(buf))
        chosen = json.loads(buf.value.decode()).get("gather_chosen")
        for _ in range(args.warmup):
            _lib.check(L.gli_eri_batch_run(plan, cptr, None), "plan run")
        L.gli_synchronize()
        torch.cuda.synchronize()
        p0 = torch.cuda.Event(enable_timing=True)
        p1 = torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        for _ in range(args.steps):
            _lib.check(L.gli_eri_batch_run(plan, cptr, None), "plan run")
        p1.record(stream)
        L.gli_synchronize()
        torch.cuda.synchronize()
        plan_ms = p0.elapsed_time(p1) / args.steps
        pv = np.empty(num)
        _lib.check(L.gli_memcpy(pv.ctypes.data, d_pl.ptr, 8 * num, 1), "copy plan values")
        L.gli_eri_batch_destroy(plan)
        del d_pl
        plan_line = {"ms_per_step": plan_ms, "value": num / (plan_ms / 1e3), "plan_create_ms": create_ms,
                     "gather": chosen, "max_abs_err_vs_single_call": float(np.max(np.abs(pv - vals)))}
        del pv
        log("one-grid plan replay: %.3f ms/step (plan creation %.0f ms), gather %s" % (plan_ms, create_ms, chosen))

    # the mixed-size batch of grids (5^3 K29, 7^3 K43, 9^3 K57 seed 1, 11^3 K57) as one plan replay
    batch_line = None
    if not args.no_batch:
        batch_line = batch_measure(L, args, stream)
        batch_line.pop("autotune", None)
        if args.method != "mixed":
            # the same batch with the split-precision numerics of the precision study (5-digit BF16 slices
            # on MFMA, exact FP32 group sums, FP64 combination)
            bm = batch_measure(L, args, stream, "mixed")
            batch_line["mixed"] = {"method": "mixed", "dtype": "f64+split-bf16", "ms_per_step": bm["ms_per_step"],
                                   "value": bm["value"], "plan_create_ms": bm["plan_create_ms"],
                                   "max_abs_err_vs_fp64": max(e["max_abs_err_vs_fp64"] for e in bm["errors"]),
                                   "note": "slower than the FP64 plan: the split products and their digit passes cost "
                                           "more than FP64 MFMA on MI355X"}

    # the reference entry points with host arrays: indices in and values out over PCIe, the host
    # copies included (reported beside `value`, never as it)
    host_api = None
    if not args.no_host_api:
        hg = _abi.Grid.of(((-(args.n - 1) // 2, args.n),) * 3)
        gd = _abi.GaussletData()
        gd.coefficients = c.ctypes.data_as(_abi.P_double)
        gd.indices.istart = -(len(c) - 1) // 2
        gd.indices.num = len(c)
        _lib.check(L.gli_set_eri_method(method), "method")
        t_enum, t_val = [], []
        diff = None
        for rep in range(2):   # the first call of each also allocates the library's transfer buffers
            ri = _abi.SparseEriIndices()
            t0 = time.perf_counter()
            L.enumerate_symmetry_reduced_eri_indices(C.byref(hg), C.byref(ri))
            t_enum.append((time.perf_counter() - t0) * 1e3)
            L.delete_sparse_eri_indices(C.byref(ri))
            res = _abi.SparseEri()
            t0 = time.perf_counter()
            L.compute_sparse_eri_gausslet_integrals(C.byref(gd), C.byref(hg), host_idx.ctypes.data_as(_abi.P_glong), num,
                                                    args.tol, C.byref(res))
            t_val.append((time.perf_counter() - t0) * 1e3)
            if rep == 1:
                hv = np.ctypeslib.as_array(res.integral_values, shape=(num,))
                diff = float(np.max(np.abs(hv - vals)))
                del hv
            L.delete_sparse_eri_gausslet_integrals(C.byref(res))
        _lib.check(L.gli_set_eri_method(0), "method")
        host_api = {"enumerate_ms": t_enum[1], "values_ms": t_val[1], "values_elements_per_s": num / (t_val[1] / 1e3),
                    "max_abs_diff_vs_device_path": diff,
                    "note": "enumerate_symmetry_reduced_eri_indices and compute_sparse_eri_gausslet_integrals with "
                            "host arrays (2nd call), wall clock"}
        log("host API: enumerate %.1f ms, values %.1f ms (diff vs device path %.3g)" % (t_enum[1], t_val[1], diff))

    # the largest single-GPU case (13^3): methods and their errors against exact FP64, and ERI-DA; after
    # the host-array calls, whose wall clock it could otherwise disturb
    large_line = None
    if not args.no_large:
        large_line = large_measure(L, args, stream)
        torch.cuda.empty_cache()

    # roofline of the dominant kernel
    workload = "eri_%dx%dx%d_K%d_seed%d%s" % (args.n, args.n, args.n, args.K, args.seed, "_asym" if args.asymmetric else "")
    dom = max(phases.items(), key=lambda kv: kv[1]) if phases else ("step", step_ms)
    name, dom_ms = dom
    M = 6 * (args.n - 1) + 2 * (args.K - 1)
    Mp = (M + 1 + 15) // 16 * 16
    C1 = L.gli_eri_key_count(args.n, int(np.array_equal(c, c[::-1])))
    Cp = (C1 + 63) // 64 * 64
    if name in ("gather", "gather_bucketed", "gather_classed"):
        # the library names the phase after the gather kernel it launched
        kernel = "k_" + name
        achieved = 16.0 * num / (dom_ms / 1e3) / 1e9   # index in + value out, bytes per element
        traffic = measured_traffic(kernel, workload)
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": ("profiles/%s_%s_traffic.md: rocprofv3 FETCH_SIZE + WRITE_SIZE of one dispatch, "
                                   "calibrated with a copy kernel of known size (a committed profile, not measured in "
                                   "this run)" % (kernel, workload)) if traffic is not None else None,
                "kernel": kernel, "algorithmic_bytes_per_launch": 16 * num}
    else:
        flops = {"mode1": 2.0 * Mp * Mp * Cp * Mp, "mode2": 2.0 * Mp * Cp * Cp * Mp / 2,
                 "mode3": 2.0 * Cp * Cp * Cp * Mp / 6}.get(name, 0.0)
        achieved = flops / (dom_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFS, "traffic": None, "kernel": "k_mode_product:" + name}
    roof["kernel_ms"] = dom_ms

    cpu = None
    cpu_err = None
    if not args.no_cpu:
        log("CPU baseline ...")
        cpu, pos, cpu_vals = cpu_baseline(c, args.n, host_idx, args.tol, args.cpu_seconds, args.seed,
                                          "%d^3" % args.n)
        cpu_err = float(np.max(np.abs(vals[pos] - cpu_vals)))
        log("cpu:", cpu, "max abs err vs cpu", cpu_err)

    value = num / (step_ms / 1e3)
    line = {
        "metric": "unique ERI elements/s at 11^3 grid, mixed vs FP64, err<=1e-10 vs FP64",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.method],
        "data": "synthetic (seeded gausslet coefficients)",
        "config": {"workload": workload,
                   "grid": args.n, "K": args.K, "seed": args.seed, "tol": args.tol, "method": args.method,
                   "unique_elements": num, "axis_keys": C1, "tuning": tuning},
        "roofline": roof,
        "cpu_baseline": cpu,
        "phases_ms": phases,
        "enumeration_ms": enum_ms,
        "host_api": host_api,
        "fp64_direct": {"elements_per_s": direct_rate, "sample": len(s_idx), "ms": direct_ms,
                        "reference_form_flops_per_element": w_ref.value / len(s_idx),
                        "executed_flops_per_element": w_exe.value / len(s_idx),
                        "achieved_tflops_reference_form": w_ref.value / (direct_ms / 1e3) / 1e12,
                        "achieved_tflops_executed": w_exe.value / (direct_ms / 1e3) / 1e12,
                        "frac_fp64_vector_peak_executed": w_exe.value / (direct_ms / 1e3) / 1e12 / FP64_PEAK_TFS},
        "max_abs_err_vs_fp64_direct": err_abs,
        "max_rel_err_vs_fp64_direct": err_rel,
        "max_abs_err_vs_cpu_reference": cpu_err,
        "speedup_vs_fp64_direct": value / direct_rate,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "speedup_vs_cpu_host_api": (host_api["values_elements_per_s"] / cpu["value"]) if (cpu and host_api) else None,
        "compare": compare,
        "mixed_vs_fp64": mixed_vs_fp64(compare, args.method, value),
        "plan_replay": plan_line,
        "batch": batch_line,
        "large": large_line,
    }
    L.gli_device_free(d_idx)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
