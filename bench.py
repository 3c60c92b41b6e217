#!/usr/bin/env python3
"""Benchmark of the split-operator hot path on one MI355X.

Workload (the headline configuration): 2D split-operator evolution, O2
(Strang), 1024 x 1024 periodic grid on [-150,150)^2, dt = 0.2, complex128;
potential = 16 seeded Gaussian bumps plus an imaginary absorbing layer;
Gaussian packet sigma=8 at (-60,0), k=(0.8,0).  One "step" = one split step
of the whole grid.  psi, V and the phase tables are resident in HBM before
the timed region starts.

Prints one JSON line: grid-pt-steps/s, the per-kernel HBM roofline of the
dominant kernel (algorithmic bytes / HIP-event-measured launch duration) and
the CPU baseline (the FP64 oracle port of the reference CpuSeq+KISS path timed
on this host's threads).  Further sections in the same line:
* fp32_mixed: the FP32-mixed form batched (16 x 1024^2) and at 2048^2 (alone,
  4 batched), against FP64 alone and batched, and its psi error after 1000
  steps on both grids;
* born_3d: 3D Born series, 128^3 x 32 energies (energy-iterations/s,
  amplitudes against complex128 runs);
* serving: the request loop's solves/s, p50/p99 latency and FP64 spot checks."""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("2D split-op grid-pt-steps/s FP64 vs FP32 @L2err≤1e-5; "
          "served solves/s + p99")
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per grid point per launch (complex128 psi read + write,
# plus the potential factor read by the row sweep)
ROW_B_PER_PT = {64: 48.0, 32: 32.0}   # FP32-mixed reads an FP64 (hi+lo) factor
COL_B_PER_PT = {64: 32.0, 32: 16.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_job(N, prob, precision):
    from qusim_amd import _native as Nn
    n = prob["n"]
    psi = Nn.DeviceArray.from_host(prob["psi"])
    vre = Nn.DeviceArray.from_host(np.ascontiguousarray(prob["V"].real))
    vim = Nn.DeviceArray.from_host(np.ascontiguousarray(prob["V"].imag))
    d = Nn.SplitDesc()
    d.dim, d.nx, d.ny = 2, n, n
    d.x0, d.x1, d.y0, d.y1 = prob["x0"], prob["x1"], prob["x0"], prob["x1"]
    d.dt_re, d.dt_im = prob["dt"], 0.0
    d.mass = d.hbar = 1.0
    d.method, d.boundary, d.precision = 2, 0, precision
    p = ctypes.c_void_p()
    Nn.call("qsim_hip_split_create", ctypes.byref(p), ctypes.byref(d), vre.ptr, vim.ptr, psi.ptr)
    return p.value, (psi, vre, vim)


def autotune(job, reps=20):
    """pick the fastest launch variants of the row / column sweeps"""
    from qusim_amd import _native as Nn
    vr, vc = ctypes.c_int(), ctypes.c_int()
    ru, cu = ctypes.c_double(), ctypes.c_double()
    Nn.call("qsim_hip_split_autotune", job, reps, ctypes.byref(vr), ctypes.byref(vc),
            ctypes.byref(ru), ctypes.byref(cu), None)
    xs = ctypes.c_int()
    Nn.call("qsim_hip_split_get_schedule", job, None, None, ctypes.byref(xs))
    return {"row_variant": vr.value, "col_variant": vc.value, "xsplit": xs.value,
            "pair_us": cu.value}  # one alternating row + column sweep pair


def time_steps(job, steps, warmup):
    from qusim_amd import _native as Nn
    if warmup:
        Nn.call("qsim_hip_split_steps", job, warmup, None)
    Nn.synchronize()
    ev0, ev1 = Nn.Event(), Nn.Event()
    t0 = time.perf_counter()
    ev0.record()
    Nn.call("qsim_hip_split_steps", job, steps, None)
    ev1.record()
    Nn.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, ev0.elapsed_ms(ev1) / 1e3


def kernel_times(job, steps):
    """HIP-event time of every sweep launch of `steps` steps, by kind:
    0 row sweep (MID), 1 column sweep, 2 1D kernel, 3 first row sweep (E_V,
    FFT_y), 4 last row sweep (IFFT_y, E_V)."""
    from qusim_amd import _native as Nn
    ms = (ctypes.c_double * 5)()
    cnt = (ctypes.c_int * 5)()
    Nn.call("qsim_hip_split_steps_timed", job, steps, ms, cnt, None)
    return list(ms), list(cnt)


# (kind, name, algorithmic bytes per grid point by precision, launches per step)
# FP32-mixed: work is complex64 (8 B), psi complex128 (16 B), E_V FP64 (16 B)
SWEEP_KINDS = ((0, "row_sweep", {64: 48.0, 32: 32.0}, 1),          # work r+w, E_V r
               (1, "col_sweep", {64: 32.0, 32: 16.0}, 1),          # work r+w
               (3, "row_sweep_first", {64: 48.0, 32: 40.0}, 0),    # psi, E_V r; work w
               (4, "row_sweep_last", {64: 48.0, 32: 40.0}, 0))     # work, E_V r; psi w


def roofline(ms, cnt, npts, precision):
    """Per-kernel average launch time and algorithmic GB/s.  The dominant
    kernel is the one with the largest time per step (average launch time x
    launches per step); the first and last row sweeps run once per call."""
    per = {}
    for k, name, bpp, per_step in SWEEP_KINDS:
        if cnt[k]:
            avg_s = ms[k] / cnt[k] / 1e3
            b = bpp[precision] * npts
            per[name] = {"avg_us": avg_s * 1e6, "launches": cnt[k], "bytes_per_launch": b,
                         "achieved_GBs": b / avg_s / 1e9,
                         "frac": b / avg_s / 1e9 / HBM_PEAK_GBS,
                         "us_per_step": avg_s * 1e6 * per_step}
    dom = max(per, key=lambda k: per[k]["us_per_step"])
    return dom, per


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def host_threads(cap=16):
    """threads for the CPU baseline: this process's CPU set, at most `cap`
    (one GPU's share of the host)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


def traffic_from_profiles(name, schedule):
    """HBM bytes per launch of sweep `name` ("row_sweep" or "col_sweep") under
    the launch schedule (row variant, column variant, x-split) from the
    committed PMC summary, when that kernel was measured there; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        import re
        t = json.load(open(path))
        pat = re.compile(dict(probe_kernels(*schedule))[name])
        for kname, ent in t["by_kernel"].items():
            if pat.search(kname):
                return ent["bytes_per_launch"]
        return None
    except Exception:
        return None


# kernels the PMC summary reports: (key, substring of the rocprof kernel name)
PROBE_COPY_N = 1 << 20          # complex128 elements per calibration copy (16 MiB each way)
# launch schedules of the traffic probe (row variant, column variant, x-split
# factor): the ones the autotuner picks for the C2 job on MI355X
PROBE_SCHEDULES = ((6, 0, 2), (6, 1, 2), (6, 4, 1), (6, 5, 1), (6, 1, 1), (0, 0, 4))


def probe_kernels(rv, cv, xs):
    """(key, regular expression of the rocprof kernel name) of the C2 middle
    row sweep and column sweep under a launch schedule.  x-split row sweeps
    are instantiated per radix plan: row variants 1 and 5 run plan 0, 3 runs
    2, 7 runs 6."""
    if xs == 1:
        return (("row_sweep", r"k_row_sweep<double, 1024, 1, %d, (true|false)>" % rv),
                ("col_sweep", r"k_col_sweep<double, 1024, %d, (true|false), 1>" % cv))
    plan = {1: 0, 5: 0, 3: 2, 7: 6}.get(rv, rv)
    return (("row_sweep", r"k_row_sweep_x<double, 1024, 1, %d, (true|false), %d>" % (plan, xs)),
            ("col_sweep", r"k_col_sweep<double, %d, %d, (true|false), %d>" % (1024 // xs, cv, xs)))


def batch_probe(steps, njobs, variants, n=1024):
    """the FP32-mixed batched workload alone, for rocprofv3: njobs n^2 jobs
    (the 1024^2 construction, or the 2048^2 dt = 0.1 one) advanced together
    with fixed launch variants (no autotuner trials in the trace)"""
    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c2_problem, c3_problem
    Nn.init(0)
    prob = c2_problem() if n == 1024 else c3_problem(n)
    jobs = [make_job(Nn, prob, 32) for _ in range(njobs)]
    ptrs = [j[0] for j in jobs]
    for p_ in ptrs:
        Nn.call("qsim_hip_split_set_variants", p_, variants[0], variants[1])
    arr = (ctypes.c_void_p * njobs)(*ptrs)
    b = ctypes.c_void_p()
    Nn.call("qsim_hip_split_batch_create", ctypes.byref(b), njobs)
    Nn.call("qsim_hip_split_batch_steps", b, arr, njobs, steps, None)
    Nn.synchronize()
    Nn.call("qsim_hip_split_batch_destroy", b)
    for p_ in ptrs:
        Nn.call("qsim_hip_split_destroy", p_)
    log("batch probe: %d FP32-mixed %d^2 jobs, %d steps, variants %s" % (njobs, n, steps, variants))


def fft_probe(reps):
    """FourierTransform plans alone, for rocprofv3: `reps` forward 2D/3D
    transforms each of 1024^2 and 2048^2 (complex128 and complex64) and 128^3
    (both), out of place, from a constant input; no timing."""
    from qusim_amd import _native as Nn
    from qusim_amd.fft import FourierTransform
    Nn.init(0)
    for dims in ((1024, 1024), (2048, 2048), (128, 128, 128)):
        for prec in (64, 32):
            dt = np.complex128 if prec == 64 else np.complex64
            a = Nn.DeviceArray.from_host(np.ones(dims, dt))
            b = Nn.DeviceArray(dims, dt)
            f = FourierTransform(dims, False, prec)
            for _ in range(reps):
                f.transform(a, b)
            Nn.synchronize()
            a.close()
            b.close()
    log("fft probe: %d transforms per shape and precision" % reps)


def traffic_probe(steps):
    """Workload of the rocprofv3 --pmc passes: 3 calibration copies of
    PROBE_COPY_N complex128 (known bytes), then `steps` C2 steps with fixed
    launch variants.  No timing, no autotuning."""
    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c2_problem
    Nn.init(0)
    a = Nn.DeviceArray.from_host(np.zeros(PROBE_COPY_N, np.complex128))
    b = Nn.DeviceArray.from_host(np.zeros(PROBE_COPY_N, np.complex128))
    for _ in range(3):
        Nn.call("qsim_hip_copy", b.ptr, a.ptr, PROBE_COPY_N, None)
    Nn.synchronize()
    job, keep = make_job(Nn, c2_problem(), 64)
    for rv, cv, xs in PROBE_SCHEDULES:
        Nn.call("qsim_hip_split_set_variants", job, 0, 0)
        Nn.call("qsim_hip_split_set_xsplit", job, xs)
        Nn.call("qsim_hip_split_set_variants", job, rv, cv)
        Nn.call("qsim_hip_split_steps", job, steps, None)
        Nn.synchronize()
    Nn.call("qsim_hip_split_destroy", job)
    log("traffic probe: 3 copies of %d complex128, %d C2 steps under each schedule %s"
        % (PROBE_COPY_N, steps, PROBE_SCHEDULES))


def _pmc_rows(d):
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % d)
    rows = []
    for f in files:
        rows += list(csv.DictReader(open(f)))
    return rows


def _pmc_avg(rows, counter, sub):
    """average over dispatches of kernels whose name contains `sub` (the
    counter summed over its instances within one dispatch)."""
    per = {}
    for r in rows:
        if r["Counter_Name"] != counter or sub not in r["Kernel_Name"]:
            continue
        key = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    return (sum(per.values()) / len(per), len(per)) if per else (None, 0)


def pmc_summary(fetch_dir, write_dir):
    """Turn the FETCH_SIZE and WRITE_SIZE passes into profiles/traffic.json:
    each counter (KiB per dispatch) scaled by what it reads for the copy
    kernel of known size, for every middle row sweep and column sweep of the
    probe's schedules (by kernel name)."""
    import re
    fr, wr = _pmc_rows(fetch_dir), _pmc_rows(write_dir)
    copy_kib = PROBE_COPY_N * 16 / 1024.0
    cf, nf = _pmc_avg(fr, "FETCH_SIZE", "qsim_hip_copy")
    cw, nw = _pmc_avg(wr, "WRITE_SIZE", "qsim_hip_copy")
    out = {"calibration": {"copy_bytes_each_way": copy_kib * 1024, "copy_fetch_kib": cf,
                           "copy_write_kib": cw, "fetch_factor": copy_kib / cf,
                           "write_factor": copy_kib / cw, "dispatches": [nf, nw]},
           "schedules": [list(x) for x in PROBE_SCHEDULES], "by_kernel": {}}
    names = sorted({re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", "") for r in fr}
                   | {re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", "") for r in wr})
    for kname in names:
        if re.search(r"k_row_sweep(_x)?<double, 1024, 1,", kname):
            alg = ROW_B_PER_PT[64] * (1 << 20)
        elif "k_col_sweep<double" in kname:
            alg = COL_B_PER_PT[64] * (1 << 20)
        else:
            continue
        f, n1 = _pmc_avg(fr, "FETCH_SIZE", kname + "(")
        w, n2 = _pmc_avg(wr, "WRITE_SIZE", kname + "(")
        if f is None or w is None:
            continue
        fb = f * 1024 * out["calibration"]["fetch_factor"]
        wb = w * 1024 * out["calibration"]["write_factor"]
        out["by_kernel"][kname] = {"fetch_bytes": fb, "write_bytes": wb,
                                   "bytes_per_launch": fb + wb, "algorithmic_bytes": alg,
                                   "ratio": (fb + wb) / alg, "raw_fetch_kib": f,
                                   "raw_write_kib": w, "dispatches": [n1, n2]}
    path = os.path.join(ROOT, "profiles", "traffic.json")
    json.dump(out, open(path, "w"), indent=1)
    print(json.dumps(out, indent=1))


def batched_rate(prob, precision, njobs, steps=40, reps=10):
    """grid-pt-steps/s of njobs same-shape jobs advanced together (one launch
    per sweep for all of them), variants from the batched autotuner"""
    from qusim_amd import _native as Nn
    jobs = [make_job(Nn, prob, precision) for _ in range(njobs)]
    ptrs = [j[0] for j in jobs]
    arr = (ctypes.c_void_p * njobs)(*ptrs)
    vr, vc = ctypes.c_int(), ctypes.c_int()
    ru, cu = ctypes.c_double(), ctypes.c_double()
    Nn.call("qsim_hip_split_autotune_batched", arr, njobs, reps, ctypes.byref(vr),
            ctypes.byref(vc), ctypes.byref(ru), ctypes.byref(cu), None)
    b = ctypes.c_void_p()
    Nn.call("qsim_hip_split_batch_create", ctypes.byref(b), njobs)
    try:
        Nn.call("qsim_hip_split_batch_steps", b, arr, njobs, 2, None)
        Nn.synchronize()
        e0, e1 = Nn.Event(), Nn.Event()
        e0.record()
        Nn.call("qsim_hip_split_batch_steps", b, arr, njobs, steps, None)
        e1.record()
        Nn.synchronize()
        secs = e0.elapsed_ms(e1) / 1e3
    finally:
        Nn.call("qsim_hip_split_batch_destroy", b)
        for p_ in ptrs:
            Nn.call("qsim_hip_split_destroy", p_)
    n = prob["n"]
    return {"jobs": njobs, "value": njobs * steps * n * n / secs,
            "us_per_step_per_job": secs / steps / njobs * 1e6,
            "launch_variants": [vr.value, vc.value]}


def group_section(steps=120, passes=20):
    """FP32-mixed 2D jobs of the request mix's sizes (8 x 256^2, 4 x 512^2,
    2 x 1024^2, the C2 construction at each size) advanced `steps` steps:
    one multi-size group (one launch per sweep for all jobs, advances of
    `passes` passes, with and without graph replay) against the same jobs as
    one same-size batch per size (qsim_hip_split_batch_steps), all on one
    stream.  The first advance (which captures the graph) is not timed."""
    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c2_problem
    sizes = [256] * 8 + [512] * 4 + [1024] * 2
    probs = {n: c2_problem(n=n) for n in set(sizes)}
    pts = sum(n * n for n in sizes)
    out = {"workload": "8 x 256^2 + 4 x 512^2 + 2 x 1024^2 FP32-mixed 2D jobs, %d steps" % steps}
    for use_graph in (1, 0):
        jobs = [make_job(Nn, probs[n], 32) for n in sizes]
        g = ctypes.c_void_p()
        Nn.call("qsim_hip_split_group_create", ctypes.byref(g), len(jobs), 32, 2)
        g = g.value
        for j, _ in jobs:
            Nn.call("qsim_hip_split_group_add", g, j, steps)
        Nn.call("qsim_hip_split_group_advance", g, passes, use_graph, None)
        Nn.synchronize()
        e0, e1 = Nn.Event(), Nn.Event()
        e0.record()
        left = steps + 1 - passes
        nadv = 0
        while left > 0:
            Nn.call("qsim_hip_split_group_advance", g, passes, use_graph, None)
            left -= passes
            nadv += 1
        e1.record()
        Nn.synchronize()
        secs = e0.elapsed_ms(e1) / 1e3
        Nn.call("qsim_hip_split_group_destroy", g)
        for j, _ in jobs:
            Nn.call("qsim_hip_split_destroy", j)
        out["group_graph" if use_graph else "group_direct"] = {
            "value": (steps - passes) * pts / secs, "unit": "grid-pt-steps/s",
            "advances_timed": nadv, "passes_per_advance": passes}
    jobs = [make_job(Nn, probs[n], 32) for n in sizes]
    by = {}
    for (j, _), n in zip(jobs, sizes):
        by.setdefault(n, []).append(j)
    batches = {}
    for n, js in by.items():
        b = ctypes.c_void_p()
        Nn.call("qsim_hip_split_batch_create", ctypes.byref(b), len(js))
        batches[n] = (b.value, (ctypes.c_void_p * len(js))(*js), len(js))
    for b, arr, k in batches.values():
        Nn.call("qsim_hip_split_batch_steps", b, arr, k, 2, None)
    Nn.synchronize()
    e0, e1 = Nn.Event(), Nn.Event()
    e0.record()
    for b, arr, k in batches.values():
        Nn.call("qsim_hip_split_batch_steps", b, arr, k, steps, None)
    e1.record()
    Nn.synchronize()
    secs = e0.elapsed_ms(e1) / 1e3
    for b, _, _ in batches.values():
        Nn.call("qsim_hip_split_batch_destroy", b)
    for j, _ in jobs:
        Nn.call("qsim_hip_split_destroy", j)
    out["per_size_batches"] = {"value": steps * pts / secs, "unit": "grid-pt-steps/s"}
    out["group_vs_batches"] = out["group_graph"]["value"] / out["per_size_batches"]["value"]
    return out


def large_grid_section(n=4096, steps=20):
    """The C2 construction at n^2 in complex128: psi and the E_V table are
    256 MiB each at 4096^2, beyond the 256 MB Infinity Cache, so the sweeps
    stream from HBM.  Autotuned variants, HIP-event timing, per-kernel bytes
    as for the headline."""
    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c2_problem
    prob = c2_problem(n=n)
    job, keep = make_job(Nn, prob, 64)
    tuned = autotune(job, reps=5)
    wall, gpu = time_steps(job, steps, 2)
    ms, cnt = kernel_times(job, 4)
    dom, per = roofline(ms, cnt, n * n, 64)
    Nn.call("qsim_hip_split_destroy", job)
    rate = steps * n * n / wall
    return {"workload": "C2 construction at %dx%d, complex128, dt 0.2" % (n, n),
            "value": rate, "unit": "grid-pt-steps/s", "steps": steps,
            "ms_per_step": wall / steps * 1e3, "step_frac": rate * 80.0 / 1e9 / HBM_PEAK_GBS,
            "launch_variants": tuned, "dominant_kernel": dom, "kernels": per}


def c1_section(threads):
    """C1 (1D, 2048 points, dt 0.5, 1000 O2 steps) on the GPU in complex128:
    wall time of the whole run (one fused launch), and psi, the norm and the
    transmitted probability P(x > 0) against the oracle."""
    from oracle import pyoracle as O
    from qusim_amd.api import BoundaryCondition, Evolver1D, Options, SolverMethod
    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c1_problem
    p = c1_problem()
    e = Evolver1D().init(p["psi"], False, p["dt"], False, p["V"], p["x0"], p["x1"], p["n"],
                         BoundaryCondition.Period, SolverMethod.SplittingMethodO2, 1.0, 1.0,
                         Options().Hip().Batch(p["steps"]))
    e.step()
    Nn.synchronize()
    e2 = Evolver1D().init(p["psi"], False, p["dt"], False, p["V"], p["x0"], p["x1"], p["n"],
                          BoundaryCondition.Period, SolverMethod.SplittingMethodO2, 1.0, 1.0,
                          Options().Hip().Batch(p["steps"]))
    Nn.synchronize()
    t0 = time.perf_counter()
    e2.step()
    Nn.synchronize()
    secs = time.perf_counter() - t0
    ref = O.split1d(p["psi"], p["V"], p["x0"], p["x1"], p["dt"], p["steps"], method=2)
    got = e.GetPsi()
    pt_ref = O.norm2_range_1d(ref, p["x0"], p["x1"], 0.0, p["x1"])
    return {"workload": "C1: 1D 2048 points on [-100,100), dt 0.5, 1000 O2 steps, complex128",
            "gpu_seconds": secs, "value": p["steps"] * p["n"] / secs,
            "unit": "grid-pt-steps/s",
            "psi_rel_l2_vs_oracle": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            "norm": e.Norm2(), "transmitted": e.Norm2(0.0, p["x1"]),
            "transmitted_oracle": pt_ref}


def fp32_error(prob, steps):
    """psi relative L2 error and norm drift of the FP32-mixed path against the
    complex128 path of this device after `steps` steps from the same state
    (the complex128 path's parity with the oracle is the test suite's)"""
    from qusim_amd import _native as Nn
    n = prob["n"]
    j64, keep64 = make_job(Nn, prob, 64)
    j32, keep32 = make_job(Nn, prob, 32)
    Nn.call("qsim_hip_split_steps", j64, steps, None)
    Nn.call("qsim_hip_split_steps", j32, steps, None)
    a = keep64[0].download()
    b = keep32[0].download()
    for j in (j64, j32):
        Nn.call("qsim_hip_split_destroy", j)
    dx = (prob["x1"] - prob["x0"]) / n
    n64 = float(np.sum(np.abs(a) ** 2) * dx * dx)
    n32 = float(np.sum(np.abs(b) ** 2) * dx * dx)
    return {"psi_rel_l2_vs_fp64": float(np.linalg.norm(b - a) / np.linalg.norm(a)),
            "steps": steps, "norm_drift_vs_fp64": (n32 - n64) / n64}


def transmitted(psi, prob):
    """P(x > 0) = sum over x_i >= 0 of |psi|^2 dx dy (x is the first axis)."""
    n = prob["n"]
    dx = (prob["x1"] - prob["x0"]) / n
    x = prob["x0"] + dx * np.arange(n)
    return float(np.sum(np.abs(psi[x >= 0.0]) ** 2) * dx * dx)


def c3_vs_oracle(prob, steps, threads):
    """The configuration's own check: the complex128 and the FP32-mixed GPU
    results after `steps` steps against the FP64 oracle (the CPU restatement
    of the reference, complex potential) from the same state: psi relative
    L2, norm drift and the transmitted probability P(x > 0)."""
    from oracle import pyoracle as O
    from qusim_amd import _native as Nn
    n = prob["n"]
    got = {}
    for prec in (64, 32):
        j, keep = make_job(Nn, prob, prec)
        Nn.call("qsim_hip_split_steps", j, steps, None)
        got[prec] = keep[0].download().reshape(n, n)
        Nn.call("qsim_hip_split_destroy", j)
    t0 = time.perf_counter()
    ref = O.split2d(prob["psi"], prob["V"], prob["x0"], prob["x1"], prob["x0"], prob["x1"],
                    prob["dt"], steps, method=2, complex_v=True, threads=threads)
    osecs = time.perf_counter() - t0
    dx = (prob["x1"] - prob["x0"]) / n
    nref = float(np.sum(np.abs(ref) ** 2) * dx * dx)
    pref = transmitted(ref, prob)
    out = {"steps": steps, "oracle_seconds": osecs, "oracle_threads": threads,
           "oracle_norm": nref, "oracle_transmitted": pref}
    for prec, tag in ((64, "complex128"), (32, "fp32_mixed")):
        a = got[prec]
        na = float(np.sum(np.abs(a) ** 2) * dx * dx)
        pa = transmitted(a, prob)
        out[tag] = {"psi_rel_l2": float(np.linalg.norm(a - ref) / np.linalg.norm(ref)),
                    "norm_drift": (na - nref) / nref, "transmitted": pa,
                    "transmitted_rel_err": (pa - pref) / pref}
    return out


def fp32_section(steps_err, steps_rate, fp64_plain_1024, oracle_threads):
    """FP32-mixed form (complex64 psi and FFTs, FP64-built factors applied as
    float-float pairs): the 2048^2 dt = 0.1 grid alone and batched, the
    1024^2 grid of the FP64 headline batched, and the psi error against the
    complex128 path after `steps_err` steps on both grids."""
    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c2_problem, c3_problem
    p3 = c3_problem()
    n = p3["n"]
    j64, _k64 = make_job(Nn, p3, 64)
    j32, _k32 = make_job(Nn, p3, 32)
    autotune(j64)
    tuned32 = autotune(j32)
    w32, g32 = time_steps(j32, steps_rate, 2)
    rate32 = steps_rate * n * n / g32
    w64, g64 = time_steps(j64, steps_rate, 2)
    rate64 = steps_rate * n * n / g64
    ms, cnt = kernel_times(j32, 4)
    dom, per = roofline(ms, cnt, n * n, 32)
    for j in (j64, j32):
        Nn.call("qsim_hip_split_destroy", j)
    b3 = batched_rate(p3, 32, 4)
    p2 = c2_problem()
    b2 = batched_rate(p2, 32, 16)
    b2_64 = batched_rate(p2, 64, 16)
    e3 = fp32_error(p3, steps_err)
    e2 = fp32_error(p2, steps_err)
    log("C3 against the oracle: 2048^2, %d steps (oracle on %d threads)"
        % (steps_err, oracle_threads))
    c3o = c3_vs_oracle(p3, steps_err, oracle_threads)
    return {
        "workload": "FP32-mixed split-operator O2: complex64 psi and FFTs, factors built in "
                    "FP64 and applied as float-float pairs",
        "value": b2["value"], "unit": "grid-pt-steps/s",
        "speedup_vs_fp64": b2["value"] / fp64_plain_1024,
        "speedup_basis": "16 batched 1024^2 FP32-mixed jobs vs the plain FP64 1024^2 job",
        "grid_1024": {"workload": "1024x1024, dt=0.2 (the FP64 headline grid)",
                      "fp32_batched": b2, "fp64_plain": fp64_plain_1024,
                      "fp64_batched": dict(b2_64, roofline_frac=b2_64["value"] / 1.0e11),
                      "speedup_vs_fp64_plain": b2["value"] / fp64_plain_1024,
                      "speedup_vs_fp64_batched": b2["value"] / b2_64["value"],
                      "error": e2},
        "grid_2048": {"workload": "2048x2048, dt=0.1",
                      "fp32_single": rate32, "fp32_single_variants": tuned32,
                      "fp32_batched": b3, "fp64_single": rate64,
                      "speedup_single_vs_fp64": rate32 / rate64,
                      "speedup_batched_vs_fp64": b3["value"] / rate64,
                      # against the plain FP64 rate of the 1024^2 headline job
                      "speedup_single_vs_c2_fp64": rate32 / fp64_plain_1024,
                      "speedup_batched_vs_c2_fp64": b3["value"] / fp64_plain_1024,
                      "error": e3, "error_vs_oracle": c3o,
                      "dominant_kernel": dom, "kernels": per},
    }


# Algorithmic bytes per point-iteration of a Born iteration: a pencil sweep
# (x lines: read + write of one complex array) and one fused y-z plane sweep
# reading tmp, dpsi and V and writing dpsi and tmp, with psi0 generated on the
# fly: 16 + 40 = 56 B in complex64 (112 B in complex128).  For FP32 128^2
# planes with the Vellekoop, Hao1 or identity preconditioner (C4 uses
# Vellekoop) the library runs exactly that schedule (one merged y-z plane
# pass + the outer x pass); complex128 runs the inner, two middle and the
# outer passes (176 B).
BORN_B_ALG = {64: 112.0, 32: 56.0}
BORN_B_SCHEDULE = {64: 176.0, 32: 56.0}


def born_section(max_iter=200, tol=1e-8, checks=(0, 7, 31)):
    """3D Born series, 128^3 on [-32,32)^3, 32 energies k ~ U(1,3) in one
    job: complex64 iterates with FP64 V, tables and residual, Vellekoop,
    slow 0.5, stopping at a relative update below `tol` or `max_iter`.
    Far-field amplitudes (forward and 37 degrees off axis) of a few energies
    are compared with complex128 single-energy jobs run for the same number
    of iterations."""
    from qusim_amd import _native as Nn
    from qusim_amd.api import Options, QuPerturbation3D, SolverMethod
    from qusim_amd.workloads import c4_energies, c4_potential
    n, L = 128, 32.0
    V, Es = c4_potential(n=n, L=L), c4_energies()

    def job(prec, energies):
        o = Options().Hip().Order(1).Preconditional(True).VellekoopPreconditioner().Slow(0.5)
        if prec == 32:
            o.Precision("mixed32")
        return QuPerturbation3D().init(V, -L, L, n, -L, L, n, -L, L, n, energies, 0.0, 0, 0, 1,
                                       SolverMethod.BornSerise, 1.0, 1.0, o)

    s32 = job(32, Es)
    Nn.synchronize()
    t0 = time.perf_counter()
    s32.Compute(max_iter=max_iter, tol=tol)
    Nn.synchronize()
    secs = time.perf_counter() - t0
    its = s32.iters
    rate = sum(its) / secs
    errs = []
    t64 = 0.0
    for e in checks:
        s64 = job(64, Es[e])
        Nn.synchronize()
        t1 = time.perf_counter()
        s64.Compute(max_iter=its[e], tol=0.0)
        Nn.synchronize()
        t64 += time.perf_counter() - t1
        row = {"energy": Es[e], "iterations": its[e]}
        for tag, d in (("forward", (0.0, 0.0, 1.0)), ("side", (0.6, 0.0, 0.8))):
            a64, a32 = s64.ScatteringAmplitude(*d), s32.ScatteringAmplitude(*d, e)
            row["amp_rel_err_" + tag] = abs(a32 - a64) / abs(a64)
        p64, p32 = s64.GetPsi(), s32.GetPsi(e)
        row["dpsi_rel_l2"] = float(np.linalg.norm(p32 - p64) / np.linalg.norm(p64))
        errs.append(row)
        s64.close()
    s32.close()
    gbs = rate * n ** 3 * BORN_B_ALG[32] / 1e9
    roof = HBM_PEAK_GBS * 1e9 / (BORN_B_ALG[32] * n ** 3)
    return {"workload": "3D Born series 128^3 on [-32,32)^3, 32 energies k~U(1,3), Vellekoop "
                        "slow 0.5, complex64 iterates + FP64 residual, tol %g / %d iterations"
                        % (tol, max_iter),
            "value": rate, "unit": "energy-iterations/s", "seconds": secs,
            "iterations_min_max": [min(its), max(its)],
            "residual_max": max(s32.residual), "converged": sum(r < tol for r in s32.residual),
            "fp64_single_energy_rate": sum(errs_i["iterations"] for errs_i in errs) / t64,
            "algorithmic_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
            "roofline_energy_iterations_per_s": roof,
            "bytes_per_point_iteration": BORN_B_ALG[32],
            "schedule_bytes_per_point_iteration": BORN_B_SCHEDULE[32],
            "schedule_GBs": rate * n ** 3 * BORN_B_SCHEDULE[32] / 1e9,
            "checks_vs_fp64": errs}


def serving_section(count, rate, budget_gib, spot_every=50, capacity_count=600):
    """Request loop over the seeded mix (60% 1D, 30% 2D evolutions, 10% 2D
    Born), Poisson arrivals at `rate`/s, FP32-mixed, admission by HBM
    footprint; every `spot_every`-th request re-solved alone in FP64.  The
    capacity line serves the first `capacity_count` requests all arriving at
    t = 0 (solves/s with the device never idle)."""
    from qusim_amd.serving import Server, make_requests, spot_check
    reqs = make_requests(count, rate=rate)
    cap = Server(precision=32, hbm_budget=budget_gib << 30, realtime=False)
    cst = cap.run(reqs[:capacity_count])
    keep = [r.rid for r in reqs[::spot_every]]
    srv = Server(precision=32, hbm_budget=budget_gib << 30, keep=keep)
    st = srv.run(reqs)
    st["capacity"] = {"requests": cst["requests"], "solves_per_s": cst["solves_per_s"],
                      "makespan_s": cst["makespan_s"]}
    st["offered_rate_per_s"] = rate
    st["spot_check_fp64"] = spot_check(reqs, srv.results, every=spot_every)
    st["workload"] = ("%d requests, Poisson %.0f/s: 60%% 1D N 1024..16384 x 100..2000 steps, "
                      "30%% 2D 256^2..1024^2 x 50..500 steps, 10%% 2D Born 256^2..1024^2 "
                      "(<= 200 iterations, tol 1e-6); FP32-mixed; HBM budget %d GiB"
                      % (count, rate, budget_gib))
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=50,
                    help="C2 steps per timed CPU run (seq and par; strong runs 5x)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--fp32-err-steps", type=int, default=1000)
    ap.add_argument("--no-born", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=200.0, help="request arrivals per second")
    ap.add_argument("--hbm-budget-gib", type=int, default=64)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run the workload of the rocprofv3 --pmc passes and exit")
    ap.add_argument("--probe-steps", type=int, default=6)
    ap.add_argument("--batch-probe", type=int, metavar="NJOBS",
                    help="run only NJOBS batched FP32-mixed 1024^2 jobs (for rocprofv3) and exit")
    ap.add_argument("--probe-variants", type=int, nargs=2, default=(7, 7))
    ap.add_argument("--probe-n", type=int, default=1024, choices=(1024, 2048))
    ap.add_argument("--fft-probe", action="store_true",
                    help="run only FourierTransform plans of 1024^2, 2048^2, 128^3 (rocprofv3)")
    ap.add_argument("--born-probe", action="store_true",
                    help="run only the C4 Born section (for rocprofv3) and exit")
    ap.add_argument("--serving-probe", type=int, metavar="NREQ",
                    help="run only the request loop over NREQ requests (for rocprofv3) and exit")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc outputs (CPU)")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.traffic_probe:
        traffic_probe(args.probe_steps)
        return
    if args.batch_probe:
        batch_probe(args.probe_steps, args.batch_probe, tuple(args.probe_variants), args.probe_n)
        return
    if args.fft_probe:
        fft_probe(max(args.probe_steps, 1))
        return
    if args.born_probe:
        from qusim_amd import _native as Nn
        Nn.init(0)
        b = born_section()
        print(json.dumps(b), flush=True)
        return
    if args.serving_probe:
        from qusim_amd import _native as Nn
        Nn.init(0)
        st = serving_section(args.serving_probe, args.rate, args.hbm_budget_gib,
                             capacity_count=min(600, args.serving_probe))
        st.pop("spot_check_fp64", None)
        print(json.dumps(st), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU and has no multi-GPU path: requests and energies "
              "would shard across GPUs with no exchange, and one grid beyond a GPU's memory "
              "as y-slabs with two all-to-all transposes per step",
              file=sys.stderr)
        sys.exit(2)

    from qusim_amd import _native as Nn
    from qusim_amd.workloads import c2_problem
    Nn.init(0)
    info = Nn.device_info()
    log("device: %s, %d CUs" % (info["name"], info["cu"]))
    prob = c2_problem()
    n = prob["n"]
    npts = n * n
    job, keep = make_job(Nn, prob, 64)
    tuned = autotune(job)
    log("autotuned variants: %s" % tuned)
    log("timing %d steps of 1024^2 (warmup %d)" % (args.steps, args.warmup))
    wall, gpu = time_steps(job, args.steps, args.warmup)
    value = args.steps * npts / wall
    log("C2 FP64: %.3e grid-pt-steps/s (%.3f ms/step)" % (value, wall / args.steps * 1e3))
    ms, cnt = kernel_times(job, max(4, args.steps // 4))
    dom, per = roofline(ms, cnt, npts, 64)
    achieved = per[dom]["achieved_GBs"]
    step_gbs = 80.0 * npts / (gpu / args.steps) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D split-operator O2 step, 1024x1024 periodic grid on "
                               "[-150,150)^2, dt=0.2, complex128, 16 seeded Gaussian bumps "
                               "+ imaginary absorbing layer",
                   "grid": [n, n], "precision": "complex128", "method": "Strang O2",
                   "device": info["name"], "launch_variants": tuned},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles(
                         dom, [tuned["row_variant"], tuned["col_variant"], tuned["xsplit"]]),
                     "step_algorithmic_GBs": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS,
                     "kernels": per},
    }
    Nn.call("qsim_hip_split_destroy", job)
    def section(key, fn):
        """a secondary section: an exception is recorded in its field and the
        headline line is still printed"""
        try:
            out[key] = fn()
            return out[key]
        except Exception as e:  # noqa: BLE001 -- reported in the JSON line
            import traceback
            traceback.print_exc()
            out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            return None

    log("multi-size group section")
    gs = section("group_2d", group_section)
    if gs:
        log("group of 14 2D jobs (256^2..1024^2): %.3e grid-pt-steps/s with graphs, %.3e direct, "
            "%.3e as per-size batches" % (gs["group_graph"]["value"], gs["group_direct"]["value"],
                                          gs["per_size_batches"]["value"]))
    log("4096^2 complex128 (beyond the Infinity Cache)")
    g = section("grid_4096", large_grid_section)
    if g:
        log("4096^2 FP64: %.3e grid-pt-steps/s (%.3f ms/step), dominant %s %.2f of 8 TB/s"
            % (g["value"], g["ms_per_step"], g["dominant_kernel"],
               g["kernels"][g["dominant_kernel"]]["frac"]))
    log("C1 section (1D 2048, 1000 steps)")
    c = section("c1", lambda: c1_section(host_threads()))
    if c:
        log("C1: %.4f s on the GPU, psi vs oracle %.2e, P(x>0) %.10f (oracle %.10f)"
            % (c["gpu_seconds"], c["psi_rel_l2_vs_oracle"], c["transmitted"],
               c["transmitted_oracle"]))
    if not args.no_fp32:
        log("FP32-mixed section (2048^2 alone and batched, 1024^2 batched)")
        f = section("fp32_mixed", lambda: fp32_section(args.fp32_err_steps, 50, value,
                                                       host_threads()))
        if f:
            eo = f["grid_2048"]["error_vs_oracle"]
            log("FP32-mixed: 1024^2 x16 batched %.3e (%.2fx FP64 1024^2), 2048^2 %.3e alone, "
                "%.3e x4 batched; psi error after %d steps vs the oracle %.2e (FP32, 2048^2), "
                "%.2e (complex128), transmitted %.6f vs %.6f"
                % (f["value"], f["speedup_vs_fp64"], f["grid_2048"]["fp32_single"],
                   f["grid_2048"]["fp32_batched"]["value"], args.fp32_err_steps,
                   eo["fp32_mixed"]["psi_rel_l2"], eo["complex128"]["psi_rel_l2"],
                   eo["fp32_mixed"]["transmitted"], eo["oracle_transmitted"]))
    if not args.no_born:
        log("Born 128^3 x 32 energies section")
        b = section("born_3d", born_section)
        if b:
            log("Born: %.0f energy-iterations/s" % b["value"])
    if not args.no_serving:
        log("serving section: %d requests at %.0f/s" % (args.requests, args.rate))
        v = section("serving",
                    lambda: serving_section(args.requests, args.rate, args.hbm_budget_gib))
        if v:
            log("serving: %.1f solves/s, p50 %.3f s, p99 %.3f s"
                % (v["solves_per_s"], v["latency_p50_s"], v["latency_p99_s"]))
    if not args.no_cpu:
        nt = host_threads()
        log("CPU baseline: the oracle port, seq / par / strong on up to %d threads" % nt)
        section("cpu_baseline", lambda: cpu_section(prob, args.cpu_steps, nt))
    print(json.dumps(out), flush=True)


def cpu_section(prob, steps, nt, runs=5):
    """The FP64 oracle port of the reference's CPU path, timed on this host
    (set-up excluded, one warm-up run, then the median of `runs`):
    * C2 (the headline workload) in three modes: seq (1 thread, the
      reference's CpuSeq), par (the reference's CpuPar shape: the pointwise
      Device ops threaded over nt cores, the KISS FFT serial) and strong (the
      FFT lines threaded too; not the reference);
    * C4: energy-iterations/s of one 128^3 energy (Vellekoop, slow 0.5), seq
      and par;
    * C1: the full 2048-point, 1000-step run on one thread."""
    from oracle import pyoracle as O
    from qusim_amd.workloads import c1_problem, c4_energies, c4_potential
    p = prob
    n = p["n"]
    a = (p["psi"], p["V"], p["x0"], p["x1"], p["x0"], p["x1"], p["dt"])
    modes = {}
    for name, tp, tf, st in (("seq", 1, 1, steps), ("par", nt, 1, steps),
                             ("strong", nt, nt, 5 * steps)):
        O.split2d_time(*a, max(1, st // 5), complex_v=True, threads=tp, fft_threads=tf)
        secs = [O.split2d_time(*a, st, complex_v=True, threads=tp, fft_threads=tf)
                for _ in range(runs)]
        med = float(np.median(secs))
        modes[name] = {"value": st * n * n / med, "unit": "grid-pt-steps/s",
                       "threads_pointwise": tp, "threads_fft": tf, "steps_per_run": st,
                       "runs_s": secs}
    V4, E4 = c4_potential(n=128, L=32.0), c4_energies()[0]
    c4 = {}
    for name, tp, tf in (("seq", 1, 1), ("par", nt, 1)):
        its = 2
        secs = [O.born_time(V4, (-32.0,) * 3, (32.0,) * 3, E4, (0, 0, 1), its, True, 1, 0.5,
                            threads=tp, fft_threads=tf) for _ in range(3)]
        c4[name] = {"value": its / float(np.median(secs)), "unit": "energy-iterations/s",
                    "threads_pointwise": tp, "threads_fft": tf, "iterations_per_run": its,
                    "runs_s": secs}
    p1 = c1_problem()
    O.split1d(p1["psi"], p1["V"], p1["x0"], p1["x1"], p1["dt"], 10)
    c1s = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.split1d(p1["psi"], p1["V"], p1["x0"], p1["x1"], p1["dt"], p1["steps"])
        c1s.append(time.perf_counter() - t0)
    c1m = float(np.median(c1s))
    par = modes["par"]
    return {
        "value": par["value"], "unit": "grid-pt-steps/s", "cores": nt, "kind": "port",
        "sample": "C2 1024^2 complex-V O2 steps, %d per run, median of %d after a warm-up; "
                  "FP64 oracle port of the reference's unfused op order (Mul, FFT2, Mul, IFFT2, "
                  "Scale) with KISS-style mixed-radix FFTs, in the reference's CpuPar shape "
                  "(pointwise ops on %d threads, FFT serial)" % (steps, runs, nt),
        "cpu_model": cpu_model(), "modes": modes,
        "c4_born_128": c4,
        "c1_full_run": {"seconds": c1m, "value": p1["steps"] * p1["psi"].size / c1m,
                        "unit": "grid-pt-steps/s", "threads": 1, "runs_s": c1s,
                        "workload": "2048 points, O2, dt 0.5, 1000 steps (CpuSeq)"}}


if __name__ == "__main__":
    main()
