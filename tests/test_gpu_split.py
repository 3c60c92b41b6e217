"""Split-operator evolution on the GPU (fused sweeps and the unfused
reference-order path) against the FP64 oracle (CpuSeq + KISS restatement).

Bar: complex128 psi within 1e-12 relative L2 (the path's stated tolerance);
FP32-mixed within the tolerances written beside each case."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def gauss2d(n0, n1, L0=(-30.0, 30.0), L1=(-20.0, 25.0)):
    x = L0[0] + (L0[1] - L0[0]) / n0 * np.arange(n0)
    y = L1[0] + (L1[1] - L1[0]) / n1 * np.arange(n1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    psi = np.exp(0.8j * X - ((X + 5) ** 2 + Y ** 2) / (4 * 9.0))
    V = 0.5 * np.exp(-(X ** 2 + Y ** 2) / 8) - 0.05j * np.exp(-((X - 3) ** 2 + Y ** 2) / 30)
    return psi, V, L0, L1


def run_1d(psi, V, x0, x1, dt, steps, method, boundary=0, precision="fp64", complex_v=False,
           batch=1):
    from qusim_amd.api import BoundaryCondition, Evolver1D, Options, SolverMethod
    o = Options().Hip().Batch(batch)
    if precision != "fp64":
        o.Precision(precision)
    if complex_v:
        o.ComplexPotential(True)
    e = Evolver1D().init(psi, False, dt, False, V, x0, x1, psi.size,
                         BoundaryCondition.Period if boundary == 0 else BoundaryCondition.InfiniteWall,
                         SolverMethod.SplittingMethodO2 if method == 2 else SolverMethod.SplittingMethodO4,
                         1.0, 1.0, o)
    e.steps(steps // batch)
    return e


def run_2d(psi, V, L0, L1, dt, steps, method, precision="fp64", complex_v=False):
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    o = Options().Hip().Batch(steps)
    if precision != "fp64":
        o.Precision(precision)
    if complex_v:
        o.ComplexPotential(True)
    e = Evolver2D().init(psi, False, dt, False, V, L0[0], L0[1], psi.shape[0], L1[0], L1[1],
                         psi.shape[1], BoundaryCondition.Period,
                         SolverMethod.SplittingMethodO2 if method == 2 else SolverMethod.SplittingMethodO4,
                         1.0, 1.0, o)
    e.step()
    return e


def test_c1_tunnelling_matches_oracle(oracle):
    """C1: N=2048, dt=0.5, 1000 steps, O2 -- psi, norm and P(x>0)."""
    from qusim_amd.workloads import c1_problem
    p = c1_problem()
    e = run_1d(p["psi"], p["V"], p["x0"], p["x1"], p["dt"], p["steps"], 2, batch=50)
    got = e.GetPsi()
    ref = oracle.split1d(p["psi"], p["V"], p["x0"], p["x1"], p["dt"], p["steps"], method=2)
    assert rel_l2(got, ref) < 1e-12
    dx = 200.0 / 2048
    assert abs(e.Norm2() - oracle.norm2(ref) * dx) < 1e-12
    assert abs(e.Norm2(0.0, 100.0) - oracle.norm2_range_1d(ref, -100, 100, 0, 100)) < 1e-12


@pytest.mark.parametrize("n,method", [(256, 4), (4096, 2), (64, 2), (8192, 2), (16384, 2),
                                      (16384, 4)])
def test_1d_fused_matches_oracle(n, method, oracle):
    """fused persistent 1D kernel up to 8192 (FP64); 16384 runs the unfused
    order on the four-step FFT"""
    x = -50 + 100.0 / n * np.arange(n)
    psi = np.exp(1.3j * x - (x + 10) ** 2 / 20)
    V = 0.4 * np.exp(-x ** 2 / 3)
    got = run_1d(psi, V, -50.0, 50.0, 0.05, 120, method).GetPsi()
    ref = oracle.split1d(psi, V, -50.0, 50.0, 0.05, 120, method=method)
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("n,boundary,method", [(1000, 0, 2), (1000, 1, 4), (512, 1, 2)])
def test_1d_unfused_paths_match_oracle(n, boundary, method, oracle):
    """non-power-of-two length (generic FFT) and the InfiniteWall sine path"""
    x = -30 + 60.0 / n * np.arange(n)
    psi = np.exp(1.0j * x - (x + 5) ** 2 / 10)
    V = 0.3 * np.exp(-x ** 2 / 2)
    got = run_1d(psi, V, -30.0, 30.0, 0.02, 60, method, boundary=boundary).GetPsi()
    ref = oracle.split1d(psi, V, -30.0, 30.0, 0.02, 60, method=method, boundary=boundary)
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("shape", [(128, 64), (64, 256), (512, 512)])
@pytest.mark.parametrize("method", [2, 4])
def test_2d_fused_matches_oracle(shape, method, oracle):
    psi, V, L0, L1 = gauss2d(*shape)
    got = run_2d(psi, V, L0, L1, 0.05, 40, method).GetPsi()
    ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], 0.05, 40, method=method)
    assert rel_l2(got, ref) < 1e-12


def test_2d_complex_potential_and_complex_dt(oracle):
    psi, V, L0, L1 = gauss2d(128, 128)
    dt = 0.05 - 0.01j
    got = run_2d(psi, V, L0, L1, dt, 30, 2, complex_v=True).GetPsi()
    ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], dt, 30, method=2, complex_v=True)
    assert rel_l2(got, ref) < 1e-12


def test_2d_unfused_matches_oracle(oracle):
    psi, V, L0, L1 = gauss2d(96, 80)
    for m in (2, 4):
        got = run_2d(psi, V, L0, L1, 0.05, 25, m).GetPsi()
        ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], 0.05, 25, method=m)
        assert rel_l2(got, ref) < 1e-12


def test_c2_shape_prefix_matches_oracle(oracle):
    """C2 at full size (1024^2, dt=0.2, complex V) for 20 steps."""
    from qusim_amd.workloads import c2_problem
    p = c2_problem()
    L = (p["x0"], p["x1"])
    got = run_2d(p["psi"], p["V"], L, L, p["dt"], 20, 2, complex_v=True).GetPsi()
    ref = oracle.split2d(p["psi"], p["V"], L[0], L[1], L[0], L[1], p["dt"], 20, method=2,
                         complex_v=True, threads=16)
    assert rel_l2(got, ref) < 1e-12


def test_fp32_mixed_error_bounds(oracle):
    psi, V, L0, L1 = gauss2d(256, 256)
    got = run_2d(psi, V, L0, L1, 0.05, 200, 2, precision="mixed32", complex_v=True).GetPsi()
    ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], 0.05, 200, method=2, complex_v=True)
    err = rel_l2(got, ref)
    print("fp32-mixed 256^2 x200 rel L2", err)
    assert err < 1e-4
    x = -50 + 100.0 / 1024 * np.arange(1024)
    p1 = np.exp(1.3j * x - (x + 10) ** 2 / 20)
    V1 = 0.4 * np.exp(-x ** 2 / 3)
    g1 = run_1d(p1, V1, -50.0, 50.0, 0.05, 400, 2, precision="mixed32").GetPsi()
    r1 = oracle.split1d(p1, V1, -50.0, 50.0, 0.05, 400, method=2)
    err1 = rel_l2(g1, r1)
    print("fp32-mixed 1D 1024 x400 rel L2", err1)
    assert err1 < 1e-4


def test_batched_equals_alone():
    """Each job's result is bitwise the same alone and inside a batched launch."""
    from qusim_amd.api import advance_batched
    outs_alone, evs = [], []
    for k in range(3):
        psi, V, L0, L1 = gauss2d(128, 128)
        V = V * (1 + 0.3 * k)
        a = run_2d(psi, V, L0, L1, 0.05 * (k + 1), 10, 2, complex_v=True)
        outs_alone.append(a.GetPsi())
        b = run_2d(psi, V, L0, L1, 0.05 * (k + 1), 0, 2, complex_v=True)
        evs.append(b)
    advance_batched(evs, 10)
    for a, b in zip(outs_alone, evs):
        assert np.array_equal(a, b.GetPsi())


def test_batched_autotune_sets_one_pair_on_every_job(oracle):
    """the batched autotuner leaves the jobs' states untouched, gives them all
    the same variants (qsim_hip_split_batch_steps refuses mixed ones) and the
    batch then matches the oracle"""
    import ctypes
    from qusim_amd._native import call
    psi, V, L0, L1 = gauss2d(64, 1024)
    evs = [run_2d(psi, V, L0, L1, 0.05, 0, 2, complex_v=True) for _ in range(3)]
    arr = (ctypes.c_void_p * 3)(*[e._job for e in evs])
    vr, vc = ctypes.c_int(-1), ctypes.c_int(-1)
    ru, cu = ctypes.c_double(), ctypes.c_double()
    call("qsim_hip_split_autotune_batched", arr, 3, 2, ctypes.byref(vr), ctypes.byref(vc),
         ctypes.byref(ru), ctypes.byref(cu), None)
    assert 0 <= vr.value and vc.value == 0 and ru.value > 0
    b = ctypes.c_void_p()
    call("qsim_hip_split_batch_create", ctypes.byref(b), 3)
    call("qsim_hip_split_batch_steps", b, arr, 3, 6, None)
    call("qsim_hip_split_batch_destroy", b)
    ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], 0.05, 6, method=2, complex_v=True)
    for e in evs:
        assert rel_l2(e.GetPsi(), ref) < 1e-12


def test_fp32_mixed_1000_steps_within_1e5_of_fp64():
    """FP32-mixed against the complex128 device path (itself pinned to the
    oracle) after 1000 steps of the 1024^2-construction problem at 512^2: the
    phase factors' float-float parts enter before the last rounding, so
    near-identity factors lose nothing the same way every step"""
    from qusim_amd.workloads import c2_problem
    p = c2_problem(n=512)
    L = (p["x0"], p["x1"])
    a = run_2d(p["psi"], p["V"], L, L, p["dt"], 1000, 2, complex_v=True).GetPsi()
    b = run_2d(p["psi"], p["V"], L, L, p["dt"], 1000, 2, precision="mixed32",
               complex_v=True).GetPsi()
    err = rel_l2(b, a)
    print("fp32-mixed 512^2 x1000 rel L2 vs complex128", err)
    assert err < 1e-5


def test_norm_conserved_for_real_potential():
    psi, V, L0, L1 = gauss2d(256, 256)
    e = run_2d(psi, V.real, L0, L1, 0.1, 0, 2)
    n0 = e.Norm2()
    e.steps(1)
    for _ in range(5):
        from qusim_amd._native import call
        call("qsim_hip_split_steps", e._job, 100, None)
    assert abs(e.Norm2() / n0 - 1) < 1e-12


@pytest.mark.parametrize("n", [8192, 16384])
def test_1d_fp32_long_lines(n, oracle):
    """FP32-mixed persistent 1D kernel at 8192 and 16384 points (one
    workgroup holds the line): error vs the FP64 oracle after 500 steps."""
    x = -100 + 200.0 / n * np.arange(n)
    psi = np.exp(1.0j * x - (x + 30) ** 2 / 100)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * (200.0 / n))
    V = 0.6 * np.exp(-x ** 2 / 2)
    e = run_1d(psi, V, -100.0, 100.0, 0.1, 500, 2, precision="mixed32", batch=100)
    ref = oracle.split1d(psi, V, -100.0, 100.0, 0.1, 500, method=2)
    err = rel_l2(e.GetPsi(), ref)
    print("1D FP32 n=%d 500 steps rel L2 %.2e" % (n, err))
    assert err < 1e-5


@pytest.mark.parametrize("variant", range(8))
@pytest.mark.parametrize("shape", [(1024, 64), (64, 1024), (2048, 64), (64, 2048)],
                         ids=lambda s: "x".join(map(str, s)))
def test_launch_variants_match_oracle(shape, variant, oracle):
    """every row / column launch variant of the 1024 and 2048 sweeps (radix
    plans of three and four passes) against the oracle; FP32-mixed against
    the FP64 device result"""
    from qusim_amd._native import call
    psi, V, L0, L1 = gauss2d(*shape)
    outs = {}
    for prec in ("fp64", "mixed32"):
        e = run_2d(psi, V, L0, L1, 0.05, 0, 2, precision=prec, complex_v=True)
        call("qsim_hip_split_set_variants", e._job, variant, variant)
        call("qsim_hip_split_steps", e._job, 6, None)
        outs[prec] = e.GetPsi()
    ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], 0.05, 6, method=2, complex_v=True)
    assert rel_l2(outs["fp64"], ref) < 1e-12
    assert rel_l2(outs["mixed32"], outs["fp64"]) < 1e-5


def _evolver(dim, precision, method, dt=0.05):
    from qusim_amd.api import BoundaryCondition, Evolver1D, Evolver2D, Options, SolverMethod
    o = Options().Hip().ComplexPotential(True)
    if precision != "fp64":
        o.Precision(precision)
    sm = SolverMethod.SplittingMethodO2 if method == 2 else SolverMethod.SplittingMethodO4
    if dim == 2:
        psi, V, L0, L1 = gauss2d(128, 64)
        return Evolver2D().init(psi, True, dt, False, V, L0[0], L0[1], 128, L1[0], L1[1], 64,
                                BoundaryCondition.Period, sm, 1.0, 1.0, o)
    x = -40 + 80.0 / 1024 * np.arange(1024)
    psi = np.exp(1j * x - (x + 10) ** 2 / 16)
    V = 0.6 * np.exp(-x ** 2 / 2) - 0.02j * np.exp(-(x - 5) ** 2 / 50)
    return Evolver1D().init(psi, True, dt, False, V, -40.0, 40.0, 1024,
                            BoundaryCondition.Period, sm, 1.0, 1.0, o)


@pytest.mark.parametrize("dim,precision,method",
                         [(2, "fp64", 2), (2, "mixed32", 2), (2, "fp64", 4), (1, "fp64", 2),
                          (1, "mixed32", 4)])
def test_snapshot_resume_is_bitwise(tmp_path, dim, precision, method):
    """10 steps, SaveSnapshot, LoadSnapshot into a fresh evolver, 15 steps:
    psi and Time() equal one evolver's 10 + 15 steps bit for bit."""
    path = str(tmp_path / "psi.qpsi")
    a = _evolver(dim, precision, method)
    a.steps(10)
    a.SaveSnapshot(path)
    a.steps(15)
    b = _evolver(dim, precision, method)
    b.LoadSnapshot(path)
    assert b.fStep == 10
    b.steps(15)
    assert b.Time() == a.Time()
    assert np.array_equal(b.GetPsi().view(np.uint64), a.GetPsi().view(np.uint64))
    a.close()
    b.close()


def test_snapshot_of_another_setup_is_refused(tmp_path):
    path = str(tmp_path / "psi.qpsi")
    a = _evolver(2, "fp64", 2)
    a.steps(2)
    a.SaveSnapshot(path)
    with pytest.raises(ValueError, match="dt"):
        _evolver(2, "fp64", 2, dt=0.1).LoadSnapshot(path)
    with pytest.raises(ValueError, match="method"):
        _evolver(2, "fp64", 4).LoadSnapshot(path)
    with pytest.raises(ValueError, match="grid"):
        _evolver(1, "fp64", 2).LoadSnapshot(path)
    # the precision may differ: psi is complex128 in both forms
    c = _evolver(2, "mixed32", 2)
    c.LoadSnapshot(path)
    assert np.array_equal(c.GetPsi(), a.GetPsi())


def test_roctx_ranges_on_leave_results_unchanged():
    """With QSIM_ROCTX=1 the entry points push and pop roctx ranges (no
    profiler attached here): the same steps give the same bits."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prog = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from test_gpu_split import _evolver\n"
        "e = _evolver(2, 'fp64', 2); e.steps(7)\n"
        "sys.stdout.buffer.write(e.GetPsi().tobytes())\n" % (root, os.path.join(root, "tests")))
    outs = []
    for flag in ("0", "1"):
        r = subprocess.run([sys.executable, "-c", prog], capture_output=True, timeout=120,
                           env=dict(os.environ, QSIM_ROCTX=flag))
        assert r.returncode == 0, r.stderr.decode()[-2000:]
        outs.append(r.stdout)
    assert len(outs[0]) == 128 * 64 * 16 and outs[0] == outs[1]


def test_c2_full_run_matches_oracle(oracle):
    """C2 as configured: 1024^2 on [-150,150)^2, complex V with the absorbing
    layer, dt = 0.2, all 500 steps, against the FP64 oracle (16 threads)."""
    from qusim_amd.workloads import c2_problem
    p = c2_problem()
    assert p["steps"] == 500 and p["psi"].shape == (1024, 1024)
    L = (p["x0"], p["x1"])
    got = run_2d(p["psi"], p["V"], L, L, p["dt"], p["steps"], 2, complex_v=True).GetPsi()
    ref = oracle.split2d(p["psi"], p["V"], L[0], L[1], L[0], L[1], p["dt"], p["steps"],
                         method=2, complex_v=True, threads=16)
    err = rel_l2(got, ref)
    print("C2 1024^2 x500 rel L2 vs oracle", err)
    assert err < 1e-12


def test_c3_full_run_against_oracle(oracle):
    """C3 as configured: 2048^2, dt = 0.1, 1000 steps, against the FP64 oracle
    run on the host's threads (complex potential, the absorbing layer): the
    complex128 path within 1e-12 relative L2; the FP32-mixed path within 1e-5
    in psi, in the norm and in the transmitted probability P(x > 0)."""
    import os
    from qusim_amd.workloads import c3_problem
    p = c3_problem()
    assert p["steps"] == 1000 and p["psi"].shape == (2048, 2048)
    L = (p["x0"], p["x1"])
    a = run_2d(p["psi"], p["V"], L, L, p["dt"], p["steps"], 2, complex_v=True).GetPsi()
    b = run_2d(p["psi"], p["V"], L, L, p["dt"], p["steps"], 2, precision="mixed32",
               complex_v=True).GetPsi()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ref = oracle.split2d(p["psi"], p["V"], L[0], L[1], L[0], L[1], p["dt"], p["steps"],
                         method=2, complex_v=True, threads=threads)
    n = p["n"]
    dx = (L[1] - L[0]) / n
    right = L[0] + dx * np.arange(n) >= 0.0

    def norm_and_t(q):
        w = np.abs(q) ** 2 * dx * dx
        return float(np.sum(w)), float(np.sum(w[right]))

    nr, tr = norm_and_t(ref)
    na, ta = norm_and_t(a)
    nb, tb = norm_and_t(b)
    e64, e32 = rel_l2(a, ref), rel_l2(b, ref)
    print("C3 2048^2 x1000 vs oracle: complex128 %.2e, FP32-mixed %.2e, norm drift %.2e, "
          "P(x>0) %.8f vs %.8f" % (e64, e32, nb / nr - 1, tb, tr))
    assert e64 < 1e-12 and abs(na / nr - 1) < 1e-12 and abs(ta / tr - 1) < 1e-12
    assert e32 < 1e-5 and abs(nb / nr - 1) < 1e-5 and abs(tb / tr - 1) < 1e-5


def test_graph_replay_is_bitwise_and_its_cache_bounded():
    """qsim_hip_split_use_graph: replayed captures give the bits of direct
    launches, for more step counts than the per-job cache keeps (8), again
    after set_variants drops the captures, and after turning graphs off."""
    from qusim_amd import _native as N
    a, b = _evolver(2, "fp64", 2), _evolver(2, "fp64", 2)
    N.call("qsim_hip_split_use_graph", b._job, 1)
    for k in (3, 1, 4, 2, 5, 6, 7, 8, 9, 10, 3, 1):
        N.call("qsim_hip_split_steps", a._job, k, None)
        N.call("qsim_hip_split_steps", b._job, k, None)
    assert np.array_equal(a.GetPsi(), b.GetPsi())
    for v in ((6, 6), (4, 4)):
        N.call("qsim_hip_split_set_variants", a._job, *v)
        N.call("qsim_hip_split_set_variants", b._job, *v)
        N.call("qsim_hip_split_steps", a._job, 5, None)
        N.call("qsim_hip_split_steps", b._job, 5, None)
        assert np.array_equal(a.GetPsi(), b.GetPsi())
    N.call("qsim_hip_split_use_graph", b._job, 0)
    N.call("qsim_hip_split_steps", a._job, 2, None)
    N.call("qsim_hip_split_steps", b._job, 2, None)
    assert np.array_equal(a.GetPsi(), b.GetPsi())


@pytest.mark.parametrize("xsplit", [2, 4])
@pytest.mark.parametrize("shape,method,precision",
                         [((1024, 1024), 2, "fp64"), ((256, 512), 4, "fp64"),
                          ((512, 128), 2, "fp64"), ((2048, 64), 2, "fp64"),
                          ((1024, 256), 2, "mixed32"), ((256, 2048), 4, "mixed32"),
                          ((256, 4096), 2, "fp64"), ((512, 4096), 4, "mixed32")])
def test_xsplit_schedule_matches_oracle(shape, method, precision, xsplit, oracle):
    """x axis split into R-point transforms across rows (lane exchanges in the
    row sweep) and nx/R-point column transforms: complex128 within 1e-12 of
    the oracle, FP32-mixed within 1e-5, for every row variant the group
    fits (rows of 4096 points hold the R rows of a group in each thread)."""
    from qusim_amd import _native as N
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    psi, V, L0, L1 = gauss2d(*shape)
    steps = 12
    ref = oracle.split2d(psi, V, L0[0], L0[1], L1[0], L1[1], 0.05, steps, method=method,
                         complex_v=True, threads=8)
    norm = np.sqrt(np.sum(np.abs(psi) ** 2) * (L0[1] - L0[0]) / shape[0]
                   * (L1[1] - L1[0]) / shape[1])
    variants = (range(8) if shape[1] in (1024, 2048)
                else [0, 1] if shape[1] in (128, 256, 512, 4096) else [0])
    for rv in variants:
        o = Options().Hip().Batch(steps).ComplexPotential(True)
        if precision != "fp64":
            o.Precision(precision)
        e = Evolver2D().init(psi, False, 0.05, False, V, L0[0], L0[1], shape[0], L1[0], L1[1],
                             shape[1], BoundaryCondition.Period,
                             SolverMethod.SplittingMethodO2 if method == 2
                             else SolverMethod.SplittingMethodO4, 1.0, 1.0, o)
        try:
            N.call("qsim_hip_split_set_xsplit", e._job, xsplit)
            N.call("qsim_hip_split_set_variants", e._job, rv, 0)
        except N.QsimError:
            continue  # this row variant's group does not fit one workgroup
        e.step()
        err = rel_l2(e.GetPsi(), ref)
        assert err < (1e-12 if precision == "fp64" else 1e-5), (rv, err)


def test_xsplit_limits_and_autotune():
    from qusim_amd import _native as N
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    psi, V, L0, L1 = gauss2d(128, 64)
    e = Evolver2D().init(psi, True, 0.05, False, V, L0[0], L0[1], 128, L1[0], L1[1], 64,
                         BoundaryCondition.Period, SolverMethod.SplittingMethodO2, 1.0, 1.0,
                         Options().Hip())
    with pytest.raises(N.QsimError, match="1, 2 or 4"):
        N.call("qsim_hip_split_set_xsplit", e._job, 3)
    with pytest.raises(N.QsimError, match="nx / R"):
        N.call("qsim_hip_split_set_xsplit", e._job, 4)   # 128 / 4 = 32 < 64
    N.call("qsim_hip_split_set_xsplit", e._job, 2)
    a = e.GetPsi()
    vr, vc, xs = ctypes_ints(3)
    N.call("qsim_hip_split_autotune", e._job, 2, ctypes.byref(vr), ctypes.byref(vc), None, None,
           None)
    N.call("qsim_hip_split_get_schedule", e._job, ctypes.byref(vr), ctypes.byref(vc),
           ctypes.byref(xs))
    assert xs.value in (1, 2)
    assert np.array_equal(e.GetPsi(), a)  # autotuning leaves the state alone


def ctypes_ints(n):
    return [ctypes.c_int() for _ in range(n)]
