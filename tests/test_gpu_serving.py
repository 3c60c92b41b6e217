"""The request loop on the GPU: every request completes, results are bitwise
the same alone and in the mix, admission never exceeds the HBM budget, and
the FP32-mixed results agree with FP64 re-solves within the stated bounds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_stream(count, rate, seed):
    """a stream of the benchmark mix with shorter runs (the sizes are kept)"""
    from qusim_amd.serving import make_requests
    reqs = make_requests(count, rate=rate, seed=seed)
    for r in reqs:
        r.work = max(10, r.work // 10) if r.kind != "born2d" else 24
    return reqs


def test_mix_equals_alone_bitwise():
    from qusim_amd.serving import Server, solve_alone
    reqs = _small_stream(30, 3000.0, 11)
    s = Server(precision=32, keep=[r.rid for r in reqs])
    st = s.run(reqs)
    print(st)
    assert st["completed"] == len(reqs) and st["errors"] == 0
    picked = {}
    for r in reqs:  # two of each kind
        picked.setdefault(r.kind, [])
        if len(picked[r.kind]) < 2:
            picked[r.kind].append(r)
    for kind, rs in picked.items():
        for r in rs:
            alone = solve_alone(r, precision=32)
            mixed = s.results[r.rid]
            if kind == "born2d":
                assert alone["iterations"] == mixed["iterations"]
                assert alone["dsigma"] == mixed["dsigma"]
                assert np.array_equal(alone["dpsi"], mixed["dpsi"])
            else:
                assert np.array_equal(alone["psi"], mixed["psi"]), (kind, r)
                assert alone["norm"] == mixed["norm"]


def test_admission_respects_budget():
    from qusim_amd.serving import Server, footprint
    reqs = _small_stream(24, 1e6, 12)
    budget = 2 * max(footprint(r, 32) for r in reqs)
    s = Server(precision=32, hbm_budget=budget, realtime=False)
    st = s.run(reqs)
    assert st["completed"] == len(reqs)
    assert s.peak_used <= budget


def test_footprint_bounds_allocation():
    """what a job allocates stays within its footprint (device free memory
    before and after creating it)"""
    from qusim_amd import _native as N
    from qusim_amd.serving import BornJob, EvoJob, Request, _Staging, footprint, make_requests
    N.init(0)
    proto = {r.kind: r for r in make_requests(200, seed=13)}
    for kind, n in (("evo1d", 16384), ("evo2d", 1024), ("born2d", 1024)):
        p = proto[kind]
        r = Request(0, kind, n, 10, 0.0, p.params)
        for prec in (32, 64):
            staging = _Staging()
            if kind != "born2d":
                staging.get(r.points())
            # a first job builds the per-size twiddle tables, shared by all
            warm = BornJob(r, prec) if kind == "born2d" else EvoJob(r, prec, staging)
            warm.close()
            N.synchronize()
            free0 = N.device_info()["mem_free"]
            job = BornJob(r, prec) if kind == "born2d" else EvoJob(r, prec, staging)
            N.synchronize()
            used = free0 - N.device_info()["mem_free"]
            # staging (16 B/pt) was allocated before: it is inside the footprint
            assert used <= footprint(r, prec), (kind, prec, used, footprint(r, prec))
            job.close()
            staging.close()


def test_fp32_results_against_fp64():
    from qusim_amd.serving import Server, compare, solve_alone
    reqs = _small_stream(12, 1e6, 14)
    s = Server(precision=32, realtime=False, keep=[r.rid for r in reqs])
    s.run(reqs)
    for r in reqs:
        ref = solve_alone(r, precision=64)
        err = compare(r.kind, s.results[r.rid], ref)
        print(r, "FP32 vs FP64 %.2e" % err)
        assert err < (1e-4 if r.kind == "born2d" else 1e-5)
