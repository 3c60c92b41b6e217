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


def test_group_lane_result_is_one_call_of_the_whole_run():
    """A 2D evolution in the request loop runs its whole step count as one
    sweep program of the multi-size group: its psi equals one
    qsim_hip_split_steps call of all its steps on the job alone."""
    from qusim_amd import _native as N
    from qusim_amd.serving import EvoJob, Server, _Staging, make_requests
    reqs = [r for r in make_requests(120, rate=1e6, seed=21) if r.kind == "evo2d"][:4]
    for r in reqs:
        r.work = max(10, r.work // 5)
    s = Server(precision=32, realtime=False, keep=[r.rid for r in reqs])
    st = s.run(reqs)
    assert st["completed"] == len(reqs)
    staging = _Staging()
    for r in reqs:
        j = EvoJob(r, 32, staging)
        N.call("qsim_hip_split_steps", j.job, r.work, None)
        N.synchronize()
        assert np.array_equal(j.psi.download(), s.results[r.rid]["psi"]), (r.n, r.work)
        j.close()
    staging.close()


def test_full_size_mix_spot_checks():
    """300 requests of the benchmark mix at full size (steps and iterations
    as drawn), all at t = 0: every request completes and every 50th,
    re-solved alone in FP64, agrees within the stated bounds (evolutions
    1e-5 psi relative L2, Born 1e-4 of the forward cross section)."""
    from qusim_amd.serving import Server, make_requests, spot_check
    reqs = make_requests(300, rate=1e6, seed=22)
    keep = [r.rid for r in reqs[::50]]
    s = Server(precision=32, realtime=False, keep=keep)
    st = s.run(reqs)
    print({k: st[k] for k in ("completed", "solves_per_s", "latency_p50_s", "latency_p99_s")})
    assert st["completed"] == len(reqs) and st["errors"] == 0
    # the Born workers create and free jobs while the main thread captures
    # group advances: no capture may be invalidated by them
    assert st["graph_capture_failures"] == 0
    sc = spot_check(reqs, s.results, every=50)
    print(sc)
    assert sum(v["checked"] for v in sc.values()) == len(keep)
    for kind, v in sc.items():
        if v["checked"]:
            assert v["max"] < (1e-4 if kind == "born2d" else 1e-5), (kind, v)
