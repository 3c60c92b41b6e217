"""bench.py's host-side bookkeeping (CPU): the roofline of the per-kernel
timings, the lookup of measured HBM traffic by the run's launch schedule, and
the refusal of a multi-GPU run."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_roofline_picks_the_largest_time_per_step():
    npts = 1 << 20
    # kinds: 0 MID row, 1 column, 2 1D, 3 FIRST row, 4 LAST row (ms, launches)
    ms = [13.0e-3 * 10, 11.0e-3 * 10, 0.0, 50.0e-3, 9.0e-3]
    cnt = [10, 10, 0, 1, 1]
    dom, per = bench.roofline(ms, cnt, npts, 64)
    # the FIRST sweep is the slowest launch but runs once per call
    assert dom == "row_sweep"
    assert per["row_sweep"]["avg_us"] == pytest.approx(13.0)
    assert per["row_sweep"]["achieved_GBs"] == pytest.approx(48.0 * npts / 13.0e-6 / 1e9)
    assert per["col_sweep"]["frac"] == pytest.approx(32.0 * npts / 11.0e-6 / 1e9 / 8000.0)
    assert per["row_sweep_first"]["us_per_step"] == 0.0
    ms[1] = 14.0e-3 * 10
    assert bench.roofline(ms, cnt, npts, 64)[0] == "col_sweep"


def test_probe_kernel_names_follow_the_schedule():
    import re
    rows = dict(bench.probe_kernels(6, 1, 2))
    assert re.search(rows["row_sweep"], "qsim::k_row_sweep_x<double, 1024, 1, 6, true, 2>")
    assert not re.search(rows["row_sweep"], "qsim::k_row_sweep_x<double, 1024, 1, 6, true, 4>")
    assert re.search(rows["col_sweep"], "qsim::k_col_sweep<double, 512, 1, true, 2>")
    # x-split row variants 1, 3, 5, 7 run the kernels of plans 0, 2, 0, 6
    assert "1024, 1, 6," in dict(bench.probe_kernels(7, 0, 2))["row_sweep"]
    plain = dict(bench.probe_kernels(4, 5, 1))
    assert re.search(plain["row_sweep"], "qsim::k_row_sweep<double, 1024, 1, 4, true>")
    assert re.search(plain["col_sweep"], "qsim::k_col_sweep<double, 1024, 5, true, 1>")


def test_committed_traffic_covers_the_probed_schedules():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    t = json.load(open(path))
    assert 1.9 < t["calibration"]["fetch_factor"] < 2.1
    for rv, cv, xs in t["schedules"]:
        for name in ("row_sweep", "col_sweep"):
            b = bench.traffic_from_profiles(name, (rv, cv, xs))
            alg = (48.0 if name == "row_sweep" else 32.0) * (1 << 20)
            assert b is not None and 0.95 * alg < b < 1.25 * alg, (rv, cv, xs, name, b)
    # a schedule no probe ran has no traffic figure
    assert bench.traffic_from_profiles("col_sweep", (6, 3, 1)) is None


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
