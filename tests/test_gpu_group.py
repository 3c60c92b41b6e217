"""Multi-size job groups (qsim_hip_split_group_*): fused 2D jobs of
different grid sizes advanced together, one launch per sweep for all, with
jobs joining and leaving between advances and graph-replayed advances.  Each
job's result must equal the same job run alone (qsim_hip_split_steps with
its step count) bit for bit, and the oracle to the path's tolerance."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _job(shape, seed, precision, method):
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    rng = np.random.default_rng(seed)
    nx, ny = shape
    x = -20 + 40.0 / nx * np.arange(nx)
    y = -15 + 30.0 / ny * np.arange(ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    c = rng.uniform(-4, 4, 2)
    psi = np.exp(1j * rng.uniform(0.3, 1.0) * X - ((X - c[0]) ** 2 + (Y - c[1]) ** 2) / 12.0)
    V = rng.uniform(0.1, 0.6) * np.exp(-(X ** 2 + Y ** 2) / 8) - 0.03j * (X / 20) ** 2
    o = Options().Hip().ComplexPotential(True)
    if precision != "fp64":
        o.Precision(precision)
    e = Evolver2D().init(psi, True, 0.05, False, V, -20.0, 20.0, nx, -15.0, 15.0, ny,
                         BoundaryCondition.Period,
                         SolverMethod.SplittingMethodO2 if method == 2
                         else SolverMethod.SplittingMethodO4, 1.0, 1.0, o)
    return e, psi, V


def _status(g, job):
    from qusim_amd import _native as N
    f, left = ctypes.c_int(), ctypes.c_longlong()
    N.call("qsim_hip_split_group_status", g, job, ctypes.byref(f), ctypes.byref(left))
    return bool(f.value), left.value


@pytest.mark.parametrize("precision,method,use_graph",
                         [("mixed32", 2, 1), ("fp64", 2, 0), ("fp64", 4, 1)])
def test_group_of_uneven_sizes_equals_each_job_alone(precision, method, use_graph, oracle):
    from qusim_amd import _native as N
    # (shape, steps, advance at which the job joins)
    plan = [((256, 256), 23, 0), ((512, 256), 9, 0), ((1024, 1024), 17, 1),
            ((128, 2048), 12, 2), ((64, 64), 30, 2), ((2048, 128), 5, 3)]
    g = ctypes.c_void_p()
    N.call("qsim_hip_split_group_create", ctypes.byref(g), 16, 64 if precision == "fp64" else 32,
           method)
    g = g.value
    jobs, alone = [], []
    try:
        for i, (shape, steps, _) in enumerate(plan):
            jobs.append(_job(shape, i, precision, method))
            alone.append(_job(shape, i, precision, method))
        K, adv, active, done = 7, 0, [], set()
        while len(done) < len(plan):
            for i, (_, steps, at) in enumerate(plan):
                if at == adv:
                    N.call("qsim_hip_split_group_add", g, jobs[i][0]._job, steps)
                    active.append(i)
            N.call("qsim_hip_split_group_advance", g, K, use_graph, None)
            adv += 1
            for i in list(active):
                fin, _ = _status(g, jobs[i][0]._job)
                if fin:
                    N.call("qsim_hip_split_group_remove", g, jobs[i][0]._job)
                    active.remove(i)
                    done.add(i)
            assert adv < 100
        N.call("qsim_hip_synchronize", None)
        nj, ns, ng, nf = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong())
        N.call("qsim_hip_split_group_info", g, ctypes.byref(nj), ctypes.byref(ns),
               ctypes.byref(ng), ctypes.byref(nf))
        assert nj.value == 0 and nf.value == 0
        assert (ng.value >= 1) == bool(use_graph)
        for i, (shape, steps, _) in enumerate(plan):
            N.call("qsim_hip_split_steps", alone[i][0]._job, steps, None)
            got, want = jobs[i][0].GetPsi(), alone[i][0].GetPsi()
            assert np.array_equal(got, want), (shape, steps)
            if precision == "fp64" and shape[0] * shape[1] <= 256 * 512:
                e, psi, V = jobs[i]
                p0 = psi / np.sqrt(np.sum(np.abs(psi) ** 2) * (40.0 / shape[0])
                                   * (30.0 / shape[1]))
                ref = oracle.split2d(p0, V, -20.0, 20.0, -15.0, 15.0, 0.05, steps, method=method,
                                     complex_v=True)
                assert rel_l2(got, ref) < 1e-12, shape
    finally:
        N.call("qsim_hip_split_group_destroy", g)


def test_group_refuses_what_it_cannot_run():
    from qusim_amd import _native as N
    g = ctypes.c_void_p()
    N.call("qsim_hip_split_group_create", ctypes.byref(g), 2, 64, 2)
    g = g.value
    try:
        e, _, _ = _job((256, 256), 0, "fp64", 2)
        f, _, _ = _job((128, 128), 1, "mixed32", 2)
        with pytest.raises(N.QsimError, match="precision and method"):
            N.call("qsim_hip_split_group_add", g, f._job, 5)
        big, _, _ = _job((1024, 1024), 2, "fp64", 2)
        N.call("qsim_hip_split_set_variants", big._job, 6, 6)
        with pytest.raises(N.QsimError, match="default launch schedule"):
            N.call("qsim_hip_split_group_add", g, big._job, 5)
        N.call("qsim_hip_split_group_add", g, e._job, 5)
        with pytest.raises(N.QsimError, match="already"):
            N.call("qsim_hip_split_group_add", g, e._job, 5)
        with pytest.raises(N.QsimError, match="at least one step"):
            N.call("qsim_hip_split_group_add", g, big._job, 0)
        N.call("qsim_hip_split_group_advance", g, 3, 0, None)
        fin, left = _status(g, e._job)
        assert not fin and left == 5 + 1 - 3
        N.call("qsim_hip_split_group_advance", g, 3, 1, None)
        fin, left = _status(g, e._job)
        assert fin and left == 0
    finally:
        N.call("qsim_hip_split_group_destroy", g)


def _info(g):
    from qusim_amd import _native as N
    v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()]
    N.call("qsim_hip_split_group_info", g, *[ctypes.byref(x) for x in v])
    return [x.value for x in v]


def test_group_captures_follow_the_shape_set():
    """Captured advances are cached by their launch configuration: when a
    shape set comes back (a 512^2 job joins a 256^2 one and leaves again), the
    group replays its earlier capture instead of capturing again, and the
    results stay bitwise those of direct launches."""
    from qusim_amd import _native as N
    groups, jobs = [], []
    for use_graph in (1, 0):
        g = ctypes.c_void_p()
        N.call("qsim_hip_split_group_create", ctypes.byref(g), 4, 32, 2)
        groups.append(g.value)
        jobs.append((_job((256, 256), 5, "mixed32", 2)[0], _job((512, 512), 6, "mixed32", 2)[0]))
    try:
        seen = []
        for gi, use_graph in enumerate((1, 0)):
            g, (a, b) = groups[gi], jobs[gi]
            N.call("qsim_hip_split_group_add", g, a._job, 40)
            N.call("qsim_hip_split_group_advance", g, 5, use_graph, None)
            N.call("qsim_hip_split_group_add", g, b._job, 6)
            adv = 0
            while not _status(g, b._job)[0]:
                N.call("qsim_hip_split_group_advance", g, 5, use_graph, None)
                adv += 1
                assert adv < 20
            N.call("qsim_hip_split_group_remove", g, b._job)
            while not _status(g, a._job)[0]:
                N.call("qsim_hip_split_group_advance", g, 5, use_graph, None)
                adv += 1
                assert adv < 40
            N.call("qsim_hip_synchronize", None)
            seen.append(_info(g))
        # the graph group holds two captures ({256^2} and {256^2, 512^2}),
        # none failed; the direct group none
        assert seen[0][2] == 2 and seen[0][3] == 0, seen
        assert seen[1][2] == 0, seen
        assert np.array_equal(jobs[0][0].GetPsi(), jobs[1][0].GetPsi())
        assert np.array_equal(jobs[0][1].GetPsi(), jobs[1][1].GetPsi())
    finally:
        for g in groups:
            N.call("qsim_hip_split_group_destroy", g)


def test_job_graph_replays_on_two_streams():
    """A job's captured steps replayed on alternating streams give the bits
    of direct launches (each replay follows the last one)."""
    from qusim_amd import _native as N
    a, _, _ = _job((256, 256), 7, "fp64", 2)
    b, _, _ = _job((256, 256), 7, "fp64", 2)
    streams = []
    for _ in range(2):
        p = ctypes.c_void_p()
        N.call("qsim_hip_stream_create", ctypes.byref(p))
        streams.append(p.value)
    try:
        N.call("qsim_hip_split_use_graph", b._job, 1)
        for i in range(6):
            N.call("qsim_hip_split_steps", a._job, 3, None)
            N.call("qsim_hip_split_steps", b._job, 3, streams[i % 2])
        N.call("qsim_hip_synchronize", None)
        for st in streams:
            N.call("qsim_hip_synchronize", st)
        assert np.array_equal(a.GetPsi(), b.GetPsi())
    finally:
        for st in streams:
            N.call("qsim_hip_stream_destroy", st)
