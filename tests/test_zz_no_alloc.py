"""qsim_hip_split_steps_batched_noalloc (descriptor table in a static
per-GPU device array) against qsim_hip_split_steps_batched, through the C ABI:
bitwise equal results on two calls in a row, the second with fewer and
smaller jobs, plus its limits."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(shape, count, precision="fp64", seed=0):
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nx, ny = shape
        x = -20 + 40.0 / nx * np.arange(nx)
        y = -20 + 40.0 / ny * np.arange(ny)
        X, Y = np.meshgrid(x, y, indexing="ij")
        c = rng.uniform(-5, 5, 2)
        psi = np.exp(0.7j * X - ((X - c[0]) ** 2 + (Y - c[1]) ** 2) / 20.0)
        V = rng.uniform(0.1, 0.6) * np.exp(-(X ** 2 + Y ** 2) / 10) - 0.02j * (X / 20) ** 2
        o = Options().Hip().Batch(1).ComplexPotential(True)
        if precision != "fp64":
            o.Precision(precision)
        out.append(Evolver2D().init(psi, True, 0.05, False, V, -20.0, 20.0, nx, -20.0, 20.0,
                                    ny, BoundaryCondition.Period,
                                    SolverMethod.SplittingMethodO2, 1.0, 1.0, o))
    return out


def _advance(fn, evs, nsteps):
    from qusim_amd import _native as N
    arr = (ctypes.c_void_p * len(evs))(*[e._job for e in evs])
    N.call(fn, arr, len(evs), nsteps, None)
    N.call("qsim_hip_synchronize", None)


@pytest.mark.parametrize("precision", ["fp64", "mixed32"])
def test_noalloc_equals_batched_on_two_calls(precision):
    for shape, count, steps in (((256, 256), 6, 7), ((128, 64), 3, 5)):
        a = _jobs(shape, count, precision, seed=count)
        b = _jobs(shape, count, precision, seed=count)
        _advance("qsim_hip_split_steps_batched", a, steps)
        _advance("qsim_hip_split_steps_batched_noalloc", b, steps)
        for ea, eb in zip(a, b):
            assert np.array_equal(ea.GetPsi(), eb.GetPsi()), (shape, count)


def test_noalloc_mixed_shapes_and_limits():
    from qusim_amd import _native as N
    a = _jobs((128, 128), 2, seed=3) + _jobs((64, 256), 2, seed=4)
    b = _jobs((128, 128), 2, seed=3) + _jobs((64, 256), 2, seed=4)
    _advance("qsim_hip_split_steps_batched", a, 4)
    _advance("qsim_hip_split_steps_batched_noalloc", b, 4)
    for ea, eb in zip(a, b):
        assert np.array_equal(ea.GetPsi(), eb.GetPsi())
    arr = (ctypes.c_void_p * 513)(*([b[0]._job] * 513))
    with pytest.raises(N.QsimError, match="512"):
        N.call("qsim_hip_split_steps_batched_noalloc", arr, 513, 1, None)
    # an unfused job (non-power-of-two axis) is refused, and the library
    # stays usable
    from qusim_amd.api import BoundaryCondition, Evolver2D, Options, SolverMethod
    psi = np.ones((48, 40), np.complex128)
    u = Evolver2D().init(psi, True, 0.05, False, np.zeros((48, 40)), -5.0, 5.0, 48, -5.0, 5.0,
                         40, BoundaryCondition.Period, SolverMethod.SplittingMethodO2, 1.0, 1.0,
                         Options().Hip())
    arr2 = (ctypes.c_void_p * 2)(b[0]._job, u._job)
    with pytest.raises(N.QsimError, match="fused"):
        N.call("qsim_hip_split_steps_batched_noalloc", arr2, 2, 1, None)
    _advance("qsim_hip_split_steps_batched_noalloc", b[:2], 1)
    _advance("qsim_hip_split_steps_batched", a[:2], 1)
    assert np.array_equal(a[0].GetPsi(), b[0].GetPsi())
