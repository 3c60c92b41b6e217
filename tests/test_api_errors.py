"""Argument errors of the host API that the reference raises before any
device work (src/DeviceType.h:21, src/ScatteringSolverImpl.cpp:126-130,
src/PerturbationOptions.cpp:14-28): they raise here too, on the host, before
the library touches a GPU.  No GPU needed."""
import numpy as np
import pytest

from qusim_amd.api import (BoundaryCondition, Evolver1D, Evolver2D, Options, QuPerturbation2D,
                           QuPerturbation3D, SolverMethod)


def _evolve2d(opts, solver=SolverMethod.SplittingMethodO2, b=BoundaryCondition.Period):
    psi = np.ones((64, 64), np.complex128)
    return Evolver2D().init(psi, True, 0.1, False, np.zeros((64, 64)), -1, 1, 64, -1, 1, 64, b,
                            solver, 1.0, 1.0, opts)


def _born3d(V, opts, met=SolverMethod.BornSerise):
    return QuPerturbation3D().init(V, -16, 16, 32, -16, 16, 32, -16, 16, 32, 3.0, 0.0, 0, 0, 1,
                                   met, 1.0, 1.0, opts)


@pytest.mark.parametrize("dev", ["CPU_SEQ", "CPU_PAR", "GPU_CUDA"])
def test_other_reference_devices_are_refused(dev):
    with pytest.raises(ValueError, match="not provided"):
        _evolve2d(Options().SetString("device", dev))


def test_unknown_device_string():
    with pytest.raises(ValueError, match="invalid device"):
        _evolve2d(Options().SetString("device", "GPU_OPENCL"))
    with pytest.raises(ValueError, match="invalid device"):
        _born3d(np.zeros((32, 32, 32)), Options().SetString("device", "TPU"))


def test_invalid_precision():
    with pytest.raises(ValueError, match="invalid precision"):
        Options().Precision("fp16")


def test_unsupported_solver_and_boundary():
    with pytest.raises(RuntimeError, match="unsupported solver"):
        x = np.zeros(64)
        Evolver1D().init(np.ones(64, np.complex128), True, 0.1, False, x, -1, 1, 64,
                         BoundaryCondition.Period, SolverMethod.BornSerise, 1.0, 1.0,
                         Options().Hip())
    with pytest.raises(RuntimeError, match="unsupported boundary"):
        _evolve2d(Options().Hip(), b=BoundaryCondition.InfiniteWall)


def test_nan_in_the_potential():
    V = np.zeros((32, 32, 32), np.complex128)
    V[3, 4, 5] = complex(0.0, np.nan)
    with pytest.raises(RuntimeError, match="nan in function"):
        _born3d(V, Options().Hip())
    V2 = np.zeros((64, 64))
    V2[0, 0] = np.nan
    with pytest.raises(RuntimeError, match="nan in function"):
        QuPerturbation2D().init(V2, -16, 16, 64, -16, 16, 64, 3.0, 0.0, 1, 0,
                                SolverMethod.BornSerise, 1.0, 1.0, Options().Hip())


def test_born_method_and_preconditioner_names():
    with pytest.raises(RuntimeError, match="not support method"):
        _born3d(np.zeros((32, 32, 32)), Options().Hip(), met=SolverMethod.SplittingMethodO2)
    with pytest.raises(RuntimeError, match="Unknown preconditioner"):
        _born3d(np.zeros((32, 32, 32)),
                Options().Hip().Preconditional(True).SetString("preconditioner", "Jacobi"))
