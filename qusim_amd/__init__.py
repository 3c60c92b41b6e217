"""qusim_amd -- MI355X (gfx950) device for QuSim's split-operator evolution
and Born-series scattering path.

Host-side mirror of the reference interface (include/QuSim.h):
Options (with .Hip()), Evolver1D / Evolver2D, QuPerturbation2D / 3D, and the
FourierTransform plug-in, all executing through libqsim_hip.so's C ABI."""
from . import _native  # noqa: F401
from .fft import FourierTransform, fftn  # noqa: F401
