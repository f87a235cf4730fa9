"""MI355X-native ptychographic reconstruction (ePIE / mPIE / multiPIE / e3PIE).

Drop-in for the reference's engine API: ``Engines.ePIE(reconstruction,
experimentalData, params, monitor).reconstruct()`` with the same ``Params``
fields and the same 6D (nlambda, nosm, npsm, nslice, N, N) layout; the
per-position loop runs as hand-written gfx950 HIP kernels behind the C ABI in
include/pty_hip.h (``libptyhip.so``)."""

from . import Engines, Operators  # noqa: F401
from .ExperimentalData import ExperimentalData  # noqa: F401
from .Monitor import DummyMonitor  # noqa: F401
from .Params import Params  # noqa: F401
from .Reconstruction import Reconstruction  # noqa: F401

__version__ = "0.1.0"
