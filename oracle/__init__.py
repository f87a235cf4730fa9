"""CPU oracle for the per-position ptychography loop.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference's hot path (PtyLab 0.3.1: Engines/ePIE.py, mPIE.py, multiPIE.py,
e3PIE.py, Engines/BaseEngine.py, Operators/Operators.py, utils/utils.py,
Reconstruction/Reconstruction.py, ExperimentalData/ExperimentalData.py,
utils/initializationFunctions.py).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and there only as
the checker or as the timed CPU baseline.  The product package ``ptylab_amd``
never imports it and has no CPU fallback through it.

Parity status of each function is listed in ``ptyoracle.PINNED``.
"""

from .ptyoracle import *  # noqa: F401,F403
