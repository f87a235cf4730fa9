"""e3PIE (Engines/e3PIE.py): multislice PIE, Maiden, Humphry & Rodenburg,
JOSA A 2012.  The slice-to-slice angular-spectrum transfer function is built
on the host in float64 from the probe geometry (e3PIE.py:61-69) and used in
complex64 on the GPU."""

import numpy as np

from ..Operators import aspw_transfer_function
from .BaseEngine import BaseEngine


class e3PIE(BaseEngine):
    ENGINE = "e3PIE"

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()

    def initializeReconstructionParams(self):
        """e3PIE.py:46-69."""
        self.betaProbe = 0.25
        self.betaObject = 0.25
        self.numIterations = 50
        r = self.reconstruction
        H = aspw_transfer_function(r.dz, r.wavelength / r.refrIndex, r.Np, r.Lp)
        r.H = np.fft.ifftshift(H)

    def _transfer_function(self):
        if self.params.propagatorType.lower() != "fraunhofer":
            raise NotImplementedError("e3PIE runs with the Fraunhofer detector propagator on the GPU path")
        return np.ascontiguousarray(self.reconstruction.H, dtype=np.complex64)

    def _engine_coefficients(self):
        return dict(betaObject=self.betaObject, betaProbe=self.betaProbe, sliceBeta=0.9)

    def reconstruct(self):
        """e3PIE.py:71-190; objectProd as the product of the slices (:176)."""
        super().reconstruct()
        self.reconstruction.objectProd = np.prod(self.reconstruction.object, 3)
