"""multiPIE (Engines/multiPIE.py): mPIE with the probe update summed over
the wavelength and object-mode axes (multiPIE.py:190-205) and its own
momentum coefficients betaM/stepM (multiPIE.py:125-151)."""

import numpy as np

from .BaseEngine import BaseEngine


class multiPIE(BaseEngine):
    ENGINE = "multiPIE"

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()
        self.params.momentumAcceleration = True

    def initializeReconstructionParams(self):
        """multiPIE.py:46-66."""
        self.betaProbe = 0.25
        self.betaObject = 0.25
        self.alphaProbe = 0.1
        self.alphaObject = 0.1
        self.betaM = 0.3
        self.stepM = 0.7
        self.numIterations = 50
        r = self.reconstruction
        r.initializeObjectMomentum()
        r.initializeProbeMomentum()
        r.objectBuffer = np.array(r.object, copy=True)
        r.probeBuffer = np.array(r.probe, copy=True)

    def _momentum_enabled(self):
        return True

    def _engine_coefficients(self):
        return dict(
            betaObject=self.betaObject, betaProbe=self.betaProbe,
            alphaObject=self.alphaObject, alphaProbe=self.alphaProbe,
            momFriction=self.stepM, momFeedback=self.betaM,
        )

    def reconstruct(self):
        """multiPIE.py:68-123."""
        return super().reconstruct()
