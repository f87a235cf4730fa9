"""pcPIE (Engines/pcPIE.py): mPIE-type updates and momentum (stepM friction,
betaM feedback) with the position correction of BaseEngine.positionCorrection
after every position when params.positionCorrectionSwitch is set
(pcPIE.py:71-122, BaseEngine.py:1241-1323 and 1400-1438)."""

import numpy as np

from .BaseEngine import BaseEngine


class pcPIE(BaseEngine):
    ENGINE = "pcPIE"
    _POSITION_CORRECTION = True

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()
        r = self.reconstruction
        # pcPIE.py:44-51
        r.initializeObjectMomentum()
        r.initializeProbeMomentum()
        r.objectBuffer = np.array(r.object, copy=True)
        r.probeBuffer = np.array(r.probe, copy=True)
        self.params.momentumAcceleration = True
        self.name = "pcPIE"

    def initializeReconstructionParams(self):
        """pcPIE.py:54-69."""
        self.betaProbe = 0.25
        self.betaObject = 0.25
        self.alphaProbe = 0.1
        self.alphaObject = 0.1
        self.betaM = 0.3
        self.stepM = 0.7
        self.numIterations = 50

    def _momentum_enabled(self):
        return True

    def _engine_coefficients(self):
        # the object and probe updates are mPIE's (pcPIE.py:154-202 with
        # weight 1); the momentum step is mPIE's with stepM and betaM
        return dict(betaObject=self.betaObject, betaProbe=self.betaProbe,
                    alphaObject=self.alphaObject, alphaProbe=self.alphaProbe,
                    momFriction=self.stepM, momFeedback=self.betaM, probeWeight=1.0)
