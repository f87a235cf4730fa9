"""mPIE (Engines/mPIE.py): momentum-accelerated PIE, Maiden, Johnson & Li,
Optica 2017."""

import numpy as np

from .BaseEngine import BaseEngine


class mPIE(BaseEngine):
    ENGINE = "mPIE"
    _POSITION_CORRECTION = True

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()
        self.params.momentumAcceleration = True
        self.name = "mPIE"

    def initializeReconstructionParams(self):
        """mPIE.py:73-94."""
        self.betaProbe = 0.25
        self.betaObject = 0.25
        self.alphaProbe = 0.1
        self.alphaObject = 0.1
        self.feedbackM = 0.3
        self.frictionM = 0.7
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
            momFriction=self.frictionM, momFeedback=self.feedbackM, probeWeight=1.0,
        )

    def reconstruct(self, experimentalData=None, reconstruction=None, vis_after_each_iteration=None):
        """mPIE.py:96-194.  ``vis_after_each_iteration(loop, reconstruction)``
        is called after every iteration with the live state synchronised."""
        if vis_after_each_iteration is None:
            return super().reconstruct(experimentalData, reconstruction)
        if experimentalData is not None:
            self.experimentalData = experimentalData
        if reconstruction is not None:
            self.reconstruction = reconstruction
        self._prepareReconstruction()
        plan, _ = self._begin()
        done = 0
        try:
            for loop in range(int(self.numIterations)):
                plan.run(loop, loop + 1)
                done = loop + 1
                self.applyConstraints(loop)
                if callable(vis_after_each_iteration):
                    plan.check()
                    vis_after_each_iteration(loop, self.reconstruction)
        finally:
            self._end(done)
