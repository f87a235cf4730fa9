"""ePIE (Engines/ePIE.py): extended PIE, Maiden & Rodenburg 2009."""

from .BaseEngine import BaseEngine


class ePIE(BaseEngine):
    ENGINE = "ePIE"

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()

    def initializeReconstructionParams(self):
        """ePIE.py:35-42."""
        self.betaProbe = 0.25
        self.betaObject = 0.25
        self.numIterations = 50

    def reconstruct(self, experimentalData=None):
        """Run to completion (ePIE.py:44-51)."""
        return super().reconstruct(experimentalData)

    def reconstruct_stepwise(self, experimentalData=None):
        """Generator variant (ePIE.py:53-126): yields ``(iteration,
        positionLoop)`` after every scan position.  While it runs,
        ``reconstruction.object``/``probe`` are the live GPU tensors; after the
        last yield they are NumPy arrays."""
        if experimentalData is not None:
            self.reconstruction.data = experimentalData
            self.experimentalData = experimentalData
        self._prepareReconstruction()
        plan, _ = self._begin()
        F = self.experimentalData.numFrames
        done = 0
        try:
            for loop in range(int(self.numIterations)):
                for j in range(F):
                    plan.run_positions(loop, j, j + 1)
                    yield loop, j
                done = loop + 1
                self.applyConstraints(loop)
        finally:
            self._end(done)
