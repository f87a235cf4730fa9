"""qNewton (Engines/qNewton.py): a quasi-Newton PIE variant whose object
step weights conj(P) by |P| / max(sum |P|) over (|P|^2 + regObject) and whose
probe step takes Omax over the whole object after the object step
(qNewton.py:106-140)."""

from .BaseEngine import BaseEngine


class qNewton(BaseEngine):
    ENGINE = "qNewton"

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()
        self.name = "qNewton"

    def initializeReconstructionParams(self):
        """qNewton.py:45-53."""
        self.betaProbe = 1
        self.betaObject = 1
        self.regObject = 1
        self.regProbe = 1
        self.numIterations = 50

    def _engine_coefficients(self):
        # the C ABI carries the two regularisers in the alpha slots
        return dict(betaObject=self.betaObject, betaProbe=self.betaProbe,
                    alphaObject=self.regObject, alphaProbe=self.regProbe)
