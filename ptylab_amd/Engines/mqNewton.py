"""mqNewton (Engines/mqNewton.py): qNewton with adaptive momentum after every
scan position.  The position step is qNewton's (object step, Omax over the
updated object, probe step; qNewton.py:106-140 = mqNewton.py:197-235); then
the object and the probe each take an ADAM step on the gradient buffer -
current with the squared 2-norm of the whole gradient as the second moment
(mqNewton.py:112-113, 160-174, 176-195).  Only the default momentum_method
"ADAM" is implemented."""

import numpy as np

from .BaseEngine import BaseEngine


class mqNewton(BaseEngine):
    ENGINE = "mqNewton"
    _MOMENTUM_TRIGGERS = False

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()
        r = self.reconstruction
        # mqNewton.py:42-48
        r.initializeObjectMomentum()
        r.initializeProbeMomentum()
        r.objectBuffer = np.array(r.object, copy=True)
        r.probeBuffer = np.array(r.probe, copy=True)
        self.params.momentumAcceleration = True
        self.name = "mqNewton"

    def initializeReconstructionParams(self):
        """mqNewton.py:50-66."""
        self.betaProbe = 1
        self.betaObject = 1
        self.regObject = 1
        self.regProbe = 1
        self.beta1 = 0.5
        self.beta2 = 0.5
        self.betaProbe_m = 0.25
        self.betaObject_m = 0.25
        self.feedbackM = 0.3
        self.frictionM = 0.7
        self.momentum_method = "ADAM"
        self.numIterations = 50

    def _momentum_enabled(self):
        return True

    def _engine_coefficients(self):
        # the C ABI carries the regularisers in the alpha slots and the ADAM
        # coefficients in the momentum slots (include/pty_hip.h, PTY_MQNEWTON)
        return dict(betaObject=self.betaObject, betaProbe=self.betaProbe,
                    alphaObject=self.regObject, alphaProbe=self.regProbe,
                    momFriction=self.beta1, momFeedback=self.beta2,
                    probeWeight=self.betaObject_m, sliceBeta=self.betaProbe_m)

    def _prepareReconstruction(self):
        if self.momentum_method != "ADAM":
            raise NotImplementedError(f"mqNewton momentum_method {self.momentum_method!r} is not implemented "
                                      "by the MI355X engines (ADAM is)")
        r = self.reconstruction
        # initializeAdaptiveMomentum (mqNewton.py:70-84) seeds the second
        # moments with a copy of the first; the GPU keeps them as one number
        # each, which is exact when the first moments start at zero
        for name in ("objectMomentum", "probeMomentum"):
            m = getattr(r, name, None)
            if m is not None and np.any(np.asarray(m) != 0):
                raise NotImplementedError("mqNewton resumed with a nonzero momentum is not implemented by the MI355X "
                                          "engines")
        super()._prepareReconstruction()
