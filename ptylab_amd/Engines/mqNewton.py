"""mqNewton (Engines/mqNewton.py): qNewton with momentum after every scan
position.  The position step is qNewton's (object step, Omax over the updated
object, probe step; qNewton.py:106-140 = mqNewton.py:226-251); then the object
and the probe each take a step of the selected momentum engine on the
gradient buffer - current (mqNewton.py:121-122, 184-224): "ADAM"
(:142-153, the default) and "NADAM" (:155-174) with the squared 2-norm of the
whole gradient as the second moment, or plain "momentum" (:176-182)."""

import numpy as np

from .. import _native
from .BaseEngine import BaseEngine


class mqNewton(BaseEngine):
    ENGINE = "mqNewton"
    _MOMENTUM_TRIGGERS = False

    def __init__(self, reconstruction, experimentalData, params, monitor):
        super().__init__(reconstruction, experimentalData, params, monitor)
        self.initializeReconstructionParams()
        r = self.reconstruction
        # mqNewton.py:43-48
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
        # the C ABI carries the regularisers in the alpha slots and the
        # momentum coefficients in the momentum slots (include/pty_hip.h,
        # PTY_MQNEWTON): beta1 / beta2, or frictionM for plain momentum
        plain = self.momentum_method == "momentum"
        return dict(betaObject=self.betaObject, betaProbe=self.betaProbe,
                    alphaObject=self.regObject, alphaProbe=self.regProbe,
                    momFriction=self.frictionM if plain else self.beta1, momFeedback=self.beta2,
                    probeWeight=self.betaObject_m, sliceBeta=self.betaProbe_m)

    def _instance_flags(self):
        return {"ADAM": 0, "NADAM": _native.PTY_FLAG_MQN_NADAM,
                "momentum": _native.PTY_FLAG_MQN_MOMENTUM}[self.momentum_method]

    def _prepareReconstruction(self):
        if self.momentum_method not in ("ADAM", "NADAM", "momentum"):
            # getattr(mqNewton, momentum_method) (mqNewton.py:69) fails the same way
            raise AttributeError(f"type object 'mqNewton' has no attribute {self.momentum_method!r}")
        r = self.reconstruction
        if self.momentum_method == "momentum":
            # initializeAdaptiveMomentum (mqNewton.py:68-79) creates the second
            # moments for ADAM and NADAM only, and the momentum updates read them
            # (mqNewton.py:190-198), so on a reconstruction without them the
            # reference stops with this error at the first position
            for name in ("objectMomentum_v", "probeMomentum_v"):
                if not hasattr(r, name):
                    raise AttributeError(f"'Reconstruction' object has no attribute {name!r}")
        else:
            # initializeAdaptiveMomentum seeds the second moments with a copy of
            # the first; the GPU keeps them as one number each, which is exact
            # when the first moments start at zero
            for name in ("objectMomentum", "probeMomentum"):
                m = getattr(r, name, None)
                if m is not None and np.any(np.asarray(m) != 0):
                    raise NotImplementedError("mqNewton resumed with a nonzero momentum is not implemented by the "
                                              "MI355X engines")
        super()._prepareReconstruction()
