"""Reconstruction switches, mirroring the reference's flat ``Params`` object
(Params/Params.py:41-195): same field names and defaults.  Only the fields the
hot path reads change behaviour here; every switch the MI355X engines do not
implement is rejected with NotImplementedError when an engine starts
(``BaseEngine._check_supported``), the way the reference raises for switches
it does not implement (BaseEngine.py:862, :1491)."""

import numpy as np


class Params:
    def __init__(self):
        # TV options (Params.py:45-49)
        self.positionCorrectionSwitch_radius = 1
        self.objectTVfreq = 5
        self.objectTVregSwitch = False
        self.objectTVregStepSize = 1e-3
        self.weigh_probe_updates_by_intensity = False
        self.dump_obj = False
        # computation switches (Params.py:58-72)
        self.fftshiftSwitch = False
        self.fftshiftFlag = False
        self.FourierMaskSwitch = False
        self.CPSCswitch = False
        self.CPSCupsamplingFactor = None
        self.intensityConstraint = "standard"
        self.propagatorType = "Fraunhofer"
        self.momentumAcceleration = False
        self.adaptiveMomentumAcceleration = False
        self._gpuSwitch = True
        self.saveMemory = False
        self.probeUpdateStart = 1
        self.objectUpdateStart = 1
        self.positionOrder = "random"
        # applyConstraints switches (Params.py:77-133)
        self.orthogonalizationSwitch = False
        self.orthogonalizationFrequency = 10
        self.objectSmoothenessSwitch = False
        self.objectSmoothenessWidth = 2
        self.objectSmoothnessAleph = 1e-2
        self.absObjectSwitch = False
        self.absObjectBeta = 1e-2
        self.objectContrastSwitch = False
        self.probeSmoothenessSwitch = False
        self.probeSmoothnessAleph = 5e-2
        self.probeSmoothenessWidth = 3
        self.probeBoundary = False
        self.absorbingProbeBoundary = False
        self.absorbingProbeBoundaryAleph = 5e-2
        self.probePowerCorrectionSwitch = False
        self.probeSpectralPowerCorrectionSwitch = False
        self.modulusEnforcedProbeSwitch = False
        self.absProbeSwitch = False
        self.absProbeBeta = 1e-2
        self.couplingSwitch = False
        self.couplingAleph = 50e-2
        self.binaryProbeSwitch = False
        self.binaryProbeThreshold = 0.1
        self.binaryProbeAleph = 10e-2
        self.backgroundModeSwitch = False
        self.comStabilizationSwitch = False
        self.PSDestimationSwitch = False
        self.positionCorrectionSwitch = False
        self.adaptiveDenoisingSwitch = False
        self.l2reg = False
        self.l2reg_probe_aleph = 0.01
        self.l2reg_object_aleph = 0.001
        # autofocus (Params.py:141-170)
        self.TV_autofocus = False
        self.TV_autofocus_metric = "TV"
        self.TV_autofocus_intensityonly = False
        self.TV_autofocus_stepsize = 5
        self.TV_autofocus_aleph = 0.01
        self.TV_autofocus_roi = [0.4, 0.6]
        self.TV_autofocus_range_dof = 11
        self.TV_autofocus_friction = 0.7
        self.TV_autofocus_what = "object"
        self.TV_autofocus_run_every = 3
        self.TV_autofocus_min_z = None
        self.TV_autofocus_max_z = None
        self.TV_autofocus_nplanes = 11
        self.map_position_to_z_change = False
        # OPR (Params.py:156-190)
        self.OPRP = False
        self.OPR_modes = np.array([0])
        self.OPR_subspace = 4
        self.OPR_alpha = 0.05
        self.OPR_tv_freq = 5
        self.OPR_tv = False
        self.OPR_tsvd_type = "gram"
        self.OPR_orthogonalize_modes = True
        self.OPR_fast_orthogonalization = True
        self.OPR_neighbor_constraint = False
        self.SHG_probe = False

    @property
    def propagator(self):
        """Alias for propagatorType (Params.py:197-199)."""
        return self.propagatorType

    @property
    def gpuSwitch(self):
        return self._gpuSwitch

    @gpuSwitch.setter
    def gpuSwitch(self, value):
        # The MI355X engines always run on the GPU; False is accepted for
        # source compatibility and recorded, but there is no CPU path.
        self._gpuSwitch = bool(value)
