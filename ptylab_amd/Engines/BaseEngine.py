"""Shared engine machinery (Engines/BaseEngine.py), MI355X edition.

An engine keeps references to the reconstruction, data, params and monitor
(BaseEngine.py:78-94).  ``reconstruct()`` does, in order:

1. the host-side preparation of ``_prepareReconstruction`` that the hot path
   depends on (fftshift of the data, error containers), and a check that no
   switch outside the GPU path is on;
2. the random schedule: the position order of every iteration
   (setPositionOrder, BaseEngine.py:637-665) and, for mPIE/multiPIE, the
   momentum triggers ``np.random.rand(1) > 0.95`` after every position
   (mPIE.py:170-172).  Neither depends on the data, so they are drawn up
   front from NumPy's global RNG in exactly the reference's order and the RNG
   state afterwards is the reference's;
3. upload of the state to torch tensors on the GPU, one ``pty_plan`` for the
   reconstruction, one persistent launch per iteration, the per-iteration
   constraints between launches;
4. download: object, probe, momenta and errors are NumPy again
   (ePIE.py:123-126).
"""

import logging

import numpy as np

from .. import _native
from ..Operators import (asp_tf_fftorder, aspw_transfer_function, polychrome_tf_fftorder, quadratic_phase,
                         scaled_asp_kernel_arrays, scaled_polychrome_kernel_arrays)

# switches whose non-default value changes the per-position math or the
# per-iteration constraints; the GPU engines implement the defaults and the
# constraints orthogonalization, probe power correction, COM stabilisation,
# probe boundary and absorbing probe boundary; for the others they raise
# NotImplementedError, as the reference does for the switches it does not
# implement (BaseEngine.py:862, :1491)
_UNSUPPORTED_TRUE = [
    "FourierMaskSwitch", "CPSCswitch", "backgroundModeSwitch", "adaptiveDenoisingSwitch",
    "objectTVregSwitch", "OPRP", "l2reg",
    "modulusEnforcedProbeSwitch", "probeSpectralPowerCorrectionSwitch", "PSDestimationSwitch",
    "probeSmoothenessSwitch", "objectSmoothenessSwitch", "absObjectSwitch", "absProbeSwitch",
    "objectContrastSwitch", "couplingSwitch", "binaryProbeSwitch", "TV_autofocus",
    "map_position_to_z_change", "weigh_probe_updates_by_intensity",
]


class BaseEngine:
    ENGINE = None
    # mPIE / multiPIE draw one rand(1) per position for their momentum step
    # (mPIE.py:170-172); engines with momentum at every position draw none
    _MOMENTUM_TRIGGERS = True
    # engines whose position loop calls positionCorrection (mPIE.py:163-166,
    # pcPIE.py:107-108); with the switch on the others only recentre
    _POSITION_CORRECTION = False

    def __init__(self, reconstruction, experimentalData, params, monitor):
        self.betaObject = 0.25
        self.reconstruction = reconstruction
        self.experimentalData = experimentalData
        self.params = params
        self.monitor = monitor
        self.monitor.reconstruction = reconstruction
        # workgroups cooperating on one reconstruction (0: chosen by the
        # library from Np); an MI355X-specific knob with no reference twin
        self.workgroups = 0
        self._T = None  # device tensors while a plan runs this engine
        self._plan = None
        self.logger = logging.getLogger(self.ENGINE or "BaseEngine")

    # -- preparation -------------------------------------------------------
    def _check_supported(self):
        p = self.params
        for name in _UNSUPPORTED_TRUE:
            if bool(getattr(p, name, False)):
                raise NotImplementedError(f"params.{name} is not implemented by the MI355X engines")
        if p.intensityConstraint != "standard":
            if p.intensityConstraint in ("fluctuation", "exponential", "poission", "interferometric"):
                raise NotImplementedError(f"intensityConstraint {p.intensityConstraint!r} is not implemented")
            raise ValueError("intensity constraint not properly specified!")
        if p.positionOrder not in ("sequential", "random", "NA"):
            raise ValueError("position order not properly set")
        kind = p.propagatorType.lower()
        if kind == "twosteppolychrome":
            # the reference's detector2object returns the same back-propagation
            # as esw and eswUpdate for it (Operators.py:285-310), so its
            # engines never change object or probe; not run here
            raise NotImplementedError("twoStepPolychrome in the engines is not implemented by the MI355X engines "
                                      "(Operators.object2detector / detector2object support it)")
        if kind not in ("fraunhofer", "asp", "fresnel", "polychromeasp", "scaledasp", "scaledpolychromeasp"):
            raise NotImplementedError(f"propagatorType {p.propagatorType!r} is not implemented by the MI355X engines")
        if kind in ("asp", "polychromeasp") and p.fftshiftSwitch:
            raise ValueError("ASP propagatorType works only with fftshiftSwitch = False")
        if kind in ("scaledasp", "scaledpolychromeasp") and p.fftshiftSwitch:
            raise ValueError(f"{p.propagatorType} propagatorType works only with fftshiftSwitch = False!")
        if kind == "fresnel" and p.fftshiftSwitch:
            raise NotImplementedError("Fresnel with fftshiftSwitch=True is not implemented by the MI355X engines")
        if self.experimentalData.operationMode != "CPM":
            raise NotImplementedError("only CPM is implemented")
        if p.positionCorrectionSwitch and self._POSITION_CORRECTION:
            r = self.reconstruction
            if r.nlambda != 1 or r.nosm != 1 or r.nslice != 1:
                raise NotImplementedError("position correction needs nlambda = nosm = nslice = 1")

    def _checkFFT(self):
        """Shift the data once for fftshiftSwitch (BaseEngine.py:405-455)."""
        d = self.experimentalData
        if self.params.fftshiftSwitch:
            if not self.params.fftshiftFlag:
                d.ptychogram = np.fft.ifftshift(d.ptychogram, axes=(-1, -2))
                self.params.fftshiftFlag = 1
        elif self.params.fftshiftFlag:
            d.ptychogram = np.fft.fftshift(d.ptychogram, axes=(-1, -2))
            self.params.fftshiftFlag = 0

    def _initializeErrors(self):
        """BaseEngine._initializeErrors (BaseEngine.py:199-225): with
        saveMemory the detector error is the scalar 0, otherwise a
        (numFrames, Nd, Nd) array (kept from an earlier run)."""
        r = self.reconstruction
        if self.params.saveMemory:
            r.detectorError = 0
        elif not hasattr(r, "detectorError"):
            r.detectorError = np.zeros((self.experimentalData.numFrames, r.Nd, r.Nd))
        if not hasattr(r, "errorAtPos"):
            r.errorAtPos = np.zeros((self.experimentalData.numFrames, 1), dtype=np.float32)
        if not hasattr(r, "error"):
            r.error = []

    def _prepareReconstruction(self):
        self._check_supported()
        if not self.params.gpuSwitch:
            self.logger.warning("params.gpuSwitch is False: the MI355X engines run on the GPU regardless")
        self._checkFFT()
        self._initializeErrors()
        self.monitor.initializeMonitors()

    # -- schedule -----------------------------------------------------------
    def _position_order(self, n_errors):
        F = self.experimentalData.numFrames
        order = self.params.positionOrder
        if order == "sequential":
            return np.arange(F)
        if order == "random":
            idx = np.arange(F)
            if n_errors >= 2:
                np.random.shuffle(idx)
            return idx
        pos = self.reconstruction.positions
        rows = pos[:, 0] - np.mean(pos[:, 0])
        cols = pos[:, 1] - np.mean(pos[:, 1])
        return np.argsort(np.sqrt(rows ** 2 + cols ** 2))

    def _schedule(self, n_iter, momentum):
        F = self.experimentalData.numFrames
        n_err = len(getattr(self.reconstruction, "error", []))
        orders = np.empty((n_iter, F), dtype=np.int32)
        masks = np.zeros((n_iter, F), dtype=np.uint8)
        for it in range(n_iter):
            orders[it] = self._position_order(n_err + it)
            if momentum:
                masks[it] = np.random.rand(F) > 0.95
        return orders, masks

    # -- engine hooks ---------------------------------------------------------
    def _engine_coefficients(self):
        return dict(betaObject=self.betaObject, betaProbe=self.betaProbe)

    def _transfer_function(self):
        r = self.reconstruction
        kind = self.params.propagatorType.lower()
        if kind == "asp":
            if r.nlambda > 1:
                raise ValueError("For multi-wavelength, set polychromeASP instead of ASP")
            return asp_tf_fftorder(r.zo, r.wavelength, r.Np, r.Lp)
        if kind == "polychromeasp":
            sd = np.atleast_1d(r.spectralDensity)
            if len(sd) != r.nlambda:
                raise ValueError(f"polychromeASP: spectralDensity has {len(sd)} wavelengths, nlambda is {r.nlambda}")
            return polychrome_tf_fftorder(r.zo, sd, r.Np, r.Lp)
        if kind == "fresnel":
            return quadratic_phase(r.zo, r.wavelength, r.Np, r.dxp)
        if kind == "scaledasp":
            if r.nlambda > 1:
                raise ValueError("For multi-wavelength, scaledPolychromeASP needs to be used instead of scaledASP")
            return scaled_asp_kernel_arrays(r.zo, r.wavelength, r.Np, r.dxo, r.dxd)
        if kind == "scaledpolychromeasp":
            sd = np.atleast_1d(r.spectralDensity)
            if len(sd) != r.nlambda:
                raise ValueError(f"scaledPolychromeASP: spectralDensity has {len(sd)} wavelengths, nlambda is {r.nlambda}")
            return scaled_polychrome_kernel_arrays(r.zo, sd, r.Np, r.dxo, r.dxd)
        return None

    def _propagator_id(self):
        kind = self.params.propagatorType.lower()
        if kind == "asp":
            return _native.PTY_ASP
        if kind == "polychromeasp":
            return _native.PTY_POLYCHROME_ASP
        if kind == "fresnel":
            return _native.PTY_FRESNEL
        if kind == "scaledasp":
            return _native.PTY_SCALED_ASP
        if kind == "scaledpolychromeasp":
            return _native.PTY_SCALED_POLYCHROME_ASP
        return _native.PTY_FRAUNHOFER_SHIFTED if self.params.fftshiftSwitch else _native.PTY_FRAUNHOFER

    # -- device side ------------------------------------------------------------
    def _upload(self, n_iter, momentum):
        import torch

        dev = torch.device("cuda")
        r, d = self.reconstruction, self.experimentalData
        T = {}

        def c64(a):
            return torch.as_tensor(np.ascontiguousarray(np.asarray(a), dtype=np.complex64), device=dev).contiguous()

        T["object"] = c64(r.object)
        T["probe"] = c64(r.probe)
        if self.params.probePowerCorrectionSwitch:
            # _initialProbePowerCorrection (BaseEngine.py:227-235), before the
            # mPIE buffers copy the probe (mPIE.py:101-105)
            self._probe_power_correction(T)
        if tuple(T["object"].shape) != tuple(r.shape_O):
            raise ValueError(f"object shape {tuple(T['object'].shape)} != {r.shape_O}")
        if tuple(T["probe"].shape) != tuple(r.shape_P):
            raise ValueError(f"probe shape {tuple(T['probe'].shape)} != {r.shape_P}")
        if momentum:
            T["objectBuffer"] = T["object"].clone()
            T["probeBuffer"] = T["probe"].clone()
            T["objectMomentum"] = c64(getattr(r, "objectMomentum", np.zeros(r.shape_O, np.complex64)))
            T["probeMomentum"] = c64(getattr(r, "probeMomentum", np.zeros(r.shape_P, np.complex64)))
        T["ptychogram"] = torch.as_tensor(np.ascontiguousarray(d.ptychogram, dtype=np.float32), device=dev)
        T["positions"] = torch.as_tensor(np.ascontiguousarray(r.positions, dtype=np.int32), device=dev)
        T["energy"] = torch.as_tensor(np.ascontiguousarray(np.asarray(d.energyAtPos, dtype=np.float32)), device=dev)
        orders, masks = self._schedule(n_iter, momentum and self._MOMENTUM_TRIGGERS)
        self._n_err0 = len(getattr(r, "error", []))
        if self.params.positionCorrectionSwitch and self._POSITION_CORRECTION:
            # the search directions D start at zero (_initializePCParameters,
            # BaseEngine.py:178-197)
            T["pc_direction"] = torch.zeros((d.numFrames, 2), dtype=torch.float64, device=dev)
        self._orders, self._masks = orders, masks
        T["order"] = torch.as_tensor(orders, device=dev)
        T["mask"] = torch.as_tensor(masks, device=dev)
        F = d.numFrames
        if not self.params.saveMemory:
            # reconstruction.detectorError of the last iteration (getRMSD,
            # BaseEngine.py:851-868), written by the kernels in natural
            # detector order
            T["detector_error"] = torch.zeros((F, int(r.Np), int(r.Np)), dtype=torch.float32, device=dev)
        T["err_at_pos"] = torch.zeros((n_iter, F), dtype=torch.float64, device=dev)
        T["err_iter"] = torch.zeros((n_iter,), dtype=torch.float64, device=dev)
        tf = self._transfer_function()
        if tf is not None:
            T["tf"] = torch.as_tensor(np.ascontiguousarray(tf, dtype=np.complex64), device=dev)
        p = self.params
        if p.probeBoundary or p.absorbingProbeBoundary:
            T["probeWindow"] = torch.as_tensor(self._probe_window(), device=dev)
        if p.comStabilizationSwitch is not None and p.comStabilizationSwitch is not False:
            T["Xp"] = torch.as_tensor(np.ascontiguousarray(r.Xp, dtype=np.float32), device=dev)
            T["Yp"] = torch.as_tensor(np.ascontiguousarray(r.Yp, dtype=np.float32), device=dev)
        return T

    def _instance(self, T, n_iter, momentum):
        r = self.reconstruction
        inst = _native.PtyInstance()
        inst.engine = _native.ENGINE_IDS[self.ENGINE]
        inst.propagator = self._propagator_id()
        inst.nlambda, inst.nosm, inst.npsm, inst.nslice = int(r.nlambda), int(r.nosm), int(r.npsm), int(r.nslice)
        inst.Np, inst.No, inst.F = int(r.Np), int(r.No), int(self.experimentalData.numFrames)
        inst.n_iter = int(n_iter)
        co = self._engine_coefficients()
        inst.betaObject = float(co.get("betaObject", 0.25))
        inst.betaProbe = float(co.get("betaProbe", 0.25))
        inst.alphaObject = float(co.get("alphaObject", 0.1))
        inst.alphaProbe = float(co.get("alphaProbe", 0.1))
        inst.momFriction = float(co.get("momFriction", 0.7))
        inst.momFeedback = float(co.get("momFeedback", 0.3))
        inst.probeWeight = float(co.get("probeWeight", 1.0))
        inst.sliceBeta = float(co.get("sliceBeta", 0.9))
        ptr = lambda k: T[k].data_ptr() if k in T else None  # noqa: E731
        inst.object, inst.probe = ptr("object"), ptr("probe")
        inst.objectBuffer, inst.objectMomentum = ptr("objectBuffer"), ptr("objectMomentum")
        inst.probeBuffer, inst.probeMomentum = ptr("probeBuffer"), ptr("probeMomentum")
        inst.ptychogram, inst.positions = ptr("ptychogram"), ptr("positions")
        inst.order = ptr("order")
        inst.momentum_mask = ptr("mask") if momentum else None
        inst.energy, inst.tf = ptr("energy"), ptr("tf")
        inst.err_at_pos, inst.err_iter = ptr("err_at_pos"), ptr("err_iter")
        inst.pc_direction = ptr("pc_direction")
        inst.detector_error = ptr("detector_error")
        # positionCorrection runs once len(error) > startAtIteration = 1
        inst.pc_start = max(0, 2 - int(getattr(self, "_n_err0", 0)))
        inst.pc_radius = int(getattr(self.params, "positionCorrectionSwitch_radius", 1))
        inst.workgroups = int(self.workgroups)
        inst.flags = self._instance_flags()
        return inst

    def _instance_flags(self):
        """PTY_FLAG_* of this engine's instance (mqNewton's momentum method)."""
        return 0

    def _momentum_enabled(self):
        return False

    def _stage(self):
        """Upload the state and describe the instance; no plan yet.
        Returns (pty_instance, tensors)."""
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("the MI355X engines need a HIP GPU (torch.cuda.is_available() is False)")
        n_iter = int(self.numIterations)
        momentum = self._momentum_enabled()
        T = self._upload(n_iter, momentum)
        inst = self._instance(T, n_iter, momentum)
        self._T = T
        r = self.reconstruction
        # expose the live device state like the reference's GPU path does
        r.object, r.probe = T["object"], T["probe"]
        if momentum:
            r.objectBuffer, r.probeBuffer = T["objectBuffer"], T["probeBuffer"]
            r.objectMomentum, r.probeMomentum = T["objectMomentum"], T["probeMomentum"]
        return inst, T

    def _begin(self):
        """Upload, schedule, plan.  Returns (plan, tensors)."""
        import torch

        inst, T = self._stage()
        plan = _native.Plan([inst], T, torch.cuda.current_stream())
        self._plan = plan
        return plan, T

    def _collect(self, n_done):
        """Errors and state back to NumPy (ePIE.py:123-126) after the plan
        that ran this engine has been checked."""
        T = self._T
        r = self.reconstruction
        if self.params.saveMemory:
            self._collect_save_memory_errors(n_done)
        else:
            errs = T["err_iter"][:n_done].cpu().numpy()
            r.error = np.append(np.asarray(r.error, dtype=np.float64), errs)
            if n_done:
                r.errorAtPos = T["err_at_pos"][n_done - 1].cpu().numpy()
                r.detectorError = T["detector_error"].cpu().numpy().astype(np.float64)
        r._move_data_to_cpu()
        if "pc_direction" in T:
            self.D = T["pc_direction"].cpu().numpy()  # the search directions, as the reference's self.D
        for k in ("object", "probe", "objectBuffer", "probeBuffer", "objectMomentum", "probeMomentum"):
            if k in T:
                setattr(r, k, T[k].cpu().numpy())
        self._T = None

    def _collect_save_memory_errors(self, n_done):
        """Errors as the reference reports them with saveMemory
        (BaseEngine.py:819-841, 856-866): errorAtPos stays the (numFrames, 1)
        array that getRMSD fills with sum |I_k - Iest|, and dividing it by the
        (numFrames,) energies broadcasts to a (numFrames, numFrames) array,
        so every iteration's error is sum_k e_k * sum_j 1 / (E_j + 1e-20) and
        errorAtPos ends as that matrix."""
        T, r = self._T, self.reconstruction
        if not n_done:
            return
        en = np.asarray(self.experimentalData.energyAtPos, dtype=np.float32) + np.float32(1e-20)
        # the kernels store e_k / (E_k + 1e-20); undo the division in FP64
        e = T["err_at_pos"][:n_done].cpu().numpy() * en.astype(np.float64)[None, :]
        inv = float(np.sum(1.0 / en.astype(np.float64)))
        r.error = np.append(np.asarray(r.error, dtype=np.float64), e.sum(axis=1) * inv)
        last = e[-1].astype(np.float32)[:, None]
        r.errorAtPos = last / en
        r.detectorError = 0

    def _end(self, n_done):
        """Check the device flag, then errors and state back to NumPy."""
        plan = self._plan
        plan.check()
        self._collect(n_done)
        plan.close()
        self._plan = None

    # -- per-iteration constraints (BaseEngine.applyConstraints :1440-1612) -----
    def _has_constraints(self):
        p = self.params
        return bool(p.orthogonalizationSwitch or p.probePowerCorrectionSwitch or p.probeBoundary
                    or p.absorbingProbeBoundary or p.positionCorrectionSwitch
                    or (p.comStabilizationSwitch is not None and p.comStabilizationSwitch is not False))

    def applyConstraints(self, loop):
        """In the reference's order: orthogonalization, probe power
        correction, COM stabilisation, probe boundary, absorbing probe
        boundary (BaseEngine.py:1456-1506), position correction update
        (:1590-1591)."""
        p = self.params
        if p.orthogonalizationSwitch and loop % p.orthogonalizationFrequency == 0:
            self.orthogonalization()
        if p.probePowerCorrectionSwitch:
            self._probe_power_correction(self._T)
        if p.comStabilizationSwitch is not None and p.comStabilizationSwitch is not False:
            if loop % int(p.comStabilizationSwitch) == 0:
                self.comStabilization()
        if p.probeBoundary:
            self._window_probe(1.0)
        if p.absorbingProbeBoundary:
            self._window_probe(float(p.absorbingProbeBoundaryAleph))
        if p.positionCorrectionSwitch:
            self._position_correction_update(loop)

    def _position_correction_update(self, loop):
        """positionCorrectionUpdate (BaseEngine.py:1400-1438, CPM): once two
        errors exist, encoder_corrected -= adaptStep D dxo (adaptStep = 1),
        recentred on the measured encoder's mean; the new pixel positions go
        to the device for the next launch."""
        import torch

        r, T = self.reconstruction, self._T
        if self._n_err0 + loop + 1 <= 1:  # len(error) > startAtIteration (= 1)
            return
        D = T["pc_direction"].cpu().numpy() if "pc_direction" in T else np.zeros((self.experimentalData.numFrames, 2))
        new = r.encoder_corrected - 1 * D * r.dxo
        new = new - new.mean(axis=0, keepdims=True)
        new = new + self.experimentalData.encoder.mean(axis=0, keepdims=True)
        r.encoder_corrected = new
        pos = np.asarray(r.positions)
        if pos.min() < 0 or pos.max() + r.Np > r.No:
            raise ValueError("position correction moved a patch outside the object")
        T["positions"].copy_(torch.as_tensor(np.ascontiguousarray(pos, dtype=np.int32)))
        if self.params.positionOrder == "NA" and loop + 1 < T["order"].shape[0]:
            # setPositionOrder "NA" (BaseEngine.py:655-663) sorts by distance
            # from the centre of the positions at the start of every
            # iteration: the next iteration's order follows the moved positions
            nxt = self._position_order(self._n_err0 + loop + 1)
            self._orders[loop + 1] = nxt
            T["order"][loop + 1].copy_(torch.as_tensor(np.ascontiguousarray(nxt, dtype=np.int32)))

    def _probe_window(self):
        """_probeWindow (BaseEngine.py:237-266) in float64, cast to float32:
        a super-Gaussian of FWHM 3/4 of the probe window (order 10), replaced
        by a disc of 1.2 entrance-pupil diameters when probeBoundary is on."""
        r = self.reconstruction
        Xp, Yp = r.Xp, r.Yp
        w = np.exp(-(((Xp ** 2 + Yp ** 2) / (2 * (3 / 4 * r.Np * r.dxp / 2.355) ** 2)) ** 10))
        if self.params.probeBoundary:
            d = self.experimentalData.entrancePupilDiameter
            w = (Xp ** 2 + Yp ** 2) < ((d + d * 0.2) / 2) ** 2
        return np.ascontiguousarray(w, dtype=np.float32)

    def _probe_power_correction(self, T):
        """probe / sqrt(sum |probe|^2) * maxProbePower (BaseEngine.py:
        1462-1469) with the sum as an FP64 reduction on the GPU."""
        import torch

        probe = T["probe"]
        power = torch.empty(1, dtype=torch.float64, device=probe.device)
        lib, h = _native.load(), _native._stream_handle(torch.cuda.current_stream())
        n = probe.numel()
        _native.check(lib.pty_field_power(probe.data_ptr(), n, power.data_ptr(), h))
        _native.check(lib.pty_scale_to_power(probe.data_ptr(), n, power.data_ptr(),
                                             float(self.experimentalData.maxProbePower), h))

    def _window_probe(self, aleph):
        import torch

        T, r = self._T, self.reconstruction
        probe = T["probe"]
        nf = probe.numel() // (r.Np * r.Np)
        _native.check(_native.load().pty_window_fields(probe.data_ptr(), nf, int(r.Np), T["probeWindow"].data_ptr(),
                                                       aleph, _native._stream_handle(torch.cuda.current_stream())))

    def comStabilization(self):
        """Centre of mass of |P| of mode 0, last slice (BaseEngine.py:
        1724-1745, FP64 moments on the GPU); when it lies more than one pixel
        off centre, roll probe and object -- and the mPIE/multiPIE buffers and
        momenta -- back by the rounded offset (:1747-1776)."""
        import torch

        T, r = self._T, self.reconstruction
        lib, h = _native.load(), _native._stream_handle(torch.cuda.current_stream())
        mom = torch.empty(3, dtype=torch.float64, device=T["probe"].device)
        p0 = T["probe"][0, 0, 0, -1].contiguous()
        _native.check(lib.pty_field_moments(p0.data_ptr(), int(r.Np), T["Xp"].data_ptr(), T["Yp"].data_ptr(),
                                            mom.data_ptr(), h))
        sx, sy, s0 = (float(v) for v in mom.cpu().numpy())
        demon = s0 * r.dxp
        xc, yc = int(np.around(sx / demon)), int(np.around(sy / demon))
        if xc ** 2 + yc ** 2 <= 1:
            return 0, 0
        keys = ["probe", "object"]
        if self._momentum_enabled():
            keys += ["probeMomentum", "probeBuffer", "objectMomentum", "objectBuffer"]
        for k in keys:
            x = T[k]
            n = x.shape[-1]
            scratch = torch.empty_like(x)
            _native.check(lib.pty_roll_fields(x.data_ptr(), scratch.data_ptr(), x.numel() // (n * n), n, n,
                                              -yc, -xc, h))
        torch.cuda.current_stream().synchronize()
        return yc, xc

    def orthogonalization(self):
        """'snapShots' orthogonalisation of the probe (npsm > 1) or object
        modes (BaseEngine.py:1614-1722, utils.py:130-202): the Gram matrix is a
        complex64 GEMM on the GPU, its eigendecomposition runs on the host as
        in the reference, the mixing is a GEMM on the GPU; buffer and momentum
        follow the same mixing when momentum is on."""
        import torch

        r = self.reconstruction
        T = self._T
        if r.npsm > 1:
            key, n, npx, sel = "probe", r.npsm, r.Np, lambda X, il, isl: X[il, 0, :, isl]
            extra = ("probeBuffer", "probeMomentum")
        elif r.nosm > 1:
            key, n, npx, sel = "object", r.nosm, r.No, lambda X, il, isl: X[il, :, 0, isl]
            extra = ("objectBuffer", "objectMomentum")
        else:
            return
        X = T[key]
        for il in range(r.nlambda):
            for isl in range(r.nslice):
                p2 = sel(X, il, isl).reshape(n, npx * npx)
                # Gram matrix G[i, j] = sum_x conj(p_i) p_j as an elementwise
                # product and a reduction, accumulated in complex128 (a library
                # GEMM with K = Np^2 and an n x n output ran on one workgroup)
                q = p2.to(torch.complex128)
                G = (q.conj()[:, None, :] * q[None, :, :]).sum(dim=-1)
                w, V = np.linalg.eigh(G.cpu().numpy())
                order = np.argsort(w)[::-1]
                s = np.sqrt(np.clip(w[order].real, 0.0, None))
                ev = s ** 2 / np.sum(s ** 2)
                Vt = torch.as_tensor(np.ascontiguousarray(V[:, order].T, dtype=np.complex64), device=X.device)
                sel(X, il, isl).copy_(torch.matmul(Vt, p2).reshape(n, npx, npx))
                if key == "probe":
                    r.purityProbe = float(np.sqrt(np.sum(ev ** 2)))
                    r.purityProbeHist.append(r.purityProbe)
                else:
                    r.purityObject = float(np.sqrt(np.sum(ev ** 2)))
                if self._momentum_enabled():
                    for k in extra:
                        q = sel(T[k], il, isl).reshape(n, npx * npx)
                        sel(T[k], il, isl).copy_(torch.matmul(Vt, q).reshape(n, npx, npx))

    # -- public API ---------------------------------------------------------------
    def reconstruct(self, experimentalData=None, reconstruction=None):
        if experimentalData is not None:
            self.experimentalData = experimentalData
            self.reconstruction.data = experimentalData
        if reconstruction is not None:
            self.reconstruction = reconstruction
        self._prepareReconstruction()
        plan, _ = self._begin()
        n_iter = int(self.numIterations)
        done = 0
        try:
            orth = self._has_constraints()
            if not orth:
                # no host work between iterations: one launch per iteration,
                # no synchronisation until the end
                for loop in range(n_iter):
                    plan.run(loop, loop + 1)
                    done = loop + 1
            else:
                for loop in range(n_iter):
                    plan.run(loop, loop + 1)
                    done = loop + 1
                    self.applyConstraints(loop)
        finally:
            self._end(done)
