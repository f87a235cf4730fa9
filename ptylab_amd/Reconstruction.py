"""Reconstruction state, mirroring Reconstruction/Reconstruction.py.

Holds the 6D object (nlambda, nosm, 1, nslice, No, No) and probe
(nlambda, 1, npsm, nslice, Np, Np) in complex64, the geometry derived from
the experimental data, and the scan positions in pixels.  Initial guesses
draw from NumPy's global RNG in the reference's order (object, then probe),
so a seeded run starts from the same arrays as the reference
(Reconstruction.py:346-413, utils/initializationFunctions.py:10-62).
While an engine runs the fields live on the GPU as torch tensors; afterwards
they are NumPy arrays again, as after the reference's _move_data_to_cpu
(ePIE.py:123-126)."""

import logging

import numpy as np
from pathlib import Path

from .utils import circ


def calculate_pixel_positions(encoder_corrected, dxo, No, Np, asint):
    """round(encoder / dxo) + No//2 - Np//2 (Reconstruction.py:21-31)."""
    positions = np.round(encoder_corrected / dxo)
    positions = positions + No // 2 - Np // 2
    if asint:
        positions = positions.astype(int)
    return positions


def initialProbeOrObject(shape, type_of_init, data):
    """'ones' / 'circ' initial guesses with 1e-3 uniform noise from the global
    RNG (initializationFunctions.py:10-62)."""
    if isinstance(type_of_init, np.ndarray):
        return type_of_init
    supported = ["circ", "rand", "gaussian", "ones"]
    if type_of_init not in supported:
        raise NotImplementedError(f"Got {type_of_init} for shape. Supported shapes are: {supported}")
    if type_of_init == "ones":
        return np.ones(shape) + 0.001 * np.random.rand(*shape)
    if type_of_init == "circ":
        pupil = circ(data.Xp, data.Yp, data.data.entrancePupilDiameter)
        field = np.ones(shape, dtype=np.complex64) + 0.001 * np.random.rand(*shape)
        field *= pupil
        return field
    if type_of_init == "rand":
        return np.random.rand(*shape) * np.exp(1j * 2 * np.pi * np.random.rand(*shape))
    # gaussian
    sigma = data.data.entrancePupilDiameter / 2.355
    g = np.exp(-(data.Xp ** 2 + data.Yp ** 2) / (2 * sigma ** 2))
    return np.ones(shape) * g + 0.001 * np.random.rand(*shape)


class Reconstruction:
    listOfReconstructionPropertiesCPM = ["wavelength", "dxd", "theta", "spectralDensity", "entrancePupilDiameter"]

    def __init__(self, data, params):
        self.logger = logging.getLogger("Reconstruction")
        self.data = data
        self.params = params
        self.zMomentum = 0
        self.encoder_corrected = None
        for key in self.listOfReconstructionPropertiesCPM:
            setattr(self, key, getattr(data, key))
        self._zo = None
        self.zo = data.zo
        self.encoder_corrected = np.array(data.encoder, copy=True)
        self.computeParameters()
        self.initializeSettings()
        self.possible_GPU_fields = [
            "probe", "object", "probeBuffer", "objectBuffer", "probeMomentum", "objectMomentum", "H",
        ]

    # geometry (Reconstruction.py:145-204, 579-710) -------------------------
    @property
    def zo(self):
        return self._zo

    @zo.setter
    def zo(self, value):
        self._zo = value
        self.dxp = self.wavelength * self._zo / self.Ld

    def computeParameters(self):
        if self.data.entrancePupilDiameter is None:
            self.data.entrancePupilDiameter = self.Lp / 3
        if self.spectralDensity is None:
            self.spectralDensity = np.atleast_1d(self.wavelength)
        if not hasattr(self, "No"):
            self.No = self.Np * 2 ** 2
            rng_px = np.max(self.positions, axis=0) - np.min(self.positions, axis=0)
            rng_px = np.max(rng_px) + self.Np * 2
            if rng_px % 2 == 1:
                rng_px += 1
            self.No = np.max([self.Np, rng_px])

    def initializeSettings(self):
        self.nlambda = 1
        self.nosm = 1
        self.npsm = 1
        self.nslice = 1
        self.purityProbe = 1.0
        self.purityObject = 1.0
        self.purityProbeHist = []
        self.positions0 = self.positions.copy()
        self.initialProbe = "circ"
        self.initialObject = "ones"

    @property
    def Nd(self):
        return self.data.ptychogram.shape[1]

    @property
    def Ld(self):
        return self.Nd * self.dxd

    @property
    def Np(self):
        return self.Nd

    @property
    def Lp(self):
        return self.Np * self.dxp

    @property
    def xp(self):
        return np.linspace(-self.Np / 2, self.Np / 2, int(self.Np)) * self.dxp

    @property
    def Xp(self):
        return np.meshgrid(self.xp, self.xp)[0]

    @property
    def Yp(self):
        return np.meshgrid(self.xp, self.xp)[1]

    @property
    def dxo(self):
        return self.dxp

    @property
    def Lo(self):
        return self.No * self.dxo

    @property
    def positions(self):
        return calculate_pixel_positions(self.encoder_corrected, self.dxo, self.No, self.Np, asint=True)

    @property
    def shape_O(self):
        return (self.nlambda, self.nosm, 1, self.nslice, int(self.No), int(self.No))

    @property
    def shape_P(self):
        return (self.nlambda, 1, self.npsm, self.nslice, int(self.Np), int(self.Np))

    # initial guesses (Reconstruction.py:346-420) ----------------------------
    def initializeObjectProbe(self, force=True):
        self.initializeObject(force=force)
        self.initializeProbe(force=force)
        self.object = self.initialGuessObject.copy()
        self.probe = self.initialGuessProbe.copy()

    def initializeObject(self, type_of_init=None, force=True):
        if type_of_init is not None:
            self.initialObject = type_of_init
        self.initialGuessObject = initialProbeOrObject(self.shape_O, self.initialObject, self).astype(np.complex64)

    def initializeProbe(self, force=False):
        if self.data.entrancePupilDiameter is None:
            self.data.entrancePupilDiameter = self.Lp / 3
        self.initialGuessProbe = initialProbeOrObject(self.shape_P, self.initialProbe, self).astype(np.complex64)

    def initializeObjectMomentum(self):
        self.objectMomentum = np.zeros_like(self.initialGuessObject)

    def initializeProbeMomentum(self):
        self.probeMomentum = np.zeros_like(self.initialGuessProbe)

    # results on disk (Reconstruction.py:495-576) ------------------------------
    _SAVE_TYPES = ["all", "object", "probe"]

    def _results(self, type, squeeze):
        if type not in self._SAVE_TYPES:
            raise NotImplementedError(f"Only {self._SAVE_TYPES} are allowed keywords for type")
        sq = np.squeeze if squeeze else (lambda x: x)
        c64 = lambda x: np.asarray(x, dtype=np.complex64)  # noqa: E731
        f32 = lambda x: np.asarray(x, dtype=np.float32)  # noqa: E731
        if type == "probe":
            return {"probe": c64(sq(np.asarray(self.probe)))}
        if type == "object":
            return {"object": c64(sq(np.asarray(self.object)))}
        out = {
            "probe": c64(self.probe),
            "object": c64(self.object),
            "error": f32(getattr(self, "error", [])),
            "zo": f32(self._zo),
            "wavelength": f32(self.wavelength),
            "dxp": f32(self.dxp),
            "purityProbe": f32(self.purityProbe),
            "purityObject": f32(self.purityObject),
            "I object": f32(np.abs(np.asarray(self.object))),
            "I probe": f32(np.abs(np.asarray(self.probe))),
            "encoder_corrected": np.asarray(self.encoder_corrected),
        }
        if getattr(self, "theta", None) is not None:
            out["theta"] = f32(self.theta)
        return out

    def saveResults(self, fileName="recent", type="all", squeeze=False):
        """Save the reconstruction with the reference's dataset names and
        dtypes: an HDF5 file when h5py is importable and the name does not end
        in .npz, else an .npz archive (".npz" appended when missing).  Returns
        the path written."""
        data = self._results(type, squeeze)
        self._move_data_to_cpu()
        path = Path(fileName)
        h5 = path.suffix != ".npz"
        if h5:
            try:
                import h5py
            except ImportError:
                h5 = False
        if h5:
            with h5py.File(path, "w") as hf:
                for k, v in data.items():
                    hf.create_dataset(k, data=v)
            return path
        if path.suffix != ".npz":
            path = path.with_name(path.name + ".npz")
        np.savez(path, **data)
        return path

    def load(self, filename):
        """Load what saveResults wrote (Reconstruction.py:495-507)."""
        path = Path(filename)
        if path.suffix == ".npz":
            with np.load(path, allow_pickle=False) as z:
                src = {k: z[k] for k in z.files}
        else:
            import h5py

            with h5py.File(path, "r") as hf:
                src = {k: np.array(hf[k]) for k in hf.keys()}
        for k in ("probe", "object", "error", "wavelength", "dxp", "purityProbe", "purityObject"):
            setattr(self, k, np.array(src[k]))
        # the zo setter recomputes dxp from the loaded wavelength, as the reference's does
        self.zo = np.array(src["zo"])
        if "theta" in src:
            self.theta = np.array(src["theta"])

    # host/device movement -----------------------------------------------------
    def _move_data_to_cpu(self):
        """torch tensors -> NumPy for every field in possible_GPU_fields."""
        try:
            import torch
        except ImportError:  # pragma: no cover
            return
        for name in self.possible_GPU_fields:
            v = getattr(self, name, None)
            if isinstance(v, torch.Tensor):
                setattr(self, name, v.detach().cpu().numpy())
