"""Measured data container, mirroring ExperimentalData/ExperimentalData.py.

Only the CPM fields the hot path reads are kept.  Data is set in memory (the
reference's manual path: set fields, then ``_setData()``,
example_scripts/example_mulit_slice.py:234-249) or loaded from an HDF5 file
with the reference's root-dataset names when h5py is importable
(ExperimentalData.py:86-149, io/readHdf5.py:27-89)."""

import logging
from pathlib import Path

import numpy as np


def _scalify(v):
    """hdf5 scalars come back as 0-d / 1x1 arrays (io/readHdf5.py:14-24)."""
    v = np.asarray(v).squeeze()
    try:
        return v.item()
    except ValueError:
        return v


class ExperimentalData:
    requiredFieldsCPM = ["ptychogram", "wavelength", "encoder", "dxd", "zo"]
    optionalFieldsCPM = ["entrancePupilDiameter", "spectralDensity", "theta", "emptyBeam"]

    def __init__(self, filename=None, operationMode="CPM"):
        self.logger = logging.getLogger("ExperimentalData")
        if operationMode != "CPM":
            raise NotImplementedError("only operationMode='CPM' is implemented (FPM is not)")
        self.operationMode = operationMode
        self.ptychogram = None
        self.encoder = None
        self.wavelength = None
        self.dxd = None
        self.zo = None
        self.entrancePupilDiameter = None
        self.spectralDensity = None
        self.theta = None
        self.emptyBeam = None
        if filename is not None:
            self.loadData(filename)

    def loadData(self, filename):
        """Load the reference's HDF5 layout (root datasets ptychogram, encoder,
        dxd, zo, wavelength[, entrancePupilDiameter, ..., orientation]) or an
        .npz with the same keys; "test:nodata" sets the reference's minimal
        stub (ExperimentalData.py:101-116)."""
        if str(filename) == "test:nodata":
            self.filename = filename
            self.ptychogram = np.zeros((1, 16, 16), dtype=np.float32)
            self.encoder = np.zeros((1, 2), dtype=np.float64)
            self.wavelength, self.dxd, self.zo = 500e-9, 6.5e-6, 0.1
            for k in self.optionalFieldsCPM:
                setattr(self, k, None)
            self._setData()
            return
        path = Path(filename)
        self.filename = filename
        if not path.exists():
            raise FileNotFoundError(f"Could not find file {path}.")
        if path.suffix == ".npz":
            with np.load(path, allow_pickle=False) as z:
                src = {k: z[k] for k in z.files}
        else:
            try:
                import h5py
            except ImportError as e:
                raise ImportError("h5py is required to read HDF5 datasets") from e
            with h5py.File(path, "r") as hf:
                src = {k: np.array(hf[k]) for k in hf.keys()}
        for k in self.requiredFieldsCPM:
            if k not in src:
                raise KeyError("hdf5 file misses key %s" % k)
            setattr(self, k, _scalify(src[k]))
        for k in self.optionalFieldsCPM:
            setattr(self, k, _scalify(src[k]) if k in src else None)
        enc = np.asarray(self.encoder)
        if enc.shape[0] < enc.shape[1]:
            # io/readHdf5.py:68-75: (2, F) encoders are transposed and centred
            self.encoder = enc.T - enc.T.mean(axis=0, keepdims=True)
        self._setData()
        # last, as in ExperimentalData.py:147-149: it may change the frames
        orientation = None
        if "orientation" in src:
            orientation = _scalify(np.asarray(src["orientation"]).ravel()[0])
        self.setOrientation(orientation)

    def setOrientation(self, orientation, force_contiguous=True):
        """ptypy's orientation codes 0-7 (ExperimentalData.py:192-232): bit 2
        transposes the frames, then bit 0 flips columns and bit 1 rows."""
        if orientation is None:
            return
        if not isinstance(orientation, int):
            raise TypeError("Orientation value is not valid.")
        if orientation == 0:
            return
        if not 0 < orientation <= 7:
            raise ValueError(f"Orientation {orientation} is not implemented")
        if orientation & 4:
            self.ptychogram = np.transpose(self.ptychogram, (0, 2, 1))
        if orientation & 1:
            self.ptychogram = np.flip(self.ptychogram, axis=-1)
        if orientation & 2:
            self.ptychogram = np.flip(self.ptychogram, axis=-2)
        if force_contiguous:
            self.ptychogram = np.ascontiguousarray(self.ptychogram)

    def _setData(self):
        """Derived fields (ExperimentalData.py:234-250)."""
        self.ptychogram = np.asarray(self.ptychogram)
        self.Nd = self.ptychogram.shape[-1]
        self.xd = np.linspace(-self.Nd / 2, self.Nd / 2, int(self.Nd)) * self.dxd
        self.Xd, self.Yd = np.meshgrid(self.xd, self.xd)
        self.Ld = self.Nd * self.dxd
        self.numFrames = self.ptychogram.shape[0]
        self.energyAtPos = np.sum(abs(self.ptychogram), (-1, -2))
        self.maxProbePower = np.sqrt(np.max(np.sum(self.ptychogram, (-1, -2))))

    def relative_intensity(self, index):
        """ExperimentalData.py:274-288."""
        if not hasattr(self, "_relative_intensity"):
            ri = self.ptychogram.mean((-2, -1))
            self._relative_intensity = ri / (ri.mean() + 2 * ri.std())
        return self._relative_intensity[index]
