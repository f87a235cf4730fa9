"""Inputs of the reference's regression goldens, rebuilt with NumPy alone.

The reference synthesises its regression dataset in-test
(tests/regression/conftest.py:17-58): a 4x4 jittered raster, ND=32 low-passed
speckle patterns, seed 20240607, written to HDF5 with the ptychogram and the
encoder stored as float32 and the scalars read back as Python floats
(io/readHdf5.py:14-24).  This module regenerates exactly those arrays so both
the oracle and the GPU engines can be run on the inputs the goldens were
recorded with; the protocol (seed before init and before reconstruct,
sequential order) follows tests/regression/test_engine_regression.py:60-112.
"""

from pathlib import Path

import numpy as np

ND = 32
GRID = 4
SEED = 20240607
REFERENCE_GOLDEN_DIR = Path("/root/reference/tests/regression/data")

# name -> engine, propagator, (nlambda, nosm, npsm, nslice), iterations
# (test_engine_regression.py:47-57, in-scope rows only)
ENGINE_CASES = {
    "epie_fraunhofer": ("ePIE", "Fraunhofer", (1, 1, 1, 1), 3),
    "epie_asp": ("ePIE", "ASP", (1, 1, 1, 1), 3),
    "epie_fresnel": ("ePIE", "Fresnel", (1, 1, 1, 1), 3),
    "epie_polychrome": ("ePIE", "polychromeASP", (3, 1, 1, 1), 2),
    "epie_mixed_state": ("ePIE", "Fraunhofer", (1, 2, 3, 1), 3),
    "mpie_single": ("mPIE", "Fraunhofer", (1, 1, 1, 1), 3),
    "mpie_mixed_state": ("mPIE", "Fraunhofer", (1, 2, 3, 1), 3),
    "qnewton_single": ("qNewton", "Fraunhofer", (1, 1, 1, 1), 3),
    "e3pie_multislice": ("e3PIE", "Fraunhofer", (1, 1, 1, 3), 2),
}


def spectral_density(wavelength, nlambda):
    """Wavelengths of a polychromatic case (test_engine_regression.py:75-77)."""
    return float(wavelength) * np.linspace(0.98, 1.02, nlambda)


def regression_dataset(nd=ND, grid=GRID, seed=SEED):
    """Return dict(ptychogram, encoder, dxd, zo, wavelength, entrancePupilDiameter)."""
    rng = np.random.default_rng(seed)
    n_frames = grid * grid
    step = 3e-6
    coords = (np.arange(grid) - (grid - 1) / 2) * step
    yy, xx = np.meshgrid(coords, coords, indexing="ij")
    encoder = np.stack([yy.ravel(), xx.ravel()], axis=1)
    encoder = encoder + rng.normal(0, step / 50, encoder.shape)
    speckle = rng.random((n_frames, nd, nd))
    spectrum = np.fft.fftshift(np.fft.fft2(speckle), axes=(-2, -1))
    ky = np.fft.fftshift(np.fft.fftfreq(nd))
    mask = (ky[:, None] ** 2 + ky[None, :] ** 2) < 0.25 ** 2
    ptychogram = np.abs(np.fft.ifft2(np.fft.ifftshift(spectrum * mask, axes=(-2, -1)))) ** 2
    ptychogram = (ptychogram / ptychogram.max()).astype(np.float32)
    return {
        "ptychogram": ptychogram,
        "encoder": encoder.astype(np.float32),
        "dxd": 75e-6,
        "zo": 0.05,
        "wavelength": 632.8e-9,
        "entrancePupilDiameter": 400e-6,
    }
