"""Synthetic CPM datasets for the benchmark and test workloads.

Each generator returns the fields an ExperimentalData needs (ptychogram,
encoder, wavelength, dxd, zo, entrancePupilDiameter) plus the ground truth
(object, probe, pixel positions).  Objects are smoothed random complex
fields, scans are jittered Fermat spirals (utils.GenerateNonUniformFermat,
the reference's scanGrids.py:7-23), intensities are Poisson samples of the
forward model scaled to a photon peak.  The forward models are the ones the
engines invert: the centred unitary FFT (Fraunhofer), the band-limited
angular spectrum (ASP) and, for multislice, the slice chain with an ASP step
between slices.  Everything here runs once on the host with NumPy.
"""

from dataclasses import dataclass

import numpy as np
from scipy.ndimage import gaussian_filter

from .Operators import asp_tf_fftorder, aspw_transfer_function
from .utils import GenerateNonUniformFermat, circ

WAVELENGTH = 632.8e-9


@dataclass
class Workload:
    ptychogram: np.ndarray
    encoder: np.ndarray
    wavelength: float
    dxd: float
    zo: float
    entrancePupilDiameter: float
    object_true: np.ndarray
    probe_true: np.ndarray
    positions: np.ndarray
    No: int
    Np: int

    def as_dict(self):
        return dict(ptychogram=self.ptychogram, encoder=self.encoder, wavelength=self.wavelength, dxd=self.dxd,
                    zo=self.zo, entrancePupilDiameter=self.entrancePupilDiameter)


def fft2c(x):
    ax = (-2, -1)
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=ax), norm="ortho"), axes=ax)


def smooth_random_object(No, rng, sigma=3.0, amp=(0.5, 1.0), phase=(-np.pi, np.pi)):
    """amplitude U[amp] and phase U[phase], each Gaussian-smoothed (sigma px)."""
    a = gaussian_filter(rng.uniform(*amp, size=(No, No)), sigma, mode="reflect")
    p = gaussian_filter(rng.uniform(*phase, size=(No, No)), sigma, mode="reflect")
    return (a * np.exp(1j * p)).astype(np.complex64)


def fermat_scan(n, Np, No, step_px, rng, jitter=1):
    """Pixel corners of a Fermat spiral whose median nearest-neighbour
    distance is ~step_px, with U{-jitter..jitter} integer jitter, centred in
    the object; the radius is found by bisection."""
    def med_nn(R):
        r, c = GenerateNonUniformFermat(n, R, 1)
        pts = np.stack([r, c], 1)
        d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
        np.fill_diagonal(d, np.inf)
        return np.median(d.min(1))

    lo, hi = 1.0, float(No)
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if med_nn(mid) < step_px:
            lo = mid
        else:
            hi = mid
    R = 0.5 * (lo + hi)
    rows, cols = GenerateNonUniformFermat(n, R, 1)
    px = np.round(np.stack([rows, cols], 1)).astype(np.int64)
    if jitter:
        px = px + rng.integers(-jitter, jitter + 1, size=px.shape)
    lim = (No - Np) // 2
    if np.abs(px).max() > lim:
        raise ValueError(f"scan radius {np.abs(px).max()} px does not fit an object of {No} px for Np={Np}")
    return px


def _encoder_from_pixels(px, dxo):
    """Encoder in metres such that round(encoder/dxo) reproduces px."""
    enc = px.astype(np.float64) * dxo
    assert np.array_equal(np.round(enc / dxo).astype(np.int64), px)
    return enc


def aperture_probe(Np, radius_px, dxp, wavelength, curvature=1.0):
    """circ(r < radius) x exp(i pi curvature (r/radius)^2), propagated by one
    Fresnel length a^2/lambda with the band-limited ASP (complex64)."""
    x = (np.arange(Np) - Np / 2) * dxp
    X, Y = np.meshgrid(x, x)
    a = radius_px * dxp
    R2 = X ** 2 + Y ** 2
    p0 = circ(X, Y, 2 * a) * np.exp(1j * np.pi * curvature * R2 / a ** 2)
    zF = a ** 2 / wavelength
    H = np.fft.ifftshift(aspw_transfer_function(zF, wavelength, Np, Np * dxp))
    p = np.fft.ifft2(np.fft.fft2(p0) * H)
    return p.astype(np.complex64)


def hermite_gauss_modes(Np, n_modes, waist_px, radius_px, dxp, wavelength):
    """Mode k: Hermite-Gaussian (m, n) in the order (0,0), (1,0), (0,1),
    (1,1), (2,0), ... times the aperture probe; then orthogonalised through
    the Gram matrix (snapshot method) with decreasing powers."""
    base = aperture_probe(Np, radius_px, dxp, wavelength)
    u = (np.arange(Np) - Np / 2) / waist_px
    Ux, Uy = np.meshgrid(u, u)
    orders = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (1, 2)]
    herm = [lambda t: np.ones_like(t), lambda t: 2 * t, lambda t: 4 * t ** 2 - 2]
    modes = []
    for k in range(n_modes):
        m, n = orders[k]
        hg = herm[m](Ux) * herm[n](Uy) * np.exp(-(Ux ** 2 + Uy ** 2) / 2)
        modes.append(base * hg)
    p = np.array(modes, dtype=np.complex128).reshape(n_modes, -1)
    G = p.conj() @ p.T
    w, V = np.linalg.eigh(G)
    order = np.argsort(w)[::-1]
    q = (V[:, order].T @ p)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    weights = 0.6 ** np.arange(n_modes)
    q *= np.sqrt(weights / weights.sum())[:, None] * np.linalg.norm(base)
    return q.reshape(n_modes, Np, Np).astype(np.complex64)


def _poisson(I, peak, rng):
    I = I * (peak / I.max())
    return rng.poisson(I).astype(np.float32)


def cpm_fraunhofer(Np=64, No=256, F=100, seed=0, sigma=3.0, probe_radius_px=None, step_frac=0.25, peak=1e4,
                   dxo=4e-6, zo=0.05, wavelength=WAVELENGTH):
    """Config "plain ePIE" family: smoothed random object, aperture probe with
    quadratic phase propagated one Fresnel length, jittered Fermat spiral,
    I = |fft2c(P O)|^2 scaled to `peak` photons, Poisson noise."""
    rng = np.random.default_rng(seed)
    radius = probe_radius_px if probe_radius_px is not None else Np / 4
    obj = smooth_random_object(No, np.random.default_rng(seed), sigma=sigma)
    probe = aperture_probe(Np, radius, dxo, wavelength)
    px = fermat_scan(F, Np, No, step_frac * Np, np.random.default_rng(seed + 1))
    corners = px + No // 2 - Np // 2
    I = np.empty((F, Np, Np), dtype=np.float64)
    for f in range(F):
        r, c = corners[f]
        I[f] = np.abs(fft2c(probe * obj[r:r + Np, c:c + Np])) ** 2
    ptycho = _poisson(I, peak, np.random.default_rng(seed + 2))
    dxd = wavelength * zo / (Np * dxo)
    return Workload(ptycho, _encoder_from_pixels(px, dxo), wavelength, dxd, zo, 2 * radius * dxo,
                    obj[None, None, None, None], probe[None, None, None, None], corners, No, Np)


def cpm_asp_modes(Np=256, No=1024, F=1000, npsm=4, seed=10, peak=1e5, dxo=4e-6, zo=0.05, wavelength=WAVELENGTH,
                  probe_radius_px=64, sigma=3.0, step_px=None):
    """Config "mPIE, ASP, mixed-state probe": Hermite-Gaussian x aperture
    modes (orthogonalised), near-field ASP at zo, incoherent mode sum."""
    obj = smooth_random_object(No, np.random.default_rng(seed), sigma=sigma)
    probe = hermite_gauss_modes(Np, npsm, probe_radius_px / 2, probe_radius_px, dxo, wavelength)
    lim = (No - Np) // 2 - 1
    if step_px is None:
        # the densest spiral that fits: radius ~ lim for F points
        step_px = 0.9 * lim * np.sqrt(np.pi / F)
    px = fermat_scan(F, Np, No, step_px, np.random.default_rng(seed + 1))
    corners = px + No // 2 - Np // 2
    H = asp_tf_fftorder(zo, wavelength, Np, Np * dxo)
    I = np.empty((F, Np, Np), dtype=np.float32)
    for f in range(F):
        r, c = corners[f]
        ex = probe * obj[r:r + Np, c:c + Np][None]
        I[f] = np.sum(np.abs(np.fft.ifft2(np.fft.fft2(ex) * H)) ** 2, axis=0)
    ptycho = _poisson(I.astype(np.float64), peak, np.random.default_rng(seed + 2))
    dxd = wavelength * zo / (Np * dxo)
    return Workload(ptycho, _encoder_from_pixels(px, dxo), wavelength, dxd, zo, 2 * probe_radius_px * dxo,
                    obj[None, None, None, None], probe[None, None, :, None], corners, No, Np)


def cpm_multislice(Np=128, No=512, F=400, nslice=4, dz=5e-6, seed=20, peak=1e4, dxo=1e-6, zo=0.05,
                   wavelength=WAVELENGTH, phase_max=0.3, sigma=2.0):
    """Config "multislice": nslice weak-phase slices exp(i G_sigma(U[0, phase_max])),
    ASP over dz between slices, Fraunhofer to the detector."""
    rng = np.random.default_rng(seed)
    slices = np.stack([np.exp(1j * gaussian_filter(rng.uniform(0, phase_max, (No, No)), sigma, mode="reflect"))
                       for _ in range(nslice)]).astype(np.complex64)
    probe0 = aperture_probe(Np, Np / 4, dxo, wavelength)
    px = fermat_scan(F, Np, No, 0.25 * Np * 0.5, np.random.default_rng(seed + 1))
    corners = px + No // 2 - Np // 2
    H = np.fft.ifftshift(aspw_transfer_function(dz, wavelength, Np, Np * dxo)).astype(np.complex64)
    I = np.empty((F, Np, Np), dtype=np.float64)
    for f in range(F):
        r, c = corners[f]
        w = probe0 * slices[0, r:r + Np, c:c + Np]
        for s in range(1, nslice):
            w = np.fft.ifft2(np.fft.fft2(w) * H) * slices[s, r:r + Np, c:c + Np]
        I[f] = np.abs(fft2c(w)) ** 2
    ptycho = _poisson(I, peak, np.random.default_rng(seed + 2))
    dxd = wavelength * zo / (Np * dxo)
    return Workload(ptycho, _encoder_from_pixels(px, dxo), wavelength, dxd, zo, Np / 2 * dxo,
                    slices[None, None, None], probe0[None, None, None, None], corners, No, Np)


def batched_instance(i, seed0=1000, nps=(64, 128, 256), fmin=100, fmax=800, inputs="asp_modes"):
    """Instance i of the batched configuration: Np uniform in `nps`,
    No = 4 Np, F uniform in [fmin, fmax], all drawn from seed seed0 + i.
    inputs="asp_modes": generated as the single mPIE ASP workload
    (cpm_asp_modes: 4 Hermite-Gaussian x aperture modes of radius Np/4,
    near-field ASP at 50 mm, 1e5 photons); inputs="fraunhofer": the plain-ePIE
    family (cpm_fraunhofer, one mode)."""
    rng = np.random.default_rng(seed0 + i)
    Np = int(rng.choice(nps))
    F = int(rng.integers(fmin, fmax + 1))
    No = 4 * Np
    lim = (No - Np) // 2 - 1
    step = min(0.25 * Np, 0.9 * lim * np.sqrt(np.pi / F))
    if inputs == "asp_modes":
        return cpm_asp_modes(Np=Np, No=No, F=F, npsm=4, seed=seed0 + i, probe_radius_px=Np // 4, step_px=step)
    if inputs == "fraunhofer":
        return cpm_fraunhofer(Np=Np, No=No, F=F, seed=seed0 + i, step_frac=step / Np)
    raise ValueError(f"unknown batched inputs {inputs!r}")


def batched_instances(n=64, seed0=1000, nps=(64, 128, 256), fmin=100, fmax=800, workers=1, inputs="asp_modes",
                      progress=None):
    """Config "batched": n instances (batched_instance(i), i < n).  With
    workers > 1 they are generated by a pool of forked processes; call it
    before the process initialises the GPU.  ``progress(i)``: called after
    instance i when they are generated in this process."""
    args = [(i, seed0, nps, fmin, fmax, inputs) for i in range(n)]
    if workers <= 1:
        out = []
        for a in args:
            out.append(batched_instance(*a))
            if progress is not None:
                progress(a[0])
        return out
    import multiprocessing as mp

    with mp.get_context("fork").Pool(workers) as pool:
        return pool.starmap(batched_instance, args, chunksize=1)
