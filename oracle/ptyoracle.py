"""NumPy restatement of the reference's per-position reconstruction loop.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg, never by
the product package.

Every function cites the reference file:line it restates (paths relative to
the PtyLab 0.3.1 source tree).  The restatement keeps the reference's NumPy
operations and dtype flow (NEP 50 promotion, complex64 state, float64 error
sums, complex128 slice transfer function in e3PIE) so that it reproduces the
reference's regression goldens to rtol 1e-5; ``precision="c128"`` promotes the
whole state to complex128 for the "error against a double-precision run"
figure.

Parity pins (what checks each part; the tests live in ``tests/``):

* ``fft2c``/``ifft2c``/Fraunhofer: the reference golden
  ``propagator_fraunhofer.npz``, the unitary round trip, a brute-force DFT.
* ``aspw_transfer``/ASP: golden ``propagator_asp.npz``; z=0 identity; the
  closed-form band limit; |H|=1 inside the band.
* Fresnel (``quad_phase``) and polychromeASP (``polychrome_transfer_stack``):
  goldens ``propagator_fresnel.npz``, ``propagator_polychromeasp.npz``,
  ``epie_fresnel.npz``, ``epie_polychrome.npz``.
* ePIE/mPIE engines (single and mixed state): goldens ``epie_fraunhofer``,
  ``epie_asp``, ``epie_mixed_state``, ``mpie_single``, ``mpie_mixed_state``.
* e3PIE slice chain: golden ``e3pie_multislice``.
* multiPIE: no golden exists; pinned by the identity multiPIE == mPIE when
  nlambda == nosm == 1 (the two probe updates then sum over the same axes) and
  by duplicated wavelengths / object modes / slices reducing to one channel
  (which mPIE's probe sum fails).
* ``fftshiftSwitch``: the shifted-data, unshifted-transform run equals the
  centred one to rounding.
* mqNewton ADAM / NADAM / momentum: the beta_m = 0 reduction to qNewton and
  each engine's closed-form first steps.
* position correction: pcPIE == mPIE without it; known rolls give the
  expected directions (direct and Fourier branches); the Fourier correlation
  equals a brute-force circular cross-correlation of the mean-free patches.
* detectorError / saveMemory: the per-pixel errors sum to the error curve; the
  saveMemory curve is the reference's broadcast sum.
* ``positionOrder='random'``: no golden; pinned by the RNG-consumption identity
  with ``np.random.shuffle`` (tested) -- otherwise "parity unpinned".
* ``orthogonalize_modes``: pinned by orthogonality, descending eigenvalues and
  span preservation (the reference's own invariants,
  tests/Engines/test_orthogonalization.py).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

__all__ = [
    "fft2c", "ifft2c", "circ", "aspw_transfer", "asp_transfer_stack",
    "quad_phase", "polychrome_transfer_stack", "scaled_asp_phases", "two_step_polychrome", "propagate", "pixel_positions", "Geometry", "initial_guess",
    "energy_at_pos", "position_order", "orthogonalize_modes", "OracleRun",
    "nrmse", "GEMMA",
]

GEMMA = 1e-10  # zero-division guard of intensityProjection, BaseEngine.py:877


# ---------------------------------------------------------------------------
# FFT primitives -- utils/utils.py:12-47
# ---------------------------------------------------------------------------
def fft2c(field, shifted=False):
    """Unitary centred 2D FFT over the last two axes (utils.py:12-28).

    ``shifted=True`` is the reference's ``fftshiftSwitch`` (no shifts)."""
    if shifted:
        return np.fft.fft2(field, norm="ortho")
    ax = (-2, -1)
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(field, axes=ax), norm="ortho"), axes=ax)


def ifft2c(field, shifted=False):
    """Unitary centred inverse 2D FFT (utils.py:31-47)."""
    if shifted:
        return np.fft.ifft2(field, norm="ortho")
    ax = (-2, -1)
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(field, axes=ax), norm="ortho"), axes=ax)


def circ(x, y, D):
    """Binary disc of diameter D on a meshgrid (utils.py:50-59)."""
    return (x ** 2 + y ** 2) < (D / 2) ** 2


# ---------------------------------------------------------------------------
# Angular spectrum transfer function -- Operators.py:587-703, 861-893
# ---------------------------------------------------------------------------
def aspw_transfer(z, wavelength, N, L, bandlimit=True):
    """Band-limited ASP transfer function, float64/complex128, centred
    frequency order (Operators.py:631-684, complexexp :687-703)."""
    z = float(z)
    wavelength = float(wavelength)
    L = float(L)
    N = int(N)
    a_z = abs(z)
    k = 2 * np.pi / wavelength
    f = np.arange(-N / 2, N / 2) / L
    Fx, Fy = np.meshgrid(f, f)
    f_max = L / (wavelength * np.sqrt(L ** 2 + 4 * a_z ** 2))
    W = np.array(circ(Fx, Fy, 2 * f_max))
    arg = 1 - (Fx * wavelength) ** 2 - (Fy * wavelength) ** 2
    keep = arg > 0
    if not bandlimit:
        keep = 0 * keep + 1
    arg = np.clip(arg, 0, np.inf)
    phase = k * a_z * np.sqrt(arg)
    H = np.array(keep * (np.cos(phase) + 1j * np.sin(phase)))
    if z < 0:
        H = H.conj()
    return H * W


def asp_transfer_stack(nosm, npsm, Np, z, wavelength, Lp, nlambda=1):
    """The (nlambda, nosm, npsm, 1, Np, Np) complex64 stack the ASP
    propagator multiplies with (Operators.py:861-893)."""
    if nlambda > 1:
        raise ValueError("For multi-wavelength, polychromeASP needs to be used instead of ASP")
    one = aspw_transfer(z, wavelength, Np, Lp)
    stack = np.empty((nlambda, nosm, npsm, 1, Np, Np), dtype=np.complex64)
    stack[...] = one.astype(np.complex64)
    return stack


def quad_phase(z, wavelength, Np, dxp):
    """exp(i pi / (lambda z) (x^2 + y^2)) on linspace(-Np/2, Np/2, Np) * dxp,
    complex128 (_propagation_kernels.py:16-36)."""
    x = np.linspace(-Np / 2, Np / 2, int(Np)) * dxp
    X, Y = np.meshgrid(x, x)
    return np.exp(1.0j * np.pi / (wavelength * z) * (X ** 2 + Y ** 2))


def scaled_asp_phases(z, wavelength, N, dx, dq):
    """Q1 (source plane) and Q2 (centred frequency grid) of the scaled
    angular spectrum method with band limits, float64/complex128
    (Operators.py:706-760, scaledASP with bandlimit=True, exactSolution=False).
    dx: source sampling (dxo), dq: destination sampling (dxd)."""
    k = 2 * np.pi / wavelength
    N = int(N)
    x1 = np.arange(-N // 2, N // 2) * dx
    X1, Y1 = np.meshgrid(x1, x1)
    r1sq = X1 ** 2 + Y1 ** 2
    f = np.arange(-N // 2, N // 2) / (N * dx)
    FX, FY = np.meshgrid(f, f)
    fsq = FX ** 2 + FY ** 2
    m = dq / dx
    Q1 = np.exp(1.0j * k / 2 * (1 - m) / z * r1sq)
    Q2 = np.exp(-1.0j * np.pi ** 2 * 2 * z / m / k * fsq)
    if m != 1:
        r1sq_max = wavelength * z / (2 * dx * (1 - m))
        Q1 = Q1 * np.array(circ(X1, Y1, 2 * r1sq_max))
    fsq_max = m / (2 * z * wavelength * (1 / (N * dx)))
    Q2 = Q2 * np.array(circ(FX, FY, 2 * fsq_max))
    return Q1, Q2


def two_step_polychrome(geo):
    """Transfer functions and quadratic phase of the two-step polychromatic
    propagator (Operators.py:1033-1079), complex128 as built: wavelength l
    first propagates z0 (1 - lambda_0 / lambda_l) by ASP (centred grid), then
    the Fresnel quadratic phase of lambda_0 and a far-field transform."""
    sd = np.atleast_1d(geo.spectralDensity)
    nl = len(sd)
    T = np.empty((nl, geo.nosm, geo.npsm, 1, int(geo.Np), int(geo.Np)), dtype=np.complex128)
    for il in range(nl):
        T[il] = aspw_transfer(geo.zo * (1 - sd[0] / sd[il]), sd[il], geo.Np, geo.Lp)
    Q = quad_phase(geo.zo, sd[0], geo.Np, geo.dxp)
    return T, Q


def polychrome_transfer_stack(nosm, npsm, Np, z, spectral_density, Lp):
    """The (nlambda, nosm, npsm, 1, Np, Np) stack of per-wavelength ASP
    transfer functions, centred, complex128 as built (Operators.py:910-952)."""
    nl = len(spectral_density)
    stack = np.empty((nl, nosm, npsm, 1, Np, Np), dtype=np.complex128)
    for il in range(nl):
        stack[il] = aspw_transfer(z, spectral_density[il], Np, Lp)
    return stack


def propagate(fields, kind, geo, inverse=False, shifted=False, tf_cache=None, pure=False):
    """object2detector / detector2object (Operators.py:561-584, dispatch
    :1082-1103) for Fraunhofer (:27-74), Fresnel (:77-152), ASP (:155-230) and
    polychromeASP (:458-558), scaledASP (:313-370), scaledPolychromeASP (:378-455) and
    twoStepPolychrome (:233-310).  pure=True keeps complex64 fields complex64
    (the quadratic phase cast to complex64, as a complex64 GPU path does);
    the default follows NumPy's promotion with the reference's complex128
    quadratic phase."""
    kind = kind.lower()
    if kind == "fraunhofer":
        return ifft2c(fields, shifted) if inverse else fft2c(fields, shifted)
    if kind == "fresnel":
        Q = quad_phase(geo.zo, geo.wavelength, geo.Np, geo.dxp)
        if pure and fields.dtype == np.complex64:
            Q = Q.astype(np.complex64)
        if inverse:
            return ifft2c(fields, shifted) * Q.conj()
        return fft2c(fields * Q, shifted)
    if kind == "asp":
        if shifted:
            raise ValueError("ASP propagator only works with fftshiftswitch == False")
        key = (geo.nosm, geo.npsm, geo.Np, geo.zo, geo.wavelength, geo.Lp, geo.nlambda)
        tf = None if tf_cache is None else tf_cache.get(key)
        if tf is None:
            tf = asp_transfer_stack(geo.nosm, geo.npsm, geo.Np, geo.zo, geo.wavelength, geo.Lp, geo.nlambda)
            tf = np.fft.ifftshift(tf, axes=(-2, -1))
            if tf_cache is not None:
                tf_cache[key] = tf
        if inverse:
            tf = tf.conj()
        if fields.dtype == np.complex128:
            tf = tf.astype(np.complex128)
        return ifft2c(fft2c(fields, True) * tf, True)
    if kind == "polychromeasp":
        if shifted:
            raise ValueError("ASP propagatorType works only with fftshiftSwitch = False!")
        key = ("poly", geo.nosm, geo.npsm, geo.Np, geo.zo, tuple(geo.spectralDensity), geo.Lp)
        tf = None if tf_cache is None else tf_cache.get(key)
        if tf is None:
            # built in complex128 and returned as is on the CPU path (:938-952)
            tf = polychrome_transfer_stack(geo.nosm, geo.npsm, geo.Np, geo.zo, geo.spectralDensity, geo.Lp)
            if tf_cache is not None:
                tf_cache[key] = tf
        if pure and fields.dtype == np.complex64:
            tf = tf.astype(np.complex64)
        if inverse:
            tf = tf.conj()
        return ifft2c(fft2c(fields) * tf)
    if kind == "scaledasp":
        if shifted:
            raise ValueError("scaledASP propagatorType works only with fftshiftSwitch = False!")
        if geo.nlambda > 1:
            raise ValueError("For multi-wavelength, scaledPolychromeASP needs to be used instead of scaledASP")
        # stored as complex64 (Operators.py:955-987), so the chain stays in
        # the fields' precision
        Q1, Q2 = scaled_asp_phases(geo.zo, geo.wavelength, geo.Np, geo.dxo, geo.dxd)
        Q1, Q2 = Q1.astype(np.complex64), Q2.astype(np.complex64)
        if inverse:
            return ifft2c(fft2c(fields) * Q2.conj()) * Q1.conj()
        return ifft2c(fft2c(fields * Q1) * Q2)
    if kind == "scaledpolychromeasp":
        if shifted:
            raise ValueError("scaledPolychromeASP propagatorType works only with fftshiftSwitch = False!")
        # per wavelength of the spectral density, each pair as scaledASP
        # builds it, stored complex64 (Operators.py:990-1032)
        sd = np.atleast_1d(geo.spectralDensity)
        shape = (len(sd), 1, 1, 1, int(geo.Np), int(geo.Np))
        Q1 = np.empty(shape, np.complex64)
        Q2 = np.empty(shape, np.complex64)
        for il, wl in enumerate(sd):
            q1, q2 = scaled_asp_phases(geo.zo, float(wl), geo.Np, geo.dxo, geo.dxd)
            Q1[il, 0, 0, 0], Q2[il, 0, 0, 0] = q1, q2
        if inverse:
            return ifft2c(fft2c(fields) * Q2.conj()) * Q1.conj()
        return ifft2c(fft2c(fields * Q1) * Q2)
    if kind == "twosteppolychrome":
        if shifted:
            raise ValueError("twoStepPolychrome propagatorType works only with fftshiftSwitch = False!")
        T, Q = two_step_polychrome(geo)
        if pure and fields.dtype == np.complex64:
            T, Q = T.astype(np.complex64), Q.astype(np.complex64)
        if inverse:
            return ifft2c(fft2c(fields * Q.conj()) * T.conj())
        return fft2c(ifft2c(fft2c(fields) * T) * Q)
    raise NotImplementedError(f"propagator {kind!r} is outside the oracle's scope")


# ---------------------------------------------------------------------------
# Geometry -- Reconstruction.py:21-31, 145-204, 579-710
# ---------------------------------------------------------------------------
def pixel_positions(encoder, dxo, No, Np):
    """round(encoder/dxo) + No//2 - Np//2 as int (Reconstruction.py:21-31)."""
    positions = np.round(encoder / dxo)
    positions = positions + No // 2 - Np // 2
    return positions.astype(int)


@dataclass
class Geometry:
    """Derived CPM geometry (Reconstruction.py:145-204, 579-710)."""
    ptychogram: np.ndarray
    encoder: np.ndarray
    wavelength: float
    dxd: float
    zo: float
    entrancePupilDiameter: Optional[float] = None
    No: Optional[int] = None
    nlambda: int = 1
    nosm: int = 1
    npsm: int = 1
    nslice: int = 1
    spectralDensity: Optional[np.ndarray] = None

    def __post_init__(self):
        if self.spectralDensity is None:
            # Reconstruction.py:166-170
            self.spectralDensity = np.atleast_1d(self.wavelength)
        self.Nd = self.ptychogram.shape[-1]
        self.Ld = self.Nd * self.dxd
        self.dxp = self.wavelength * self.zo / self.Ld
        self.Np = self.Nd
        self.Lp = self.Np * self.dxp
        self.dxo = self.dxp
        if self.entrancePupilDiameter is None:
            self.entrancePupilDiameter = self.Lp / 3
        if self.No is None:
            # Reconstruction.py:193-204
            provisional = pixel_positions(self.encoder, self.dxo, self.Np * 2 ** 2, self.Np)
            rng_px = np.max(provisional, axis=0) - np.min(provisional, axis=0)
            rng_px = np.max(rng_px) + self.Np * 2
            if rng_px % 2 == 1:
                rng_px += 1
            self.No = np.max([self.Np, rng_px])
        self.positions = pixel_positions(self.encoder, self.dxo, self.No, self.Np)
        xp = np.linspace(-self.Np / 2, self.Np / 2, int(self.Np)) * self.dxp
        self.Xp, self.Yp = np.meshgrid(xp, xp)

    @property
    def shape_O(self):
        return (self.nlambda, self.nosm, 1, self.nslice, int(self.No), int(self.No))

    @property
    def shape_P(self):
        return (self.nlambda, 1, self.npsm, self.nslice, int(self.Np), int(self.Np))


def initial_guess(shape, kind, geo):
    """'ones' / 'circ' initial guesses with 1e-3 uniform noise drawn from the
    global NumPy RNG (initializationFunctions.py:10-62; cast to complex64 as in
    Reconstruction.py:356-413)."""
    if kind == "ones":
        out = np.ones(shape) + 0.001 * np.random.rand(*shape)
    elif kind == "circ":
        pupil = circ(geo.Xp, geo.Yp, geo.entrancePupilDiameter)
        out = np.ones(shape, dtype=np.complex64) + 0.001 * np.random.rand(*shape)
        out *= pupil
    else:
        raise NotImplementedError(kind)
    return out.astype(np.complex64)


def energy_at_pos(ptychogram):
    """Per-frame energy (ExperimentalData.py:248)."""
    return np.sum(abs(ptychogram), (-1, -2))


def position_order(kind, n_frames, n_errors, positions=None):
    """setPositionOrder (BaseEngine.py:637-665); consumes the global RNG exactly
    as the reference (one ``np.random.shuffle`` per iteration once two errors
    exist)."""
    if kind == "sequential":
        return np.arange(n_frames)
    if kind == "random":
        idx = np.arange(n_frames)
        if n_errors >= 2:
            np.random.shuffle(idx)
        return idx
    if kind == "NA":
        rows = positions[:, 0] - np.mean(positions[:, 0])
        cols = positions[:, 1] - np.mean(positions[:, 1])
        return np.argsort(np.sqrt(rows ** 2 + cols ** 2))
    raise ValueError("position order not properly set")


def orthogonalize_modes(p):
    """'snapShots' mode orthogonalisation (utils.py:130-202): Gram matrix,
    host eigh, descending order, modes = V^T p."""
    p2 = p.reshape(p.shape[0], p.shape[1] * p.shape[2])
    G = np.dot(p2.conj(), np.transpose(p2))
    w, V = np.linalg.eigh(G)
    order = np.argsort(w)[::-1]
    s = np.sqrt(np.clip(w[order].real, 0.0, None))
    ev = s ** 2 / np.sum(s ** 2)
    V = np.ascontiguousarray(V[:, order].T)
    modes = np.dot(V, p2).reshape(p.shape)
    return modes, ev, V


def nrmse(rec, true):
    """min over complex gamma of ||gamma*rec - true|| / ||true|| (project metric)."""
    rec = np.asarray(rec, dtype=np.complex128).ravel()
    true = np.asarray(true, dtype=np.complex128).ravel()
    den = np.vdot(rec, rec)
    gamma = np.vdot(rec, true) / den if abs(den) > 0 else 0.0
    return float(np.linalg.norm(gamma * rec - true) / np.linalg.norm(true))


# ---------------------------------------------------------------------------
# Engines -- ePIE.py:53-160, mPIE.py:73-270, multiPIE.py:46-205, e3PIE.py:46-225,
# BaseEngine.py:819-1012
# ---------------------------------------------------------------------------
@dataclass
class OracleRun:
    """State and settings of one reconstruction, mirroring the reference's
    Reconstruction/Engine attributes that the hot path reads and writes."""
    geo: Geometry
    object: np.ndarray
    probe: np.ndarray
    engine: str = "ePIE"
    propagator: str = "Fraunhofer"
    fftshiftSwitch: bool = False
    positionOrder: str = "random"
    numIterations: int = 50
    betaObject: Optional[float] = None  # ePIE/mPIE 0.25, qNewton 1 (qNewton.py:45-53)
    betaProbe: Optional[float] = None
    regObject: float = 1.0   # qNewton regularisers (qNewton.py:51-52)
    regProbe: float = 1.0
    momentum_method: str = "ADAM"  # mqNewton: "ADAM", "NADAM" or "momentum" (mqNewton.py:65)
    beta1: float = 0.5        # mqNewton ADAM coefficients (mqNewton.py:57-60)
    beta2: float = 0.5
    betaObject_m: float = 0.25
    betaProbe_m: float = 0.25
    alphaObject: float = 0.1
    alphaProbe: float = 0.1
    feedbackM: float = 0.3   # mPIE feedback / multiPIE betaM
    frictionM: float = 0.7   # mPIE friction / multiPIE stepM
    dz: float = 0.0
    refrIndex: float = 1.0
    orthogonalizationSwitch: bool = False
    orthogonalizationFrequency: int = 10
    probePowerCorrectionSwitch: bool = False
    comStabilizationSwitch: object = False   # False or the period in iterations
    probeBoundary: bool = False
    absorbingProbeBoundary: bool = False
    absorbingProbeBoundaryAleph: float = 5e-2
    positionCorrectionSwitch: bool = False
    positionCorrectionSwitch_radius: int = 1
    saveMemory: bool = False
    precision: str = "c64"
    error: List[float] = field(default_factory=list)

    def __post_init__(self):
        """precision: "c64" (the reference's CPU dtype flow), "c64pure"
        (complex64 everywhere, including e3PIE's slice transfer function, as
        the reference's GPU path casts it), "c128" (double precision)."""
        g = self.geo
        qn = self.engine in ("qNewton", "mqNewton")
        if self.betaObject is None:
            self.betaObject = 1 if qn else 0.25
        if self.betaProbe is None:
            self.betaProbe = 1 if qn else 0.25
        c = np.complex128 if self.precision == "c128" else np.complex64
        self.cdtype = c
        self.object = np.asarray(self.object).astype(c)
        self.probe = np.asarray(self.probe).astype(c)
        self.ptychogram = g.ptychogram.astype(np.float64) if self.precision == "c128" else g.ptychogram
        if self.fftshiftSwitch:
            # _checkFFT (BaseEngine.py:405-422): the detector data are
            # ifftshifted once so the unshifted transforms see them in FFT order
            self.ptychogram = np.fft.ifftshift(self.ptychogram, axes=(-1, -2))
        self.energy = energy_at_pos(self.ptychogram)
        self.err_at_pos_sum = np.zeros(self.ptychogram.shape[0])
        self.errorAtPos = None
        # _initializeErrors (BaseEngine.py:199-225)
        F, Nd = self.ptychogram.shape[0], self.ptychogram.shape[-1]
        if self.saveMemory:
            self.detectorError = 0
            self.errorAtPos = np.zeros((F, 1), dtype=np.float32)
        else:
            self.detectorError = np.zeros((F, Nd, Nd))
        self._tf = {}
        self.schedule = []  # (order, momentum triggers) per iteration, for inspection
        # Reconstruction.encoder_corrected and the positions derived from it
        # (Reconstruction.py:675-710); position correction moves them
        self.encoder_corrected = np.array(g.encoder, copy=True)
        self.positions = g.positions.copy()
        self.D = np.zeros((self.ptychogram.shape[0], 2))  # _initializePCParameters (BaseEngine.py:178-197)
        if self.engine in ("mPIE", "multiPIE", "mqNewton", "pcPIE"):
            # mPIE.py:87-94 / multiPIE.py:61-66 / mqNewton.py:43-44 / pcPIE.py:44-51
            self.objectMomentum = np.zeros_like(self.object)
            self.probeMomentum = np.zeros_like(self.probe)
        if self.engine in ("mqNewton", "pcPIE"):
            # buffers at construction (mqNewton.py:46-47, pcPIE.py:49-51)
            self.objectBuffer = self.object.copy()
            self.probeBuffer = self.probe.copy()
        if self.engine == "e3PIE":
            # e3PIE.py:61-69
            H = aspw_transfer(self.dz, g.wavelength / self.refrIndex, g.Np, g.Lp)
            self.H = np.fft.ifftshift(H)

    # -- shared pieces ------------------------------------------------------
    def _fwd(self, x):
        return propagate(x, self.propagator, self.geo, False, self.fftshiftSwitch, self._tf,
                         pure=self.precision == "c64pure")

    def _bwd(self, x):
        return propagate(x, self.propagator, self.geo, True, self.fftshiftSwitch, self._tf,
                         pure=self.precision == "c64pure")

    def _intensity_projection(self, k, esw):
        """BaseEngine.intensityProjection 'standard' branch (:870-1012) with
        getRMSD (:843-868)."""
        ESW = self._fwd(esw)
        Iest = np.sum(np.abs(ESW) ** 2, axis=(0, 1, 2))[-1]
        Imeas = self.ptychogram[k]
        det_err = abs(Imeas - Iest)
        self.err_at_pos_sum[k] = np.sum(np.abs(det_err.astype(np.float64)))
        # getRMSD (BaseEngine.py:851-868)
        if self.saveMemory:
            self.errorAtPos[k] = np.sum(det_err)
        else:
            self.detectorError[k] = det_err
        frac = np.sqrt(Imeas / (Iest + GEMMA))
        ESW = ESW * frac
        return self._bwd(ESW)

    def _error_metrics(self):
        """getErrorMetrics (BaseEngine.py:819-841).  With saveMemory the
        (numFrames, 1) errorAtPos divided by the (numFrames,) energies
        broadcasts to a (numFrames, numFrames) array, whose sum is the error."""
        if self.saveMemory:
            self.errorAtPos = self.errorAtPos / (self.energy + 1e-20)
        else:
            self.errorAtPos = self.err_at_pos_sum / (self.energy + 1e-20)
        self.error.append(float(np.sum(self.errorAtPos)))

    @property
    def max_probe_power(self):
        """ExperimentalData.maxProbePower (ExperimentalData.py:250)."""
        return np.sqrt(np.max(np.sum(self.geo.ptychogram, (-1, -2))))

    def _probe_power_correction(self):
        """BaseEngine.py:1462-1469 (and _initialProbePowerCorrection :227-235)."""
        P = self.probe
        self.probe = P / np.sqrt(np.sum(P * P.conj())) * self.max_probe_power

    def _initial_probe_power_correction(self):
        if self.probePowerCorrectionSwitch:
            self._probe_power_correction()

    def _probe_window(self):
        """_probeWindow (BaseEngine.py:237-266): the super-Gaussian absorbing
        window, replaced by a circle of 1.2 entrance-pupil diameters when
        probeBoundary is on; float64 (float32 in c64pure)."""
        g = self.geo
        w = np.exp(-(((g.Xp ** 2 + g.Yp ** 2) / (2 * (3 / 4 * g.Np * g.dxp / 2.355) ** 2)) ** 10))
        if self.probeBoundary:
            w = circ(g.Xp, g.Yp, g.entrancePupilDiameter + g.entrancePupilDiameter * 0.2)
        if self.precision == "c64pure":
            w = np.asarray(w, dtype=np.float32)
        self.probeWindow = w

    def _com_stabilization(self):
        """comStabilization (BaseEngine.py:1724-1776): centre of mass of |P|
        of the first mode, last slice; roll probe and object (and the mPIE
        buffers and momenta) by the rounded offset when it exceeds one pixel."""
        g = self.geo
        P2 = abs(self.probe[0, 0, 0, -1])
        demon = np.sum(P2) * g.dxp
        xc = int(np.around(np.sum(np.array(g.Xp, np.float32) * P2) / demon))
        yc = int(np.around(np.sum(np.array(g.Yp, np.float32) * P2) / demon))
        if xc ** 2 + yc ** 2 > 1:
            roll = lambda a: np.roll(a, (-yc, -xc), axis=(-2, -1))  # noqa: E731
            self.probe = roll(self.probe)
            self.object = roll(self.object)
            if self.engine in ("mPIE", "multiPIE", "pcPIE", "mqNewton"):  # params.momentumAcceleration
                self.probeMomentum, self.probeBuffer = roll(self.probeMomentum), roll(self.probeBuffer)
                self.objectMomentum, self.objectBuffer = roll(self.objectMomentum), roll(self.objectBuffer)
        return yc, xc

    def _apply_constraints(self, loop):
        """applyConstraints (BaseEngine.py:1440-1612): orthogonalization
        (:1614-1722), probe power correction (:1462-1469), COM stabilisation
        (:1486-1491, 1724-1776), probe boundary (:1493-1494), absorbing probe
        boundary (:1496-1506), in the reference's order."""
        self._apply_orthogonalization(loop)
        if self.probePowerCorrectionSwitch:
            self._probe_power_correction()
        if self.comStabilizationSwitch is not None and self.comStabilizationSwitch is not False:
            if loop % int(self.comStabilizationSwitch) == 0:
                self._com_stabilization()
        if self.probeBoundary:
            self.probe *= self.probeWindow
        if self.absorbingProbeBoundary:
            a = self.absorbingProbeBoundaryAleph
            self.probe = (1 - a) * self.probe + a * self.probe * self.probeWindow
        if self.positionCorrectionSwitch:
            self._position_correction_update()

    def _apply_orthogonalization(self, loop):
        if self.orthogonalizationSwitch and loop % self.orthogonalizationFrequency == 0:
            g = self.geo
            if g.npsm > 1:
                for il in range(g.nlambda):
                    for isl in range(g.nslice):
                        modes, ev, V = orthogonalize_modes(self.probe[il, 0, :, isl])
                        self.probe[il, 0, :, isl] = modes
                        if self.engine in ("mPIE", "multiPIE", "pcPIE", "mqNewton"):
                            for buf in (self.probeBuffer, self.probeMomentum):
                                p = buf[il, 0, :, isl].reshape(g.npsm, g.Np ** 2)
                                buf[il, 0, :, isl] = (V @ p).reshape(g.npsm, g.Np, g.Np)
            elif g.nosm > 1:
                for il in range(g.nlambda):
                    for isl in range(g.nslice):
                        modes, ev, V = orthogonalize_modes(self.object[il, :, 0, isl])
                        self.object[il, :, 0, isl] = modes
                        if self.engine in ("mPIE", "multiPIE", "pcPIE", "mqNewton"):
                            for buf in (self.objectBuffer, self.objectMomentum):
                                p = buf[il, :, 0, isl].reshape(g.nosm, g.No ** 2)
                                buf[il, :, 0, isl] = (V @ p).reshape(g.nosm, g.No, g.No)

    # -- engines ----------------------------------------------------------------
    def reconstruct(self, positions_limit=None):
        """Run ``numIterations`` iterations; ``positions_limit`` truncates each
        iteration's position loop (used only to time a bounded CPU sample)."""
        g = self.geo
        self._initial_probe_power_correction()
        self._probe_window()
        if self.engine in ("mPIE", "multiPIE"):
            # mPIE.py:104-105
            self.objectBuffer = self.object.copy()
            self.probeBuffer = self.probe.copy()
        if self.engine == "mqNewton":
            # initializeAdaptiveMomentum (mqNewton.py:68-79)
            self.objectMomentum_v = self.objectMomentum.copy()
            self.probeMomentum_v = self.probeMomentum.copy()
        F = self.ptychogram.shape[0]
        for loop in range(self.numIterations):
            order = position_order(self.positionOrder, F, len(self.error), self.positions)
            trig = []
            for j, k in enumerate(order):
                if positions_limit is not None and j >= positions_limit:
                    break
                if self.engine == "e3PIE":
                    self._e3pie_step(k)
                else:
                    patch = self._pie_step(k)
                    if self.positionCorrectionSwitch and self.engine in ("mPIE", "pcPIE"):
                        self._position_correction(k, patch)
                if self.engine == "mqNewton":
                    self._adam_momentum(loop)
                if self.engine in ("mPIE", "multiPIE", "pcPIE"):
                    t = bool(np.random.rand(1) > 0.95)
                    trig.append(t)
                    if t:
                        self._momentum()
            self.schedule.append((np.array(order), np.array(trig, dtype=bool)))
            self._error_metrics()
            self._apply_constraints(loop)
        return self

    def _pie_step(self, k):
        g = self.geo
        Np = int(g.Np)
        row, col = self.positions[k]
        sy = slice(row, row + Np)
        sx = slice(col, col + Np)
        patch = self.object[..., sy, sx].copy()
        esw = patch * self.probe
        esw_upd = self._intensity_projection(k, esw)
        DELTA = esw_upd - esw
        P = self.probe
        if self.engine in ("qNewton", "mqNewton"):
            # qNewton.py:106-140: the object step first, then the probe step with
            # Omax taken over the whole, already updated, object
            Pmax = np.max(np.sum(np.abs(P), axis=(0, 1, 2, 3)))
            fracO = np.abs(P) / Pmax * P.conj() / (np.abs(P) ** 2 + self.regObject)
            self.object[..., sy, sx] = patch + self.betaObject * np.sum(fracO * DELTA, axis=(0, 2, 3), keepdims=True)
            Omax = np.max(np.sum(np.abs(self.object), axis=(0, 1, 2, 3)))
            fracP = np.abs(patch) / Omax * patch.conj() / (np.abs(patch) ** 2 + self.regProbe)
            self.probe = P + self.betaProbe * np.sum(fracP * DELTA, axis=(0, 1, 3), keepdims=True)
        elif self.engine == "ePIE":
            # ePIE.py:128-160
            fracO = P.conj() / np.max(np.sum(np.abs(P) ** 2, axis=(0, 1, 2, 3)))
            self.object[..., sy, sx] = patch + self.betaObject * np.sum(fracO * DELTA, axis=(0, 2, 3), keepdims=True)
            fracP = patch.conj() / np.max(np.sum(np.abs(patch) ** 2, axis=(0, 1, 2, 3)))
            self.probe = P + self.betaProbe * np.sum(fracP * DELTA, axis=(0, 1, 3), keepdims=True)
        else:
            # mPIE.py:226-270 / multiPIE.py:153-205 (CPM branch)
            absP2 = np.abs(P) ** 2
            Pmax = np.max(np.sum(absP2, axis=(0, 1, 2, 3)), axis=(-1, -2))
            fracO = P.conj() / (self.alphaObject * Pmax + (1 - self.alphaObject) * absP2)
            self.object[..., sy, sx] = patch + self.betaObject * np.sum(fracO * DELTA, axis=2, keepdims=True)
            absO2 = np.abs(patch) ** 2
            Omax = np.max(np.sum(absO2, axis=(0, 1, 2, 3)), axis=(-1, -2))
            fracP = patch.conj() / (self.alphaProbe * Omax + (1 - self.alphaProbe) * absO2)
            if self.engine in ("mPIE", "pcPIE"):
                weight = 1
                self.probe = P + weight * self.betaProbe * np.sum(fracP * DELTA, axis=1, keepdims=True)
            else:
                self.probe = P + self.betaProbe * np.sum(fracP * DELTA, axis=(0, 1), keepdims=True)
        return patch

    _PC_ROW = np.array([-1, -1, -1, 0, 0, 0, 1, 1, 1])  # BaseEngine.py:191-194
    _PC_COL = np.array([-1, 0, 1, -1, 0, 1, -1, 0, 1])

    def _position_correction(self, k, patch):
        """positionCorrection (BaseEngine.py:1241-1323), CPM, after the update
        of position k: the direct branch (radius < 2, 9 one-pixel shifts) or
        the Fourier branch (radius >= 2, :1277-1302)."""
        if len(self.error) <= 1:  # startAtIteration = 1
            return
        Np = int(self.geo.Np)
        row, col = self.positions[k]
        sy, sx = slice(row, row + Np), slice(col, col + Np)
        O = self.object
        radius = int(self.positionCorrectionSwitch_radius)
        if radius < 2:
            rowShifts, colShifts = self._PC_ROW, self._PC_COL
            cc = np.zeros((len(rowShifts), 1))
            for s in range(len(rowShifts)):
                t = np.roll(patch, rowShifts[s], axis=-2)
                sh = np.roll(t, colShifts[s], axis=-1)
                # a complex sum stored into a float array keeps its real part
                cc[s] = np.squeeze(np.sum(sh.conj() * O[..., sy, sx], axis=(-2, -1))).real
            betaGrad, r = 1000, 3
        else:
            ss = slice(-radius, radius + 1)
            rowShifts, colShifts = np.mgrid[ss, ss]
            rowShifts, colShifts = rowShifts.flatten(), colShifts.flatten()
            FT_O = np.fft.fft2(O[..., sy, sx] - O[..., sy, sx].mean())
            FT_Op = np.fft.fft2(patch - O.mean())
            xcor = np.fft.ifft2(FT_O * FT_Op.conj())
            xcor = abs(np.fft.fftshift(xcor))
            N = xcor.shape[-1]
            sw = slice(N // 2 - radius, N // 2 + radius + 1)
            cc = xcor[..., sw, sw].flatten()
            betaGrad, r = 5, 10
        cc = abs(cc)
        normFactor = np.sum(patch.conj() * patch, axis=(-2, -1)).real
        grad_x = betaGrad * np.sum((cc.T - np.mean(cc)) / normFactor * colShifts)
        grad_y = betaGrad * np.sum((cc.T - np.mean(cc)) / normFactor * rowShifts)
        if abs(grad_x) > r:
            grad_x = r * grad_x / abs(grad_x)
        if abs(grad_y) > r:
            grad_y = r * grad_y / abs(grad_y)
        delta_p = 0.5 * np.array([grad_y, grad_x])
        self.D[k, :] = delta_p + 0.9 * self.D[k, :]

    def _position_correction_update(self):
        """positionCorrectionUpdate (BaseEngine.py:1400-1438), CPM."""
        if len(self.error) <= 1:
            return
        g = self.geo
        new = self.encoder_corrected - 1 * self.D * g.dxo
        new = new - new.mean(axis=0, keepdims=True)
        new = new + g.encoder.mean(axis=0, keepdims=True)
        self.encoder_corrected = new
        self.positions = pixel_positions(new, g.dxo, g.No, g.Np)

    def _momentum(self):
        """objectMomentumUpdate/probeMomentumUpdate (mPIE.py:196-224,
        multiPIE.py:125-151)."""
        grad = self.objectBuffer - self.object
        self.objectMomentum = grad + self.frictionM * self.objectMomentum
        self.object = self.object - self.feedbackM * self.objectMomentum
        self.objectBuffer = self.object.copy()
        grad = self.probeBuffer - self.probe
        self.probeMomentum = grad + self.frictionM * self.probeMomentum
        self.probe = self.probe - self.feedbackM * self.probeMomentum
        self.probeBuffer = self.probe.copy()

    def _adam(self, grad, mt, vt, itr):
        """mqNewton.ADAM (mqNewton.py:142-153): the second moment takes the
        squared 2-norm of the whole flattened gradient."""
        beta1_scale = 1 - self.beta1 ** itr
        beta2_scale = 1 - self.beta2 ** itr
        mt = self.beta1 * mt + (1 - self.beta1) * grad
        vt = self.beta2 * vt + (1 - self.beta2) * np.linalg.norm(grad.flatten().squeeze(), 2) ** 2
        m_hat = mt / beta1_scale
        v_hat = vt / beta2_scale
        return m_hat / (v_hat ** 0.5 + 1e-8), mt, vt

    def _nadam(self, grad, mt, vt, itr):
        """mqNewton.NADAM (mqNewton.py:155-174): ADAM's moments with the
        Nesterov step (beta1 m_hat + (1 - beta1) g / (1 - beta1^itr)) /
        (sqrt(v_hat) + 1e-8)."""
        beta1_scale = 1 - self.beta1 ** itr
        beta2_scale = 1 - self.beta2 ** itr
        norm_sq = np.linalg.norm(grad.flatten(), 2) ** 2
        mt = self.beta1 * mt + (1 - self.beta1) * grad
        vt = self.beta2 * vt + (1 - self.beta2) * norm_sq
        m_hat = mt / beta1_scale
        v_hat = vt / beta2_scale
        update = (self.beta1 * m_hat + grad * (1 - self.beta1) / beta1_scale) / (v_hat ** 0.5 + 1e-8)
        return update, mt, vt

    def _plain_momentum(self, grad, mt, vt, itr):
        """mqNewton.momentum (mqNewton.py:176-182): m = g + frictionM m, the
        step is m; the second moment is passed through untouched."""
        mt = grad + self.frictionM * mt
        return mt, mt, vt

    def _momentum_engine(self):
        return {"ADAM": self._adam, "NADAM": self._nadam, "momentum": self._plain_momentum}[self.momentum_method]

    def _adam_momentum(self, loop):
        """objectMomentumUpdate / probeMomentumUpdate (mqNewton.py:184-224)
        after every position, with itr = loop + 1, through the selected
        momentum engine."""
        step = self._momentum_engine()
        grad = self.objectBuffer - self.object
        upd, self.objectMomentum, self.objectMomentum_v = step(grad, self.objectMomentum, self.objectMomentum_v,
                                                               loop + 1)
        self.object = self.object - self.betaObject_m * upd
        self.objectBuffer = self.object.copy()
        grad = self.probeBuffer - self.probe
        upd, self.probeMomentum, self.probeMomentum_v = step(grad, self.probeMomentum, self.probeMomentum_v, loop + 1)
        self.probe = self.probe - self.betaProbe_m * upd
        self.probeBuffer = self.probe.copy()

    def _e3pie_step(self, k):
        """e3PIE.py:89-173 plus its update rules :192-225."""
        g = self.geo
        Np = int(g.Np)
        nS = g.nslice
        row, col = g.positions[k]
        sy = slice(row, row + Np)
        sx = slice(col, col + Np)
        if not hasattr(self, "esw"):
            self.esw = self.probe.copy()  # e3PIE.py:77
        if self.precision == "c64":
            H = self.H  # complex128, as the reference's CPU path (e3PIE.py:61-69)
        elif self.precision == "c64pure":
            H = self.H.astype(np.complex64)  # the reference's GPU dtype policy (gpuUtils.py:62-70)
        else:
            H = self.H.astype(np.complex128)
        patch = self.object[..., sy, sx].copy()
        self.esw[:, :, :, 0] = patch[:, :, :, 0] * self.probe[:, :, :, 0]
        for s in range(1, nS):
            self.probe[:, :, :, s] = np.fft.ifft2(np.fft.fft2(self.esw[:, :, :, s - 1]) * H)
            self.esw[:, :, :, s] = self.probe[:, :, :, s] * patch[:, :, :, s]
        esw_upd = self._intensity_projection(k, self.esw)
        DELTA = (esw_upd - self.esw)[:, :, :, -1]

        def obj_upd(op, d, lp):
            frac = lp.conj() / np.max(np.sum(np.abs(lp) ** 2, axis=(0, 1, 2)))
            return op + self.betaObject * np.sum(frac * d, axis=(0, 2), keepdims=True)

        def probe_upd(op, d, lp, beth):
            frac = op.conj() / np.max(np.sum(np.abs(op) ** 2, axis=(0, 1, 2)))
            return lp + beth * np.sum(frac * d, axis=(0, 1), keepdims=True)

        for t in range(nS - 1):
            s = nS - 1 - t
            self.object[..., s, sy, sx] = obj_upd(patch[:, :, :, s], DELTA, self.probe[:, :, :, s])
            self.probe[:, :, :, s] = probe_upd(patch[:, :, :, s], DELTA, self.probe[:, :, :, s], 0.9)
            DELTA = np.fft.ifft2(np.fft.fft2(self.probe[:, :, :, s]) * H.conj()) - self.esw[:, :, :, s - 1]
        self.object[..., 0, sy, sx] = obj_upd(patch[:, :, :, 0], DELTA, self.probe[:, :, :, 0])
        self.probe[:, :, :, 0] = probe_upd(patch[:, :, :, 0], DELTA, self.probe[:, :, :, 0], self.betaProbe)


PINNED = {
    "fft2c": "golden propagator_fraunhofer + unitary round trip + brute-force DFT",
    "ifft2c": "golden propagator_fraunhofer + round trip",
    "aspw_transfer": "golden propagator_asp + z=0 identity + closed-form band limit",
    "propagate": "goldens propagator_fraunhofer / propagator_asp",
    "pixel_positions": "engine goldens (positions drive the patch indexing)",
    "initial_guess": "engine goldens (seeded init is their input)",
    "energy_at_pos": "engine goldens (error vectors)",
    "position_order": "sequential: goldens; random: np.random.shuffle identity test",
    "OracleRun.ePIE": "goldens epie_fraunhofer, epie_asp, epie_mixed_state",
    "OracleRun.mPIE": "goldens mpie_single, mpie_mixed_state",
    "OracleRun.multiPIE": "identity with mPIE at nlambda=nosm=1 (no golden exists)",
    "OracleRun.e3PIE": "golden e3pie_multislice",
    "OracleRun.qNewton": "golden qnewton_single",
    "OracleRun.pcPIE": "equals mPIE without position correction; the correction's gradient points back along a known shift",
    "OracleRun.mqNewton": "equals qNewton with betaObject_m = betaProbe_m = 0; ADAM's first step is g / (|g| + 1e-8)",
    "propagate(fresnel/polychromeasp)": "goldens propagator_fresnel / propagator_polychromeasp, epie_fresnel, epie_polychrome",
    "orthogonalize_modes": "orthogonality + descending eigenvalues + span invariants",
    "nrmse": "closed form: invariant to a global complex scale; zero at gamma*true",
}
