"""Propagators of the hot path (Operators/Operators.py).

``object2detector`` / ``detector2object`` keep the reference's signature
``(fields, params, reconstruction) -> (esw, propagated)`` (Operators.py:561-584)
and run the Fraunhofer, Fresnel, ASP, polychromeASP and scaledASP propagators
on the GPU through ``pty_propagate`` (Fresnel multiplies by the quadratic phase
around a Fraunhofer transform, polychromeASP runs one ASP per wavelength,
scaledASP is an ASP between two checkerboarded phase multiplies).  Transfer
functions and the quadratic phase are built on the host in float64 exactly as
the reference builds them (Operators.py:631-684, 861-893, 910-952,
_propagation_kernels.py:16-36) and cast to complex64 once; they are cached per
geometry like the reference's lru_cache."""

from functools import lru_cache

import numpy as np

from . import _native
from .utils import circ

PROPAGATORS = {"fraunhofer", "asp", "fresnel", "polychromeasp", "scaledasp", "scaledpolychromeasp", "twosteppolychrome"}


@lru_cache(maxsize=8)
def _aspw_tf_cached(z, wavelength, N, L, bandlimit):
    a_z = abs(z)
    k = 2 * np.pi / wavelength
    f = np.arange(-N / 2, N / 2) / L
    Fx, Fy = np.meshgrid(f, f)
    f_max = L / (wavelength * np.sqrt(L ** 2 + 4 * a_z ** 2))
    inside = np.array(circ(Fx, Fy, 2 * f_max))
    arg = 1 - (Fx * wavelength) ** 2 - (Fy * wavelength) ** 2
    keep = arg > 0
    if not bandlimit:
        keep = 0 * keep + 1
    arg = np.clip(arg, 0, np.inf)
    phase = k * a_z * np.sqrt(arg)
    H = np.array(keep * (np.cos(phase) + 1j * np.sin(phase)))
    if z < 0:
        H = H.conj()
    out = H * inside
    out.setflags(write=False)
    return out


def aspw_transfer_function(z, wavelength, N, L, bandlimit=True):
    """Band-limited angular-spectrum transfer function (Matsushima & Shimobaba,
    Opt. Express 2009), complex128, centred frequency order
    (Operators.py:631-684)."""
    return _aspw_tf_cached(float(z), float(wavelength), int(N), float(L), bool(bandlimit))


def asp_tf_fftorder(z, wavelength, N, L):
    """complex64 transfer function in FFT frequency order, as the ASP
    propagator multiplies it (Operators.py:206-213 with the complex64 stack of
    :861-893)."""
    return np.ascontiguousarray(np.fft.ifftshift(aspw_transfer_function(z, wavelength, N, L).astype(np.complex64)))


@lru_cache(maxsize=8)
def _quad_phase_cached(z, wavelength, Np, dxp):
    x = np.linspace(-Np / 2, Np / 2, int(Np)) * dxp
    X, Y = np.meshgrid(x, x)
    out = np.exp(1.0j * np.pi / (wavelength * z) * (X ** 2 + Y ** 2))
    out.setflags(write=False)
    return out


def quadratic_phase(z, wavelength, Np, dxp):
    """Fresnel quadratic phase exp(i pi (x^2 + y^2) / (lambda z)) on
    linspace(-Np/2, Np/2, Np) * dxp, complex64, natural order
    (_propagation_kernels.py:16-36; the reference keeps it complex128)."""
    return np.ascontiguousarray(_quad_phase_cached(float(z), float(wavelength), int(Np), float(dxp)).astype(np.complex64))


def polychrome_tf_fftorder(z, spectral_density, N, L):
    """(nlambda, N, N) complex64 transfer functions, one per wavelength of the
    spectral density, in FFT frequency order (Operators.py:910-952)."""
    return np.ascontiguousarray(np.stack([asp_tf_fftorder(z, float(w), N, L) for w in np.atleast_1d(spectral_density)]))


@lru_cache(maxsize=8)
def _scaled_asp_cached(z, wavelength, N, dx, dq):
    k = 2 * np.pi / wavelength
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
    Q1, Q2 = Q1.astype(np.complex64), Q2.astype(np.complex64)
    Q1.setflags(write=False)
    Q2.setflags(write=False)
    return Q1, Q2


def scaled_asp_phases(z, wavelength, N, dxo, dxd):
    """Q1 (source plane) and Q2 (centred frequency grid) of the band-limited
    scaled angular spectrum method, built in float64 and stored complex64 as
    the reference stores them (Operators.py:706-760, 955-987).  Forward:
    ifft2c(fft2c(psi Q1) Q2); inverse: ifft2c(fft2c(Psi) conj Q2) conj Q1."""
    return _scaled_asp_cached(float(z), float(wavelength), int(N), float(dxo), float(dxd))


def scaled_asp_kernel_arrays(z, wavelength, N, dxo, dxd):
    """The (2, N, N) complex64 pair the kernels take for scaledASP: [0] = Q2 as
    the transfer function of unshifted transforms, [1] = Q1 times the
    checkerboard (-1)^(y+x).  For even N, fft2c(x) = c F(c x) and ifft2c(X) =
    c F^-1(c X) with c = (-1)^(y+x), so ifft2c(fft2c(psi Q1) Q2) =
    c F^-1(F(c Q1 psi) Q2): the centred grid of Q2 needs no shift, and the
    detector's checkerboard cancels in the intensity constraint."""
    Q1, Q2 = scaled_asp_phases(z, wavelength, N, dxo, dxd)
    idx = np.arange(int(N))
    c = np.where((idx[:, None] + idx[None, :]) % 2 == 0, 1, -1).astype(np.float32)
    return np.ascontiguousarray(np.stack([Q2, (Q1 * c).astype(np.complex64)]))


def scaled_polychrome_kernel_arrays(z, spectral_density, N, dxo, dxd):
    """(2 nlambda, N, N) complex64 for scaledPolychromeASP: the Q2 of every
    wavelength of the spectral density, then its Q1 (-1)^(y+x)
    (Operators.py:990-1032 builds each pair as scaledASP does)."""
    pairs = [scaled_asp_kernel_arrays(z, float(w), N, dxo, dxd) for w in np.atleast_1d(spectral_density)]
    return np.ascontiguousarray(np.concatenate([np.stack([q[0] for q in pairs]), np.stack([q[1] for q in pairs])]))


def two_step_polychrome_arrays(z, spectral_density, N, L, dxp):
    """The two-step polychromatic propagator's (nlambda, N, N) complex64
    transfer functions in FFT order -- wavelength l propagates
    z (1 - lambda_0 / lambda_l) -- and the complex64 quadratic phase of
    lambda_0 (Operators.py:1033-1079; the reference keeps both complex128)."""
    sd = np.atleast_1d(spectral_density)
    T = np.stack([asp_tf_fftorder(z * (1 - sd[0] / sd[il]), float(sd[il]), N, L) for il in range(len(sd))])
    return np.ascontiguousarray(T), quadratic_phase(z, float(sd[0]), N, dxp)


def _kind(params):
    kind = params.propagatorType.lower()
    if kind not in PROPAGATORS:
        raise NotImplementedError(f"propagator {params.propagatorType!r} is not implemented on the GPU path")
    return kind


def _native_kind(params, kind):
    if kind in ("scaledasp", "scaledpolychromeasp") and params.fftshiftSwitch:
        raise ValueError(f"{params.propagatorType} propagatorType works only with fftshiftSwitch = False!")
    if kind == "twosteppolychrome" and params.fftshiftSwitch:
        raise ValueError("twoStepPolychrome propagatorType works only with fftshiftSwitch = False!")
    if kind in ("asp", "polychromeasp", "scaledasp", "scaledpolychromeasp", "twosteppolychrome"):
        if params.fftshiftSwitch:
            raise ValueError("ASP propagator only works with fftshiftswitch == False")
        return _native.PTY_ASP
    if kind == "fresnel" and params.fftshiftSwitch:
        raise NotImplementedError("Fresnel with fftshiftSwitch=True is not implemented on the GPU path")
    return _native.PTY_FRAUNHOFER_SHIFTED if params.fftshiftSwitch else _native.PTY_FRAUNHOFER


def _propagate(fields, params, reconstruction, inverse):
    import torch

    kind = _kind(params)
    nk = _native_kind(params, kind)
    r = reconstruction
    if kind == "asp" and r.nlambda > 1:
        raise ValueError("For multi-wavelength, set polychromeASP instead of ASP")
    dev = torch.device("cuda")
    src_numpy = not isinstance(fields, torch.Tensor)
    t = torch.as_tensor(np.asarray(fields, dtype=np.complex64) if src_numpy else fields, device=dev)
    out = t.to(torch.complex64).contiguous().clone()
    stream = torch.cuda.current_stream()
    if kind == "asp":
        tf = torch.as_tensor(asp_tf_fftorder(r.zo, r.wavelength, r.Np, r.Lp), device=dev)
        _native.propagate(out, nk, inverse, tf, stream)
    elif kind == "polychromeasp":
        # fields (nlambda, ...): wavelength l propagates with its own transfer function
        tfs = torch.as_tensor(polychrome_tf_fftorder(r.zo, r.spectralDensity, r.Np, r.Lp), device=dev)
        if out.shape[0] != tfs.shape[0]:
            raise ValueError(f"polychromeASP: {out.shape[0]} wavelength planes, {tfs.shape[0]} wavelengths")
        for il in range(tfs.shape[0]):
            plane = out[il].contiguous()
            _native.propagate(plane, nk, inverse, tfs[il].contiguous(), stream)
            out[il] = plane
    elif kind in ("scaledasp", "scaledpolychromeasp"):
        if kind == "scaledasp":
            if r.nlambda > 1:
                raise ValueError("For multi-wavelength, scaledPolychromeASP needs to be used instead of scaledASP")
            arr = scaled_asp_kernel_arrays(r.zo, r.wavelength, r.Np, r.dxo, r.dxd)
        else:
            arr = scaled_polychrome_kernel_arrays(r.zo, r.spectralDensity, r.Np, r.dxo, r.dxd)
        q = torch.as_tensor(arr, device=dev)
        nl = q.shape[0] // 2
        if nl > 1 and out.shape[0] != nl:
            raise ValueError(f"scaledPolychromeASP: {out.shape[0]} wavelength planes, {nl} wavelengths")
        idx = torch.arange(int(r.Np), device=dev)
        cb = (1 - 2 * ((idx[:, None] + idx[None, :]) % 2)).to(torch.complex64)  # (-1)^(y+x)
        planes = [out] if nl == 1 else [out[il] for il in range(nl)]
        for il, plane in enumerate(planes):
            q2, q1c = q[il].contiguous(), q[nl + il]
            if inverse:
                plane = (plane * cb).contiguous()
                _native.propagate(plane, nk, True, q2, stream)
                plane = plane * q1c.conj()
            else:
                plane = (plane * q1c).contiguous()
                _native.propagate(plane, nk, False, q2, stream)
                plane = plane * cb
            if nl == 1:
                out = plane.contiguous()
            else:
                out[il] = plane
    elif kind == "twosteppolychrome":
        # forward fft2c(ifft2c(fft2c(x) T) Q), inverse ifft2c(fft2c(x conj Q)
        # conj T) as the reference writes them (Operators.py:276-282)
        T, Q = two_step_polychrome_arrays(r.zo, r.spectralDensity, r.Np, r.Lp, r.dxp)
        T, Q = torch.as_tensor(T, device=dev), torch.as_tensor(Q, device=dev)
        if out.shape[0] != T.shape[0]:
            raise ValueError(f"twoStepPolychrome: {out.shape[0]} wavelength planes, {T.shape[0]} wavelengths")
        for il in range(T.shape[0]):
            plane = out[il].contiguous()
            if inverse:
                plane = (plane * Q.conj()).contiguous()
                _native.propagate(plane, _native.PTY_ASP, True, T[il].contiguous(), stream)
            else:
                _native.propagate(plane, _native.PTY_ASP, False, T[il].contiguous(), stream)
                plane = (plane * Q).contiguous()
                _native.propagate(plane, _native.PTY_FRAUNHOFER, False, None, stream)
            out[il] = plane
    elif kind == "fresnel":
        Q = torch.as_tensor(quadratic_phase(r.zo, r.wavelength, r.Np, r.dxp), device=dev)
        if inverse:
            _native.propagate(out, nk, True, None, stream)
            out = (out * Q.conj()).contiguous()
        else:
            out = (out * Q).contiguous()
            _native.propagate(out, nk, False, None, stream)
    else:
        _native.propagate(out, nk, inverse, None, stream)
    return out.cpu().numpy() if src_numpy else out


def object2detector(fields, params, reconstruction):
    """Forward propagation; returns (reconstruction.esw, propagated)."""
    if fields is None:
        fields = reconstruction.esw
    return getattr(reconstruction, "esw", None), _propagate(fields, params, reconstruction, False)


def detector2object(fields, params, reconstruction):
    """Back propagation; returns (reconstruction.esw, propagated).  For
    twoStepPolychrome the first value is the back-propagation of
    reconstruction.ESW, as the reference's propagate_twoStepPolychrome_inv
    returns it (Operators.py:285-310)."""
    if fields is None:
        fields = reconstruction.ESW
    back = _propagate(fields, params, reconstruction, True)
    if _kind(params) == "twosteppolychrome":
        return _propagate(reconstruction.ESW, params, reconstruction, True), back
    return getattr(reconstruction, "esw", None), back
