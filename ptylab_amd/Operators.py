"""Propagators of the hot path (Operators/Operators.py).

``object2detector`` / ``detector2object`` keep the reference's signature
``(fields, params, reconstruction) -> (esw, propagated)`` (Operators.py:561-584)
and run the Fraunhofer and ASP propagators on the GPU through
``pty_propagate``.  The ASP transfer function is built on the host in float64
exactly as the reference builds it (Operators.py:631-684, 861-893) and cast to
complex64 once; it is cached per geometry like the reference's lru_cache."""

from functools import lru_cache

import numpy as np

from . import _native
from .utils import circ

PROPAGATORS = {"fraunhofer", "asp"}


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


def _kind(params):
    kind = params.propagatorType.lower()
    if kind not in PROPAGATORS:
        raise NotImplementedError(f"propagator {params.propagatorType!r} is not implemented on the GPU path")
    return kind


def _native_kind(params, kind):
    if kind == "asp":
        if params.fftshiftSwitch:
            raise ValueError("ASP propagator only works with fftshiftswitch == False")
        return _native.PTY_ASP
    return _native.PTY_FRAUNHOFER_SHIFTED if params.fftshiftSwitch else _native.PTY_FRAUNHOFER


def _propagate(fields, params, reconstruction, inverse):
    import torch

    kind = _kind(params)
    nk = _native_kind(params, kind)
    if nk == _native.PTY_ASP and reconstruction.nlambda > 1:
        raise ValueError("For multi-wavelength, set polychromeASP instead of ASP")
    dev = torch.device("cuda")
    src_numpy = not isinstance(fields, torch.Tensor)
    t = torch.as_tensor(np.asarray(fields, dtype=np.complex64) if src_numpy else fields, device=dev)
    out = t.to(torch.complex64).contiguous().clone()
    tf = None
    if nk == _native.PTY_ASP:
        tf = torch.as_tensor(
            asp_tf_fftorder(reconstruction.zo, reconstruction.wavelength, reconstruction.Np, reconstruction.Lp),
            device=dev,
        )
    _native.propagate(out, nk, inverse, tf, torch.cuda.current_stream())
    return out.cpu().numpy() if src_numpy else out


def object2detector(fields, params, reconstruction):
    """Forward propagation; returns (reconstruction.esw, propagated)."""
    if fields is None:
        fields = reconstruction.esw
    return getattr(reconstruction, "esw", None), _propagate(fields, params, reconstruction, False)


def detector2object(fields, params, reconstruction):
    """Back propagation; returns (reconstruction.esw, propagated)."""
    if fields is None:
        fields = reconstruction.ESW
    return getattr(reconstruction, "esw", None), _propagate(fields, params, reconstruction, True)
