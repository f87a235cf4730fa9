"""The wave-shuffle FFT and the standalone propagators against numpy.fft and
the oracle's fft2c/ASP."""

import numpy as np
import pytest

import oracle as O
from ptyhelpers import rel

pytestmark = pytest.mark.gpu

SIZES = [16, 32, 64, 128, 256, 512]


def _field(n, nf=3, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((nf, n, n)) + 1j * rng.standard_normal((nf, n, n))).astype(np.complex64)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2_matches_numpy(n, sign):
    import torch
    from ptylab_amd import _native

    x = _field(n)
    t = torch.as_tensor(x, device="cuda")
    _native.fft2(t, sign)
    got = t.cpu().numpy()
    want = np.fft.fft2(x.astype(np.complex128)) if sign < 0 else np.fft.ifft2(x.astype(np.complex128)) * n * n
    assert rel(got, want) < 2e-6 * np.log2(n)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("inverse", [False, True])
def test_fraunhofer_propagator_matches_fft2c(n, inverse):
    import torch
    from ptylab_amd import _native

    x = _field(n, seed=1)
    t = torch.as_tensor(x, device="cuda")
    _native.propagate(t, _native.PTY_FRAUNHOFER, inverse)
    want = O.ifft2c(x.astype(np.complex128)) if inverse else O.fft2c(x.astype(np.complex128))
    assert rel(t.cpu().numpy(), want) < 2e-6 * np.log2(n)


@pytest.mark.parametrize("n", [32, 64, 256])
def test_asp_propagator_matches_oracle(n):
    import torch
    from ptylab_amd import _native
    from ptylab_amd.Operators import asp_tf_fftorder

    x = _field(n, seed=2)
    z, wl, L = 5e-3, 632.8e-9, n * 4e-6
    tf = asp_tf_fftorder(z, wl, n, L)
    t = torch.as_tensor(x, device="cuda")
    _native.propagate(t, _native.PTY_ASP, False, torch.as_tensor(tf, device="cuda"))
    want = O.ifft2c(O.fft2c(x.astype(np.complex128), True) * tf.astype(np.complex128), True)
    assert rel(t.cpu().numpy(), want) < 2e-6 * np.log2(n)


@pytest.mark.parametrize("prop,nl", [("Fresnel", 1), ("polychromeASP", 3), ("scaledASP", 1),
                                     ("scaledPolychromeASP", 3), ("twoStepPolychrome", 3)])
def test_fresnel_and_polychrome_propagators_match_oracle(prop, nl):
    """Operators.object2detector / detector2object on the GPU against the
    oracle's complex128 propagation of the same complex64 field (the
    reference's own CPU path runs these two in complex128)."""
    import ptylab_amd as P
    from ptyhelpers import make_data
    from regression_inputs import SEED, regression_dataset, spectral_density

    d = regression_dataset()
    data = make_data(d)
    params = P.Params()
    params.propagatorType = prop
    rec = P.Reconstruction(data, params)
    rec.nlambda = nl
    if nl > 1:
        rec.spectralDensity = spectral_density(d["wavelength"], nl)
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"],
                     nlambda=nl, spectralDensity=rec.spectralDensity if nl > 1 else None)
    field = np.asarray(rec.probe)
    _, fwd = P.Operators.object2detector(field, params, rec)
    # BaseEngine.intensityProjection sets ESW before propagating back; the
    # two-step inverse also propagates it (Operators.py:285-310)
    rec.ESW = fwd
    first, back = P.Operators.detector2object(fwd, params, rec)
    if prop == "twoStepPolychrome":
        np.testing.assert_array_equal(first, back)
    want_f = O.propagate(field.astype(np.complex128), prop, geo)
    want_b = O.propagate(want_f, prop, geo, inverse=True)
    ef, eb = rel(fwd, want_f), rel(back, want_b)
    print(f"{prop}: forward {ef:.2e} back {eb:.2e}", flush=True)
    assert ef < 1e-5 and eb < 1e-5
