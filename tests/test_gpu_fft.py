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
