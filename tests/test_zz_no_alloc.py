"""pty_roll_fields_noalloc (static device scratch, no allocation) against
pty_roll_fields (caller workspace), through the C ABI: two calls in a row,
the second with a smaller input, and a size above the scratch limit
rejected before any launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _roll_both(lib, torch, x_np, dy, dx, stream):
    nf, ny, nx = x_np.shape
    a = torch.as_tensor(x_np, device="cuda")
    b = a.clone()
    ws = torch.empty_like(a)
    h = ctypes.c_void_p(stream.cuda_stream)
    assert lib.pty_roll_fields(a.data_ptr(), ws.data_ptr(), nf, ny, nx, dy, dx, h) == 0
    assert lib.pty_roll_fields_noalloc(b.data_ptr(), nf, ny, nx, dy, dx, h) == 0
    stream.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_roll_noalloc_matches_workspace_variant():
    import torch

    from ptylab_amd import _native

    lib = _native.load()
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(7)
    shapes = [((4, 256, 256), (5, -7)), ((1, 64, 96), (-3, 11))]  # second call smaller
    for (shape, (dy, dx)) in shapes:
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        print(f"roll no-alloc: shape {shape} shift {(dy, dx)}", flush=True)
        ref, got = _roll_both(lib, torch, x, dy, dx, stream)
        assert np.array_equal(got, ref)
        assert np.array_equal(got, np.roll(x, (dy, dx), axis=(-2, -1)))
    big = torch.zeros((_native.PTY_ROLL_SCRATCH_MAX_VALUES // (64 * 64) + 1, 64, 64), dtype=torch.complex64,
                      device="cuda")
    rc = lib.pty_roll_fields_noalloc(big.data_ptr(), big.shape[0], 64, 64, 1, 1, ctypes.c_void_p(stream.cuda_stream))
    assert rc == _native.PTY_EINVAL
    assert b"PTY_ROLL_SCRATCH_MAX_VALUES" in lib.pty_last_error()
