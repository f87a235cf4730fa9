"""Host logic of the learned initialiser (no GPU): stitching recovers an
object from exact patches that each carry an arbitrary complex scale, the
NRMSE is invariant to a global complex scale, the scanned-region mask, the
patch targets' normalisation."""

import numpy as np
import torch

from ptylab_amd import learned_init as LI
from ptylab_amd import synthetic as S


def _scan():
    rng = np.random.default_rng(0)
    No = 256
    obj = S.smooth_random_object(No, rng)
    px = S.fermat_scan(60, LI.NP, No, 12.0, np.random.default_rng(1))
    pos = px + No // 2 - LI.NP // 2
    return obj, pos, No


def test_stitch_recovers_object_from_arbitrarily_scaled_patches():
    obj, pos, No = _scan()
    rng = np.random.default_rng(2)
    patches = []
    for r, c in pos:
        p = obj[r + LI.OFF:r + LI.OFF + LI.PATCH, c + LI.OFF:c + LI.OFF + LI.PATCH]
        g = rng.uniform(0.5, 2.0) * np.exp(1j * rng.uniform(-np.pi, np.pi))
        patches.append(p * g)
    patches = torch.as_tensor(np.array(patches, np.complex64))
    probe = LI.probe()
    est = LI.stitch(patches, pos, probe, No)
    true = torch.as_tensor(obj)
    cover = torch.zeros((No, No), dtype=torch.bool)
    for r, c in pos:
        cover[r + LI.OFF + 16:r + LI.OFF + 48, c + LI.OFF + 16:c + LI.OFF + 48] = True
    assert float(LI.nrmse_gpu(est, true, cover)) < 1e-4
    # outside every patch the 'ones' value stays
    assert complex(est[0, 0]) == 1.0


def test_nrmse_is_scale_invariant_and_zero_at_the_truth():
    obj, pos, No = _scan()
    t = torch.as_tensor(obj)
    roi = torch.as_tensor(LI.roi_mask(pos, LI.probe(), No))
    assert float(LI.nrmse_gpu(t * (0.3 - 2j), t, roi)) < 1e-6
    noisy = t + 0.01 * torch.as_tensor(np.random.default_rng(3).standard_normal((No, No)).astype(np.complex64))
    e = float(LI.nrmse_gpu(noisy, t, roi))
    assert 0.005 < e < 0.03
    assert abs(float(LI.nrmse_gpu(noisy * 1j, t, roi)) - e) < 1e-6


def test_roi_mask_is_union_of_footprints():
    obj, pos, No = _scan()
    m = LI.roi_mask(pos, LI.probe(), No)
    I = np.abs(LI.probe()) ** 2
    fp = I > 0.01 * I.max()
    r, c = pos[0]
    assert np.all(m[r:r + LI.NP, c:c + LI.NP][fp])
    assert not m[0, 0]


def test_patch_targets_have_unit_complex_mean():
    rng = np.random.default_rng(4)
    win = torch.as_tensor((rng.random((3, LI.NP, LI.NP)) + 0.5) * np.exp(1j * rng.random((3, LI.NP, LI.NP))))
    t = LI.patch_targets(win.to(torch.complex64))
    assert t.shape == (3, LI.PATCH, LI.PATCH, 2)
    z = torch.complex(t[..., 0], t[..., 1]).mean(dim=(1, 2))
    assert torch.allclose(z, torch.ones(3, dtype=torch.complex64), atol=1e-5)
