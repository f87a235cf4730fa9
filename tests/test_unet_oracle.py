"""Pins of the float64 U-Net oracle (no GPU): the convolution in the
weight-row layout against a brute-force loop, autograd gradients against
central finite differences, Adam's first step against its closed form, and
the host-side parameter layout of ptylab_amd.unet."""

import numpy as np
import pytest
import torch

from oracle import unet_ref as R
from ptylab_amd import unet as U


def test_conv_rows_match_brute_force():
    rng = np.random.default_rng(0)
    B, H, W, cin, cin_s, cout, ks = 2, 5, 6, 3, 8, 4, 3
    x = np.zeros((B, H, W, cin_s))
    x[..., :cin] = rng.standard_normal((B, H, W, cin))
    Kp = U.kpad(ks, cin_s)
    rows = np.zeros((cout, Kp))
    rows[:, :ks * ks * cin_s] = rng.standard_normal((cout, ks * ks * cin_s))
    b = rng.standard_normal(cout)
    R.SHAPE["_t"] = (cin, cin_s, cout, ks, False)
    try:
        got = R.conv_nhwc(torch.as_tensor(x), R.weight_tensor(rows, "_t"), torch.as_tensor(b), ks).numpy()
    finally:
        del R.SHAPE["_t"]
    ref = np.zeros((B, H, W, cout))
    for n in range(B):
        for y in range(H):
            for xx in range(W):
                for co in range(cout):
                    s = b[co]
                    for ky in range(ks):
                        for kx in range(ks):
                            yy, xs = y + ky - 1, xx + kx - 1
                            if 0 <= yy < H and 0 <= xs < W:
                                for ci in range(cin_s):
                                    s += rows[co, (ky * ks + kx) * cin_s + ci] * x[n, yy, xs, ci]
                    ref[n, y, xx, co] = s
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_oracle_layers_match_the_package_architecture():
    assert [(s.name, s.cin, s.cin_s, s.cout, s.ks, s.relu) for s in U.ARCH] == [tuple(x) for x in R.LAYERS]
    assert 1.8e6 < U.n_params() < 2.1e6


def test_autograd_matches_finite_differences():
    # a 16 x 16 input keeps every level (16 -> 2 at the bottom) cheap in float64
    rng = np.random.default_rng(1)
    params = {k: (w.astype(np.float64), b.astype(np.float64) + 0.01) for k, (w, b) in U.init_params(3).items()}
    x = np.zeros((1, 16, 16, 8))
    x[..., 0] = rng.random((1, 16, 16))
    tgt = rng.standard_normal((1, 8, 8, 2))
    loss, grads, _, _ = R.loss_and_grads(params, x, tgt, mirror_bf16=False)
    for name, (r, c) in (("e0a", (5, 9)), ("e2b", (7, 100)), ("d1a", (3, 40)), ("head", (1, 17))):
        h = 1e-5
        pp = {k: (w.copy(), b.copy()) for k, (w, b) in params.items()}
        pm = {k: (w.copy(), b.copy()) for k, (w, b) in params.items()}
        pp[name][0][r, c] += h
        pm[name][0][r, c] -= h
        lp = R.loss_and_grads(pp, x, tgt, mirror_bf16=False)[0]
        lm = R.loss_and_grads(pm, x, tgt, mirror_bf16=False)[0]
        fd = (lp - lm) / (2 * h)
        assert grads[name][0][r, c] == pytest.approx(fd, rel=1e-5, abs=1e-10), name
    # bias gradient of the head
    h = 1e-5
    pp = {k: (w.copy(), b.copy()) for k, (w, b) in params.items()}
    pp["head"][1][1] += h
    pm = {k: (w.copy(), b.copy()) for k, (w, b) in params.items()}
    pm["head"][1][1] -= h
    fd = (R.loss_and_grads(pp, x, tgt, mirror_bf16=False)[0] - R.loss_and_grads(pm, x, tgt, mirror_bf16=False)[0]) / (2 * h)
    assert grads["head"][1][1] == pytest.approx(fd, rel=1e-5)


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([0.3, -2.0, 1e-3])
    w, m, v = R.adam_ref(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    np.testing.assert_allclose(w, -0.01 * np.sign(g), rtol=1e-12)


def test_padding_weights_start_at_zero():
    p = U.init_params(0)
    for s in U.ARCH:
        w, b = p[s.name]
        assert w.shape == (s.cout, s.Kp)
        assert not np.any(w[~U.valid_input_mask(s)])
        assert np.all(b == 0)
