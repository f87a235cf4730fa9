"""CPU float64 restatement of the learned initialiser's U-Net.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by ``tests/``
as the checker of the GPU network kernels, never by the product package.

The reference has no learned initialiser, so this restates the project's
own network (ptylab_amd/unet.py) with torch's CPU float64 operators
(``conv2d``, ``avg_pool2d``, nearest ``interpolate``, autograd for the
gradients): an independent implementation of the same layers, in NCHW,
against which the GPU's NHWC implicit-GEMM kernels are compared.

``mirror_bf16=True`` rounds the weights and every stored activation to
bfloat16 (round to nearest even) where the GPU stores bf16, so that forward
differences come only from fp32 accumulation order.  Gradients are always
exact float64 derivatives of that forward.

Parity pins: ``conv_nhwc`` (the weight-row layout (cout, (ky, kx, ci))) is
checked against a brute-force loop in tests/test_unet_oracle.py; the
autograd gradients against central finite differences in float64.
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as Fn

# (name, cin, cin_s, cout, ks, relu): the network of ptylab_amd/unet.py
LAYERS = (
    ("e0a", 4, 8, 32, 3, True), ("e0b", 32, 32, 32, 3, True),
    ("e1a", 32, 32, 64, 3, True), ("e1b", 64, 64, 64, 3, True),
    ("e2a", 64, 64, 128, 3, True), ("e2b", 128, 128, 128, 3, True),
    ("e3a", 128, 128, 256, 3, True), ("e3b", 256, 256, 256, 3, True),
    ("d2a", 384, 384, 128, 3, True), ("d2b", 128, 128, 128, 3, True),
    ("d1a", 192, 192, 64, 3, True), ("d1b", 64, 64, 64, 3, True),
    ("d0a", 96, 96, 32, 3, True), ("d0b", 32, 32, 32, 3, True),
    ("head", 32, 32, 2, 1, False),
)
SHAPE = {n: (cin, cin_s, cout, ks, relu) for n, cin, cin_s, cout, ks, relu in LAYERS}


def bf16(t):
    return t.to(torch.bfloat16).to(torch.float64)


def weight_tensor(w_rows, name):
    """(cout, Kp) rows in (ky, kx, ci) order -> torch (cout, cin_s, ks, ks)."""
    cin, cin_s, cout, ks, _ = SHAPE[name]
    w = torch.as_tensor(np.asarray(w_rows, dtype=np.float64))[:, :ks * ks * cin_s]
    return w.reshape(cout, ks, ks, cin_s).permute(0, 3, 1, 2).contiguous()


def conv_nhwc(x, w, b, ks):
    """x (B, H, W, C) float64 torch, w (cout, cin_s, ks, ks); 'same' padding."""
    y = Fn.conv2d(x.permute(0, 3, 1, 2), w, b, padding=ks // 2)
    return y.permute(0, 2, 3, 1)


def forward(params, x, mirror_bf16=True, requires_grad=False):
    """params {name: (w (cout, Kp), b (cout,))}; x (B, 128, 128, 8) array.
    Returns (out (B, 128, 128, 2), activations dict, leaf weight tensors)."""
    r = bf16 if mirror_bf16 else (lambda t: t)
    leaves = {}
    for name, *_ in LAYERS:
        w, b = params[name]
        wt = weight_tensor(w, name)
        if mirror_bf16:
            wt = bf16(wt)
        wt = wt.detach().requires_grad_(requires_grad)
        bt = torch.as_tensor(np.asarray(b, np.float64)).detach().requires_grad_(requires_grad)
        leaves[name] = (wt, bt)
    act = {}

    def conv(name, inp):
        cin, cin_s, cout, ks, relu = SHAPE[name]
        z = conv_nhwc(inp, *leaves[name], ks)
        if relu:
            z = torch.relu(z)
        act[name] = r(z) if name != "head" else z
        return act[name]

    def pool(t):
        return r(Fn.avg_pool2d(t.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1))

    def up(t):
        return Fn.interpolate(t.permute(0, 3, 1, 2), scale_factor=2, mode="nearest").permute(0, 2, 3, 1)

    x = torch.as_tensor(np.asarray(x, dtype=np.float64))
    h = conv("e0b", conv("e0a", x))
    s0 = h
    h = conv("e1b", conv("e1a", pool(h)))
    s1 = h
    h = conv("e2b", conv("e2a", pool(h)))
    s2 = h
    h = conv("e3b", conv("e3a", pool(h)))
    h = conv("d2b", conv("d2a", torch.cat([up(h), s2], dim=3)))
    h = conv("d1b", conv("d1a", torch.cat([up(h), s1], dim=3)))
    h = conv("d0b", conv("d0a", torch.cat([up(h), s0], dim=3)))
    out = conv("head", h)
    return out, act, leaves


def loss_and_grads(params, x, target, mirror_bf16=True):
    """MSE of the output's central crop (the target's size, centred) against
    target and the gradients {name: (dw (cout, Kp) rows, db (cout,))} in
    float64."""
    out, act, leaves = forward(params, x, mirror_bf16=mirror_bf16, requires_grad=True)
    t = torch.as_tensor(np.asarray(target, np.float64))
    o0 = (out.shape[1] - t.shape[1]) // 2
    loss = ((out[:, o0:o0 + t.shape[1], o0:o0 + t.shape[2]] - t) ** 2).mean()
    loss.backward()
    grads = {}
    for name, (wt, bt) in leaves.items():
        cin, cin_s, cout, ks, _ = SHAPE[name]
        Kp = np.asarray(params[name][0]).shape[1]
        rows = np.zeros((cout, Kp))
        rows[:, :ks * ks * cin_s] = wt.grad.permute(0, 2, 3, 1).reshape(cout, -1).numpy()
        grads[name] = (rows, bt.grad.numpy().copy())
    return float(loss.detach()), grads, out.detach().numpy(), {k: v.detach().numpy() for k, v in act.items()}


def adam_ref(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
