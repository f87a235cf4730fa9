"""The learned initialiser's gfx950 kernels against the float64 CPU oracle
(oracle/unet_ref.py): the MFMA operand layout, each layer kind on ragged
shapes, the whole network's forward output and every layer's weight and
bias gradient, the Adam step, and a short training run.

Tolerances (bf16 operands, fp32 accumulation): forward layers <= 1e-2
relative (the oracle mirrors the bf16 storage points, so only accumulation
order and rare rounding flips differ); weight gradients <= 5e-2 relative
(the GPU stores every backward activation gradient in bf16, the oracle keeps
them exact)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import unet_ref as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _h():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _bf(x):
    return torch.as_tensor(np.asarray(x, np.float32)).to(torch.bfloat16).cuda().contiguous()


def test_mfma_operand_layout():
    from ptylab_amd import unet as U

    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    B = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    D = torch.zeros((32, 32), dtype=torch.float32, device="cuda")
    Ad, Bd = _bf(A), _bf(B)  # kept alive until the kernel has run
    U._check(U.load().pty_net_mfma_probe(Ad.data_ptr(), Bd.data_ptr(), D.data_ptr(), _h()))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(D.cpu().numpy(), A @ B)


@pytest.mark.parametrize("B,H,W,cin,cin_s,cout,ks,relu,gate", [
    (2, 16, 16, 3, 8, 40, 3, True, False),
    (1, 9, 7, 24, 24, 70, 3, False, True),
    (3, 8, 8, 64, 64, 2, 1, False, False),
    (2, 32, 32, 192, 192, 64, 3, True, False),
])
def test_conv_forward_input_and_weight_gradients(B, H, W, cin, cin_s, cout, ks, relu, gate):
    from ptylab_amd import unet as U

    rng = np.random.default_rng(B * 1000 + cout)
    lib, h = U.load(), _h()
    cout_s = (cout + 7) // 8 * 8
    Kp = U.kpad(ks, cin_s)
    x = np.zeros((B, H, W, cin_s), np.float32)
    x[..., :cin] = rng.standard_normal((B, H, W, cin))
    w = np.zeros((cout, Kp), np.float32)
    for t in range(ks * ks):
        w[:, t * cin_s:t * cin_s + cin] = rng.standard_normal((cout, cin)) / np.sqrt(ks * ks * cin)
    b = rng.standard_normal(cout).astype(np.float32)
    g = np.zeros((B, H, W, cout_s), np.float32)
    g[..., :cout] = rng.standard_normal((B, H, W, cout))
    xd, wd, gd = _bf(x), _bf(w), _bf(g)
    bd = torch.as_tensor(b).cuda()
    y = torch.zeros((B, H, W, cout_s), dtype=torch.bfloat16, device="cuda")
    U._check(lib.pty_net_conv_fwd(xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), gd.data_ptr() if gate else None,
                                  y.data_ptr(), B, H, W, cin_s, cout, cout_s, ks, int(relu), 0, h))
    # oracle with the same bf16-rounded operands
    xr = R.bf16(torch.as_tensor(x, dtype=torch.float64))
    R.SHAPE["_t"] = (cin, cin_s, cout, ks, relu)
    try:
        wt = R.bf16(R.weight_tensor(w, "_t"))
    finally:
        del R.SHAPE["_t"]
    ref = R.conv_nhwc(xr, wt, torch.as_tensor(b, dtype=torch.float64), ks)
    if relu:
        ref = torch.relu(ref)
    if gate:
        ref = ref * (torch.as_tensor(g[..., :cout]).to(torch.bfloat16).double() > 0)
    got = y.float().cpu().numpy()
    assert _rel(got[..., :cout], ref.numpy()) < 1e-2
    assert not np.any(got[..., cout:])
    # weight and bias gradients for dz = g
    dz = gd
    splits = 7
    part = torch.zeros(splits * cout * Kp, dtype=torch.float32, device="cuda")
    dw = torch.zeros((cout, Kp), dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    U._check(lib.pty_net_conv_wgrad(xd.data_ptr(), dz.data_ptr(), part.data_ptr(), splits, dw.data_ptr(), db.data_ptr(),
                                    B, H, W, cin_s, cout, cout_s, ks, h))
    xt = xr.clone().requires_grad_(True)
    wtt = wt.clone().requires_grad_(True)
    bt = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    z = R.conv_nhwc(xt, wtt, bt, ks)
    gz = R.bf16(torch.as_tensor(g[..., :cout], dtype=torch.float64))
    (z * gz).sum().backward()
    dw_ref = wtt.grad.permute(0, 2, 3, 1).reshape(cout, -1).numpy()
    assert _rel(dw.cpu().numpy()[:, :ks * ks * cin_s], dw_ref) < 1e-4
    assert not np.any(dw.cpu().numpy()[:, ks * ks * cin_s:])
    assert _rel(db.cpu().numpy(), bt.grad.numpy()) < 1e-4
    # input gradient through the flipped weights
    wflip = torch.zeros((cin, U.kpad(ks, cout_s)), dtype=torch.bfloat16, device="cuda")
    U._check(lib.pty_net_weight_flip(wd.data_ptr(), cin, cin_s, cout, cout_s, ks, wflip.data_ptr(), h))
    dx = torch.zeros((B, H, W, cin_s), dtype=torch.bfloat16, device="cuda")
    U._check(lib.pty_net_conv_fwd(dz.data_ptr(), wflip.data_ptr(), None, None, dx.data_ptr(), B, H, W, cout_s, cin,
                                  cin_s, ks, 0, 0, h))
    assert _rel(dx.float().cpu().numpy()[..., :cin], xt.grad.numpy()[..., :cin]) < 1e-2


def test_pool_upsample_concat_and_their_backward():
    from ptylab_amd import unet as U

    rng = np.random.default_rng(5)
    lib, h = U.load(), _h()
    B, H, W, ca, cb = 2, 8, 12, 16, 8
    a = rng.standard_normal((B, H // 2, W // 2, ca)).astype(np.float32)
    bsk = rng.standard_normal((B, H, W, cb)).astype(np.float32)
    ad, bd = _bf(a), _bf(bsk)
    y = torch.zeros((B, H, W, ca + cb), dtype=torch.bfloat16, device="cuda")
    U._check(lib.pty_net_up_concat(ad.data_ptr(), ca, bd.data_ptr(), cb, y.data_ptr(), B, H, W, h))
    ar = ad.float().cpu().numpy()
    up = ar.repeat(2, axis=1).repeat(2, axis=2)
    np.testing.assert_array_equal(y.float().cpu().numpy(), np.concatenate([up, bd.float().cpu().numpy()], axis=3))
    p = torch.zeros((B, H // 2, W // 2, cb), dtype=torch.bfloat16, device="cuda")
    U._check(lib.pty_net_pool_fwd(bd.data_ptr(), p.data_ptr(), B, H, W, cb, h))
    br = bd.float().cpu().numpy().astype(np.float64)
    pref = 0.25 * (br[:, 0::2, 0::2] + br[:, 0::2, 1::2] + br[:, 1::2, 0::2] + br[:, 1::2, 1::2])
    assert _rel(p.float().cpu().numpy(), pref) < 4e-3
    dy = _bf(rng.standard_normal((B, H, W, ca + cb)))
    dxa = torch.zeros_like(ad)
    dxb = torch.zeros_like(bd)
    U._check(lib.pty_net_up_concat_bwd(dy.data_ptr(), ca, cb, dxa.data_ptr(), None, dxb.data_ptr(), B, H, W, h))
    d = dy.float().cpu().numpy().astype(np.float64)
    da_ref = d[:, 0::2, 0::2, :ca] + d[:, 0::2, 1::2, :ca] + d[:, 1::2, 0::2, :ca] + d[:, 1::2, 1::2, :ca]
    assert _rel(dxa.float().cpu().numpy(), da_ref) < 4e-3
    np.testing.assert_array_equal(dxb.float().cpu().numpy(), d[..., ca:])
    dp = _bf(rng.standard_normal((B, H // 2, W // 2, cb)))
    dx = torch.zeros_like(bd)
    U._check(lib.pty_net_pool_bwd(dp.data_ptr(), dx.data_ptr(), B, H, W, cb, h))
    np.testing.assert_array_equal(dx.float().cpu().numpy(),
                                  0.25 * dp.float().cpu().numpy().repeat(2, axis=1).repeat(2, axis=2))


@pytest.fixture(scope="module")
def batch2():
    from ptylab_amd import learned_init as LI

    torch.manual_seed(0)
    sim = LI.Simulator(LI.probe(), seed=11)
    return sim.batch(2)


def test_network_forward_and_gradients_against_oracle(batch2):
    from ptylab_amd import unet as U

    x, tgt = batch2
    params = U.init_params(7)
    params = {k: (w, b + np.float32(0.01) * np.arange(len(b), dtype=np.float32) / len(b)) for k, (w, b) in params.items()}
    net = U.UNet(2, params=params)
    out = net.forward(x).cpu().numpy()
    loss_gpu = float(net.loss_and_grad(tgt))
    net.backward()
    torch.cuda.synchronize()
    loss, grads, out_ref, act = R.loss_and_grads(params, x.float().cpu().numpy(), tgt.cpu().numpy())
    eo = _rel(out, out_ref)
    print(f"forward output rel {eo:.2e}; loss gpu {loss_gpu:.6e} oracle {loss:.6e}", flush=True)
    assert eo < 1e-2
    assert loss_gpu == pytest.approx(loss, rel=2e-2)
    for s in U.ARCH:
        gw = net.dw[s.name].cpu().numpy()
        gb = net.db[s.name].cpu().numpy()
        ew, eb = _rel(gw, grads[s.name][0]), _rel(gb, grads[s.name][1])
        print(f"  {s.name}: dW rel {ew:.2e}  db rel {eb:.2e}", flush=True)
        assert ew < 5e-2 and eb < 5e-2, s.name


def test_adam_step_against_oracle():
    from ptylab_amd import unet as U

    rng = np.random.default_rng(3)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    wd, gd = torch.as_tensor(w).cuda(), torch.as_tensor(g).cuda()
    m, v = torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    w16 = torch.zeros(n, dtype=torch.bfloat16, device="cuda")
    wr, mr, vr = w.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in (1, 2, 3):
        U._check(U.load().pty_net_adam(wd.data_ptr(), w16.data_ptr(), gd.data_ptr(), m.data_ptr(), v.data_ptr(), n,
                                       1e-3, 0.9, 0.999, 1e-8, step, _h()))
        wr, mr, vr = R.adam_ref(wr, g.astype(np.float64), mr, vr, step)
    np.testing.assert_allclose(wd.cpu().numpy(), wr, rtol=1e-5, atol=1e-7)
    np.testing.assert_array_equal(w16.cpu().float().numpy(), wd.cpu().to(torch.bfloat16).float().numpy())


def test_training_lowers_the_loss():
    from ptylab_amd import learned_init as LI
    from ptylab_amd import unet as U

    net = U.UNet(16, seed=1)
    losses = LI.train(net, 60, seed=5, lr=1e-3)
    first, last = float(np.mean(losses[:5])), float(np.mean(losses[-5:]))
    print(f"training loss {first:.4e} -> {last:.4e}", flush=True)
    assert np.all(np.isfinite(losses))
    assert last < 0.7 * first
