"""Learned initialiser: a U-Net maps each diffraction pattern to an estimate
of the object patch under the probe; the patches, stitched by probe-weighted
averaging, replace the reference's 'ones' initial object
(initialProbeOrObject, utils/initializationFunctions.py:10-62, used by
Reconstruction.initializeObject, Reconstruction.py:356-374) and mPIE refines
from there.  Plain uniform-init mPIE stays the fallback and the yardstick.

Workload (the learned-initialiser configuration): probe Np = 128 (the
plain-ePIE family's aperture probe, radius 32 px, quadratic phase,
propagated one Fresnel length), objects of the smoothed random
amplitude/phase type, Fraunhofer data at 1e4 photons peak with Poisson
noise.  Network input (features()): the measured amplitudes back-propagated
with the probe's far-field phase and divided by the probe, which puts a
first object estimate in real space, aligned with the output; the probe
intensity; log(1 + I) / 10.  Output: the central 64 x 64 object patch (the
illuminated disc) divided by its complex mean (the global phase and scale
are not determined by one pattern), as real and imaginary channels.

Training data come from a GPU simulator that uses the hot path's own
Fraunhofer propagator (pty_propagate) on random objects; the held-out scan
uses a different seed.  NRMSE (project-defined; the reference has none):
min over complex gamma of ||gamma O_rec - O_true|| / ||O_true|| on the
scanned region (union of the probe footprints above 1 % of peak intensity).
"""

from __future__ import annotations

import math
import sys
import time

import numpy as np

from . import _native, synthetic
from . import unet as U

WAVELENGTH = synthetic.WAVELENGTH
NP, NO, F_SCAN, RADIUS, DXO, ZO, PEAK, SIGMA = 128, 512, 500, 32, 4e-6, 0.05, 1e4, 3.0
PATCH, OFF = 64, 32  # the predicted patch: rows/cols [OFF, OFF + PATCH) of the probe window


def probe():
    """The workload's probe (complex64 NumPy)."""
    return synthetic.aperture_probe(NP, RADIUS, DXO, WAVELENGTH)


def held_out_scan(seed=4242, F=F_SCAN):
    """The held-out ptychography dataset (object seed differs from every
    training seed): plain-ePIE-family inputs at Np = 128, No = 512."""
    lim = (NO - NP) // 2 - 1
    step = min(0.25 * NP, 0.9 * lim * math.sqrt(math.pi / F))
    return synthetic.cpm_fraunhofer(Np=NP, No=NO, F=F, seed=seed, step_frac=step / NP, peak=PEAK, dxo=DXO, zo=ZO)


# ---------------------------------------------------------------------------
# GPU simulator of training pairs
# ---------------------------------------------------------------------------
class Simulator:
    """Random smoothed objects on a canvas, 128-px windows at random
    positions, exit wave P * O, Fraunhofer propagation with the hot path's
    pty_propagate kernel, intensities scaled to `peak` photons at the batch
    maximum, Poisson noise.  Yields (x (B,128,128,8) bf16, target (B,64,64,2)
    fp32) on the GPU."""

    def __init__(self, probe_np, seed=0, canvas=512, windows_per_object=16, peak=PEAK, sigma=SIGMA):
        import torch

        self.dev = torch.device("cuda")
        self.gen = torch.Generator(device=self.dev)
        self.gen.manual_seed(int(seed))
        self.P = torch.as_tensor(np.asarray(probe_np, np.complex64), device=self.dev)
        self.feat = Features(probe_np)
        self.canvas, self.wpo, self.peak = canvas, windows_per_object, peak
        n = canvas + 8 * int(sigma) * 2
        self.pad = 8 * int(sigma)
        f = torch.fft.fftfreq(n, device=self.dev, dtype=torch.float64)
        g1 = torch.exp(-2.0 * (math.pi * sigma * f) ** 2)  # FT of a unit-sum Gaussian of std sigma px
        self.G = (g1[:, None] * g1[None, :]).to(torch.complex64)
        self.n = n

    def objects(self, k):
        """k objects (k, canvas, canvas) complex64: amplitude U[0.5, 1] and
        phase U[-pi, pi], each smoothed by a Gaussian of std sigma (the
        smooth_random_object recipe; periodic smoothing on a padded canvas
        stands in for scipy's reflect mode)."""
        import torch

        u = torch.rand((2, k, self.n, self.n), generator=self.gen, device=self.dev)
        amp = 0.5 + 0.5 * u[0]
        ph = (2 * u[1] - 1) * math.pi
        sm = lambda a: torch.fft.ifft2(torch.fft.fft2(a.to(torch.complex64)) * self.G).real  # noqa: E731
        a, p = sm(amp), sm(ph)
        c0, c1 = self.pad, self.pad + self.canvas
        return torch.polar(a[:, c0:c1, c0:c1].float(), p[:, c0:c1, c0:c1].float())

    def batch(self, B):
        import torch

        k = -(-B // self.wpo)
        obj = self.objects(k)
        lim = self.canvas - NP
        pos = torch.randint(0, lim + 1, (B, 2), generator=self.gen, device=self.dev)
        idx = torch.arange(NP, device=self.dev)
        rows = (pos[:, 0:1] + idx[None, :])[:, :, None]
        cols = (pos[:, 1:2] + idx[None, :])[:, None, :]
        which = (torch.arange(B, device=self.dev) // self.wpo)[:, None, None]
        win = obj[which, rows, cols]  # (B, 128, 128)
        x, tgt = self.pairs(win)
        return x, tgt

    def pairs(self, win):
        """Training pairs from object windows (B, 128, 128) complex64."""
        import torch

        B = win.shape[0]
        exit_wave = (win * self.P[None]).contiguous()
        _native.propagate(exit_wave, _native.PTY_FRAUNHOFER, inverse=False)
        I = exit_wave.real ** 2 + exit_wave.imag ** 2
        I = I * (self.peak / I.max())
        I = torch.poisson(I, generator=self.gen)
        return self.feat(I), patch_targets(win)


class Features:
    """Network input from measured intensities (B, 128, 128) on the GPU:

    0, 1  real, imaginary of O_est = psi0 conj(P) / (|P|^2 + 0.05 max |P|^2),
          psi0 = ifft2c(sqrt(I) exp(i arg fft2c(P))) / s, s^2 = sum I / sum |P|^2
          (the measured amplitudes with the probe's far-field phase, back in
          the object plane with the hot path's Fraunhofer propagator)
    2     |P|^2 / max |P|^2
    3     log(1 + I) / 10
    """

    def __init__(self, probe_np):
        import torch

        self.P = torch.as_tensor(np.asarray(probe_np, np.complex64), device="cuda")
        far = self.P.clone().contiguous()
        _native.propagate(far, _native.PTY_FRAUNHOFER, inverse=False)
        self.phase = (far / far.abs().clamp_min(1e-30)).contiguous()
        i2 = self.P.abs() ** 2
        self.pw = float(i2.sum())
        self.den = i2 + 0.05 * i2.max()
        self.pint = (i2 / i2.max()).to(torch.bfloat16)

    def __call__(self, I):
        import torch

        B = I.shape[0]
        amp = torch.sqrt(I.clamp_min(0))
        psi = (amp * self.phase[None]).to(torch.complex64).contiguous()
        _native.propagate(psi, _native.PTY_FRAUNHOFER, inverse=True)
        s = torch.sqrt(I.sum(dim=(1, 2)) / self.pw).clamp_min(1e-12)
        est = psi * self.P.conj()[None] / self.den[None] / s[:, None, None]
        x = torch.zeros((B, NP, NP, 8), dtype=torch.bfloat16, device=I.device)
        x[..., 0] = est.real.to(torch.bfloat16)
        x[..., 1] = est.imag.to(torch.bfloat16)
        x[..., 2] = self.pint[None]
        x[..., 3] = (torch.log1p(I) / 10.0).to(torch.bfloat16)
        return x


def patch_targets(win):
    """(B, 64, 64, 2) fp32 targets: the central patch over its complex mean."""
    import torch

    c = win[:, OFF:OFF + PATCH, OFF:OFF + PATCH]
    m = c.mean(dim=(1, 2), keepdim=True)
    t = c / m
    return torch.stack([t.real, t.imag], dim=-1).float().contiguous()


def train(net, steps, seed=0, lr=1e-3, lr_final=1e-4, log_every=0, sim=None):
    """Adam on simulated pairs with a cosine learning-rate decay; returns the
    per-step losses (NumPy, synchronised at the end)."""
    import torch

    sim = sim if sim is not None else Simulator(probe(), seed=seed)
    losses = []
    for s in range(steps):
        x, t = sim.batch(net.B)
        lr_s = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * s / max(1, steps)))
        losses.append(net.train_step(x, t, lr=lr_s))
        if log_every and (s + 1) % log_every == 0:
            print(f"[train] step {s + 1}/{steps} loss {float(losses[-1]):.4e}", file=sys.stderr, flush=True)
    return torch.stack(losses).cpu().numpy()


# ---------------------------------------------------------------------------
# inference, stitching, NRMSE
# ---------------------------------------------------------------------------
def predict(net, ptychogram, probe_np):
    """Patch estimates (F, 64, 64) complex64 on the GPU, batches of net.B.
    ``ptychogram``: (F, 128, 128) intensities, NumPy or a device tensor."""
    import torch

    I = torch.as_tensor(ptychogram, dtype=torch.float32, device="cuda")
    x = Features(probe_np)(I)
    F = x.shape[0]
    out = torch.empty((F, PATCH, PATCH), dtype=torch.complex64, device="cuda")
    B = net.B
    for b0 in range(0, F, B):
        n = min(B, F - b0)
        xb = x[b0:b0 + n]
        if n < B:
            xb = torch.cat([xb, torch.zeros((B - n,) + tuple(x.shape[1:]), dtype=x.dtype, device=x.device)])
        net.forward(xb)
        out[b0:b0 + n] = net.prediction()[:n]
    return out


def stitch(patches, positions, probe_np, No):
    """Probe-weighted average of the patches into a (No, No) object.

    Every patch carries its own arbitrary complex scale, so patches are
    placed in rings of growing size around the scan centre (1, 4, 8, 16, 32,
    then 64 at a time, by distance); each patch of a ring is first scaled by
    the complex gamma that best matches it (weighted least squares) to what
    the inner rings already stitched, then the ring is added with one
    scatter-add.  Pixels no patch covers keep the 'ones' value; the result is
    scaled to unit mean amplitude on the covered region."""
    import torch

    dev = patches.device
    F = patches.shape[0]
    P = torch.as_tensor(np.asarray(probe_np, np.complex64), device=dev)
    w = (P.abs() ** 2)[OFF:OFF + PATCH, OFF:OFF + PATCH].to(torch.float32)
    w = w / w.max()
    pos = np.asarray(positions, np.int64) + OFF
    c = pos.mean(axis=0)
    order = np.argsort(((pos - c) ** 2).sum(axis=1), kind="stable")
    ar = torch.arange(PATCH, device=dev)
    pt = torch.as_tensor(pos, device=dev)
    idx = ((pt[:, 0, None, None] + ar[None, :, None]) * No + pt[:, 1, None, None] + ar[None, None, :])  # (F, 64, 64)
    num = torch.zeros(No * No, dtype=torch.complex64, device=dev)
    den = torch.zeros(No * No, dtype=torch.float32, device=dev)
    sizes, start = [1, 4, 8, 16, 32], 0
    groups = []
    for n in sizes:
        groups.append(order[start:start + n])
        start += n
    while start < F:
        groups.append(order[start:start + 64])
        start += 64
    wf = w[None]
    for g in groups:
        if len(g) == 0:
            continue
        ids = torch.as_tensor(g, device=dev)
        T = patches[ids]
        ix = idx[ids]
        nv, dv = num[ix], den[ix]
        d = (wf * dv * (T.abs() ** 2)).sum(dim=(1, 2))
        gam = (wf * T.conj() * nv).sum(dim=(1, 2)) / d.clamp_min(1e-12)
        gam = torch.where(d > 1e-6, gam, torch.ones_like(gam))
        num.index_add_(0, ix.reshape(-1), (wf * gam[:, None, None] * T).reshape(-1))
        den.index_add_(0, ix.reshape(-1), wf.expand(len(g), -1, -1).reshape(-1))
    num, den = num.view(No, No), den.view(No, No)
    cover = den > 0.05
    obj = torch.ones((No, No), dtype=torch.complex64, device=dev)
    obj[cover] = num[cover] / den[cover]
    scale = obj[cover].abs().mean()
    obj[cover] = obj[cover] / scale
    return obj


def roi_mask(positions, probe_np, No):
    """Union of the probe footprints above 1 % of peak intensity (NumPy bool)."""
    I = np.abs(np.asarray(probe_np)) ** 2
    fp = I > 0.01 * I.max()
    m = np.zeros((No, No), dtype=bool)
    Np = fp.shape[0]
    for r, c in np.asarray(positions, np.int64):
        m[r:r + Np, c:c + Np] |= fp
    return m


def nrmse_gpu(obj, true, roi):
    """min_gamma ||gamma obj - true|| / ||true|| over roi (torch, complex128 sums)."""
    import torch

    o = obj[roi].to(torch.complex128)
    t = true[roi].to(torch.complex128)
    g = (o.conj() * t).sum() / (o.abs() ** 2).sum()
    return torch.linalg.vector_norm(g * o - t) / torch.linalg.vector_norm(t)


def run_to_target(engine, true_obj, roi, target=1e-2, max_iter=300):
    """Run the engine one iteration at a time until the object NRMSE is at or
    below `target`.  Returns (iterations, seconds, nrmse curve); seconds count
    from the call (GPU work and the per-iteration NRMSE check included)."""
    import torch

    t0 = time.perf_counter()
    engine.numIterations = max_iter
    engine._prepareReconstruction()
    plan, T = engine._begin()
    curve, done, hit = [], 0, None
    try:
        for it in range(max_iter):
            plan.run(it, it + 1)
            done = it + 1
            e = float(nrmse_gpu(T["object"][0, 0, 0, 0], true_obj, roi))
            curve.append(e)
            if e <= target:
                hit = time.perf_counter() - t0
                break
    finally:
        engine._end(done)
    torch.cuda.synchronize()
    return (done if hit is not None else None), (hit if hit is not None else time.perf_counter() - t0), curve


def data_misfit(obj, ptychogram, positions, probe_np, uniform=False):
    """sum_j ||sqrt(I_j) - |fft2c(P O_j)|||^2 / sum_j I_j for an object guess
    (GPU, the hot path's Fraunhofer propagator): how well a candidate
    initial object explains the measured amplitudes with the known probe.
    ``uniform``: obj is all ones, so every exit wave is P and one transform,
    expanded over the positions, gives the same values.  ``ptychogram`` may
    be NumPy or a device tensor."""
    import torch

    I = torch.as_tensor(ptychogram, dtype=torch.float32, device=obj.device)
    P = torch.as_tensor(np.asarray(probe_np, np.complex64), device=obj.device)
    F = I.shape[0]
    if uniform:
        far = P.clone().contiguous()
        _native.propagate(far, _native.PTY_FRAUNHOFER, inverse=False)
        ex = far[None].expand(F, NP, NP)
    else:
        pos = torch.as_tensor(np.asarray(positions, np.int64), device=obj.device)
        ar = torch.arange(NP, device=obj.device)
        rows = (pos[:, 0, None] + ar[None])[:, :, None]
        cols = (pos[:, 1, None] + ar[None])[:, None, :]
        ex = (obj[rows, cols] * P[None]).contiguous()
        _native.propagate(ex, _native.PTY_FRAUNHOFER, inverse=False)
    s = torch.sqrt(I.sum() / (ex.abs() ** 2).sum())  # the data's global scale
    return float(((torch.sqrt(I.clamp_min(0)) - s * ex.abs()) ** 2).sum() / I.sum())


def initialize_object(reconstruction, experimentalData, net, probe_np, stride=1):
    """Replace reconstruction.object by the learned initial object when it
    explains the data better than the uniform guess does (data_misfit with
    the known probe, over every position); otherwise keep the uniform 'ones'
    object.  ``stride``: predict and stitch the patches of every stride-th
    position only (the scan overlaps many times over).  Returns a dict:
    choice ('learned' or 'ones'), both misfits and the seconds spent.  The
    object keeps the reconstruction's 6D shape and complex64 dtype."""
    import torch

    t0 = time.perf_counter()
    pos = np.asarray(reconstruction.positions, np.int64)
    No = int(reconstruction.No)
    I = torch.as_tensor(np.asarray(experimentalData.ptychogram, np.float32), device="cuda")
    sel = np.arange(0, pos.shape[0], max(1, int(stride)))
    patches = predict(net, I[torch.as_tensor(sel, device="cuda")] if stride > 1 else I, probe_np)
    obj = stitch(patches, pos[sel], probe_np, No)
    ones = torch.ones_like(obj)
    e_l = data_misfit(obj, I, pos, probe_np)
    e_o = data_misfit(ones, I, pos, probe_np, uniform=True)
    choice = "learned" if e_l < e_o else "ones"
    if choice == "learned":
        o = obj.cpu().numpy().astype(np.complex64)
        reconstruction.object = np.broadcast_to(o, reconstruction.shape_O).copy()
    torch.cuda.synchronize()
    return dict(choice=choice, misfit_learned=e_l, misfit_ones=e_o, stride=int(stride),
                seconds=time.perf_counter() - t0)
