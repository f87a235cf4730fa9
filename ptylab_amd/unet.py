"""U-Net of the learned initialiser on the gfx950 kernels of libptynet.so.

Architecture (input: 4 feature channels on the 128 x 128 probe window;
output: the object over the window, real and imaginary; the loss uses the
central 64 x 64 crop, the illuminated disc):

    level  resolution  convolutions (3x3, ReLU)        then
    e0     128^2       4 -> 32, 32 -> 32    (skip s0)   pool 2x2
    e1     64^2        32 -> 64, 64 -> 64   (skip s1)   pool
    e2     32^2        64 -> 128, 128 -> 128 (skip s2)  pool
    e3     16^2        128 -> 256, 256 -> 256
    d2     32^2        up(e3) | s2 : 384 -> 128, 128 -> 128
    d1     64^2        up(d2) | s1 : 192 -> 64, 64 -> 64
    d0     128^2       up(d1) | s0 : 96 -> 32, 32 -> 32
    head   128^2       1x1 conv 32 -> 2, no activation

1.95 M parameters.  Activations and gradients are bf16 NHWC, weights have
fp32 masters and bf16 copies; all convolutions run as MFMA implicit GEMMs.
Input channels (see learned_init.features): the measured amplitudes
back-propagated with the probe's far-field phase and divided by the probe
(real, imaginary), the probe intensity, log(1 + I) / 10; channels 4..7 are
zero padding.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libptynet.so"

_lib = None


def load(path=None):
    """Load libptynet.so (after torch, so both use torch's HIP runtime)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import ptylab_amd.build as b; b.build()'`")
    import torch  # noqa: F401

    lib = ctypes.CDLL(str(p))
    vp, i32, sz, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_float
    lib.pty_net_last_error.restype = ctypes.c_char_p
    sig = {
        "pty_net_kpad": [i32, i32],
        "pty_net_mfma_probe": [vp, vp, vp, vp],
        "pty_net_conv_fwd": [vp, vp, vp, vp, vp] + [i32] * 9 + [vp],
        "pty_net_conv_wgrad": [vp, vp, vp, i32, vp, vp] + [i32] * 7 + [vp],
        "pty_net_weight_flip": [vp, i32, i32, i32, i32, i32, vp, vp],
        "pty_net_pool_fwd": [vp, vp, i32, i32, i32, i32, vp],
        "pty_net_pool_bwd": [vp, vp, i32, i32, i32, i32, vp],
        "pty_net_up_concat": [vp, i32, vp, i32, vp, i32, i32, i32, vp],
        "pty_net_up_concat_bwd": [vp, i32, i32, vp, vp, vp, i32, i32, i32, vp],
        "pty_net_add_gate": [vp, vp, vp, vp, sz, vp],
        "pty_net_mse": [vp, vp] + [i32] * 8 + [vp, vp, i32, vp],
        "pty_net_adam": [vp, vp, vp, vp, vp, sz, f32, f32, f32, f32, i32, vp],
        "pty_net_cast_bf16": [vp, vp, sz, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


EXPORTED = (
    "pty_net_last_error", "pty_net_kpad", "pty_net_mfma_probe", "pty_net_conv_fwd", "pty_net_conv_wgrad",
    "pty_net_weight_flip", "pty_net_pool_fwd", "pty_net_pool_bwd", "pty_net_up_concat", "pty_net_up_concat_bwd",
    "pty_net_add_gate", "pty_net_mse", "pty_net_adam", "pty_net_cast_bf16",
)


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"{load().pty_net_last_error().decode()} (code {rc})")


def kpad(ks, cin_s):
    """Weight row length: ks*ks*cin_s plus one spare column, rounded up to 32."""
    return (ks * ks * cin_s + 1 + 31) // 32 * 32


@dataclass(frozen=True)
class ConvSpec:
    name: str
    cin: int      # real input channels
    cin_s: int    # input channel stride
    cout: int
    ks: int
    res: int      # spatial size (square)
    relu: bool = True

    @property
    def cout_s(self):
        return (self.cout + 7) // 8 * 8

    @property
    def Kp(self):
        return kpad(self.ks, self.cin_s)


IN_RES, IN_CH, N_FEATURES = 128, 8, 4
CROP_OFF, CROP = 32, 64  # the loss window inside the 128 x 128 output

ARCH = (
    ConvSpec("e0a", N_FEATURES, 8, 32, 3, 128), ConvSpec("e0b", 32, 32, 32, 3, 128),
    ConvSpec("e1a", 32, 32, 64, 3, 64), ConvSpec("e1b", 64, 64, 64, 3, 64),
    ConvSpec("e2a", 64, 64, 128, 3, 32), ConvSpec("e2b", 128, 128, 128, 3, 32),
    ConvSpec("e3a", 128, 128, 256, 3, 16), ConvSpec("e3b", 256, 256, 256, 3, 16),
    ConvSpec("d2a", 384, 384, 128, 3, 32), ConvSpec("d2b", 128, 128, 128, 3, 32),
    ConvSpec("d1a", 192, 192, 64, 3, 64), ConvSpec("d1b", 64, 64, 64, 3, 64),
    ConvSpec("d0a", 96, 96, 32, 3, 128), ConvSpec("d0b", 32, 32, 32, 3, 128),
    ConvSpec("head", 32, 32, 2, 1, 128, relu=False),
)
SPEC = {s.name: s for s in ARCH}


def valid_input_mask(spec):
    """(cout, Kp) bool: the weight entries that multiply a real input channel."""
    m = np.zeros((spec.cout, spec.Kp), dtype=bool)
    real = np.arange(spec.cin_s) < spec.cin
    for tap in range(spec.ks * spec.ks):
        m[:, tap * spec.cin_s:(tap + 1) * spec.cin_s] = real[None, :]
    return m


def init_params(seed=0):
    """He-normal weights (std sqrt(2 / fan_in)) on the real input channels,
    zero padding entries, zero biases: {name: (w (cout, Kp) f32, b (cout,) f32)}."""
    rng = np.random.default_rng(seed)
    out = {}
    for s in ARCH:
        fan_in = s.ks * s.ks * s.cin
        w = rng.standard_normal((s.cout, s.Kp)).astype(np.float32) * np.float32(np.sqrt(2.0 / fan_in))
        if not s.relu:
            w *= np.float32(0.1)
        w[~valid_input_mask(s)] = 0.0
        out[s.name] = (w, np.zeros(s.cout, np.float32))
    return out


def n_params():
    return int(sum(valid_input_mask(s).sum() + s.cout for s in ARCH))


class UNet:
    """Parameters, activations and gradients of one batch size on the GPU."""

    def __init__(self, batch, params=None, seed=0, device="cuda"):
        import torch

        self.lib = load()
        self.B = int(batch)
        self.dev = torch.device(device)
        self.stream = torch.cuda.current_stream()
        params = params if params is not None else init_params(seed)
        z16 = lambda *s: torch.zeros(*s, dtype=torch.bfloat16, device=self.dev)  # noqa: E731
        z32 = lambda *s: torch.zeros(*s, dtype=torch.float32, device=self.dev)  # noqa: E731
        self.w32, self.w16, self.wt16, self.b32 = {}, {}, {}, {}
        self.dw, self.db, self.mw, self.vw, self.mb, self.vb = {}, {}, {}, {}, {}, {}
        for s in ARCH:
            w, b = params[s.name]
            self.w32[s.name] = torch.as_tensor(np.ascontiguousarray(w, np.float32), device=self.dev).contiguous()
            self.b32[s.name] = torch.as_tensor(np.ascontiguousarray(b, np.float32), device=self.dev).contiguous()
            self.w16[s.name] = self.w32[s.name].to(torch.bfloat16).contiguous()
            self.wt16[s.name] = z16(s.cin, kpad(s.ks, s.cout_s))
            for d, shape in ((self.dw, w.shape), (self.mw, w.shape), (self.vw, w.shape)):
                d[s.name] = z32(*shape)
            for d in (self.db, self.mb, self.vb):
                d[s.name] = z32(s.cout)
        B = self.B
        A = {}
        A["x"] = z16(B, 128, 128, 8)
        for s in ARCH[:-1]:
            A[s.name] = z16(B, s.res, s.res, s.cout_s)
        A["p0"], A["p1"], A["p2"] = z16(B, 64, 64, 32), z16(B, 32, 32, 64), z16(B, 16, 16, 128)
        A["u2"], A["u1"], A["u0"] = z16(B, 32, 32, 384), z16(B, 64, 64, 192), z16(B, 128, 128, 96)
        A["out"] = z32(B, 128, 128, 2)
        self.A = A
        G = {k: torch.zeros_like(v) for k, v in A.items() if k not in ("x", "out")}
        G["out"] = z16(B, 128, 128, 8)
        G["skip0"], G["skip1"], G["skip2"] = z16(B, 128, 128, 32), z16(B, 64, 64, 64), z16(B, 32, 32, 128)
        G["pool"] = z16(B, 128, 128, 32)
        self.G = G
        self.splits = {s.name: self._splits(s) for s in ARCH}
        self.part = z32(max(self.splits[s.name] * s.cout * s.Kp for s in ARCH))
        self.partial = torch.zeros(512, dtype=torch.float64, device=self.dev)
        self.step_count = 0

    # -- helpers ------------------------------------------------------------------
    def _splits(self, s):
        tiles = -(-s.Kp // 64) * -(-s.cout // 64)
        M = self.B * s.res * s.res
        return int(max(1, min(M // 256, max(1, 2048 // tiles))))

    def _h(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def _conv(self, name, x, y, gate=None, w16=None, cin_s=None, cout=None, cout_s=None, relu=None, out_f32=False):
        s = SPEC[name]
        _check(self.lib.pty_net_conv_fwd(
            x.data_ptr(), (w16 if w16 is not None else self.w16[name]).data_ptr(),
            None if w16 is not None else self.b32[name].data_ptr(),
            None if gate is None else gate.data_ptr(), y.data_ptr(), self.B, s.res, s.res,
            cin_s if cin_s is not None else s.cin_s, cout if cout is not None else s.cout,
            cout_s if cout_s is not None else s.cout_s, s.ks,
            int(s.relu if relu is None else relu), int(out_f32), self._h()))

    # -- forward ------------------------------------------------------------------
    def forward(self, x=None):
        """x: (B, 128, 128, 8) bf16 (or None: self.A['x']).  Returns the
        (B, 64, 64, 2) fp32 output tensor."""
        A, L, h, B = self.A, self.lib, self._h(), self.B
        if x is not None:
            A["x"].copy_(x)
        self._conv("e0a", A["x"], A["e0a"])
        self._conv("e0b", A["e0a"], A["e0b"])
        _check(L.pty_net_pool_fwd(A["e0b"].data_ptr(), A["p0"].data_ptr(), B, 128, 128, 32, h))
        self._conv("e1a", A["p0"], A["e1a"])
        self._conv("e1b", A["e1a"], A["e1b"])
        _check(L.pty_net_pool_fwd(A["e1b"].data_ptr(), A["p1"].data_ptr(), B, 64, 64, 64, h))
        self._conv("e2a", A["p1"], A["e2a"])
        self._conv("e2b", A["e2a"], A["e2b"])
        _check(L.pty_net_pool_fwd(A["e2b"].data_ptr(), A["p2"].data_ptr(), B, 32, 32, 128, h))
        self._conv("e3a", A["p2"], A["e3a"])
        self._conv("e3b", A["e3a"], A["e3b"])
        _check(L.pty_net_up_concat(A["e3b"].data_ptr(), 256, A["e2b"].data_ptr(), 128, A["u2"].data_ptr(), B, 32, 32, h))
        self._conv("d2a", A["u2"], A["d2a"])
        self._conv("d2b", A["d2a"], A["d2b"])
        _check(L.pty_net_up_concat(A["d2b"].data_ptr(), 128, A["e1b"].data_ptr(), 64, A["u1"].data_ptr(), B, 64, 64, h))
        self._conv("d1a", A["u1"], A["d1a"])
        self._conv("d1b", A["d1a"], A["d1b"])
        _check(L.pty_net_up_concat(A["d1b"].data_ptr(), 64, A["e0b"].data_ptr(), 32, A["u0"].data_ptr(), B, 128, 128, h))
        self._conv("d0a", A["u0"], A["d0a"])
        self._conv("d0b", A["d0a"], A["d0b"])
        self._conv("head", A["d0b"], A["out"], cout_s=2, out_f32=True)
        return A["out"]

    def prediction(self):
        """The central crop of the last output, (B, 64, 64) complex64."""
        import torch

        o = self.A["out"][:, CROP_OFF:CROP_OFF + CROP, CROP_OFF:CROP_OFF + CROP]
        return torch.complex(o[..., 0], o[..., 1])

    # -- backward -----------------------------------------------------------------
    def _wgrad(self, name, x, dz):
        s = SPEC[name]
        _check(self.lib.pty_net_conv_wgrad(
            x.data_ptr(), dz.data_ptr(), self.part.data_ptr(), self.splits[name], self.dw[name].data_ptr(),
            self.db[name].data_ptr(), self.B, s.res, s.res, s.cin_s, s.cout, s.cout_s, s.ks, self._h()))

    def _dgrad(self, name, dz, dx, gate):
        """dx = input gradient of layer `name` (gated by gate > 0 if given)."""
        s = SPEC[name]
        _check(self.lib.pty_net_weight_flip(self.w16[name].data_ptr(), s.cin, s.cin_s, s.cout, s.cout_s, s.ks,
                                            self.wt16[name].data_ptr(), self._h()))
        self._conv(name, dz, dx, gate=gate, w16=self.wt16[name], cin_s=s.cout_s, cout=s.cin, cout_s=s.cin_s,
                   relu=False)

    def loss_and_grad(self, target):
        """MSE of the central 64 x 64 crop against target (B, 64, 64, 2) fp32;
        fills G['out'].  Returns the loss as a 0-d FP64 tensor (no host
        synchronisation)."""
        n = self.B * CROP * CROP * 2
        nb = min(512, max(1, n // 2048))
        _check(self.lib.pty_net_mse(self.A["out"].data_ptr(), target.data_ptr(), self.B, IN_RES, IN_RES, CROP_OFF,
                                    CROP_OFF, CROP, CROP, 8, self.G["out"].data_ptr(), self.partial.data_ptr(), nb,
                                    self._h()))
        return self.partial[:nb].sum() / n

    def backward(self):
        A, G, L, h, B = self.A, self.G, self.lib, self._h(), self.B
        self._wgrad("head", A["d0b"], G["out"])
        self._dgrad("head", G["out"], G["d0b"], A["d0b"])
        self._wgrad("d0b", A["d0a"], G["d0b"])
        self._dgrad("d0b", G["d0b"], G["d0a"], A["d0a"])
        self._wgrad("d0a", A["u0"], G["d0a"])
        self._dgrad("d0a", G["d0a"], G["u0"], A["u0"])
        _check(L.pty_net_up_concat_bwd(G["u0"].data_ptr(), 64, 32, G["d1b"].data_ptr(), None, G["skip0"].data_ptr(),
                                       B, 128, 128, h))
        self._wgrad("d1b", A["d1a"], G["d1b"])
        self._dgrad("d1b", G["d1b"], G["d1a"], A["d1a"])
        self._wgrad("d1a", A["u1"], G["d1a"])
        self._dgrad("d1a", G["d1a"], G["u1"], A["u1"])
        _check(L.pty_net_up_concat_bwd(G["u1"].data_ptr(), 128, 64, G["d2b"].data_ptr(), None, G["skip1"].data_ptr(),
                                       B, 64, 64, h))
        self._wgrad("d2b", A["d2a"], G["d2b"])
        self._dgrad("d2b", G["d2b"], G["d2a"], A["d2a"])
        self._wgrad("d2a", A["u2"], G["d2a"])
        self._dgrad("d2a", G["d2a"], G["u2"], A["u2"])
        _check(L.pty_net_up_concat_bwd(G["u2"].data_ptr(), 256, 128, G["e3b"].data_ptr(), None, G["skip2"].data_ptr(),
                                       B, 32, 32, h))
        self._wgrad("e3b", A["e3a"], G["e3b"])
        self._dgrad("e3b", G["e3b"], G["e3a"], A["e3a"])
        self._wgrad("e3a", A["p2"], G["e3a"])
        self._dgrad("e3a", G["e3a"], G["p2"], None)
        _check(L.pty_net_pool_bwd(G["p2"].data_ptr(), G["pool"].data_ptr(), B, 32, 32, 128, h))
        n = B * 32 * 32 * 128
        _check(L.pty_net_add_gate(G["pool"].data_ptr(), G["skip2"].data_ptr(), A["e2b"].data_ptr(), G["e2b"].data_ptr(),
                                  n, h))
        self._wgrad("e2b", A["e2a"], G["e2b"])
        self._dgrad("e2b", G["e2b"], G["e2a"], A["e2a"])
        self._wgrad("e2a", A["p1"], G["e2a"])
        self._dgrad("e2a", G["e2a"], G["p1"], None)
        _check(L.pty_net_pool_bwd(G["p1"].data_ptr(), G["pool"].data_ptr(), B, 64, 64, 64, h))
        n = B * 64 * 64 * 64
        _check(L.pty_net_add_gate(G["pool"].data_ptr(), G["skip1"].data_ptr(), A["e1b"].data_ptr(), G["e1b"].data_ptr(),
                                  n, h))
        self._wgrad("e1b", A["e1a"], G["e1b"])
        self._dgrad("e1b", G["e1b"], G["e1a"], A["e1a"])
        self._wgrad("e1a", A["p0"], G["e1a"])
        self._dgrad("e1a", G["e1a"], G["p0"], None)
        _check(L.pty_net_pool_bwd(G["p0"].data_ptr(), G["pool"].data_ptr(), B, 128, 128, 32, h))
        n = B * 128 * 128 * 32
        _check(L.pty_net_add_gate(G["pool"].data_ptr(), G["skip0"].data_ptr(), A["e0b"].data_ptr(), G["e0b"].data_ptr(),
                                  n, h))
        self._wgrad("e0b", A["e0a"], G["e0b"])
        self._dgrad("e0b", G["e0b"], G["e0a"], A["e0a"])
        self._wgrad("e0a", A["x"], G["e0a"])

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        h = self._h()
        for s in ARCH:
            n = s.cout * s.Kp
            _check(self.lib.pty_net_adam(self.w32[s.name].data_ptr(), self.w16[s.name].data_ptr(),
                                         self.dw[s.name].data_ptr(), self.mw[s.name].data_ptr(),
                                         self.vw[s.name].data_ptr(), n, lr, b1, b2, eps, self.step_count, h))
            _check(self.lib.pty_net_adam(self.b32[s.name].data_ptr(), None, self.db[s.name].data_ptr(),
                                         self.mb[s.name].data_ptr(), self.vb[s.name].data_ptr(), s.cout, lr, b1, b2,
                                         eps, self.step_count, h))

    def train_step(self, x, target, lr=1e-3):
        self.forward(x)
        loss = self.loss_and_grad(target)
        self.backward()
        self.adam_step(lr)
        return loss

    def params(self):
        """{name: (w (cout, Kp) f32, b (cout,) f32)} as NumPy."""
        return {s.name: (self.w32[s.name].cpu().numpy(), self.b32[s.name].cpu().numpy()) for s in ARCH}
