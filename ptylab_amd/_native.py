"""ctypes binding of ``libptyhip.so`` (declarations in include/pty_hip.h).

The library is built in-tree (``ptylab_amd/libptyhip.so``) by
``ptylab_amd.build.build()``.  There is no fallback: if the library is
missing or a GPU is absent every entry point raises.
"""

from __future__ import annotations

import atexit
import ctypes
import os
import sys
import weakref
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libptyhip.so"

PTY_ABI_VERSION = 4
PTY_FLAG_BATCHED = 1
PTY_FLAG_MQN_NADAM = 2
PTY_FLAG_MQN_MOMENTUM = 4
# error codes (include/pty_hip.h)
PTY_OK, PTY_EINVAL, PTY_ENOMEM, PTY_EDEVICE, PTY_ETIMEDOUT = 0, -22, -12, -5, -110
PTY_EPIE, PTY_MPIE, PTY_MULTIPIE, PTY_E3PIE, PTY_QNEWTON, PTY_MQNEWTON = 0, 1, 2, 3, 4, 5
PTY_FRAUNHOFER, PTY_FRAUNHOFER_SHIFTED, PTY_ASP, PTY_FRESNEL, PTY_POLYCHROME_ASP, PTY_SCALED_ASP = 0, 1, 2, 3, 4, 5
PTY_SCALED_POLYCHROME_ASP = 6
PTY_ETIMEDOUT = -110

ENGINE_IDS = {"ePIE": PTY_EPIE, "mPIE": PTY_MPIE, "multiPIE": PTY_MULTIPIE, "e3PIE": PTY_E3PIE, "qNewton": PTY_QNEWTON,
              "mqNewton": PTY_MQNEWTON, "pcPIE": PTY_MPIE}

EXPORTED = (
    "pty_abi_version",
    "pty_last_error",
    "pty_plan_create",
    "pty_plan_run",
    "pty_plan_run_positions",
    "pty_plan_check",
    "pty_plan_workgroups",
    "pty_plan_use_graph",
    "pty_coresident_workgroups",
    "pty_plan_destroy",
    "pty_propagate",
    "pty_fft2",
    "pty_field_power",
    "pty_scale_to_power",
    "pty_window_fields",
    "pty_field_moments",
    "pty_roll_fields",
    "pty_roll_fields_noalloc",
)
PTY_ROLL_SCRATCH_MAX_VALUES = 2 * 1024 * 1024  # include/pty_hip.h


class PtyInstance(ctypes.Structure):
    """Mirror of ``pty_instance`` (include/pty_hip.h)."""

    _fields_ = [
        ("engine", ctypes.c_int32),
        ("propagator", ctypes.c_int32),
        ("nlambda", ctypes.c_int32),
        ("nosm", ctypes.c_int32),
        ("npsm", ctypes.c_int32),
        ("nslice", ctypes.c_int32),
        ("Np", ctypes.c_int32),
        ("No", ctypes.c_int32),
        ("F", ctypes.c_int32),
        ("n_iter", ctypes.c_int32),
        ("betaObject", ctypes.c_float),
        ("betaProbe", ctypes.c_float),
        ("alphaObject", ctypes.c_float),
        ("alphaProbe", ctypes.c_float),
        ("momFriction", ctypes.c_float),
        ("momFeedback", ctypes.c_float),
        ("probeWeight", ctypes.c_float),
        ("sliceBeta", ctypes.c_float),
        ("object", ctypes.c_void_p),
        ("probe", ctypes.c_void_p),
        ("objectBuffer", ctypes.c_void_p),
        ("objectMomentum", ctypes.c_void_p),
        ("probeBuffer", ctypes.c_void_p),
        ("probeMomentum", ctypes.c_void_p),
        ("ptychogram", ctypes.c_void_p),
        ("positions", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("momentum_mask", ctypes.c_void_p),
        ("energy", ctypes.c_void_p),
        ("tf", ctypes.c_void_p),
        ("err_at_pos", ctypes.c_void_p),
        ("err_iter", ctypes.c_void_p),
        ("workgroups", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("pc_direction", ctypes.c_void_p),
        ("pc_start", ctypes.c_int32),
        ("pc_radius", ctypes.c_int32),
        ("detector_error", ctypes.c_void_p),
    ]


class NativeError(RuntimeError):
    """A C-ABI call returned a nonzero code; the message is pty_last_error()."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


_lib = None


def load(path: os.PathLike | None = None):
    """Load and type the shared library (idempotent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -c 'import ptylab_amd.build as b; b.build()'`"
        )
    # torch ships its own libamdhip64.so.7; loading torch first makes the
    # library bind to that same HIP runtime (one runtime per process), so
    # device pointers and streams from torch are valid in the library
    import torch  # noqa: F401

    lib = ctypes.CDLL(str(p))
    for name in EXPORTED:
        getattr(lib, name)  # raises AttributeError if absent
    vp, i32 = ctypes.c_void_p, ctypes.c_int32
    lib.pty_abi_version.restype = ctypes.c_int
    lib.pty_last_error.restype = ctypes.c_char_p
    lib.pty_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(PtyInstance), i32, vp]
    lib.pty_plan_run.argtypes = [vp, i32, i32, vp]
    lib.pty_plan_run_positions.argtypes = [vp, i32, i32, i32, vp]
    lib.pty_plan_check.argtypes = [vp, vp]
    lib.pty_plan_workgroups.argtypes = [vp, ctypes.POINTER(i32)]
    lib.pty_plan_use_graph.argtypes = [vp, i32]
    lib.pty_coresident_workgroups.argtypes = [i32, i32, i32, i32, ctypes.POINTER(i32)]
    lib.pty_plan_destroy.argtypes = [vp]
    lib.pty_propagate.argtypes = [vp, i32, i32, i32, i32, vp, vp]
    lib.pty_fft2.argtypes = [vp, i32, i32, i32, vp]
    sz, f32 = ctypes.c_size_t, ctypes.c_float
    lib.pty_field_power.argtypes = [vp, sz, vp, vp]
    lib.pty_scale_to_power.argtypes = [vp, sz, vp, f32, vp]
    lib.pty_window_fields.argtypes = [vp, i32, i32, vp, f32, vp]
    lib.pty_field_moments.argtypes = [vp, i32, vp, vp, vp, vp]
    lib.pty_roll_fields.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp]
    lib.pty_roll_fields_noalloc.argtypes = [vp, i32, i32, i32, i32, i32, vp]
    for name in EXPORTED[2:]:
        getattr(lib, name).restype = ctypes.c_int
    if lib.pty_abi_version() != PTY_ABI_VERSION:
        raise RuntimeError(f"libptyhip ABI {lib.pty_abi_version()} != {PTY_ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib


def check(code):
    if code != 0:
        msg = load().pty_last_error()
        raise NativeError(code, msg.decode() if msg else "unknown error")


def _stream_handle(stream):
    return ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)


def _require_cuda(t):
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA (HIP) torch tensor")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")


def propagate(fields, propagator: int, inverse: bool, tf=None, stream=None):
    """In-place pty_propagate on a (..., Np, Np) complex64 CUDA tensor."""
    import torch

    _require_cuda(fields)
    if fields.dtype != torch.complex64:
        raise TypeError("fields must be complex64")
    Np = fields.shape[-1]
    nf = fields.numel() // (Np * Np)
    tfp = None
    if tf is not None:
        _require_cuda(tf)
        tfp = tf.data_ptr()
    check(load().pty_propagate(fields.data_ptr(), nf, Np, propagator, int(bool(inverse)), tfp, _stream_handle(stream)))
    return fields


def fft2(fields, sign: int, stream=None):
    """In-place unnormalised batched 2D DFT (pty_fft2)."""
    _require_cuda(fields)
    Np = fields.shape[-1]
    nf = fields.numel() // (Np * Np)
    check(load().pty_fft2(fields.data_ptr(), nf, Np, int(sign), _stream_handle(stream)))
    return fields


def coresident_workgroups(Np, engine, propagator, batched):
    """Workgroups of the kernel serving such an instance that fit on the
    device at once (pty_coresident_workgroups)."""
    n = ctypes.c_int32()
    check(load().pty_coresident_workgroups(int(Np), int(engine), int(propagator), int(bool(batched)), ctypes.byref(n)))
    return n.value


_live_plans = weakref.WeakSet()


@atexit.register
def _close_live_plans():
    """Destroy the plans still open when the interpreter exits, while the
    HIP runtime (and any profiler layered on it) is still up: Python's
    atexit handlers run before the C-level exit handlers of the loaded
    libraries, so no HIP call is left for module teardown."""
    for plan in list(_live_plans):
        plan.close()


class Plan:
    """RAII wrapper of a pty_plan; keeps the tensors it points to alive.
    Close it explicitly (``close()``); plans still open at interpreter exit
    are closed by an atexit hook, and a plan collected during interpreter
    finalisation is left to the process exit (no HIP call then)."""

    def __init__(self, instances, keepalive, stream=None):
        self._lib = load()
        self.stream = stream
        arr = (PtyInstance * len(instances))(*instances)
        self._instances = arr
        self._keep = keepalive
        h = ctypes.c_void_p()
        check(self._lib.pty_plan_create(ctypes.byref(h), arr, len(instances), _stream_handle(stream)))
        self._h = h
        _live_plans.add(self)

    def run(self, it_begin, it_end):
        check(self._lib.pty_plan_run(self._h, int(it_begin), int(it_end), _stream_handle(self.stream)))

    def run_positions(self, it, j_begin, j_end):
        check(self._lib.pty_plan_run_positions(self._h, int(it), int(j_begin), int(j_end), _stream_handle(self.stream)))

    def check(self):
        check(self._lib.pty_plan_check(self._h, _stream_handle(self.stream)))

    def workgroups(self):
        n = ctypes.c_int32()
        check(self._lib.pty_plan_workgroups(self._h, ctypes.byref(n)))
        return n.value

    def use_graph(self, enable=True):
        """Replay a per-iteration HIP graph instead of direct launches."""
        check(self._lib.pty_plan_use_graph(self._h, int(bool(enable))))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pty_plan_destroy(self._h)
            self._h = None
        _live_plans.discard(self)

    def __del__(self, _finalizing=sys.is_finalizing):
        # module globals (sys included) may already be None during
        # interpreter finalisation, so the check is bound at definition
        if not getattr(self, "_h", None) or _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
