"""pty_plan_create and pty_plan_run reject malformed instances and ranges
with PTY_EINVAL and a message, before any kernel is launched (host-side
validation of include/pty_hip.h's contract; each case corrupts one field of
an otherwise valid ePIE instance)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NP, NO, F = 32, 64, 4


def _valid():
    from ptylab_amd import _native

    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(0)
    T = {
        "object": torch.ones(1, 1, 1, 1, NO, NO, dtype=torch.complex64, device=dev),
        "probe": torch.complex(torch.rand(1, 1, 1, 1, NP, NP, generator=g), torch.zeros(1, 1, 1, 1, NP, NP)).to(dev),
        "ptychogram": torch.rand(F, NP, NP, generator=g).to(dev),
        "positions": torch.tensor([[0, 0], [8, 8], [16, 0], [32, 32]], dtype=torch.int32, device=dev),
        "order": torch.arange(F, dtype=torch.int32, device=dev)[None].contiguous(),
        "energy": torch.full((F,), float(NP * NP), device=dev),
        "err_at_pos": torch.zeros((1, F), dtype=torch.float64, device=dev),
        "err_iter": torch.zeros((1,), dtype=torch.float64, device=dev),
    }
    inst = _native.PtyInstance()
    inst.engine, inst.propagator = _native.PTY_EPIE, _native.PTY_FRAUNHOFER
    inst.nlambda = inst.nosm = inst.npsm = inst.nslice = 1
    inst.Np, inst.No, inst.F, inst.n_iter = NP, NO, F, 1
    inst.betaObject = inst.betaProbe = 0.25
    inst.probeWeight, inst.sliceBeta = 1.0, 0.9
    inst.object, inst.probe = T["object"].data_ptr(), T["probe"].data_ptr()
    inst.ptychogram, inst.positions, inst.order = T["ptychogram"].data_ptr(), T["positions"].data_ptr(), T["order"].data_ptr()
    inst.energy = T["energy"].data_ptr()
    inst.err_at_pos, inst.err_iter = T["err_at_pos"].data_ptr(), T["err_iter"].data_ptr()
    inst.workgroups = 1
    return inst, T


def test_valid_instance_runs():
    from ptylab_amd import _native

    inst, T = _valid()
    plan = _native.Plan([inst], T, torch.cuda.current_stream())
    try:
        plan.run(0, 1)
        plan.check()
    finally:
        plan.close()
    assert np.isfinite(float(T["err_iter"][0]))


@pytest.mark.parametrize("field,value,message", [
    ("engine", 99, "unknown engine"),
    ("Np", 48, "unsupported"),
    ("npsm", 0, "mode counts"),
    ("propagator", 2, "ASP needs a transfer function"),
    ("flags", 0x40, "unknown flags"),
])
def test_plan_create_rejects_a_bad_field(field, value, message):
    from ptylab_amd import _native

    inst, T = _valid()
    setattr(inst, field, value)
    with pytest.raises(_native.NativeError, match=message) as e:
        _native.Plan([inst], T, torch.cuda.current_stream())
    assert e.value.code == _native.PTY_EINVAL


def test_plan_create_rejects_a_patch_outside_the_object_and_a_bad_order():
    from ptylab_amd import _native

    inst, T = _valid()
    T["positions"][3, 1] = NO - NP + 1
    with pytest.raises(_native.NativeError, match="outside the 64-pixel object"):
        _native.Plan([inst], T, torch.cuda.current_stream())
    inst, T = _valid()
    T["order"][0, 2] = F
    with pytest.raises(_native.NativeError, match="order entry"):
        _native.Plan([inst], T, torch.cuda.current_stream())


def test_plan_run_rejects_bad_iteration_ranges():
    from ptylab_amd import _native

    inst, T = _valid()
    plan = _native.Plan([inst], T, torch.cuda.current_stream())
    try:
        with pytest.raises(_native.NativeError, match="n_iter"):
            plan.run(0, 2)
        with pytest.raises(_native.NativeError, match="bad iteration range"):
            plan.run(1, 0)
        plan.run(0, 1)
        plan.check()
    finally:
        plan.close()
