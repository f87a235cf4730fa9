"""What the trained network gains over the method without it, on the
held-out scan of the learned-initialiser workload (Np = 128, No = 512, 500
positions, a seed no training batch uses): the stitched initial object is
closer to the truth than the uniform one, and mPIE started from it is ahead
of mPIE started from the uniform object after its first iterations.  The
network is trained here on 76 800 simulated patterns (the benchmark uses
200 000)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import ptylab_amd as P
    from ptylab_amd import learned_init as LI
    from ptylab_amd import unet as U

    net = U.UNet(64, seed=0)
    losses = LI.train(net, 1200, seed=100, lr=1e-3, lr_final=5e-5)
    w = LI.held_out_scan()
    probe = LI.probe()
    true = torch.as_tensor(w.object_true[0, 0, 0, 0], device="cuda")
    roi = torch.as_tensor(LI.roi_mask(w.positions, probe, w.No), device="cuda")

    def engine(learned):
        data = P.ExperimentalData()
        for k, v in w.as_dict().items():
            setattr(data, k, v)
        data._setData()
        params = P.Params()
        params.propagatorType = "Fraunhofer"
        params.positionOrder = "random"
        rec = P.Reconstruction(data, params)
        rec.No = w.No
        np.random.seed(1)
        rec.initializeObjectProbe()
        info = LI.initialize_object(rec, data, net, probe) if learned else None
        eng = P.Engines.mPIE(rec, data, params, P.DummyMonitor())
        np.random.seed(2)
        return eng, info

    return dict(net=net, losses=losses, w=w, probe=probe, true=true, roi=roi, engine=engine)


def test_training_converges_below_the_constant_guess(setup):
    losses = setup["losses"]
    print(f"loss first {losses[:20].mean():.4e} last {losses[-50:].mean():.4e}", flush=True)
    # predicting the patch mean (1 + 0i) scores the targets' variance, ~1.45e-2
    assert losses[-50:].mean() < 1.0e-2


def test_learned_initial_object_is_closer_than_uniform(setup):
    from ptylab_amd import learned_init as LI

    eng, info = setup["engine"](True)
    obj = torch.as_tensor(eng.reconstruction.object[0, 0, 0, 0], device="cuda")
    e_l = float(LI.nrmse_gpu(obj, setup["true"], setup["roi"]))
    e_o = float(LI.nrmse_gpu(torch.ones_like(obj), setup["true"], setup["roi"]))
    print(f"initial NRMSE learned {e_l:.4f} uniform {e_o:.4f}; {info}", flush=True)
    assert info["choice"] == "learned"
    assert e_l < 0.85 * e_o


def test_mpie_from_the_learned_init_is_ahead(setup):
    from ptylab_amd import learned_init as LI

    curves = {}
    for learned in (False, True):
        eng, _ = setup["engine"](learned)
        _, _, curve = LI.run_to_target(eng, setup["true"], setup["roi"], target=0.0, max_iter=12)
        curves[learned] = np.array(curve)
    print("plain  ", " ".join(f"{x:.3f}" for x in curves[False]), flush=True)
    print("learned", " ".join(f"{x:.3f}" for x in curves[True]), flush=True)
    assert curves[True][0] < curves[False][0]
    assert curves[True][1] < curves[False][1]
    # both settle on the same noise floor
    assert abs(curves[True][-1] - curves[False][-1]) < 0.01


def test_uniform_misfit_shortcut_equals_the_general_form():
    """The fallback check's misfit of the all-ones object: one transform of
    the probe expanded over the positions gives the same value as the
    per-position exit waves."""
    from ptylab_amd import learned_init as LI

    w = LI.held_out_scan(F=40)
    probe = LI.probe()
    ones = torch.ones((w.No, w.No), dtype=torch.complex64, device="cuda")
    pos = np.asarray(w.positions, np.int64)
    a = LI.data_misfit(ones, w.ptychogram, pos, probe)
    b = LI.data_misfit(ones, w.ptychogram, pos, probe, uniform=True)
    assert a == pytest.approx(b, rel=1e-6)
