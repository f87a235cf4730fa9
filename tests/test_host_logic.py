"""Host-side logic of the product package against the oracle and the
reference's documented behaviour (no GPU needed)."""

import numpy as np
import pytest

import oracle as O
import ptylab_amd as P
from ptyhelpers import make_data
from regression_inputs import SEED, regression_dataset


def test_params_defaults_match_reference():
    p = P.Params()
    # Params.py:41-195 defaults of the fields the hot path reads
    assert p.propagatorType == "Fraunhofer" and p.positionOrder == "random"
    assert p.intensityConstraint == "standard" and p.fftshiftSwitch is False
    assert p.orthogonalizationSwitch is False and p.orthogonalizationFrequency == 10
    assert p.saveMemory is False and p.probePowerCorrectionSwitch is False


def test_geometry_positions_and_init_match_oracle():
    d = regression_dataset()
    data = make_data(d)
    rec = P.Reconstruction(data, P.Params())
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"])
    assert rec.No == geo.No and rec.Np == geo.Np
    assert np.array_equal(rec.positions, geo.positions)
    assert rec.dxp == geo.dxp
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    np.random.seed(SEED)
    obj = O.initial_guess(geo.shape_O, "ones", geo)
    pr = O.initial_guess(geo.shape_P, "circ", geo)
    assert np.array_equal(rec.object, obj) and np.array_equal(rec.probe, pr)
    assert np.array_equal(data.energyAtPos, O.energy_at_pos(d["ptychogram"]))


def test_asp_transfer_function_matches_oracle():
    from ptylab_amd.Operators import asp_tf_fftorder, aspw_transfer_function

    for z in (5e-2, -1e-3, 0.0):
        a = aspw_transfer_function(z, 632.8e-9, 64, 64 * 4e-6)
        b = O.aspw_transfer(z, 632.8e-9, 64, 64 * 4e-6)
        assert np.array_equal(a, b)
    stack = O.asp_transfer_stack(1, 1, 64, 5e-2, 632.8e-9, 64 * 4e-6)
    assert np.array_equal(asp_tf_fftorder(5e-2, 632.8e-9, 64, 64 * 4e-6),
                          np.fft.ifftshift(stack, axes=(-2, -1))[0, 0, 0, 0])


@pytest.mark.parametrize("engine", ["ePIE", "mPIE"])
@pytest.mark.parametrize("order", ["sequential", "random"])
def test_schedule_replays_reference_rng_consumption(engine, order):
    """The host draws every iteration's order and the mPIE momentum triggers
    up front; the draws and the RNG state afterwards equal the reference's
    interleaved calls (one shuffle per iteration once two errors exist, one
    rand(1) per position)."""
    d = regression_dataset()
    data = make_data(d)
    params = P.Params()
    params.positionOrder = order
    rec = P.Reconstruction(data, params)
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    eng = getattr(P.Engines, engine)(rec, data, params, P.DummyMonitor())
    F, n_iter = data.numFrames, 5
    np.random.seed(7)
    orders, masks = eng._schedule(n_iter, engine == "mPIE")
    state_after = np.random.get_state()[1].copy()
    np.random.seed(7)
    for it in range(n_iter):
        ref = O.position_order(order, F, it)
        assert np.array_equal(orders[it], ref)
        if engine == "mPIE":
            trig = [bool(np.random.rand(1) > 0.95) for _ in range(F)]
            assert np.array_equal(masks[it].astype(bool), trig)
    assert np.array_equal(np.random.get_state()[1], state_after)


def test_unsupported_switches_raise():
    d = regression_dataset()
    data = make_data(d)
    params = P.Params()
    params.backgroundModeSwitch = True
    rec = P.Reconstruction(data, params)
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    eng = P.Engines.ePIE(rec, data, params, P.DummyMonitor())
    with pytest.raises(NotImplementedError):
        eng._prepareReconstruction()
    params.backgroundModeSwitch = False
    params.propagatorType = "ASP"
    params.fftshiftSwitch = True
    with pytest.raises(ValueError):
        eng._prepareReconstruction()
    params.propagatorType = "polychromeASP"
    with pytest.raises(ValueError):
        eng._prepareReconstruction()
    params.propagatorType = "Fresnel"
    with pytest.raises(NotImplementedError):
        eng._prepareReconstruction()
    params.propagatorType = "scaledASP"
    with pytest.raises(ValueError):
        eng._prepareReconstruction()
    params.fftshiftSwitch = False
    params.propagatorType = "twoStepPolychrome"
    with pytest.raises(NotImplementedError):
        eng._prepareReconstruction()
    params.propagatorType = "Fraunhofer"
    params.intensityConstraint = "bogus"
    with pytest.raises(ValueError):
        eng._prepareReconstruction()


def test_reconstruct_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    d = regression_dataset()
    data = make_data(d)
    params = P.Params()
    rec = P.Reconstruction(data, params)
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    eng = P.Engines.ePIE(rec, data, params, P.DummyMonitor())
    eng.numIterations = 1
    with pytest.raises(RuntimeError, match="GPU"):
        eng.reconstruct()


def test_synthetic_encoder_round_trip():
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=64, No=256, F=100, seed=0)
    data = make_data(w.as_dict())
    rec = P.Reconstruction(data, P.Params())
    rec.No = w.No
    assert np.array_equal(rec.positions, w.positions)
    assert w.positions.min() >= 0 and w.positions.max() + 64 <= 256


def test_fresnel_phase_and_polychrome_transfer_functions_match_oracle():
    from ptylab_amd import Operators as Op

    d = regression_dataset()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"])
    Q = Op.quadratic_phase(geo.zo, geo.wavelength, geo.Np, geo.dxp)
    assert Q.dtype == np.complex64
    np.testing.assert_array_equal(Q, O.quad_phase(geo.zo, geo.wavelength, geo.Np, geo.dxp).astype(np.complex64))
    sd = float(d["wavelength"]) * np.linspace(0.98, 1.02, 3)
    tfs = Op.polychrome_tf_fftorder(geo.zo, sd, geo.Np, geo.Lp)
    ref = O.polychrome_transfer_stack(1, 1, geo.Np, geo.zo, sd, geo.Lp)[:, 0, 0, 0]
    np.testing.assert_array_equal(tfs, np.fft.ifftshift(ref, axes=(-2, -1)).astype(np.complex64))
    # distinct wavelengths give distinct transfer functions
    assert not np.array_equal(tfs[0], tfs[2])


@pytest.mark.parametrize("kind", ["all", "object", "probe"])
def test_save_results_round_trip(tmp_path, kind):
    """saveResults writes the reference's dataset names and dtypes
    (Reconstruction.py:509-566); without h5py the archive is an .npz."""
    d = regression_dataset()
    data = make_data(d)
    rec = P.Reconstruction(data, P.Params())
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    rec.object, rec.probe = rec.initialGuessObject, rec.initialGuessProbe
    rec.error = [3.0, 2.0]
    path = rec.saveResults(str(tmp_path / "recent.npz"), type=kind, squeeze=True)
    with np.load(path, allow_pickle=False) as z:
        keys = set(z.files)
        if kind == "all":
            assert {"probe", "object", "error", "zo", "wavelength", "dxp", "purityProbe", "purityObject",
                    "I object", "I probe", "encoder_corrected"} <= keys
            assert z["object"].dtype == np.complex64 and z["error"].dtype == np.float32
            np.testing.assert_array_equal(z["I probe"], np.abs(rec.probe).astype(np.float32))
        else:
            assert keys == {kind}
            np.testing.assert_array_equal(z[kind], np.squeeze(getattr(rec, kind)).astype(np.complex64))
    if kind == "all":
        other = P.Reconstruction(data, P.Params())
        other.load(path)
        np.testing.assert_array_equal(other.object, rec.object)
        np.testing.assert_array_equal(other.probe, rec.probe)
        np.testing.assert_array_equal(other.error, np.float32([3.0, 2.0]))
        assert np.isclose(other.dxp, rec.dxp, rtol=1e-6)


def test_save_results_rejects_unknown_type(tmp_path):
    d = regression_dataset()
    rec = P.Reconstruction(make_data(d), P.Params())
    with pytest.raises(NotImplementedError):
        rec.saveResults(str(tmp_path / "x.npz"), type="everything")


def test_load_data_npz_transposed_encoder_and_orientation(tmp_path):
    """loadData with the reference's root names: a (2, F) encoder is
    transposed and centred (io/readHdf5.py:68-75) and the orientation code
    flips / transposes the frames (ExperimentalData.py:192-232)."""
    rng = np.random.default_rng(0)
    frames = rng.random((3, 8, 8)).astype(np.float32)
    enc = rng.random((2, 3))
    np.savez(tmp_path / "d.npz", ptychogram=frames, encoder=enc, dxd=np.float64(5e-6), zo=np.float64(0.05),
             wavelength=np.float64(632.8e-9), orientation=np.array([5]))
    data = P.ExperimentalData()
    data.loadData(tmp_path / "d.npz")
    np.testing.assert_allclose(data.encoder, enc.T - enc.T.mean(axis=0, keepdims=True))
    # orientation 5: transpose, then flip columns
    np.testing.assert_array_equal(data.ptychogram, np.flip(np.transpose(frames, (0, 2, 1)), axis=-1))
    assert data.numFrames == 3 and data.Nd == 8
    for code, want in [(1, frames[:, :, ::-1]), (2, frames[:, ::-1, :]), (3, frames[:, ::-1, ::-1]),
                       (4, frames.transpose(0, 2, 1)), (6, frames.transpose(0, 2, 1)[:, ::-1, :]),
                       (7, frames.transpose(0, 2, 1)[:, ::-1, ::-1])]:
        data.ptychogram = frames
        data.setOrientation(code)
        np.testing.assert_array_equal(data.ptychogram, want)
    with pytest.raises(ValueError):
        data.setOrientation(8)
    stub = P.ExperimentalData()
    stub.loadData("test:nodata")
    assert stub.ptychogram.shape == (1, 16, 16) and stub.zo == 0.1


def test_scaled_asp_phases_and_checkerboard_folding():
    """The host builds scaledASP's Q1, Q2 as the oracle does (which matches
    the reference's propagator_scaledasp golden), and the kernels' form
    c F^-1(F(c Q1 psi) Q2) with c = (-1)^(y+x) and unshifted transforms equals
    ifft2c(fft2c(psi Q1) Q2) (and the inverse), the identity the GPU path
    relies on."""
    from ptylab_amd import Operators as Op

    d = regression_dataset()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"])
    Q1, Q2 = Op.scaled_asp_phases(geo.zo, geo.wavelength, geo.Np, geo.dxo, geo.dxd)
    R1, R2 = O.scaled_asp_phases(geo.zo, geo.wavelength, geo.Np, geo.dxo, geo.dxd)
    np.testing.assert_array_equal(Q1, R1.astype(np.complex64))
    np.testing.assert_array_equal(Q2, R2.astype(np.complex64))
    pair = Op.scaled_asp_kernel_arrays(geo.zo, geo.wavelength, geo.Np, geo.dxo, geo.dxd)
    n = geo.Np
    c = (-1.0) ** np.add.outer(np.arange(n), np.arange(n))
    rng = np.random.default_rng(0)
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q1, q2 = Q1.astype(np.complex128), Q2.astype(np.complex128)
    want_f = O.ifft2c(O.fft2c(psi * q1) * q2)
    got_f = c * np.fft.ifft2(np.fft.fft2(pair[1].astype(np.complex128) * psi) * pair[0])
    assert np.max(np.abs(got_f - want_f)) < 1e-12 * np.max(np.abs(want_f))
    want_b = O.ifft2c(O.fft2c(psi) * q2.conj()) * q1.conj()
    got_b = np.fft.ifft2(np.fft.fft2(c * psi) * pair[0].conj()) * pair[1].astype(np.complex128).conj()
    assert np.max(np.abs(got_b - want_b)) < 1e-12 * np.max(np.abs(want_b))


def test_batched_instances_follow_the_named_inputs():
    """The batched benchmark's instances: generated as the single mPIE ASP
    workload (4 probe modes of radius Np/4) by default, as the plain-ePIE
    family on request; Np, F and the seed per instance do not depend on the
    inputs chosen."""
    from ptylab_amd import synthetic as S

    a = S.batched_instance(3, nps=(32,), fmin=5, fmax=8)
    b = S.batched_instance(3, nps=(32,), fmin=5, fmax=8, inputs="fraunhofer")
    assert a.Np == b.Np == 32 and a.No == b.No == 128
    assert a.ptychogram.shape == b.ptychogram.shape
    assert a.probe_true.shape[2] == 4 and b.probe_true.shape[2] == 1
    with pytest.raises(ValueError):
        S.batched_instance(3, nps=(32,), fmin=5, fmax=8, inputs="other")


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and says why,
    before touching a device."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_mqnewton_momentum_methods_host_side():
    """mqNewton's momentum_method (mqNewton.py:65-79): ADAM and NADAM run; the
    plain "momentum" engine reads second moments that only ADAM / NADAM
    create, so on a fresh reconstruction it raises the reference's
    AttributeError, and runs once they exist; an unknown name fails the
    reference's getattr lookup.  Each method maps to its instance flag."""
    d = regression_dataset()
    data = make_data(d)
    params = P.Params()
    rec = P.Reconstruction(data, params)
    np.random.seed(SEED)
    rec.initializeObjectProbe()
    eng = P.Engines.mqNewton(rec, data, params, P.DummyMonitor())
    assert eng._instance_flags() == 0
    eng.momentum_method = "NADAM"
    eng._prepareReconstruction()
    assert eng._instance_flags() == P._native.PTY_FLAG_MQN_NADAM
    eng.momentum_method = "momentum"
    assert eng._instance_flags() == P._native.PTY_FLAG_MQN_MOMENTUM
    assert eng._engine_coefficients()["momFriction"] == eng.frictionM
    with pytest.raises(AttributeError, match="objectMomentum_v"):
        eng._prepareReconstruction()
    rec.objectMomentum_v = np.zeros_like(rec.object)
    rec.probeMomentum_v = np.zeros_like(rec.probe)
    eng._prepareReconstruction()
    eng.momentum_method = "Adagrad"
    with pytest.raises(AttributeError):
        eng._prepareReconstruction()
