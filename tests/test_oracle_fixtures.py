"""The oracle against the committed fixtures and against what the
mathematics fixes (closed forms, invariants, brute force on tiny inputs)."""

from pathlib import Path

import numpy as np
import pytest

import oracle as O
from golden.make_goldens import engine_case, propagator_case
from regression_inputs import ENGINE_CASES

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", list(ENGINE_CASES))
def test_oracle_engine_matches_committed_fixture(name):
    got = engine_case(name)
    ref = np.load(GOLDEN / f"oracle_{name}.npz")
    for k in ("object", "probe", "error"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-6, atol=1e-9, err_msg=f"{name}:{k}")


@pytest.mark.parametrize("prop", ["fraunhofer", "asp", "fresnel", "polychromeasp", "scaledasp", "scaledpolychromeasp",
                                  "twosteppolychrome"])
def test_oracle_propagator_matches_committed_fixture(prop):
    got = propagator_case(prop)
    ref = np.load(GOLDEN / f"oracle_propagator_{prop}.npz")
    for k in ("forward", "back"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-6, atol=1e-9)


def _dft2_brute(x):
    n = x.shape[-1]
    k = np.arange(n)
    Wm = np.exp(-2j * np.pi * np.outer(k, k) / n)
    return Wm @ x @ Wm.T


@pytest.mark.parametrize("n", [4, 8, 12])
def test_fft2c_is_the_shifted_unitary_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    want = np.fft.fftshift(_dft2_brute(np.fft.ifftshift(x))) / n
    np.testing.assert_allclose(O.fft2c(x), want, atol=1e-12)
    np.testing.assert_allclose(O.ifft2c(O.fft2c(x)), x, atol=1e-12)
    assert np.isclose(np.sum(abs(O.fft2c(x)) ** 2), np.sum(abs(x) ** 2))


@pytest.mark.parametrize("n", [16, 32, 64])
def test_checkerboard_identity_used_by_the_kernels(n):
    """For n % 4 == 0 the centred transform equals c * F(c * x) with
    c = (-1)^(y+x): the shifts fold into sign flips, so the GPU path never
    permutes data for fft2c / ifft2c."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    c = (-1.0) ** (yy + xx)
    np.testing.assert_allclose(O.fft2c(x), c * np.fft.fft2(c * x, norm="ortho"), atol=1e-12)
    np.testing.assert_allclose(O.ifft2c(x), c * np.fft.ifft2(c * x, norm="ortho"), atol=1e-12)


def test_asp_at_zero_distance_is_identity_and_band_limit_closed_form():
    N, L, wl = 64, 64 * 4e-6, 632.8e-9
    H0 = O.aspw_transfer(0.0, wl, N, L)
    np.testing.assert_allclose(H0, np.ones((N, N)), atol=0)
    z = 5e-2
    H = O.aspw_transfer(z, wl, N, L)
    f = np.arange(-N / 2, N / 2) / L
    Fx, Fy = np.meshgrid(f, f)
    fmax = L / (wl * np.sqrt(L ** 2 + 4 * z ** 2))
    inside = Fx ** 2 + Fy ** 2 < fmax ** 2
    assert np.array_equal(np.abs(H) > 0.5, inside)
    np.testing.assert_allclose(np.abs(H[inside]), 1.0, atol=1e-12)
    np.testing.assert_allclose(O.aspw_transfer(-z, wl, N, L), H.conj(), atol=0)


def test_multipie_equals_mpie_without_wavelength_or_object_modes():
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=32, No=96, F=12, seed=3, step_frac=0.3)
    d = w.as_dict()
    runs = []
    for eng in ("mPIE", "multiPIE"):
        geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"],
                         d["entrancePupilDiameter"], No=w.No, npsm=2)
        np.random.seed(4)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        r = O.OracleRun(geo, obj, pr, engine=eng, numIterations=3)
        np.random.seed(5)
        runs.append(r.reconstruct())
    assert np.array_equal(runs[0].object, runs[1].object)
    assert np.array_equal(runs[0].probe, runs[1].probe)


def test_orthogonalize_modes_invariants():
    rng = np.random.default_rng(0)
    p = (rng.standard_normal((4, 16, 16)) + 1j * rng.standard_normal((4, 16, 16)))
    modes, ev, V = O.orthogonalize_modes(p)
    m2 = modes.reshape(4, -1)
    G = m2.conj() @ m2.T
    assert np.abs(G - np.diag(np.diag(G))).max() < 1e-9 * np.abs(np.diag(G)).max()
    d = np.diag(G).real
    assert np.all(np.diff(d) <= 1e-9 * d.max())
    np.testing.assert_allclose(ev.sum(), 1.0)
    p2 = p.reshape(4, -1)
    Q, _ = np.linalg.qr(m2.T)
    np.testing.assert_allclose(Q @ (Q.conj().T @ p2.T), p2.T, atol=1e-9)


def test_nrmse_closed_form():
    rng = np.random.default_rng(1)
    t = rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))
    assert O.nrmse(t * (0.3 - 2j), t) < 1e-12
    noise = 1e-3 * (rng.standard_normal((20, 20)))
    assert 1e-5 < O.nrmse(t + noise, t) < 1e-2


def _c0_run(**kw):
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=64, No=256, F=30, seed=0)
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No)
    np.random.seed(1)
    obj = O.initial_guess(geo.shape_O, "ones", geo)
    pr = O.initial_guess(geo.shape_P, "circ", geo)
    return O.OracleRun(geo, obj, pr, engine="ePIE", numIterations=1, **kw), geo


def test_probe_power_correction_reaches_the_measured_power():
    run, geo = _c0_run(probePowerCorrectionSwitch=True)
    run._initial_probe_power_correction()
    p = np.sum(np.abs(run.probe.astype(np.complex128)) ** 2)
    assert p == pytest.approx(float(run.max_probe_power) ** 2, rel=1e-5)


def test_com_stabilization_finds_and_undoes_a_known_shift():
    run, geo = _c0_run(comStabilizationSwitch=1)
    g = np.exp(-((geo.Xp / (8 * geo.dxp)) ** 2 + (geo.Yp / (8 * geo.dxp)) ** 2)).astype(np.complex64)
    shifted = np.roll(g, (5, -7), axis=(0, 1))
    run.probe[0, 0, 0, 0] = shifted
    obj0 = run.object.copy()
    yc, xc = run._com_stabilization()
    assert (yc, xc) == (5, -7)
    np.testing.assert_array_equal(run.probe[0, 0, 0, 0], g)
    np.testing.assert_array_equal(run.object, np.roll(obj0, (-5, 7), axis=(-2, -1)))


def test_probe_windows():
    run, geo = _c0_run(probeBoundary=True)
    run._probe_window()
    d = geo.entrancePupilDiameter * 1.2
    np.testing.assert_array_equal(run.probeWindow, geo.Xp ** 2 + geo.Yp ** 2 < (d / 2) ** 2)
    run, geo = _c0_run(absorbingProbeBoundary=True)
    run._probe_window()
    w = run.probeWindow
    assert w[geo.Np // 2, geo.Np // 2] == pytest.approx(1.0) and w[0, 0] < 1e-6
    # super-Gaussian of order 10: flat to 1e-3 inside half its width
    r = np.sqrt(geo.Xp ** 2 + geo.Yp ** 2) / (3 / 4 * geo.Np * geo.dxp / 2.355)
    assert np.all(w[r < 0.5] > 1 - 1e-3)


def test_mqnewton_without_momentum_steps_is_qnewton():
    """mqNewton with betaObject_m = betaProbe_m = 0 leaves its ADAM steps
    without effect, so it reproduces qNewton (whose oracle matches the
    reference's qnewton_single golden) bit for bit."""
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=64, No=256, F=30, seed=0)
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No)
    runs = []
    for eng, kw in (("qNewton", {}), ("mqNewton", dict(betaObject_m=0.0, betaProbe_m=0.0))):
        np.random.seed(1)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        np.random.seed(5)
        runs.append(O.OracleRun(geo, obj, pr, engine=eng, numIterations=2, **kw).reconstruct())
    a, b = runs
    np.testing.assert_array_equal(a.object, b.object)
    np.testing.assert_array_equal(a.probe, b.probe)
    np.testing.assert_array_equal(a.error, b.error)


def test_adam_closed_forms():
    """ADAM's first step is g / (|g| + 1e-8); a gradient repeated at the next
    iteration gives the same step again (the bias corrections undo the
    moving averages)."""
    run, geo = _c0_run()
    rng = np.random.default_rng(0)
    g = (rng.standard_normal((1, 1, 1, 1, 16, 16)) + 1j * rng.standard_normal((1, 1, 1, 1, 16, 16))).astype(np.complex64)
    z = np.zeros_like(g)
    upd, m, v = run._adam(g, z, z.copy(), 1)
    want = g.astype(np.complex128) / (np.linalg.norm(g.astype(np.complex128).ravel()) + 1e-8)
    np.testing.assert_allclose(upd, want, rtol=2e-6, atol=1e-9)
    upd2, _, _ = run._adam(g, m, v, 2)
    np.testing.assert_allclose(upd2, want, rtol=2e-6, atol=1e-9)


def _c0_pair(engine, **kw):
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=64, No=256, F=30, seed=0)
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No)
    np.random.seed(1)
    obj = O.initial_guess(geo.shape_O, "ones", geo)
    pr = O.initial_guess(geo.shape_P, "circ", geo)
    np.random.seed(5)
    return O.OracleRun(geo, obj, pr, engine=engine, numIterations=3, **kw), geo


def test_pcpie_without_position_correction_is_mpie():
    """pcPIE's updates and momentum are mPIE's with stepM / betaM in the
    friction / feedback roles (pcPIE.py:122-202): without the position
    correction the two reconstructions agree bit for bit (mPIE's oracle
    matches the reference's mpie goldens)."""
    a = _c0_pair("mPIE")[0].reconstruct()
    b = _c0_pair("pcPIE")[0].reconstruct()
    np.testing.assert_array_equal(a.object, b.object)
    np.testing.assert_array_equal(a.probe, b.probe)
    np.testing.assert_array_equal(a.error, b.error)


def test_position_correction_points_along_a_known_shift():
    """A separable patch a(y) b(x) with b symmetric (b(-x) = b(x)): when the
    updated patch is the pre-update one rolled down by one row, the
    correlations of the +1 and -1 column shifts are equal (the column
    direction vanishes) and the row direction is positive, clipped at 3
    pixels and damped by daleth = 0.5 (BaseEngine.py:1305-1323); the
    transposed case gives a negative column direction for a roll left."""
    run, geo = _c0_pair("pcPIE", positionCorrectionSwitch=True)
    run.error = [1.0, 1.0]  # len(error) > startAtIteration
    k = 0
    row, col = run.positions[k]
    Np = int(geo.Np)
    t = 2 * np.pi * np.arange(Np) / Np
    gen = 2 + np.cos(t) + 0.3 * np.sin(2 * t)
    sym = 2 + np.cos(t)
    patch = np.outer(gen, sym).astype(np.complex64)[None, None, None, None]
    run.object[..., row:row + Np, col:col + Np] = np.roll(patch, 1, axis=-2)
    run._position_correction(k, patch)
    dy, dx = run.D[k]
    assert dy == 1.5 and abs(dx) < 1e-3  # complex64 sums, gain 1000
    run.D[k] = 0.0
    patch = np.outer(sym, gen).astype(np.complex64)[None, None, None, None]
    run.object[..., row:row + Np, col:col + Np] = np.roll(patch, -1, axis=-1)
    run._position_correction(k, patch)
    dy, dx = run.D[k]
    assert dx == -1.5 and abs(dy) < 1e-3


def test_position_correction_update_moves_the_encoder():
    """positionCorrectionUpdate (BaseEngine.py:1425-1436): encoder -= D dxo,
    recentred on the measured encoder's mean; positions follow."""
    run, geo = _c0_pair("pcPIE", positionCorrectionSwitch=True)
    run.error = [1.0, 1.0]
    run.D[:] = 0.0
    run.D[3] = (2.0, -1.0)
    enc0 = run.encoder_corrected.copy()
    run._position_correction_update()
    shift = enc0[3] - run.encoder_corrected[3]
    mean_shift = (np.array([2.0, -1.0]) * geo.dxo) / len(enc0)
    np.testing.assert_allclose(shift, np.array([2.0, -1.0]) * geo.dxo - mean_shift, rtol=1e-9)
    np.testing.assert_allclose(run.encoder_corrected.mean(axis=0), geo.encoder.mean(axis=0), rtol=1e-9, atol=1e-18)
    np.testing.assert_array_equal(run.positions, O.pixel_positions(run.encoder_corrected, geo.dxo, geo.No, geo.Np))


def _dup_channel_run(d, No, engine, ptychogram, nl=1, no=1, ns=1, **kw):
    """An oracle run whose nl wavelengths, no object modes or ns slices are
    identical copies of one channel (same seeded initial object and probe)."""
    sd = np.full(nl, float(d["wavelength"])) if nl > 1 else None
    geo = O.Geometry(ptychogram, d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"],
                     No=No, nlambda=nl, nosm=no, nslice=ns, spectralDensity=sd)
    one = O.Geometry(ptychogram, d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"], No=No)
    np.random.seed(4)
    obj = np.broadcast_to(O.initial_guess(one.shape_O, "ones", one), geo.shape_O).copy()
    pr = np.broadcast_to(O.initial_guess(one.shape_P, "circ", one), geo.shape_P).copy()
    run = O.OracleRun(geo, obj, pr, engine=engine, numIterations=3, alphaObject=1.0, alphaProbe=1.0, **kw)
    np.random.seed(5)
    return run.reconstruct()


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def test_multipie_duplicated_channels_reduce_to_one_channel():
    """multiPIE's own axes (multiPIE.py:153-205), pinned by what duplicated
    channels must give.  With alphaObject = alphaProbe = 1 the update weights
    are 1/Pmax and 1/Omax, and Pmax, Omax sum over wavelengths, object modes
    and slices (axes 0-3).  Two identical wavelengths or object modes on data
    I see Ihat = 2|Psi|^2, i.e. the same amplitude factor as one channel on
    I/2; Omax doubles and the probe increment sums over axes (0, 1), so each
    copy gets exactly the one-channel probe increment.  For wavelengths Pmax
    doubles too, which betaObject = 2 x 0.25 compensates.  Four identical
    slices see Ihat of the last slice only, so the data stay I; Pmax and Omax
    quadruple, compensated by betaObject = betaProbe = 4 x 0.25, and the
    probe sum runs over axes (0, 1) only, so every slice copy equals the
    one-slice run.  mPIE's probe sum over axis 1 alone fails the wavelength
    case by tens of percent: the check separates the two sums."""
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=32, No=96, F=12, seed=3, step_frac=0.3)
    d = w.as_dict()
    I = d["ptychogram"]
    half = _dup_channel_run(d, w.No, "multiPIE", I / 2)
    lam = _dup_channel_run(d, w.No, "multiPIE", I, nl=2, betaObject=0.5)
    osm = _dup_channel_run(d, w.No, "multiPIE", I, no=2)
    for i in range(2):
        # the only difference is the 1e-10 guard of Ihat, halved on I/2
        assert _rel(lam.object[i], half.object[0]) < 1e-5 and _rel(lam.probe[i], half.probe[0]) < 1e-5
        assert _rel(osm.object[0, i], half.object[0, 0]) < 1e-5
    assert _rel(osm.probe, half.probe) < 1e-5
    np.testing.assert_allclose(lam.error, half.error, rtol=1e-5)
    np.testing.assert_allclose(osm.error, half.error, rtol=1e-5)
    one = _dup_channel_run(d, w.No, "multiPIE", I)
    sl = _dup_channel_run(d, w.No, "multiPIE", I, ns=4, betaObject=1.0, betaProbe=1.0)
    for s in range(4):
        assert _rel(sl.object[0, 0, 0, s], one.object[0, 0, 0, 0]) < 1e-6
        assert _rel(sl.probe[0, 0, 0, s], one.probe[0, 0, 0, 0]) < 1e-6
    np.testing.assert_allclose(sl.error, one.error, rtol=1e-6)
    # the same duplication through mPIE's probe sum (axis 1 only) halves
    # every wavelength's probe increment: the identity breaks
    mp = _dup_channel_run(d, w.No, "mPIE", I, nl=2, betaObject=0.5)
    assert _rel(mp.probe[0], half.probe[0]) > 1e-2


def test_fftshift_switch_is_the_same_reconstruction():
    """fftshiftSwitch=True (BaseEngine.py:405-422, utils.py:20-47): the data
    are ifftshifted once and the transforms run without shifts.  For even Np
    the shift of the exit wave only multiplies its spectrum by (-1)^(kx+ky),
    which |.|^2 removes and the back-transform undoes, so in exact arithmetic
    the reconstruction is the unshifted one; in complex64 they agree to
    rounding.  Forgetting the data shift breaks this by orders of magnitude."""
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=32, No=96, F=12, seed=3, step_frac=0.3)
    d = w.as_dict()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"],
                     No=w.No)
    runs = []
    for sw in (False, True):
        np.random.seed(4)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        r = O.OracleRun(geo, obj, pr, engine="mPIE", numIterations=3, fftshiftSwitch=sw)
        np.random.seed(5)
        runs.append(r.reconstruct())
    assert _rel(runs[1].object, runs[0].object) < 1e-5
    assert _rel(runs[1].probe, runs[0].probe) < 1e-5
    np.testing.assert_allclose(runs[1].error, runs[0].error, rtol=1e-5)
    # the transforms alone, without the data shift, are a different problem
    np.random.seed(4)
    bad = O.OracleRun(geo, O.initial_guess(geo.shape_O, "ones", geo), O.initial_guess(geo.shape_P, "circ", geo),
                      engine="mPIE", numIterations=3, fftshiftSwitch=True)
    bad.ptychogram = np.fft.fftshift(bad.ptychogram, axes=(-1, -2))
    np.random.seed(5)
    bad.reconstruct()
    assert _rel(bad.object, runs[0].object) > 1e-2


def test_nadam_and_momentum_closed_forms():
    """NADAM (mqNewton.py:155-174): from zero moments the first step is
    (1 + beta1) g / (|g| + 1e-8), i.e. (1 + beta1) times ADAM's first step.
    Plain momentum (mqNewton.py:176-182): m1 = g1, m2 = g2 + frictionM g1,
    and the step is m."""
    run, geo = _c0_run()
    rng = np.random.default_rng(1)
    shape = (1, 1, 1, 1, 16, 16)
    g = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    g2 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    z = np.zeros_like(g)
    upd, m, v = run._nadam(g, z, z.copy(), 1)
    gd = g.astype(np.complex128)
    want = (1 + run.beta1) * gd / (np.linalg.norm(gd.ravel()) + 1e-8)
    np.testing.assert_allclose(upd, want, rtol=2e-6, atol=1e-9)
    adam, _, _ = run._adam(g, z, z.copy(), 1)
    np.testing.assert_allclose(upd, (1 + run.beta1) * adam, rtol=2e-6, atol=1e-9)
    u1, m1, v1 = run._plain_momentum(g, z, z.copy(), 1)
    np.testing.assert_array_equal(u1, g)
    np.testing.assert_array_equal(v1, z)
    u2, m2, _ = run._plain_momentum(g2, m1, v1, 2)
    np.testing.assert_allclose(u2, g2.astype(np.complex128) + run.frictionM * gd, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(u2, m2)


@pytest.mark.parametrize("method", ["NADAM", "momentum"])
def test_mqnewton_methods_without_steps_are_qnewton(method):
    """With betaObject_m = betaProbe_m = 0 every momentum engine leaves the
    object and probe alone, so mqNewton reproduces qNewton bit for bit."""
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=64, No=256, F=30, seed=0)
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No)
    runs = []
    for eng, kw in (("qNewton", {}), ("mqNewton", dict(betaObject_m=0.0, betaProbe_m=0.0, momentum_method=method))):
        np.random.seed(1)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        np.random.seed(5)
        runs.append(O.OracleRun(geo, obj, pr, engine=eng, numIterations=2, **kw).reconstruct())
    a, b = runs
    np.testing.assert_array_equal(a.object, b.object)
    np.testing.assert_array_equal(a.probe, b.probe)


def test_detector_error_and_save_memory_errors():
    """getRMSD / getErrorMetrics (BaseEngine.py:819-868).  Without saveMemory
    the per-pixel detector error of the last iteration sums, position by
    position, to the error curve's last entry (pinned by the reference's
    golden error vectors); with saveMemory the reference's (F, 1) / (F,)
    division broadcasts, so every iteration's error is the sum of the
    per-position errors times the sum of the inverse energies, and the run
    is otherwise the same."""
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=32, No=96, F=12, seed=3, step_frac=0.3)
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No)
    runs = []
    for sm in (False, True):
        np.random.seed(4)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        r = O.OracleRun(geo, obj, pr, engine="mPIE", numIterations=3, saveMemory=sm)
        np.random.seed(5)
        runs.append(r.reconstruct())
    a, b = runs
    E = O.energy_at_pos(w.ptychogram)
    per_pos = np.sum(a.detectorError, axis=(-1, -2))
    np.testing.assert_allclose(np.sum(per_pos / (E + 1e-20)), a.error[-1], rtol=1e-12)
    assert a.detectorError.shape == (12, 32, 32) and a.detectorError.dtype == np.float64
    np.testing.assert_array_equal(a.object, b.object)
    assert b.detectorError == 0 and b.errorAtPos.shape == (12, 12)
    np.testing.assert_allclose(b.error[-1], np.sum(per_pos) * np.sum(1 / (E + 1e-20)), rtol=1e-5)


def _circular_xcor_brute(a, b, radius):
    """|sum_x (a - mean a)[x + d] conj(b - mean b)[x]| over the shifts
    d in [-radius, radius]^2 (row shift slowest), by explicit loops."""
    n = a.shape[-1]
    am, bm = a - a.mean(), b - b.mean()
    out = []
    for dy in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            acc = 0j
            for y in range(n):
                for x in range(n):
                    acc += am[(y + dy) % n, (x + dx) % n] * np.conj(bm[y, x])
            out.append(abs(acc))
    return np.array(out)


def test_fourier_position_correction_is_the_circular_cross_correlation():
    """The Fourier branch (BaseEngine.py:1277-1302): |fftshift(ifft2(fft2(O -
    mean) conj(fft2(Opatch - O.mean()))))| cropped to +-radius is the circular
    cross-correlation of the mean-free patches (brute force, float64), with
    the row shift slowest as np.mgrid orders rowShifts / colShifts; the
    whole-object mean drops out (the mean-free updated patch sums to zero),
    so removing the product's zero-frequency term instead of subtracting the
    means gives the same correlation -- the form the GPU computes."""
    rng = np.random.default_rng(3)
    n, radius = 16, 3
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    want = _circular_xcor_brute(a, b, radius)
    other_mean = 0.37 - 0.21j  # any offset of the pre-update patch
    xc = np.fft.ifft2(np.fft.fft2(a - a.mean()) * np.fft.fft2(b - other_mean).conj())
    sw = slice(n // 2 - radius, n // 2 + radius + 1)
    got = abs(np.fft.fftshift(xc))[sw, sw].flatten()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    P = np.fft.fft2(a) * np.fft.fft2(b).conj()
    P[0, 0] = 0
    dc_free = abs(np.fft.fftshift(np.fft.ifft2(P)))[sw, sw].flatten()
    np.testing.assert_allclose(dc_free, want, rtol=1e-12, atol=1e-12)


def test_fourier_position_correction_points_along_a_known_shift():
    """With the updated patch the pre-update one rolled by (2, 0) the
    correlation peaks at that shift: the row direction is positive, clipped
    at 10 pixels and damped by daleth = 0.5 (BaseEngine.py:1304-1323), the
    column direction vanishes for a symmetric column profile; the transposed
    case gives a negative column direction for a roll left."""
    run, geo = _c0_pair("pcPIE", positionCorrectionSwitch=True, positionCorrectionSwitch_radius=4)
    run.error = [1.0, 1.0]
    k = 0
    row, col = run.positions[k]
    Np = int(geo.Np)
    t = 2 * np.pi * np.arange(Np) / Np
    gen = 2 + np.cos(t) + 0.3 * np.sin(2 * t)
    sym = 2 + np.cos(t)
    patch = np.outer(gen, sym).astype(np.complex64)[None, None, None, None]
    run.object[..., row:row + Np, col:col + Np] = np.roll(patch, 2, axis=-2)
    run._position_correction(k, patch)
    dy, dx = run.D[k]
    assert dy > 0 and abs(dx) < 1e-3 * abs(dy)
    run.D[k] = 0.0
    patch = np.outer(sym, gen).astype(np.complex64)[None, None, None, None]
    run.object[..., row:row + Np, col:col + Np] = np.roll(patch, -2, axis=-1)
    run._position_correction(k, patch)
    dy, dx = run.D[k]
    assert dx < 0 and abs(dy) < 1e-3 * abs(dx)
