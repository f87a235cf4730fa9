"""Batched launch: many independent reconstructions of uneven Np, position
count and iteration count in one launch per pass family, replayed as a
per-iteration HIP graph.

A batch must leave every engine exactly where its own reconstruct() would
have: the per-position arithmetic is the same whichever kernel, workgroup
size or cluster size runs it, so objects and probes are compared
bit-for-bit against separate runs (the FP64 error sums may differ in the
last bits: their partial sums group differently with the workgroup shape).
One member is also checked against the complex64 oracle at the one-iteration
bar (relative error <= 1e-5)."""

import numpy as np
import pytest

from ptyhelpers import build_package, rel, run_oracle
from ptylab_amd import synthetic as S

pytestmark = pytest.mark.gpu

# (engine, propagator, modes, Np, No, F, iterations, order, seed)
CASES = [
    ("ePIE", "Fraunhofer", (1, 1, 1, 1), 64, 256, 40, 1, "sequential", 1),
    ("mPIE", "Fraunhofer", (1, 1, 1, 1), 32, 128, 60, 4, "random", 2),
    ("mPIE", "Fraunhofer", (1, 1, 1, 1), 128, 512, 30, 3, "random", 3),
    ("ePIE", "Fraunhofer", (1, 1, 1, 1), 256, 640, 12, 2, "random", 4),
    ("mPIE", "ASP", (1, 1, 2, 1), 64, 256, 25, 4, "random", 5),
    ("ePIE", "Fraunhofer", (1, 1, 1, 1), 512, 1024, 4, 2, "sequential", 6),
]


@pytest.fixture(scope="module")
def datasets():
    out = []
    for engine, prop, modes, Np, No, F, iters, order, seed in CASES:
        if prop == "ASP":
            w = S.cpm_asp_modes(Np=Np, No=No, F=F, npsm=modes[2], seed=seed, probe_radius_px=Np // 4)
        else:
            w = S.cpm_fraunhofer(Np=Np, No=No, F=F, seed=seed)
        out.append(w)
    return out


def _engines(datasets):
    engs = []
    for (engine, prop, modes, Np, No, F, iters, order, seed), w in zip(CASES, datasets):
        engs.append(build_package(w.as_dict(), engine, prop, modes, iters, seed, order=order, No=w.No))
    return engs


def _separate(datasets):
    engs = _engines(datasets)
    np.random.seed(77)
    for e in engs:
        e.reconstruct()
    return [e.reconstruction for e in engs]


@pytest.fixture(scope="module")
def separate(datasets):
    return _separate(datasets)


def _batched(datasets, **kw):
    import ptylab_amd as P

    engs = _engines(datasets)
    np.random.seed(77)
    b = P.Engines.Batched(engs, **kw).reconstruct()
    return [e.reconstruction for e in engs], b


def _same(a, b, tag):
    for ra, rb, case in zip(a, b, CASES):
        name = f"{tag} {case[0]} Np={case[3]}"
        assert np.array_equal(np.asarray(ra.object), np.asarray(rb.object)), name
        assert np.array_equal(np.asarray(ra.probe), np.asarray(rb.probe)), name
        assert len(ra.error) == len(rb.error) == case[6], name
        np.testing.assert_allclose(np.asarray(ra.error), np.asarray(rb.error), rtol=1e-12, err_msg=name)


def test_batched_graph_equals_separate_runs(datasets, separate):
    got, b = _batched(datasets, workgroups=[1, 1, 2, 4, 2, 1], graph=True)
    _same(separate, got, "graph")


def test_batched_direct_launch_equals_separate_runs(datasets, separate):
    got, _ = _batched(datasets, workgroups=[2, 1, 1, 8, 1, 2], graph=False)
    _same(separate, got, "direct")


def test_batched_autotuned_equals_separate_runs(datasets, separate, tmp_path):
    from ptylab_amd import autotune

    table = autotune.CostTable(tmp_path / "tune.json", positions=16)
    got, b = _batched(datasets, tuner=table)
    G = b.last_workgroups
    print("autotuned cluster sizes", G, flush=True)
    assert all(g >= 1 and (g & (g - 1)) == 0 and c[3] % g == 0 for g, c in zip(G, CASES))
    assert all(v > 0 for v in table.table.values())
    _same(separate, got, "autotuned")


def test_small_probe_clusters_equal_separate_runs(datasets, separate):
    """Np = 32 on clusters of 4 and 8 workgroups, whose whole exit-wave
    workspace fits a CU's L1: the XCD-local exchange must not read a stale
    L1 line (without the L1 invalidate at its barrier it did, at 4.5e-2)."""
    for G in ([1, 8, 2, 4, 2, 1], [1, 4, 4, 8, 4, 2]):
        got, _ = _batched(datasets, workgroups=G)
        _same(separate, got, f"clusters {G}")


def test_batched_member_against_oracle(datasets):
    got, _ = _batched(datasets, workgroups=[4, 1, 1, 1, 1, 1])
    engine, prop, modes, Np, No, F, iters, order, seed = CASES[0]
    ref = run_oracle(datasets[0].as_dict(), engine, prop, modes, iters, seed, order=order, No=No)
    eo, ep = rel(got[0].object, ref.object), rel(got[0].probe, ref.probe)
    print(f"batched member vs oracle: object {eo:.2e} probe {ep:.2e}", flush=True)
    assert eo <= 1e-5 and ep <= 1e-5
    np.testing.assert_allclose(got[0].error, ref.error, rtol=1e-5)


def test_too_many_fields_for_a_workgroup_raise(datasets):
    """16 probe modes need 16 waves per column round; the batched kernel's
    workgroups have 8, so plan creation rejects the instance."""
    import ptylab_amd as P
    from ptylab_amd import _native

    w = datasets[0]
    eng = build_package(w.as_dict(), "ePIE", "Fraunhofer", (1, 1, 16, 1), 1, 1, No=w.No)
    with pytest.raises(_native.NativeError, match="exceed the 8 waves"):
        P.Engines.Batched([eng], workgroups=[1]).reconstruct()


def test_cluster_larger_than_one_xcd_raises(datasets):
    """A batched cluster runs whole on one XCD (64 co-resident workgroups of
    the batched kernel on MI355X); plan creation rejects a larger one."""
    import ptylab_amd as P
    from ptylab_amd import _native

    engs = _engines(datasets)
    G = [1, 1, 1, 128, 1, 1]
    with pytest.raises(_native.NativeError, match="of one XCD"):
        P.Engines.Batched(engs, workgroups=G).reconstruct()


def test_plan_reports_active_workgroups(datasets):
    """Blocks left idle by the XCD packing are not counted as workgroups."""
    _, b = _batched(datasets, workgroups=[1, 1, 2, 4, 2, 1], graph=True)
    assert b.last_workgroups == [1, 1, 2, 4, 2, 1]
    assert b.last_plan_workgroups == sum(b.last_workgroups)


def test_batched_asp_four_modes_equals_separate_runs():
    """The benchmark's batched inputs: mPIE with ASP and 4 probe modes at
    Np = 128 and 256 (XCD-local clusters of 4 and 16 workgroups) leave
    object and probe bit-identical to separate runs."""
    import ptylab_amd as P

    cases = [(256, 1024, 24, 2, 7, 16), (128, 512, 30, 3, 8, 4)]
    ws = [S.cpm_asp_modes(Np=Np, No=No, F=F, npsm=4, seed=seed, probe_radius_px=Np // 4)
          for Np, No, F, iters, seed, _ in cases]

    def engines():
        return [build_package(w.as_dict(), "mPIE", "ASP", (1, 1, 4, 1), c[3], c[4], order="random", No=w.No)
                for w, c in zip(ws, cases)]

    sep = engines()
    np.random.seed(78)
    for e in sep:
        e.reconstruct()
    bat = engines()
    np.random.seed(78)
    P.Engines.Batched(bat, workgroups=[c[5] for c in cases]).reconstruct()
    for a, b, c in zip(sep, bat, cases):
        ra, rb = a.reconstruction, b.reconstruction
        assert np.array_equal(np.asarray(ra.object), np.asarray(rb.object)), c
        assert np.array_equal(np.asarray(ra.probe), np.asarray(rb.probe)), c
        np.testing.assert_allclose(np.asarray(ra.error), np.asarray(rb.error), rtol=1e-12)


@pytest.mark.parametrize("engine", ["ePIE", "multiPIE"])
def test_batched_field_loop_engines_equal_separate_runs(engine):
    """The field-loop row passes (batched ASP clusters that own at least one
    line per item slot) carry each engine's own update weights: ePIE and
    multiPIE with 2 probe modes on clusters of 2 and 4 workgroups leave
    object and probe bit-identical to separate runs."""
    import ptylab_amd as P

    cases = [(64, 256, 20, 2, 11, 2), (128, 512, 16, 2, 12, 4)]
    ws = [S.cpm_asp_modes(Np=Np, No=No, F=F, npsm=2, seed=seed, probe_radius_px=Np // 4)
          for Np, No, F, iters, seed, _ in cases]

    def engines():
        return [build_package(w.as_dict(), engine, "ASP", (1, 1, 2, 1), c[3], c[4], order="random", No=w.No)
                for w, c in zip(ws, cases)]

    sep = engines()
    np.random.seed(79)
    for e in sep:
        e.reconstruct()
    bat = engines()
    np.random.seed(79)
    P.Engines.Batched(bat, workgroups=[c[5] for c in cases]).reconstruct()
    for a, b, c in zip(sep, bat, cases):
        ra, rb = a.reconstruction, b.reconstruction
        assert np.array_equal(np.asarray(ra.object), np.asarray(rb.object)), c
        assert np.array_equal(np.asarray(ra.probe), np.asarray(rb.probe)), c


def test_batched_polychrome_and_scaled_asp_equal_separate_runs():
    """Packed exchange rows with several transfer functions: polychromeASP
    (three wavelengths, each with its own band limit and live columns),
    ASP and scaledASP instances on the reference's regression inputs
    (Np = 32, a band limit of a few columns) in one batched launch leave
    object, probe and error bit-identical to separate (per-size) runs."""
    import ptylab_amd as P
    from regression_inputs import SEED, regression_dataset

    d = regression_dataset()
    cases = [("ePIE", "polychromeASP", (3, 1, 1, 1), 2), ("ePIE", "ASP", (1, 1, 1, 1), 3),
             ("mPIE", "scaledASP", (1, 1, 1, 1), 3)]

    def engines():
        return [build_package(d, e, prop, modes, iters, SEED) for e, prop, modes, iters in cases]

    sep = engines()
    np.random.seed(80)
    for e in sep:
        e.reconstruct()
    bat = engines()
    np.random.seed(80)
    P.Engines.Batched(bat).reconstruct()
    for a, b, c in zip(sep, bat, cases):
        ra, rb = a.reconstruction, b.reconstruction
        print(f"batched {c[1]}: max |dO| {np.abs(np.asarray(ra.object) - np.asarray(rb.object)).max():.2e}", flush=True)
        assert np.array_equal(np.asarray(ra.object), np.asarray(rb.object)), c
        assert np.array_equal(np.asarray(ra.probe), np.asarray(rb.probe)), c
        np.testing.assert_allclose(np.asarray(ra.error), np.asarray(rb.error), rtol=1e-12)


@pytest.mark.parametrize("prop", ["Fresnel", "polychromeASP", "scaledASP", "scaledPolychromeASP"])
def test_autotuner_measures_every_propagator(prop):
    """The autotuner times a batched instance of every propagator family on
    synthetic data (each with the transfer function or phase its plan needs)."""
    import ptylab_amd as P
    from ptylab_amd import autotune

    pid = {"Fresnel": P._native.PTY_FRESNEL, "polychromeASP": P._native.PTY_POLYCHROME_ASP,
           "scaledASP": P._native.PTY_SCALED_ASP, "scaledPolychromeASP": P._native.PTY_SCALED_POLYCHROME_ASP}[prop]
    t = autotune.measure(autotune.Shape(P._native.PTY_EPIE, pid, 32, 128, 2), 2, positions=8)
    assert np.isfinite(t) and t > 0
