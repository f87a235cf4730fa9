"""Pin the oracle against the reference's own regression goldens.

The goldens are read from the reference tree at test time (they are not
copied into this repository); the tests skip where the tree is absent
(the machines that run the GPU tests).  Tolerance is the reference's own (rtol 1e-5, atol 1e-7,
test_engine_regression.py:153); the restatement in fact matches bit for bit.
"""

import numpy as np
import pytest

import oracle as O
from regression_inputs import ENGINE_CASES, REFERENCE_GOLDEN_DIR, SEED, regression_dataset, spectral_density

needs_ref = pytest.mark.skipif(not REFERENCE_GOLDEN_DIR.exists(), reason="reference tree not mounted")


def _run(name):
    eng, prop, (nl, no, npm, ns), iters = ENGINE_CASES[name]
    d = regression_dataset()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"],
                     d["entrancePupilDiameter"], nlambda=nl, nosm=no, npsm=npm, nslice=ns,
                     spectralDensity=spectral_density(d["wavelength"], nl) if nl > 1 else None)
    np.random.seed(SEED)
    obj = O.initial_guess(geo.shape_O, "ones", geo)
    pr = O.initial_guess(geo.shape_P, "circ", geo)
    run = O.OracleRun(geo, obj, pr, engine=eng, propagator=prop, positionOrder="sequential",
                      numIterations=iters, dz=1e-4 if ns > 1 else 0.0)
    np.random.seed(SEED)
    return run.reconstruct()


@needs_ref
@pytest.mark.parametrize("name", list(ENGINE_CASES))
def test_oracle_engine_matches_reference_golden(name):
    run = _run(name)
    g = np.load(REFERENCE_GOLDEN_DIR / f"{name}.npz")
    for key, got in (("object", run.object), ("probe", run.probe), ("error", np.asarray(run.error))):
        assert got.shape == g[key].shape
        np.testing.assert_allclose(got, g[key], rtol=1e-5, atol=1e-7, err_msg=f"{name}:{key}")


@needs_ref
@pytest.mark.parametrize("prop,nl", [("fraunhofer", 1), ("asp", 1), ("fresnel", 1), ("polychromeasp", 3),
                                     ("scaledasp", 1), ("scaledpolychromeasp", 3), ("twosteppolychrome", 3)])
def test_oracle_propagator_matches_reference_golden(prop, nl):
    d = regression_dataset()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"],
                     nlambda=nl, spectralDensity=spectral_density(d["wavelength"], nl) if nl > 1 else None)
    np.random.seed(SEED)
    O.initial_guess(geo.shape_O, "ones", geo)
    probe = O.initial_guess(geo.shape_P, "circ", geo)
    fwd = O.propagate(probe.copy(), prop, geo)
    back = O.propagate(fwd, prop, geo, inverse=True)
    g = np.load(REFERENCE_GOLDEN_DIR / f"propagator_{prop}.npz")
    np.testing.assert_allclose(fwd, g["forward"], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(back, g["back"], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose([np.linalg.norm(fwd.ravel()), np.linalg.norm(back.ravel())], g["error"], rtol=1e-6)
