"""Write this repository's golden fixtures from the oracle.

The oracle reproduces the reference's own regression goldens bit for bit
(tests/test_oracle_reference.py, run where the reference tree is mounted);
the fixtures written here are the oracle's outputs on the same regenerated
inputs (tests/regression_inputs.py) plus two config-family cases, so the CPU
suite pins the oracle without the reference present.  Run from the
repository root:  python tests/golden/make_goldens.py
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

import oracle as O  # noqa: E402
from regression_inputs import ENGINE_CASES, SEED, regression_dataset, spectral_density  # noqa: E402

PROPAGATOR_CASES = {"fraunhofer": 1, "asp": 1, "fresnel": 1, "polychromeasp": 3, "scaledasp": 1, "scaledpolychromeasp": 3, "twosteppolychrome": 3}


def engine_case(name):
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
    run.reconstruct()
    return {"object": run.object, "probe": run.probe, "error": np.asarray(run.error)}


def propagator_case(prop):
    nl = PROPAGATOR_CASES[prop]
    d = regression_dataset()
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"], d["entrancePupilDiameter"],
                     nlambda=nl, spectralDensity=spectral_density(d["wavelength"], nl) if nl > 1 else None)
    np.random.seed(SEED)
    O.initial_guess(geo.shape_O, "ones", geo)
    probe = O.initial_guess(geo.shape_P, "circ", geo)
    fwd = O.propagate(probe.copy(), prop, geo)
    back = O.propagate(fwd, prop, geo, inverse=True)
    return {"forward": fwd, "back": back}


def main(only=None):
    """Write every fixture, or only the named ones (engine case names or
    propagator_<kind>)."""
    for name in ENGINE_CASES:
        if only is None or name in only:
            np.savez_compressed(HERE / f"oracle_{name}.npz", **engine_case(name))
    for prop in PROPAGATOR_CASES:
        if only is None or f"propagator_{prop}" in only:
            np.savez_compressed(HERE / f"oracle_propagator_{prop}.npz", **propagator_case(prop))
    print("written to", HERE)


if __name__ == "__main__":
    main(set(sys.argv[1:]) or None)
