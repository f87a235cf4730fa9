"""Builders shared by the tests: the same inputs for the MI355X package and
for the oracle, following the reference's regression protocol
(tests/regression/test_engine_regression.py:60-112): seed, initialise object
then probe, build the engine, seed again, reconstruct."""

import numpy as np

import oracle as O
import ptylab_amd as P


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def make_data(d):
    data = P.ExperimentalData()
    data.ptychogram = d["ptychogram"]
    data.encoder = d["encoder"]
    data.wavelength = d["wavelength"]
    data.dxd = d["dxd"]
    data.zo = d["zo"]
    data.entrancePupilDiameter = d.get("entrancePupilDiameter")
    data._setData()
    return data


def build_package(d, engine, propagator, modes, iters, seed, order="sequential", No=None, dz=None,
                  refrIndex=1.0, init=None, workgroups=0, orth=False, coeffs=None, switches=None):
    """Engine set up as the reference's scripts do, not yet reconstructed.
    ``switches``: further Params attributes (constraint switches)."""
    nl, no, npm, ns = modes
    data = make_data(d)
    params = P.Params()
    params.propagatorType = propagator
    params.positionOrder = order
    params.orthogonalizationSwitch = orth
    for k, v in (switches or {}).items():
        setattr(params, k, v)
    rec = P.Reconstruction(data, params)
    if No is not None:
        rec.No = int(No)  # user override after construction (exampleReconstructionCPM.py:79)
    rec.nlambda, rec.nosm, rec.npsm, rec.nslice = nl, no, npm, ns
    if nl > 1:
        # test_engine_regression.py:75-77
        rec.spectralDensity = float(np.atleast_1d(rec.wavelength)[0]) * np.linspace(0.98, 1.02, nl)
    if ns > 1:
        rec.dz = dz
        rec.refrIndex = refrIndex
    if init is None:
        np.random.seed(seed)
        rec.initializeObjectProbe()
    else:
        rec.object, rec.probe = init[0].copy(), init[1].copy()
        rec.initialGuessObject, rec.initialGuessProbe = init[0].copy(), init[1].copy()
    eng = getattr(P.Engines, engine)(rec, data, params, P.DummyMonitor())
    eng.numIterations = iters
    eng.workgroups = workgroups
    for k, v in (coeffs or {}).items():
        setattr(eng, k, v)
    return eng


def run_package(d, engine, propagator, modes, iters, seed, order="sequential", No=None, dz=None,
                refrIndex=1.0, init=None, workgroups=0, orth=False, coeffs=None, switches=None):
    eng = build_package(d, engine, propagator, modes, iters, seed, order=order, No=No, dz=dz,
                        refrIndex=refrIndex, init=init, workgroups=workgroups, orth=orth, coeffs=coeffs,
                        switches=switches)
    np.random.seed(seed)
    eng.reconstruct()
    return eng.reconstruction


def run_oracle(d, engine, propagator, modes, iters, seed, order="sequential", No=None, dz=0.0,
               refrIndex=1.0, init=None, precision="c64", orth=False, coeffs=None, switches=None):
    nl, no, npm, ns = modes
    sd = float(d["wavelength"]) * np.linspace(0.98, 1.02, nl) if nl > 1 else None
    geo = O.Geometry(d["ptychogram"], d["encoder"], d["wavelength"], d["dxd"], d["zo"],
                     d.get("entrancePupilDiameter"), No=No, nlambda=nl, nosm=no, npsm=npm, nslice=ns,
                     spectralDensity=sd)
    if init is None:
        np.random.seed(seed)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
    else:
        obj, pr = init
    kw = dict(coeffs or {})
    kw.update(switches or {})
    run = O.OracleRun(geo, obj.copy(), pr.copy(), engine=engine, propagator=propagator, positionOrder=order,
                      numIterations=iters, dz=dz or 0.0, refrIndex=refrIndex, precision=precision,
                      orthogonalizationSwitch=orth, **kw)
    np.random.seed(seed)
    run.reconstruct()
    return run
