"""GPU parity of the MI355X engines against the oracle (and, through the
oracle, against the reference's regression goldens).

Two criteria, both stated per test:

* golden inputs (the reference's regression dataset): the data are
  normalised to a peak of 1 while the initial probe is not, so the problem is
  ill-conditioned and the reference's own complex64 NumPy run already sits
  1e-5..3e-4 (relative Frobenius) from a complex128 run of the same
  algorithm.  The GPU run must be as close to the complex128 oracle as the
  complex64 oracle is: rel(gpu, c128) <= 2 rel(c64, c128) + 1e-6 for object
  and probe (+ 1e-5 for the error curve), and for the well-conditioned ASP
  case rel(gpu, c64) < 1e-5.  For e3PIE the complex64 comparator is the
  oracle with the slice transfer function in complex64 ("c64pure"), as the
  GPU path keeps the whole FFT chain in complex64 (the reference's CPU path
  promotes the slice chain to complex128, its CuPy path does not).  (The reference's own
  GPU-vs-CPU bound on these configs is 1e-3, test_engine_regression.py:157-175.)
* photon-scaled config inputs: relative object/probe error <= 1e-5 after one
  iteration and <= 1e-3 after 50, error curve within 1 %.
"""

import numpy as np
import pytest

from ptyhelpers import rel, run_oracle, run_package
from regression_inputs import ENGINE_CASES, SEED, regression_dataset

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(ENGINE_CASES))
def test_engine_matches_oracle_on_golden_inputs(name):
    eng, prop, modes, iters = ENGINE_CASES[name]
    d = regression_dataset()
    dz = 1e-4 if modes[3] > 1 else None
    print(f"case {name}", flush=True)
    gpu = run_package(d, eng, prop, modes, iters, SEED, dz=dz)
    # e3PIE, Fresnel and polychromeASP: the reference's CPU path promotes to
    # complex128 through a complex128 slice TF / quadratic phase / TF stack;
    # the complex64 comparator keeps those complex64 ("c64pure")
    pure = eng == "e3PIE" or prop in ("Fresnel", "polychromeASP")
    c64 = run_oracle(d, eng, prop, modes, iters, SEED, dz=dz or 0.0, precision="c64pure" if pure else "c64")
    c128 = run_oracle(d, eng, prop, modes, iters, SEED, dz=dz or 0.0, precision="c128")
    assert gpu.object.dtype == np.complex64 and gpu.probe.dtype == np.complex64
    assert gpu.object.shape == c64.object.shape and gpu.probe.shape == c64.probe.shape
    for key in ("object", "probe", "error"):
        g, a, b = np.asarray(getattr(gpu, key)), np.asarray(getattr(c64, key)), np.asarray(getattr(c128, key))
        d_gpu64, d_gpu128, d_64_128 = rel(g, a), rel(g, b), rel(a, b)
        print(f"{name} {key}: gpu-c64 {d_gpu64:.2e} gpu-c128 {d_gpu128:.2e} c64-c128 {d_64_128:.2e}", flush=True)
        floor = 1e-5 if key == "error" else 1e-6
        assert d_gpu128 <= 2.0 * d_64_128 + floor, (name, key)
        if prop in ("ASP", "polychromeASP"):
            assert d_gpu64 < 1e-5, (name, key)


@pytest.mark.parametrize("prop,nl", [("scaledASP", 1), ("scaledPolychromeASP", 3)])
def test_scaled_asp_engine_matches_oracle(prop, nl):
    """ePIE with the scaled angular spectrum propagator on the reference's
    regression dataset.  The reference holds no engine golden for scaledASP;
    the oracle's scaledASP reproduces its propagator_scaledasp golden, and
    the GPU run is held to the golden-input criterion above (its complex64
    chain keeps Q1, Q2 complex64, as the reference stores them)."""
    d = regression_dataset()
    modes = (nl, 1, 1, 1)
    print(f"case epie {prop}", flush=True)
    gpu = run_package(d, "ePIE", prop, modes, 2, SEED)
    c64 = run_oracle(d, "ePIE", prop, modes, 2, SEED, precision="c64")
    c128 = run_oracle(d, "ePIE", prop, modes, 2, SEED, precision="c128")
    for key in ("object", "probe", "error"):
        g, a, b = np.asarray(getattr(gpu, key)), np.asarray(getattr(c64, key)), np.asarray(getattr(c128, key))
        d_gpu64, d_gpu128, d_64_128 = rel(g, a), rel(g, b), rel(a, b)
        print(f"epie_{prop} {key}: gpu-c64 {d_gpu64:.2e} gpu-c128 {d_gpu128:.2e} c64-c128 {d_64_128:.2e}", flush=True)
        floor = 1e-5 if key == "error" else 1e-6
        assert d_gpu128 <= 2.0 * d_64_128 + floor, key
