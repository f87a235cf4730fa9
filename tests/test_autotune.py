"""Cluster-size allocation of the autotuner (host logic, no GPU): the
greedy doubling against brute force on small batches, capacity and
divisibility constraints, and the error for batches that cannot fit."""

import itertools

import pytest

from ptylab_amd import autotune as A


def _cost(shape, G):
    # work splits G ways; every extra workgroup adds barrier overhead
    work = shape.Np * shape.Np * shape.nF * 1e-3
    return work / G + 0.8 * (G > 1) + 0.05 * G


def _jobs(specs):
    return [A.Job(A.Shape(1, 0, Np, 4 * Np, nF), F, it) for Np, nF, F, it in specs]


def _brute(jobs, cap):
    best = None
    choices = [A.candidate_sizes(j.shape, cap) for j in jobs]
    for G in itertools.product(*choices):
        if sum(G) > cap:
            continue
        m = A.predicted_makespan(jobs, _cost, G)
        best = m if best is None else min(best, m)
    return best


@pytest.mark.parametrize("cap", [3, 4, 6, 9, 16, 40])
def test_greedy_matches_brute_force(cap):
    jobs = _jobs([(64, 1, 300, 10), (128, 1, 120, 10), (32, 2, 700, 10)])
    G = A.assign(jobs, _cost, cap)
    assert sum(G) <= cap
    assert all(j.shape.Np % g == 0 and g in A.CANDIDATES for j, g in zip(jobs, G))
    got = A.predicted_makespan(jobs, _cost, G)
    assert got == pytest.approx(_brute(jobs, cap), rel=1e-12)


def test_never_worse_than_one_workgroup_each():
    jobs = _jobs([(256, 4, 800, 100), (64, 1, 100, 100), (128, 1, 450, 50), (64, 1, 800, 100)])
    G = A.assign(jobs, _cost, 256)
    assert A.predicted_makespan(jobs, _cost, G) <= A.predicted_makespan(jobs, _cost, [1] * len(jobs))
    assert sum(G) <= 256
    # the most expensive instance gets the largest cluster
    assert G[0] == max(G)


def test_capacity_equal_to_batch_size_gives_ones():
    jobs = _jobs([(64, 1, 100, 1)] * 5)
    assert A.assign(jobs, _cost, 5) == [1] * 5


def test_too_many_instances_raise():
    with pytest.raises(ValueError):
        A.assign(_jobs([(64, 1, 100, 1)] * 5), _cost, 4)


def test_launch_keys_split_by_family_and_size():
    from ptylab_amd import _native

    a = _native.PtyInstance()
    a.engine, a.propagator, a.Np = _native.PTY_MPIE, _native.PTY_FRAUNHOFER, 128
    b = _native.PtyInstance()
    b.engine, b.propagator, b.Np = _native.PTY_EPIE, _native.PTY_FRAUNHOFER, 64
    c = _native.PtyInstance()
    c.engine, c.propagator, c.Np = _native.PTY_MPIE, _native.PTY_ASP, 64
    d = _native.PtyInstance()
    d.engine, d.propagator, d.Np = _native.PTY_EPIE, _native.PTY_FRAUNHOFER, 512
    keys = [A.launch_key(x) for x in (a, b, c, d)]
    assert keys[0] == keys[1] != keys[2]
    assert keys[3] not in keys[:3]


def test_cost_table_round_trips_json(tmp_path):
    t = A.CostTable(tmp_path / "t.json")
    s = A.Shape(1, 0, 64, 256, 1)
    t.table[s.key(2)] = 12.5
    t.save()
    assert A.CostTable(tmp_path / "t.json")(s, 2) == 12.5


def test_packing_follows_the_plan_placement():
    # largest first, each to the least-loaded bin
    assert A.packs([4, 2, 2], 2, 4)
    assert not A.packs([4, 4, 1], 2, 4)
    assert A.packs([1] * 8, 8, 1) and not A.packs([1] * 9, 8, 1)
    # powers of two up to the bin size pack whenever the total fits
    assert A.packs([2, 1, 1, 2, 2], 2, 4)


@pytest.mark.parametrize("bins,per_bin", [(2, 4), (4, 4), (8, 2), (3, 8)])
def test_greedy_with_bins_matches_brute_force(bins, per_bin):
    jobs = _jobs([(64, 1, 300, 10), (128, 1, 120, 10), (32, 2, 700, 10), (64, 1, 500, 10)])
    cap = bins * per_bin
    G = A.assign(jobs, _cost, cap, bins)
    assert A.packs(G, bins, per_bin) and max(G) <= per_bin
    best = None
    for c in itertools.product(*[A.candidate_sizes(j.shape, per_bin) for j in jobs]):
        if A.packs(list(c), bins, per_bin):
            m = A.predicted_makespan(jobs, _cost, c)
            best = m if best is None else min(best, m)
    assert A.predicted_makespan(jobs, _cost, G) == pytest.approx(best, rel=1e-12)


def test_band_of_counts_live_transfer_function_columns():
    """The cost key of an ASP instance carries the number of transfer
    function columns inside the band limit (the column passes skip the rest);
    the real band-limited ASP transfer function of the benchmark workload
    (Np = 256, z = 50 mm, dxo = 4 um) keeps 33 columns."""
    import numpy as np
    import torch

    from ptylab_amd import _native
    from ptylab_amd import synthetic as S

    inst = _native.PtyInstance()
    inst.engine, inst.propagator, inst.Np, inst.nlambda = _native.PTY_MPIE, _native.PTY_ASP, 256, 1
    H = S.asp_tf_fftorder(0.05, S.WAVELENGTH, 256, 256 * 4e-6)
    assert A.band_of(inst, {"tf": torch.as_tensor(H.astype(np.complex64))}) == 33
    assert A.band_of(inst, {"tf": torch.ones(256, 256, dtype=torch.complex64)}) == 0
    assert A.band_of(inst, None) == 0
    inst.propagator = _native.PTY_FRAUNHOFER
    assert A.band_of(inst, {"tf": torch.as_tensor(H.astype(np.complex64))}) == 0
    assert A.Shape(1, 2, 256, 1024, 4, 33).key(8) != A.Shape(1, 2, 256, 1024, 4).key(8)
