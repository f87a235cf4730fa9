"""Cluster-size autotuner for batched plans.

A batched plan runs many independent reconstructions in one persistent
launch per iteration; instance i runs on a cluster of G_i workgroups that
splits every line and column pass of a position G_i ways.  Larger clusters
cut an instance's time per position but cost more workgroups and more
barrier traffic, and the whole launch must stay co-resident: each cluster
runs whole on one of the N_XCD XCDs, so the clusters must pack into N_XCD
bins of pty_coresident_workgroups / N_XCD workgroups each, in the placement
the plan makes (largest first, each to the least-loaded XCD).  An iteration
of the batch lasts as
long as its slowest instance, so the tuner

1. measures the time per position t(G) of the batched kernel for every
   instance shape in the batch (pass family, engine, Np, No, field count) and
   every candidate cluster size G (powers of two), on synthetic data of that
   shape, with HIP events; results are cached in a JSON file;
2. assigns cluster sizes by repeatedly doubling the cluster of the instance
   with the longest predicted run time (n_iter x F x t(G)) while the
   capacity allows and the doubling shortens that instance.

The reference has no batching and no tuner; this module chooses launch
parameters only and never changes results (every cluster size gives
bit-identical reconstructions).
"""

from __future__ import annotations

import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native

CANDIDATES = (1, 2, 4, 8, 16, 32, 64)
N_XCD = 8  # XCDs of the device; a batched cluster runs whole on one of them
BATCH_MAX_NP = 256  # largest probe size of the batched kernel
_DEFAULT_CACHE = Path(os.environ.get("PTYLAB_AMD_TUNE_CACHE", Path.home() / ".cache" / "ptylab_amd" / "tune.json"))


@dataclass(frozen=True)
class Shape:
    """What the per-position cost of an instance depends on."""

    engine: int
    propagator: int
    Np: int
    No: int
    nF: int  # exit-wave fields (nlambda * nosm * npsm * nslice)
    band: int = 0  # ASP: transfer-function columns inside the band limit (0: every column)

    def key(self, G):
        b = f"/B{self.band}" if self.band else ""
        return f"e{self.engine}p{self.propagator}/Np{self.Np}/No{self.No}/nF{self.nF}{b}/G{G}"


@dataclass
class Job:
    shape: Shape
    F: int
    n_iter: int


def family(engine, propagator):
    """Pass family of the batched kernel an instance runs in."""
    asp = propagator in (_native.PTY_ASP, _native.PTY_POLYCHROME_ASP)
    if engine == _native.PTY_E3PIE:
        return 2
    if engine == _native.PTY_QNEWTON:
        return 4 if asp else 3
    if propagator in (_native.PTY_SCALED_ASP, _native.PTY_SCALED_POLYCHROME_ASP):
        return 5
    return 1 if asp else 0


def candidate_sizes(shape: Shape, capacity: int):
    """Powers of two that divide Np, at most the capacity."""
    return [g for g in CANDIDATES if g <= capacity and shape.Np % g == 0]


def packs(G, bins, per_bin):
    """Whether clusters of sizes G fit the plan's placement: stable
    largest-first order, each cluster to the least-loaded bin (lowest index
    on ties), no bin above per_bin (pty_plan_create's XCD packing)."""
    used = [0] * bins
    for i in sorted(range(len(G)), key=lambda i: -G[i]):
        x = min(range(bins), key=lambda y: used[y])
        if used[x] + G[i] > per_bin:
            return False
        used[x] += G[i]
    return True


def assign(jobs, cost, capacity, bins=1):
    """Cluster sizes minimising the batch makespan.

    ``cost(shape, G)`` is the time per position (any unit); ``capacity`` the
    workgroups that fit at once, split into ``bins`` equal bins that each
    cluster must fit whole.  Returns a list of cluster sizes, one per job,
    that pack into the bins (ValueError if even all-ones does not).
    """
    n = len(jobs)
    per_bin = capacity // bins
    if not packs([1] * n, bins, per_bin):
        raise ValueError(f"{n} instances need at least {n} workgroups; {capacity} fit on the device")
    G = [1] * n
    T = [j.n_iter * j.F * cost(j.shape, 1) for j in jobs]
    blocked = [False] * n
    while True:
        order = sorted((i for i in range(n) if not blocked[i]), key=lambda i: -T[i])
        if not order:
            break
        i = order[0]
        if T[i] < max(T):
            break  # the longest instance cannot be shortened further
        g2 = 2 * G[i]
        if g2 not in candidate_sizes(jobs[i].shape, per_bin) or not packs(G[:i] + [g2] + G[i + 1:], bins, per_bin):
            blocked[i] = True
            continue
        t2 = jobs[i].n_iter * jobs[i].F * cost(jobs[i].shape, g2)
        if t2 >= T[i]:
            blocked[i] = True
            continue
        G[i], T[i] = g2, t2
    return G


def predicted_makespan(jobs, cost, G):
    return max(j.n_iter * j.F * cost(j.shape, g) for j, g in zip(jobs, G))


class CostTable:
    """Measured microseconds per position, keyed by Shape.key(G), JSON-cached."""

    def __init__(self, path: os.PathLike | None = None, positions: int = 48):
        self.path = Path(path) if path is not None else _DEFAULT_CACHE
        self.positions = positions
        self.table = {}
        if self.path.exists():
            try:
                self.table = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self.table = {}

    def save(self):
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            self.path.write_text(json.dumps(self.table, indent=1, sort_keys=True))
        except OSError:
            pass

    def __call__(self, shape: Shape, G: int) -> float:
        k = shape.key(G)
        if k not in self.table:
            self.table[k] = measure(shape, G, self.positions)
            self.save()
        return self.table[k]


def measure(shape: Shape, G: int, positions: int = 48, seed: int = 0) -> float:
    """Microseconds per position of the batched kernel for one instance of
    this shape on a cluster of G workgroups (synthetic data, momentum
    triggers at the mPIE rate for the momentum engines)."""
    import torch

    dev = torch.device("cuda")
    rng = np.random.default_rng(seed)
    Np, No, nF = shape.Np, shape.No, shape.nF
    e3 = shape.engine == _native.PTY_E3PIE
    nslice = 2 if e3 else 1
    npsm = nF
    F = positions
    gen = torch.Generator(device="cpu").manual_seed(seed)

    def crand(*s):
        return torch.complex(torch.rand(*s, generator=gen) + 0.5, torch.rand(*s, generator=gen) - 0.5).to(dev)

    T = {}
    T["object"] = crand(1, 1, 1, nslice, No, No)
    T["probe"] = crand(1, 1, npsm, nslice, Np, Np)
    T["ptychogram"] = torch.rand(F, Np, Np, generator=gen).to(dev)
    pos = rng.integers(0, No - Np + 1, size=(F, 2)).astype(np.int32)
    T["positions"] = torch.as_tensor(pos, device=dev)
    n_iter = 2
    T["order"] = torch.as_tensor(np.tile(np.arange(F, dtype=np.int32), (n_iter, 1)), device=dev)
    momentum = shape.engine in (_native.PTY_MPIE, _native.PTY_MULTIPIE)
    T["mask"] = torch.as_tensor((rng.random((n_iter, F)) > 0.95).astype(np.uint8), device=dev)
    T["energy"] = torch.full((F,), float(Np * Np), device=dev)
    T["err_at_pos"] = torch.zeros((n_iter, F), dtype=torch.float64, device=dev)
    T["err_iter"] = torch.zeros((n_iter,), dtype=torch.float64, device=dev)
    asp_like = shape.propagator in (_native.PTY_ASP, _native.PTY_POLYCHROME_ASP)
    scaled = shape.propagator in (_native.PTY_SCALED_ASP, _native.PTY_SCALED_POLYCHROME_ASP)
    if shape.propagator == _native.PTY_FRESNEL:
        # Fresnel: the quadratic phase multiplied into the row passes
        ph = torch.rand(Np, Np, generator=gen) * 6.283
        T["tf"] = torch.polar(torch.ones(Np, Np), ph).to(torch.complex64).to(dev)
    if scaled:
        # scaled ASP: the transfer function Q2 and the phase c Q1 (one
        # wavelength here: the fields count as probe modes)
        ph = torch.rand(2, Np, Np, generator=gen) * 6.283
        T["tf"] = torch.polar(torch.ones(2, Np, Np), ph).to(torch.complex64).to(dev)
    if asp_like or e3:
        ph = torch.rand(Np, Np, generator=gen) * 6.283
        mag = torch.ones(Np, Np)
        if shape.band and not e3:
            # the column passes skip the columns outside the band limit:
            # keep `band` columns around zero frequency (FFT order)
            keep = torch.zeros(Np, dtype=torch.bool)
            keep[:(shape.band + 1) // 2] = True
            if shape.band // 2:
                keep[Np - shape.band // 2:] = True
            mag[:, ~keep] = 0.0
        T["tf"] = torch.polar(mag, ph).to(torch.complex64).to(dev)
    if momentum:
        for k in ("object", "probe"):
            T[k + "Buffer"] = T[k].clone()
            T[k + "Momentum"] = torch.zeros_like(T[k])
    inst = _native.PtyInstance()
    inst.engine, inst.propagator = shape.engine, shape.propagator
    inst.nlambda, inst.nosm, inst.npsm, inst.nslice = 1, 1, npsm, nslice
    inst.Np, inst.No, inst.F, inst.n_iter = Np, No, F, n_iter
    inst.betaObject = inst.betaProbe = 0.25
    inst.alphaObject = inst.alphaProbe = 0.1
    inst.momFriction, inst.momFeedback, inst.probeWeight, inst.sliceBeta = 0.7, 0.3, 1.0, 0.9
    p = lambda k: T[k].data_ptr() if k in T else None  # noqa: E731
    inst.object, inst.probe = p("object"), p("probe")
    inst.objectBuffer, inst.objectMomentum = p("objectBuffer"), p("objectMomentum")
    inst.probeBuffer, inst.probeMomentum = p("probeBuffer"), p("probeMomentum")
    inst.ptychogram, inst.positions, inst.order = p("ptychogram"), p("positions"), p("order")
    inst.momentum_mask = p("mask") if momentum else None
    inst.energy, inst.tf = p("energy"), p("tf")
    inst.err_at_pos, inst.err_iter = p("err_at_pos"), p("err_iter")
    inst.workgroups = G
    inst.flags = _native.PTY_FLAG_BATCHED
    stream = torch.cuda.current_stream()
    plan = _native.Plan([inst], T, stream)
    try:
        plan.use_graph(False)
        plan.run(0, 1)
        plan.check()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        plan.run(1, 2)
        e1.record(stream)
        plan.check()
        return e0.elapsed_time(e1) * 1e3 / F
    finally:
        plan.close()


def shape_of(inst, band: int = 0) -> Shape:
    e3 = inst.engine == _native.PTY_E3PIE
    nF = inst.nlambda * inst.npsm if e3 else inst.nlambda * inst.nosm * inst.npsm * inst.nslice
    return Shape(int(inst.engine), int(inst.propagator), int(inst.Np), int(inst.No), int(nF), int(band))


def band_of(inst, tensors) -> int:
    """Columns of the instance's ASP / polychromeASP transfer function (the
    widest over wavelengths) that hold a nonzero value, or 0 when every
    column does or the propagator is another: the column passes skip the
    others, so the per-position cost depends on it."""
    if inst.propagator not in (_native.PTY_ASP, _native.PTY_POLYCHROME_ASP) or inst.engine == _native.PTY_E3PIE:
        return 0
    tf = tensors.get("tf") if tensors is not None else None
    if tf is None:
        return 0
    Np = int(inst.Np)
    ntf = int(inst.nlambda) if inst.propagator == _native.PTY_POLYCHROME_ASP else 1
    h = tf.reshape(-1)[: ntf * Np * Np].reshape(ntf, Np, Np)
    n = int((h != 0).any(dim=1).sum(dim=-1).max().item())
    return n if n < Np else 0


def launch_key(inst):
    """Instances with the same key share one launch: the batched kernel of
    a pass family for Np <= 256, a per-size kernel otherwise."""
    fam = family(inst.engine, inst.propagator)
    return (fam, 0) if inst.Np <= BATCH_MAX_NP else (fam, int(inst.Np))


def tune(instances, table: CostTable | None = None, tensors=None):
    """Cluster sizes for a list of pty_instance (one batched plan).  The
    launches of a plan run one after another, each alone on the device, so
    every launch is assigned against its own co-resident capacity.
    ``tensors``: the instances' device arrays (dicts as passed to the plan),
    from which the ASP band limits are read; without them every column of a
    transfer function counts as live."""
    table = table if table is not None else CostTable()
    bands = [band_of(inst, tensors[i] if tensors is not None else None) for i, inst in enumerate(instances)]
    out = [1] * len(instances)
    launches = {}
    for i, inst in enumerate(instances):
        launches.setdefault(launch_key(inst), []).append(i)
    for (fam, np_), idx in launches.items():
        a = instances[idx[0]]
        cap = _native.coresident_workgroups(np_ or BATCH_MAX_NP, a.engine, a.propagator, np_ == 0)
        jobs = [Job(shape_of(instances[i], bands[i]), int(instances[i].F), int(instances[i].n_iter)) for i in idx]
        G = assign(jobs, table, cap, N_XCD if np_ == 0 else 1)
        for i, g in zip(idx, G):
            out[i] = g
    return out
