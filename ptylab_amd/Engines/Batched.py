"""Many independent reconstructions in one persistent launch per iteration.

The reference reconstructs one dataset per engine call; a batch of datasets
is a Python loop over ``engine.reconstruct()`` (ePIE.py:53-112,
mPIE.py:111-180), one after another.  ``Batched`` takes such a list of
engines, already set up exactly as for their own ``reconstruct()``, and runs
them together: every instance with Np <= 256 of one pass family shares one
launch of the batched kernel, replayed as a HIP graph once per iteration,
with cluster sizes from ``ptylab_amd.autotune``.  Each engine ends in the
state its own ``reconstruct()`` would have left: the NumPy global RNG is
drawn per engine in list order (the schedules of a sequential loop), and
the reconstructions are bit-identical to separate runs.
"""

from __future__ import annotations

import numpy as np

from .. import _native, autotune


class Batched:
    def __init__(self, engines, workgroups=None, graph=True, tuner=None):
        """``engines``: engine objects (ePIE, mPIE, multiPIE, e3PIE), each with
        its reconstruction, data, params and numIterations set.
        ``workgroups``: cluster size per engine (list), or None to autotune.
        ``graph``: replay a per-iteration HIP graph (True) or launch directly.
        ``tuner``: an autotune.CostTable (default: the JSON-cached table)."""
        self.engines = list(engines)
        if not self.engines:
            raise ValueError("Batched needs at least one engine")
        self.workgroups = workgroups
        self.graph = graph
        self.tuner = tuner
        self.last_workgroups = None
        self.last_plan_workgroups = None  # active workgroups of the last plan

    def reconstruct(self):
        import torch

        for e in self.engines:
            if e.params.positionCorrectionSwitch and e._POSITION_CORRECTION:
                raise NotImplementedError("position correction is not run in batched launches; call the engine's "
                                          "own reconstruct()")
            e._prepareReconstruction()
        staged = [e._stage() for e in self.engines]
        insts = [s[0] for s in staged]
        for inst in insts:
            inst.flags |= _native.PTY_FLAG_BATCHED
        if self.workgroups is not None:
            G = [int(g) for g in self.workgroups]
            if len(G) != len(insts):
                raise ValueError("one workgroup count per engine")
        else:
            G = autotune.tune(insts, self.tuner, [s[1] for s in staged])
        for inst, g in zip(insts, G):
            inst.workgroups = g
        self.last_workgroups = G
        plan = _native.Plan(insts, [s[1] for s in staged], torch.cuda.current_stream())
        plan.use_graph(self.graph)
        self.last_plan_workgroups = plan.workgroups()
        n_iter = [int(e.numIterations) for e in self.engines]
        done = 0
        try:
            for loop in range(max(n_iter)):
                plan.run(loop, loop + 1)
                done = loop + 1
                for e, n in zip(self.engines, n_iter):
                    if loop < n:
                        e.applyConstraints(loop)
            plan.check()
        finally:
            for e, n in zip(self.engines, n_iter):
                if getattr(e, "_T", None) is not None:
                    e._collect(min(done, n))
            plan.close()
        return self

    @property
    def position_updates(self):
        """Position updates of one full run (sum of F x numIterations)."""
        return int(np.sum([e.experimentalData.numFrames * int(e.numIterations) for e in self.engines]))
