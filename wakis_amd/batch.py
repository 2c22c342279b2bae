"""Batched stepping and batched wake solves over many independent solvers.

The reference runs one ``SolverFIT3D`` at a time (wakis/routines.py:231-485).
Here many device solvers (same dtype, same device) are stepped together by
one HIP launch per step (``fit_batch_*`` in include/wakis_fit.h):
each member runs the same per-tile arithmetic as in its own ``one_step``,
so its fields, wake potential and impedance are the same alone and in the
batch (with the same tile configuration).  Launch groups of ``group`` steps
can be captured into a HIP graph and replayed.
"""

from __future__ import annotations

import time

import numpy as np

from . import _lib


class BatchStepper:
    """Steps a set of device solvers together; members may change between
    calls (``set_members``)."""

    def __init__(self, solvers, graph: bool = True, group: int = 16, tile=None,
                 two_step: bool = True, kc2: int = 0, cfg2: int = -1):
        """two_step: batches whose members all suit the two-step sweep
        (FP32: nx >= 32 and ny >= 16, or nx, ny >= 24; FP64: nx >= 28, ny >= 16) advance pairs of
        steps with it (one read and write of E and H per two steps); False
        keeps the one-step launch.  kc2, cfg2: planes per z chunk and tile
        configuration (FP32 0: 32 x 16, 1: 24 x 24; FP64 2: 28 x 12) of the
        two-step sweep (0, -1: by its launch model)."""
        self.solvers = list(solvers)
        for s in self.solvers:
            s._require_device()
        self._b = _lib.Batch([s.ctx for s in self.solvers],
                             [int(bool(s.use_conductivity)) for s in self.solvers])
        self._b.set_graph(graph, group)
        self._b.set_two_step(two_step, cfg2, kc2)
        if tile is not None:
            self._b.set_tile(int(tile))

    def set_members(self, solvers):
        self.solvers = list(solvers)
        if self.solvers:
            self._b.set_members([s.ctx for s in self.solvers],
                                [int(bool(s.use_conductivity)) for s in self.solvers])

    def step(self, nsteps: int) -> float:
        """Advance every member nsteps (as nsteps calls of its one_step);
        returns the device time in ms."""
        if nsteps <= 0 or not self.solvers:
            return 0.0
        for s in self.solvers:
            if s.step_0:
                s.ctx.zero_ghosts()
                s.step_0 = False
        return self._b.step(int(nsteps))

    def tune(self, reps: int = 3):
        """Time every tile configuration on this batch and keep the fastest;
        returns {tile id: ms per step}."""
        ms = self._b.tune(reps)
        return {i: float(v) for i, v in enumerate(ms) if v > 0}

    def tile(self):
        return self._b.get_tile()

    def two_step(self):
        """(pairs by the two-step sweep on the next step call, its tile
        configuration, planes per chunk)."""
        return self._b.two_step()

    def tune_two_step(self, reps: int = 3):
        """Time the two-step launch over its tile configurations and chunk
        depths on this batch and keep the fastest; returns (cfg, kc,
        {(cfg, kc): ms per launch})."""
        return self._b.tune_two_step(reps)

    def device_bytes(self) -> int:
        """Device bytes of the batch's own tables (the members' are theirs)."""
        return self._b.device_bytes()

    def close(self):
        self._b.close()


def wakesolve_batch(solvers, wakelength=None, compute_plane="longitudinal", graph=True,
                    group=16, tile=None, tune=False, max_chunk=None, two_step=True):
    """``wakesolve`` of every solver in one batch: the same setup, device
    beam, E_z probe and wake post-processing as ``SolverFIT3D.wakesolve``,
    with the time steps of all members taken by batched launches.  A member
    leaves the batch when it has taken its own Nt steps, so each solver ends
    in exactly the state its own wakesolve would leave (with two_step, to
    FP32 rounding: see BatchStepper).  Returns the seconds spent stepping."""
    solvers = list(solvers)
    runs = [s._wake_setup(wakelength, None, compute_plane) for s in solvers]
    for s, r in zip(solvers, runs):
        s._wake_device_beam(r)
    done = [0] * len(solvers)
    active = [i for i, r in enumerate(runs) if r["Nt"] > 0]
    stepper = BatchStepper([solvers[i] for i in active] or solvers[:1], graph=graph,
                           group=group, tile=tile, two_step=two_step)
    if tune and active:
        stepper.tune()
    t0 = time.time()
    while active:
        n = min(runs[i]["Nt"] - done[i] for i in active)
        if max_chunk:
            n = min(n, int(max_chunk))
        stepper.set_members([solvers[i] for i in active])
        stepper.step(n)
        for i in active:
            done[i] += n
        active = [i for i in active if done[i] < runs[i]["Nt"]]
    elapsed = time.time() - t0
    stepper.close()
    for s, r in zip(solvers, runs):
        s.ctx.sync()
        s.simulationTime = elapsed
        s._wake_device_end(r)
        s._wake_finish(r, compute_plane)
    return elapsed
