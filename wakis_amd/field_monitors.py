"""Frequency-domain field monitor (wakis/field_monitors.py:14-131).

``FieldMonitor(frequencies)`` sums snapshots of E with exp(-2j*pi*f*t) for
each monitored frequency, t = time_index * dt (dt fixed at the first
update), exactly as the reference.  When the snapshot is a solver's device
field the complex128 accumulators live on the device: an update is one
kernel (``fit_monitor_accumulate``), and inside ``wakesolve`` the updates
run on the device after every step of a launch sequence, with the phases of
the steps computed here by numpy with the reference's expression and staged
ahead (``fit_monitor_phases``).  For a host ``Field`` (or any object with
``to_matrix`` and ``xp``) the accumulation is the reference's numpy code.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np


class FieldMonitor:
    def __init__(self, frequencies: Sequence[float]):
        self.frequencies = np.array(frequencies)
        self.time_index = 0
        self.dt = None
        self.shape = None
        self.xp = None
        self._acc = {"Ex": None, "Ey": None, "Ez": None}
        self._ctx = None  # device context holding the accumulators

    # accumulators (downloaded from the device when they live there)
    @property
    def Ex_acc(self):
        return self._component("Ex")

    @property
    def Ey_acc(self):
        return self._component("Ey")

    @property
    def Ez_acc(self):
        return self._component("Ez")

    def _component(self, key):
        if self._ctx is None:
            return self._acc[key]
        c = "xyz".index(key[1])
        return np.stack([self._ctx.monitor_read(i, c) for i in range(len(self.frequencies))])

    def _phases(self, t):
        """The nf phases at time t: wakis/field_monitors.py:99."""
        return np.array([np.exp(-2j * np.pi * f * t) for f in self.frequencies],
                        dtype=np.complex128)

    def _start(self, E, dt):
        self.dt = dt
        self.shape = (E.Nx, E.Ny, E.Nz)
        self.xp = E.xp
        ctx = getattr(E, "ctx", None)
        if ctx is not None and len(self.frequencies) > 0:
            ctx.monitor_set(len(self.frequencies))
            self._ctx = ctx
        else:
            shape = (len(self.frequencies), *self.shape)
            for key in self._acc:
                self._acc[key] = self.xp.zeros(shape, dtype=self.xp.complex128)

    def update(self, E, dt: float):
        """wakis/field_monitors.py:59-104."""
        if self.dt is None:
            self._start(E, dt)
        t = self.time_index * self.dt
        if self._ctx is not None:
            self._ctx.monitor_accumulate(self._phases(t))
        else:
            Ex = E.to_matrix("x")
            Ey = E.to_matrix("y")
            Ez = E.to_matrix("z")
            for i, f in enumerate(self.frequencies):
                phase = self.xp.exp(-2j * self.xp.pi * f * t)
                self._acc["Ex"][i] += Ex * phase
                self._acc["Ey"][i] += Ey * phase
                self._acc["Ez"][i] += Ez * phase
        self.time_index += 1

    def _stage(self, E, dt, nsteps: int):
        """Stage the next nsteps updates (after each step of the solver's next
        launch sequence) on the device; advances time_index."""
        if self.dt is None:
            self._start(E, dt)
        if self._ctx is None or self._ctx is not getattr(E, "ctx", None):
            raise ValueError("the monitor accumulates on another device context")
        rows = np.empty((nsteps, len(self.frequencies)), dtype=np.complex128)
        for s in range(nsteps):
            rows[s] = self._phases((self.time_index + s) * self.dt)
        self._ctx.monitor_phases(rows)
        self.time_index += nsteps

    def get_components(self):
        """wakis/field_monitors.py:106-131."""
        return {"Ex": self.Ex_acc, "Ey": self.Ey_acc, "Ez": self.Ez_acc}
