"""Sources: ``Beam`` (wakis/sources.py:22-110) and ``PlaneWave``
(wakis/sources.py:142-248).

``Beam.update`` evaluates the Gaussian bunch line exactly as the reference
(FP64 numpy, same expression order) and hands the Nz values to the solver's
device beam line, which the E update uses as J_z(ixs, iys, :).  With a plain
host ``Field`` solver it writes ``J`` like the reference.
"""

from __future__ import annotations

import numpy as np
from scipy.constants import c as c_light
from scipy.constants import mu_0


class Beam:
    def __init__(self, xsource=0.0, ysource=0.0, beta=1.0, q=1e-9, sigmaz=None, ti=None):
        self.xsource, self.ysource = xsource, ysource
        self.sigmaz = sigmaz
        self.q = q
        self.beta = beta
        self.v = c_light * beta
        if ti is not None:
            self.ti = ti
        else:
            self.ti = 8.548921333333334 * self.sigmaz / self.v
        self.is_first_update = True

    def _first(self, solver):
        self.ixs = np.abs(solver.x - self.xsource).argmin()
        self.iys = np.abs(solver.y - self.ysource).argmin()
        self.is_first_update = False
        self.zmin = solver.z.min()

    def line(self, solver, t):
        """J_z values on the beam line at time t (wakis/sources.py:96-110)."""
        if self.is_first_update:
            self._first(solver)
        s0 = self.zmin - self.v * self.ti
        s = solver.z - self.v * t
        profile = (1 / np.sqrt(2 * np.pi * self.sigmaz**2)
                   * np.exp(-((s - s0) ** 2) / (2 * self.sigmaz**2)))
        return self.q * self.v * profile / solver.dx[self.ixs] / solver.dy[self.iys]

    def device_params(self, solver):
        """Parameters of the in-kernel evaluation of ``line`` (FIT_BEAM_GAUSS):
        q*v, 1/sqrt(2*pi*sigmaz^2), 2*sigmaz^2, s0, v, dx[ixs], dy[iys], and
        1.0 when t = n*dt and v*t are float32 products (dt a numpy float32, as
        numpy evaluates them for a float32 solver), else 0.0."""
        if self.is_first_update:
            self._first(solver)
        return np.array([
            self.q * self.v,
            1 / np.sqrt(2 * np.pi * self.sigmaz**2),
            2 * self.sigmaz**2,
            self.zmin - self.v * self.ti,
            self.v,
            solver.dx[self.ixs],
            solver.dy[self.iys],
            1.0 if isinstance(solver.dt, np.float32) else 0.0,
        ], dtype=np.float64)

    def update(self, solver, t):
        vals = self.line(solver, t)
        if hasattr(solver, "_set_beam_line"):
            solver._set_beam_line(int(self.ixs), int(self.iys), vals)
        else:
            solver.J[self.ixs, self.iys, :, "z"] = vals


class PlaneWave:
    def __init__(self, xs=None, ys=None, zs=0, nodes=None, f=None, amplitude=1.0, beta=1.0,
                 phase=0):
        self.nodes = nodes
        self.beta = beta
        self.xs, self.ys = xs, ys
        self.zs = zs
        self.f = f
        self.amplitude = amplitude
        self.is_first_update = True
        self.phase = phase
        self.vp = self.beta * c_light
        self.w = 2 * np.pi * self.f
        self.kz = self.w / c_light
        self.tmax = np.inf
        if self.nodes is not None:
            self.tmax = self.nodes / self.f

    def update(self, solver, t):
        if self.is_first_update:
            if self.xs is None:
                self.xs = slice(0, solver.Nx)
            if self.ys is None:
                self.ys = slice(0, solver.Ny)
            self.is_first_update = False
        if t <= self.tmax:
            solver.H[self.xs, self.ys, self.zs, "y"] = -self.amplitude * np.cos(self.w * t + self.phase)
            solver.E[self.xs, self.ys, self.zs, "x"] = (
                self.amplitude * np.cos(self.w * t + self.phase) / (self.kz / (mu_0 * self.vp)))
        else:
            solver.H[self.xs, self.ys, self.zs, "y"] = 0.0
            solver.E[self.xs, self.ys, self.zs, "x"] = 0.0
