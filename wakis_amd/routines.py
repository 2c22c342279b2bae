"""Time loops ``emsolve`` and ``wakesolve`` (wakis/routines.py:17-229,
231-485) over the device solver.

``wakesolve`` keeps the reference's loop semantics (beam update, then the E_z
probe of the state after n steps, then one step; wakis/routines.py:438-448)
but runs them on the device: the bunch current is evaluated in the E-update
kernel and the E_z window is recorded into an HBM history by the same kernel,
so without a callback the whole run is one asynchronous launch sequence with
no host synchronisation per step.  The history replaces the ``Ez.h5`` file;
the keys, t axis and window of that file are kept as metadata for the
``WakeSolver``.
"""

from __future__ import annotations

import time

import numpy as np

from . import _lib
from .sources import Beam


class RoutinesMixin:
    def emsolve(self, Nt, source=None, callback=None, save=False, fields=("E",),
                components=("Abs",), save_every=1, subdomain=None, plot=False,
                plot_every=1, use_etd=False, plot3d=False, **kwargs):
        """wakis/routines.py:17-229.  ``save`` keeps the requested slices in
        ``self.saved[name]['#%05d']`` (in memory: h5py is not used here)."""
        self.Nt = Nt
        if source is not None:
            self.source = source
        if save:
            self.saved = {}
            names = []
            for f in fields:
                if len(f) == 1:
                    names += [f + c for c in components]
                else:
                    names.append(f)
            for name in names:
                self.saved[name] = {"x": self.x, "y": self.y, "z": self.z,
                                    "dx": self.grid.dx, "dy": self.grid.dy, "dz": self.grid.dz,
                                    "t": np.arange(0, Nt * self.dt, save_every * self.dt)}
            if subdomain is not None:
                xx, yy, zz = subdomain
            else:
                xx, yy, zz = slice(0, self.Nx), slice(0, self.Ny), slice(0, self.Nz)
        abc_mask, abc_conflict = self._abc_low_mask() if self.activate_abc else (0, False)
        for n in range(Nt):
            if source is not None:
                source.update(self, n * self.dt)
            if save and n % save_every == 0:
                for name in self.saved:
                    self.saved[name]["#" + str(n).zfill(5)] = getattr(self, name[0])[
                        xx, yy, zz, name[1:]]
            self.one_step()
            if abc_mask:
                self.ctx.apply_abc(abc_mask)
            if abc_conflict:
                # low+high ABC on one axis: update_abc fails after the first step
                raise KeyError("xlo")
            if callback is not None:
                callback(self, n * self.dt)

    def wakesolve(self, wakelength=None, wake=None, callback=None, compute_plane="both",
                  plot=False, plot_func=None, plot_from=None, plot_every=1, plot_until=None,
                  save_J=False, add_space=None, use_field_monitor=False, field_monitor=None,
                  use_edt=None, chunk=4096, **kwargs):
        """wakis/routines.py:231-485 on the device (see module docstring)."""
        if plot:
            raise NotImplementedError("plotting is not part of this build")
        from .field_monitors import FieldMonitor

        run = self._wake_setup(wakelength, wake, compute_plane, add_space)
        t0 = time.time()
        monitor = use_field_monitor and field_monitor is not None
        # this package's FieldMonitor accumulates on the device inside the
        # launch sequence; any other monitor object is updated from the host
        dev_monitor = monitor and callback is None and isinstance(field_monitor, FieldMonitor)
        per_step = callback is not None or (monitor and not dev_monitor)
        if per_step:
            beam = run["beam"]
            for n in range(run["Nt"]):
                beam.update(self, n * self.dt)
                self.one_step()
                if use_field_monitor and field_monitor is not None:
                    field_monitor.update(self.E, self.dt)
                if callback is not None:
                    callback(self, n * self.dt)
        else:
            self._wake_device_beam(run)
            done = 0
            while done < run["Nt"]:
                n = min(chunk, run["Nt"] - done)
                if dev_monitor:
                    field_monitor._stage(self.E, self.dt, n)
                self._advance(n)
                done += n
            self.ctx.sync()
            self._wake_device_end(run)
        self.simulationTime = time.time() - t0
        self._wake_finish(run, compute_plane)

    # -- the parts of wakesolve (shared with the batched solve) ---------------
    def _wake_setup(self, wakelength, wake, compute_plane, add_space=None):
        """Test point, Nt, beam and E_z probe of a wake run
        (wakis/routines.py:366-432)."""
        run = self._wake_plan(wakelength, wake, compute_plane, add_space)
        self._wake_attach(run)
        return run

    def _wake_plan(self, wakelength, wake, compute_plane, add_space=None):
        """The host part of _wake_setup: test point, Nt, beam and time axis
        (no device calls)."""
        if wake is not None:
            self.wake = wake
        if self.wake is None:
            raise AttributeError("Wake solver information not passed to the solver instantiation")
        if wakelength is None:
            wakelength = self.wake.wakelength
        if add_space is not None:
            self.wake.skip_cells = add_space
        self.xtest, self.ytest = self.wake.xtest, self.wake.ytest
        self.ixt = int(np.abs(self.x - self.xtest).argmin())
        self.iyt = int(np.abs(self.y - self.ytest).argmin())
        h = 0 if compute_plane.lower() == "longitudinal" else 1
        self.wake.wakelength = wakelength
        self.ti = self.wake.ti
        self.v = self.wake.v
        z = self.z
        tmax = (wakelength + self.ti * self.v + (z.max() - z.min())) / self.v
        Nt = int(tmax / self.dt)
        self.tmax, self.Nt = tmax, Nt
        beam = Beam(q=self.wake.q, sigmaz=self.wake.sigmaz, beta=self.wake.beta,
                    xsource=self.wake.xsource, ysource=self.wake.ysource)
        t_axis = np.arange(0, Nt * self.dt, self.dt)
        return {"Nt": Nt, "beam": beam, "h": h, "t": t_axis, "t0": time.time()}

    def _wake_attach(self, run):
        """The device part of _wake_setup: the E_z probe history."""
        self.ctx.wake_probe(self.ixt, self.iyt, run["h"], max(run["Nt"], 1))
        self.ctx.set_step_index(0)

    def _wake_device_beam(self, run):
        """The Gaussian bunch evaluated on the device at every step."""
        beam = run["beam"]
        params = beam.device_params(self)
        self.ctx.set_beam_gauss(int(beam.ixs), int(beam.iys), params, self.z)
        self._beam_ij = (int(beam.ixs), int(beam.iys))
        if self.step_0:
            self.ctx.zero_ghosts()
            self.step_0 = False

    def _wake_device_end(self, run):
        """Leave J as the reference leaves it: the last beam line (held by a
        non-conductive solver) or sigma*E (a conductive step overwrote it)."""
        beam, Nt = run["beam"], run["Nt"]
        if self.use_conductivity or Nt == 0:
            self.ctx.beam_off()
            self._beam_ij = None
        else:
            self._set_beam_line(int(beam.ixs), int(beam.iys), beam.line(self, (Nt - 1) * self.dt))

    def _wake_finish(self, run, compute_plane, save_logs=True):
        """Hand the device E_z history to the WakeSolver and solve
        (wakis/routines.py:456-485)."""
        h = run["h"]
        if h == 1:
            xx = slice(self.ixt - 1, self.ixt + 2)
            yy = slice(self.iyt - 1, self.iyt + 2)
        else:
            xx, yy = self.ixt, self.iyt
        meta = {
            "x": self.x[xx], "y": self.y[yy], "z": self.z[slice(0, self.Nz)],
            "dx": self.grid.dx, "dy": self.grid.dy, "dz": self.grid.dz,
            "t": run["t"], "nsteps": run["Nt"], "half": h,
        }
        self.wake.attach_device_history(self.ctx, meta)
        self.wake.solve(compute_plane=compute_plane)
        # run metadata next to the results (wakis/routines.py:481-485)
        self.logger.wakeSolver = self.wake.logger.wakeSolver
        self.logger.wakeSolver["wakelength"] = self.wake.wakelength
        self.logger.wakeSolver["simulationTime"] = time.time() - run["t0"]
        if save_logs:
            self.logger.save_logs()
