"""``WakeSolver``: longitudinal wake potential and impedance on the device.

Parameters and defaults follow wakis/wakeSolver.py:22-209 (fmax = v/(3
sigmaz), injection time ti = 8.548921333333334*sigmaz/(sqrt(beta)*v)).
The E_z history recorded by ``SolverFIT3D.wakesolve`` stays in HBM; the
integral WP(s) = sum_k Ez(z_k, t(s, z_k)) dz_k of wakis/wakeSolver.py:423-551
runs in one FP64 kernel (one thread per s sample, the reference's per-sample
order of additions, the WP carry-over between the 9 windows of
calc_long_WP_3d kept), and the impedance of wakis/wakeSolver.py:628-720 is a
direct FP64 DFT of the kept bins on the device.  The host computes only the
time-bucket table it(n, k) = int(ts/dt) with the reference's numpy
expressions, so every bucket index is identical.
"""

from __future__ import annotations

import os

import numpy as np
from scipy.constants import c as c_light

from .logger import Logger


class WakeSolver:
    def __init__(self, wakelength=None, q=1e-9, sigmaz=1e-3, beta=1.0, xsource=0.0,
                 ysource=0.0, xtest=0.0, ytest=0.0, chargedist=None, ti=None, fmax=None,
                 compute_plane="both", skip_cells=0, add_space=None, Ez_file=None, save=True,
                 results_folder="results/", verbose=1, counter_moving=False):
        self.verbose = verbose
        self.q = q
        self.sigmaz = sigmaz
        self.beta = beta
        self.v = self.beta * c_light
        self.xsource, self.ysource = xsource, ysource
        self.xtest, self.ytest = xtest, ytest
        self.chargedist = chargedist
        self.fmax = self.v / (3.0 * self.sigmaz) if fmax is None else fmax
        self.skip_cells = skip_cells if add_space is None else add_space
        self.compute_plane = compute_plane
        self.counter_moving = counter_moving
        if ti is not None:
            self.ti = ti
        else:
            self.ti = 8.548921333333334 * self.sigmaz / (np.sqrt(self.beta) * self.v)
        self.Ez_file = Ez_file
        self.wakelength = wakelength
        self.s = None
        self.lambdas = None
        self.WP = None
        self.WP_3d = None
        self.n_transverse_cells = 1
        self.WPx = self.WPy = None
        self.f = None
        self.Z = None
        self.Zx = self.Zy = None
        self.lambdaf = None
        self.save = save
        self.folder = results_folder if results_folder.endswith("/") else results_folder + "/"
        if self.Ez_file is None:
            self.Ez_file = self.folder + "Ez.h5"
        if self.save:
            os.makedirs(self.folder, exist_ok=True)
        self._ctx = None
        self._meta = None
        self.logger = Logger()
        self.assign_logs()

    def assign_logs(self):
        """The wake parameters in ``self.logger.wakeSolver``
        (wakis/wakeSolver.py:1713-1727)."""
        for key in ("ti", "q", "sigmaz", "beta", "xsource", "ysource", "xtest", "ytest",
                    "chargedist", "skip_cells"):
            self.logger.wakeSolver[key] = getattr(self, key)
        self.logger.wakeSolver["results_folder"] = self.folder

    def attach_device_history(self, ctx, meta):
        """Use the E_z history of a device context (written by wakesolve)."""
        self._ctx = ctx
        self._meta = meta
        self.xf, self.yf, self.zf = meta["x"], meta["y"], meta["z"]
        self.dx, self.dy, self.dz = meta["dx"], meta["dy"], meta["dz"]
        self.t = meta["t"]

    # -- recorded histories (wakis/wakeSolver.py:1317-1364) --------------------
    Ez_hf = None

    def read_Ez(self, filename=None, return_value=False):
        """Read a recorded E_z history (wakis/wakeSolver.py:1317-1364): the
        reference's ``Ez.h5`` layout, one dataset per step keyed ``'#%05d'``
        (a (2h+1, 2h+1, nz) window or an nz line) plus ``x, y, z, dx, dy,
        dz, t``.  An ``.npz`` archive with those keys is read with numpy
        (no pickles); an HDF5 file needs h5py.  Sets ``Ez_hf`` and the axis
        attributes; the history is moved to the device when a wake is
        computed."""
        if filename is None:
            filename = self.Ez_file
        if str(filename).endswith(".npz"):
            with np.load(filename, allow_pickle=False) as z:
                hf = {k: np.array(z[k]) for k in z.files}
        else:
            try:
                import h5py
            except ImportError as e:
                raise ImportError(
                    f"reading {filename} needs h5py; an .npz archive with the same keys "
                    "is read without it") from e
            with h5py.File(filename, "r") as f:
                hf = {k: np.array(f[k]) for k in f.keys()}
        self.Ez_hf = hf
        self.Ez_file = filename
        self._axes_from(hf)
        if return_value:
            return hf

    def _axes_from(self, hf):
        for key, attr in (("x", "xf"), ("y", "yf"), ("z", "zf"), ("dx", "dx"), ("dy", "dy"),
                          ("dz", "dz"), ("t", "t")):
            if key in hf.keys():
                setattr(self, attr, np.array(hf[key]))

    def _load_host_history(self, device=0):
        """Move the datasets of ``Ez_hf`` to a device history: the first
        min(len(t), number of datasets) steps (the reference's nt), the
        centred 3x3 window of a 3-D dataset (rows shape//2 - 1 .. shape//2 + 1,
        as wakis/wakeSolver.py:492-498 indexes it) or the line of a 1-D one."""
        from ._lib import Context

        hf = self.Ez_hf
        self._axes_from(hf)
        keys = sorted(k for k in hf.keys() if isinstance(k, str) and k.startswith("#"))
        if not keys:
            raise KeyError("no '#%05d' datasets in the E_z history")
        nt = min(len(self.t), len(keys))
        first = np.asarray(hf[keys[0]])
        if first.ndim == 3:
            h = 1
            a0, b0 = first.shape[0] // 2, first.shape[1] // 2
            if a0 < 1 or b0 < 1 or a0 + 1 >= first.shape[0] or b0 + 1 >= first.shape[1]:
                raise ValueError(f"dataset shape {first.shape}: needs a window of 3 or more")
        elif first.ndim == 1:
            h = 0
        else:
            raise ValueError(f"dataset shape {first.shape}: expected (a, b, nz) or (nz,)")
        nz = first.shape[-1]
        dtype = np.float32 if first.dtype == np.float32 else np.float64
        nw = (2 * h + 1) ** 2
        hist = np.empty((nw, nz, nt), dtype=dtype)
        for n in range(nt):
            e = np.asarray(hf[keys[n]])
            if h == 1:
                hist[:, :, n] = e[a0 - 1:a0 + 2, b0 - 1:b0 + 2, :].reshape(nw, nz)
            else:
                hist[0, :, n] = e
        side = 2 * h + 1 if h else 2
        ctx = Context(side, side, nz, dtype=dtype, device=device)
        ctx.wake_history_load(h, h, h, hist)
        self._ctx = ctx
        self._meta = {"half": h, "nsteps": nt, "x": self.xf, "y": self.yf, "z": self.zf,
                      "dx": self.dx, "dy": self.dy, "dz": self.dz, "t": self.t, "src": hf}

    def _need(self):
        stale = (self.Ez_hf is not None and self._meta is not None
                 and self._meta.get("src") is not self.Ez_hf)
        if self._ctx is None or self._meta is None or stale:
            if self.Ez_hf is None:
                if self.Ez_file is None or not os.path.exists(self.Ez_file):
                    raise RuntimeError(
                        "no E_z history: run SolverFIT3D.wakesolve first, or give a recorded "
                        "history (Ez_hf, or Ez_file / read_Ez with an .npz or .h5 file)")
                self.read_Ez()
            self._load_host_history()

    def solve(self, compute_plane=None, **kwargs):
        """wakis/wakeSolver.py:211-255."""
        if compute_plane is None:
            compute_plane = self.compute_plane
        for key, val in kwargs.items():
            setattr(self, key, val)
        if compute_plane.lower() in ["both", "transverse"]:
            self.calc_long_WP_3d()
            self.calc_trans_WP()
            self.calc_long_Z()
            self.calc_trans_Z()
        elif compute_plane.lower() == "longitudinal":
            self.calc_long_WP()
            self.calc_long_Z()

    # -- s axis and time buckets (wakis/wakeSolver.py:446-488, 517-529) ------
    def _axes(self):
        nt = len(self.t)
        dt = self.t[2] - self.t[1]
        zmax = np.max(self.zf)
        zmin = np.min(self.zf)
        if self.skip_cells != 0:
            zz = slice(self.skip_cells, -self.skip_cells)
        else:
            zz = np.s_[:]
        z = self.zf[zz]
        if self.wakelength is not None:
            wakelength = self.wakelength
        else:
            wakelength = nt * dt * self.v - (zmax - zmin) - self.ti * self.v
            self.wakelength = wakelength
        s = np.arange(-self.ti * self.v, wakelength, dt * self.v)
        return nt, dt, zmin, z, s

    def bucket_columns(self, ncols_alloc, counter=False):
        """cols[k, n]: the E_z time column the reference reads for (s_n, z_k),
        or -1.  it = int(ts/dt); it >= nt contributes nothing; a negative it
        indexes from the end of the reference's Ezt array (ncols_alloc
        columns, only the first nt of which are filled).  counter=True: the
        counter-moving time of wakis/wakeSolver.py:500-506, where term k
        samples z[-k-1] with the velocity negated and times counted from
        zmax."""
        nt_t, dt, zmin, z, s = self._axes()
        zmax = np.max(self.zf)
        nt = min(nt_t, int(self._meta["nsteps"]))
        cols = np.full((len(z), len(s)), -1, dtype=np.int32)
        for k in range(len(z)):
            if counter:
                ts = ((z[-k - 1] - s) / (-1 * self.v) - zmax / (-1 * self.v)
                      - self.t[0] + self.ti)
            else:
                ts = (z[k] + s) / self.v - zmin / self.v - self.t[0] + self.ti
            it = np.trunc(ts / dt).astype(np.int64)
            c = np.where(it < 0, it + ncols_alloc, it)
            ok = (it < nt) & (c >= 0) & (c < nt)
            cols[k, ok] = c[ok]
        return cols, s, z

    # -- wake potential --------------------------------------------------------
    def calc_long_WP_3d(self, **kwargs):
        """wakis/wakeSolver.py:423-551 (9 windows, WP carried over)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        self._need()
        if self._meta["half"] != 1:
            raise ValueError("calc_long_WP_3d needs the 3x3 probe window")
        counter = bool(self.counter_moving)
        cols, s, z = self.bucket_columns(len(self.t), counter=counter)
        dz = np.asarray(self.dz, dtype=np.float64)[:len(z)]
        if counter:
            # wakis/wakeSolver.py:500-513: Ezt[-k-1, it] * (-1 * dz[k])
            wp = self._ctx.wake_potential(np.arange(9), True, cols, self.skip_cells, -1 * dz,
                                          self.q * 1e12, zrev=True)
        else:
            wp = self._ctx.wake_potential(np.arange(9), True, cols, self.skip_cells, dz,
                                          self.q * 1e12)
        i0 = j0 = self.n_transverse_cells
        self.WP_3d = wp.reshape(2 * i0 + 1, 2 * j0 + 1, len(s))
        self.s = s
        self.WP = self.WP_3d[i0, j0, :]
        self._save_txt("WP.txt", self.s, self.WP, "s [m]", "WP [V/pC]")

    def calc_long_WP(self, Ezt=None, **kwargs):
        """wakis/wakeSolver.py:311-421 (one window; for a 3x3 history the
        reference reads window index (2, 2)).  ``Ezt`` is stored and, as in
        the reference, the sum is taken over the recorded history; this
        routine has no counter-moving branch."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        if Ezt is not None:
            self.Ezt = Ezt
        self._need()
        if self._meta["half"] == 0 and self.skip_cells != 0:
            # the reference copies the whole E_z line into an array of the
            # skip_cells-sliced length and fails (wakis/wakeSolver.py:370-389)
            raise ValueError(
                f"could not broadcast input array from shape ({len(self.zf)},) into shape "
                f"({len(self.zf) - 2 * self.skip_cells},)")
        nt = min(len(self.t), int(self._meta["nsteps"]))
        cols, s, z = self.bucket_columns(nt)
        wid = 8 if self._meta["half"] == 1 else 0
        wp = self._ctx.wake_potential([wid], False, cols, self.skip_cells, self.dz,
                                      self.q * 1e12)
        self.s = s
        self.WP = wp[0]
        self._save_txt("WP.txt", self.s, self.WP, "s [m]", "WP [V/pC]")

    def calc_trans_WP(self, **kwargs):
        """Panofsky-Wenzel transverse wake of wakis/wakeSolver.py:553-626:
        int_WP[..., n] = sum(WP_3d[..., 0:n]) * ds, then a centred difference
        over the window.  The prefix sums use a running sum (the reference uses
        numpy's pairwise sum per n; both agree to rounding)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        ds = self.s[2] - self.s[1]
        i0 = j0 = self.n_transverse_cells
        cs = np.zeros_like(self.WP_3d)
        cs[..., 1:] = np.cumsum(self.WP_3d[..., :-1], axis=-1)
        int_WP = cs * ds
        dxf = np.diff(self.xf)
        dyf = np.diff(self.yf)
        self.WPx = -(int_WP[i0 + 1, j0] - int_WP[i0 - 1, j0]) / (dxf[i0 - 1] + dxf[i0])
        self.WPy = -(int_WP[i0, j0 + 1] - int_WP[i0, j0 - 1]) / (dyf[j0 - 1] + dyf[j0])
        self._save_txt("WPx.txt", self.s, self.WPx, "s [m]", "WP [V/pC]")
        self._save_txt("WPy.txt", self.s, self.WPy, "s [m]", "WP [V/pC]")

    # -- impedance -------------------------------------------------------------
    def calc_lambdas_analytic(self, **kwargs):
        """wakis/wakeSolver.py:842-874."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        self.lambdas = (1 / (self.sigmaz * np.sqrt(2 * np.pi))
                        * np.exp(-(self.s**2) / (2 * self.sigmaz**2)))
        self._save_txt("lambda.txt", self.s, self.lambdas, "s [Hz]", "Charge distribution [C/m]")

    def calc_lambdas(self, **kwargs):
        """wakis/wakeSolver.py:787-840 (user charge distribution)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        if isinstance(self.chargedist, dict):
            keys = list(self.chargedist.keys())
            z, chargedist = self.chargedist[keys[0]], self.chargedist[keys[1]]
        elif isinstance(self.chargedist, str):
            d = np.loadtxt(self.chargedist)
            z, chargedist = d[:, 0], d[:, 1]
        else:
            chargedist = self.chargedist
            if self.z is not None and len(self.z) == len(chargedist):
                z = self.z
            elif len(self.zf) == len(chargedist):
                z = self.zf
            else:
                raise ValueError("Dimension error: check input dimensions")
        self.lambdas = np.interp(self.s, z, chargedist / self.q)

    z = None

    def _dft_setup(self, samples, fmax):
        ds = np.mean(self.s[1:] - self.s[:-1])
        fmax = self.fmax if fmax is None else fmax
        N = int((self.v / ds) // fmax * samples)
        ffft = np.fft.fftfreq(N, ds / self.v)
        mask = np.logical_and(ffft >= 0, ffft < fmax)
        K = int(mask.sum())
        if K and not mask[:K].all():
            raise RuntimeError("kept DFT bins are not the leading bins")
        return ds, N, K, ffft[mask]

    def _ctx_for_dft(self):
        """The device context the DFT runs on: the one holding the history,
        else a small context of its own (results read from files)."""
        if self._ctx is None:
            from ._lib import Context

            self._ctx = Context(2, 2, 2, dtype=np.float64)
        return self._ctx

    def calc_long_Z(self, samples=1001, fmax=None, **kwargs):
        """wakis/wakeSolver.py:628-720: Z = -(DFT_N(WP*1e12)*ds) /
        (DFT_N(lambda*v)*ds) on the bins 0 <= f < fmax."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        if self.lambdas is None and self.chargedist is not None:
            self.calc_lambdas()
        elif self.lambdas is None:
            self.calc_lambdas_analytic()
        ds, N, K, f = self._dft_setup(samples, fmax)
        Z, X, L = self._ctx_for_dft().impedance(self.WP * 1e12, self.lambdas * self.v, N, K, ds)
        self.f = f
        self.Z = Z
        self.lambdaf = L
        self.N_dft = N
        self._save_txt("Z.txt", self.f, self.Z, "f [Hz]", "Z [Ohm]")
        self._save_txt("spectrum.txt", self.f, self.lambdaf, "f [Hz]",
                       "Charge distribution spectrum [C/s]")

    def calc_trans_Z(self, samples=1001, fmax=None):
        """wakis/wakeSolver.py:722-785: Z_x,y = 1j*DFT(WP_x,y*1e12)*ds/lambdaf."""
        ds, N, K, f = self._dft_setup(samples, fmax)
        ctx = self._ctx_for_dft()
        _, Xx, L = ctx.impedance(self.WPx * 1e12, self.lambdas * self.v, N, K, ds)
        _, Xy, _ = ctx.impedance(self.WPy * 1e12, self.lambdas * self.v, N, K, ds)
        self.Zx = 1j * Xx / L
        self.Zy = 1j * Xy / L
        self.fx = self.fy = f
        self._save_txt("Zx.txt", self.fx, self.Zx, "f [Hz]", "Zx [Ohm]")
        self._save_txt("Zy.txt", self.fy, self.Zy, "f [Hz]", "Zy [Ohm]")

    def _save_txt(self, name, a, b, ha, hb):
        if self.save:
            np.savetxt(self.folder + name, np.c_[a, b],
                       header="   " + ha + " " * 20 + hb + "\n" + "-" * 48)

    # -- result files (wakis/wakeSolver.py:1366-1511) --------------------------
    def save_txt(self, f_name, x_data=None, y_data=None, x_name="X [-]", y_name="Y [-]"):
        """Two-column text file with the reference's header
        (wakis/wakeSolver.py:1422-1475); nothing is written without data."""
        if f_name.endswith(".txt"):
            f_name = f_name[:-4]
        if x_data is not None and y_data is not None:
            np.savetxt(f_name + ".txt", np.c_[x_data, y_data],
                       header="   " + x_name + " " * 20 + y_name + "\n" + "-" * 48)

    # -- stand-alone transforms (wakis/wakeSolver.py:1185-1315) -------------
    @staticmethod
    def calc_impedance_from_wake(wake, s=None, t=None, fmax=None, samples=None, verbose=True):
        """Impedance of a sampled wake function by FFT: ``[f, Z]`` for
        0 <= f < fmax (default 1/dt), Z = FFT(wake) * 2 / len(wake).
        ``wake`` may be ``[t, wake]``; ``s`` [m] gives t = s/c.  With
        ``fmax`` the wake is first resampled (linear interpolation) at
        1/(2 fmax); ``samples`` zero-pads the FFT to 2*samples points."""
        if type(wake) is list:
            t, wake = wake[0], wake[1]
        if s is not None:
            t = s / c_light
        elif t is None:
            raise AttributeError('Provide time data through parameter "t" [s] or "s" [m]')
        t = np.asarray(t)
        wake = np.asarray(wake)
        dt = np.mean(np.diff(t))
        if fmax is not None:
            ts = np.arange(t.min(), t.max(), 1 / fmax / 2)
            wake = np.interp(ts, t, wake)
            dt = np.mean(np.diff(ts))
        else:
            fmax = 1 / dt
        W = np.fft.fft(wake, n=None if samples is None else 2 * samples)
        freqs = np.fft.fftfreq(len(W), dt)
        keep = (freqs >= 0) & (freqs < fmax)
        f = freqs[keep]
        Z = W[keep] / len(wake) * 2
        if verbose:
            print(f"    * Number of samples = {len(f)}")
            print(f"    * Maximum frequency = {f.max()} Hz")
            print(f"    * Maximum resolution = {np.mean(np.diff(f))} Hz")
        return [f, Z]

    @staticmethod
    def calc_wake_from_impedance(impedance, f=None, tmax=None, samples=None, pad=0,
                                 verbose=True):
        """Wake function of a sampled impedance: ``[t, wake]`` with
        wake = Re(-FFT(Z, n=samples)) rotated left by one sample and t
        spanning [0, tmax] (default 1/df).  ``impedance`` may be ``[f, Z]``
        (any length-2 sequence is taken as that pair); with ``tmax`` Z is
        first resampled at 1/tmax on f's range."""
        if len(impedance) == 2:
            f, Z = impedance[0], impedance[1]
        elif f is None:
            raise AttributeError('Provide frequency data through parameter "f"')
        else:
            Z = impedance
        f = np.asarray(f)
        Z = np.asarray(Z)
        df = np.mean(np.diff(f))
        if tmax is not None:
            Z = np.interp(np.arange(f.min(), f.max(), 1 / tmax), f, Z)
        else:
            tmax = 1 / df
        wake = np.roll(np.real(-1 * np.fft.fft(Z, n=samples)), -1)
        t = np.linspace(0, tmax, len(wake))
        if verbose:
            print(f"    * Number of samples = {len(t)}")
            print(f"    * Maximum time = {t.max()} s")
            print(f"    * Maximum resolution = {np.mean(np.diff(t))} s")
        return [t, wake]

    def read_txt(self, txt, skiprows=2, delimiter=None, usecols=None):
        """Columns of a text file keyed by the header names "name[unit]"
        (integer keys when the header does not parse), complex when a float
        read fails (wakis/wakeSolver.py:1366-1420)."""
        try:
            load = np.loadtxt(txt, skiprows=skiprows, delimiter=delimiter, usecols=usecols)
        except Exception:
            load = np.loadtxt(txt, skiprows=skiprows, delimiter=delimiter, usecols=usecols,
                              dtype=np.complex128)
        try:
            with open(txt) as f:
                header = f.readline()
            header = header.replace(" ", "").replace("#", "").replace("\n", "").split("]")
            d = {}
            for i in range(len(load[0, :])):
                d[header[i] + "]"] = load[:, i]
        except Exception:
            d = {}
            for i in range(len(load[0, :])):
                d[i] = load[:, i]
        return d

    def load_results(self, folder=None):
        """Read the result files of a saved run back into the attributes
        (wakis/wakeSolver.py:1477-1511)."""
        if folder is None and self.folder is None:
            raise Exception("[!] Please provide a folder path")
        elif folder is None:
            folder = self.folder
        if not folder.endswith("/"):
            folder = folder + "/"
        _, self.lambdas = self.read_txt(folder + "lambda.txt").values()
        _, self.WPx = self.read_txt(folder + "WPx.txt").values()
        _, self.WPy = self.read_txt(folder + "WPy.txt").values()
        self.s, self.WP = self.read_txt(folder + "WP.txt").values()
        _, self.lambdaf = self.read_txt(folder + "spectrum.txt").values()
        _, self.Zx = self.read_txt(folder + "Zx.txt").values()
        _, self.Zy = self.read_txt(folder + "Zy.txt").values()
        self.f, self.Z = self.read_txt(folder + "Z.txt").values()
        self.f = np.abs(self.f)
        self.wakelength = self.s[-1]
        self.folder = folder
        self.Ez_file = folder + "Ez.h5"
