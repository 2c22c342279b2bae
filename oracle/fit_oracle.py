"""CPU oracle for the Wakis FIT time-stepping path (TEST INFRASTRUCTURE ONLY).

This module restates, in numpy + scipy.sparse, the arithmetic of the
reference's CPU path so that the HIP kernels of ``wakis_amd`` can be checked
against it.  It is the checker, never the thing measured or shipped: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product package never imports it.

What is restated (citations are reference files, path:line):

* ``OField``            wakis/field.py:19-191, 223-348, 350-460 (flat 3N
                        component-major array, Fortran order per component,
                        tuple indexing that returns *views* on CPU).
* ``OGrid``             wakis/gridFIT3D.py:143-185 (axes, dx/dy/dz) and
                        :263-305 ``_compute_grid`` (L, iA, tL, itA).
* ``OSolver.__init__``  wakis/solverFIT3D.py:40-376 (shift matrices, curl C,
                        BCs, materials, PML, dt, operator products) with
                        ``_apply_bc_to_C`` wakis/boundaries.py:14-188 and
                        ``_initialize_PML`` wakis/boundaries.py:190-305.
                        STL solids are passed as ready (Nx,Ny,Nz) boolean
                        masks instead of pyvista data (wakis/solverFIT3D.py:
                        438-502 restated on those masks).
* ``OSolver.one_step``  wakis/solverFIT3D.py:504-525 (same expression order:
                        ``H - (dt*M_H) @ E``; ``E + dt*(M_E @ H - ieps*J)``;
                        ``J = sigma*E``), ``_set_ghosts_to_0`` :842-862.
* ``OSolver.update_tensors`` wakis/solverFIT3D.py:378-436.
* ``get_abc/update_abc`` wakis/boundaries.py:307-404 with the CPU view
                        aliasing of wakis/field.py:250-256.
* ``emsolve/wakesolve`` wakis/routines.py:17-229, 231-485, with the HDF5
                        file replaced by an in-memory dict using the same
                        ``'#%05d'`` keys and ``x,y,z,dx,dy,dz,t`` entries.
* ``OBeam``             wakis/sources.py:22-110.
* ``OFieldMonitor``      wakis/field_monitors.py:14-131 (pinned by the
                        reference's own closed-form checks of
                        tests/test_012_field_monitors.py: zero frequency
                        sums linearly, f*dt = 1/2 sums to zero).
* ``OPlaneWave``        wakis/sources.py:142-248.
* ``OWake``             wakis/wakeSolver.py:133-209 (fmax, ti), 311-421
                        ``calc_long_WP``, 423-551 ``calc_long_WP_3d`` (the
                        WP carry-over between the 9 windows is kept;
                        its counter-moving branch 500-513),
                        553-626 ``calc_trans_WP``, 628-720 ``calc_long_Z``,
                        722-785 ``calc_trans_Z``, 842-874
                        ``calc_lambdas_analytic``.

Third-party arithmetic: scipy.sparse csc/csr mat-vec (scipy 1.15 here) and
numpy.fft (pocketfft); the reference pins neither version (setup.py:48-54).

Pinning (tests/test_oracle_*.py): reference golden vectors of
tests/test_001_pec_cubic_cavity.py:18-67 (WP, Z, cumsums), TE101 frequency of
tests/test_002_analytic_resonator.py:144, PML reflection bound of
tests/test_005_pml_boundaries.py:99, discrete div(B) conservation (S.C = 0),
brute-force loop vs vectorised WP, closed-form WP of a travelling Ez, the
counter-moving WP by the z-mirror identity (on a z axis symmetric about 0,
the counter-moving WP of the z-reversed history equals minus the co-moving
WP of the history, bit for bit, for any dz array), and
calc_long_Z and calc_trans_Z against an independent direct DFT,
calc_trans_WP against the Panofsky-Wenzel closed form of a field linear in
x and y, and ``emsolve``'s ABC view aliasing against input/output vectors
of the reference's own Field and BCsMixin (tests/golden/abc_aliasing.npz).
``_apply_bc_to_C`` (periodic/ABC metric fix-ups, PEC column and PMC row
masks with carried flags), ``_initialize_PML`` (grading and corner overwrite
order) and ``OBeam`` (beta 0.8/0.9/1.0, default and given ti) against vectors
of the reference's own boundaries.py/field.py/sources.py
(tests/golden/bc_pml_beam.npz, tests/test_oracle_pins.py); the SIBC
eps := sigma*eps_0 rule and the relaxation clamp of dt against closed forms
(tau = eps_0 for any SIBC solid and for the outer PML cell; eps_r eps_0 /
sigma without SIBC; the empty conductive set raising).
"""

from __future__ import annotations

import copy

import numpy as np
from scipy.constants import c as c_light
from scipy.constants import epsilon_0 as eps_0
from scipy.constants import mu_0
from scipy.sparse import csc_matrix, diags, hstack, vstack

# wakis/materials.py:27-44
MATERIAL_LIB = {
    "pec": [np.inf, 1.0],
    "vacuum": [1.0, 1.0],
    "dielectric": [10.0, 1.0],
    "lossy metal": [10, 1.0, 10],
    "copper": [5.8e07, 1.0, 5.8e07],
    "berillium": [2.5e07, 1.0, 2.5e07],
    "stainless steel": [1.4e06, 1.0, 1.4e06],
}


# --------------------------------------------------------------------------
# Field  (wakis/field.py)
# --------------------------------------------------------------------------
class OField:
    """Flat 3N array: x block, y block, z block; F order per block.

    wakis/field.py:19-91 (layout), 128-190 (to/from matrix),
    223-348 (indexing: CPU tuple reads return views), 350-460 (arithmetic).
    """

    def __init__(self, Nx, Ny, Nz, dtype=float, use_ones=False):
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.N = Nx * Ny * Nz
        self.dtype = dtype
        if use_ones:
            self.array = np.ones(self.N * 3, dtype=dtype, order="F")
        else:
            self.array = np.zeros(self.N * 3, dtype=dtype, order="F")

    # component blocks (field.py:78-126)
    @property
    def field_x(self):
        return self.array[0 : self.N]

    @property
    def field_y(self):
        return self.array[self.N : 2 * self.N]

    @property
    def field_z(self):
        return self.array[2 * self.N : 3 * self.N]

    @field_x.setter
    def field_x(self, value):
        if len(value.shape) > 1:
            self.from_matrix(value, "x")
        else:
            self.array[0 : self.N] = value

    @field_y.setter
    def field_y(self, value):
        if len(value.shape) > 1:
            self.from_matrix(value, "y")
        else:
            self.array[self.N : 2 * self.N] = value

    @field_z.setter
    def field_z(self, value):
        if len(value.shape) > 1:
            self.from_matrix(value, "z")
        else:
            self.array[2 * self.N : 3 * self.N] = value

    def toarray(self):
        return self.array

    def fromarray(self, array):
        self.array[:] = array

    @staticmethod
    def _comp(key):
        if key == 0 or key == "x":
            return 0
        if key == 1 or key == "y":
            return 1
        if key == 2 or key == "z":
            return 2
        raise IndexError("Component id not valid")

    def to_matrix(self, key):
        c = self._comp(key)
        return np.reshape(
            self.array[c * self.N : (c + 1) * self.N],
            (self.Nx, self.Ny, self.Nz),
            order="F",
        )

    def from_matrix(self, mat, key):
        c = self._comp(key)
        self.array[c * self.N : (c + 1) * self.N] = np.reshape(mat, self.N, order="F")

    def __getitem__(self, key):
        # field.py:239-256 (CPU branch: a view of the live array)
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            if type(key[3]) is str and key[3].lower() == "abs":
                f = np.sqrt(
                    self.to_matrix("x") ** 2
                    + self.to_matrix("y") ** 2
                    + self.to_matrix("z") ** 2
                )
                return f[key[0], key[1], key[2]]
            return self.to_matrix(key[3])[key[0], key[1], key[2]]
        if type(key) is int or type(key) is slice:
            return self.array[key]
        raise ValueError("key must be a 3-tuple or an integer")

    def __setitem__(self, key, value):
        # field.py:313-348
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            field = self.to_matrix(key[3])
            field[key[0], key[1], key[2]] = value
            self.from_matrix(field, key[3])
        else:
            self.array[key] = value

    def __mul__(self, other, dtype=None):
        # field.py:350-385
        if dtype is None:
            dtype = self.dtype
        out = OField(self.Nx, self.Ny, self.Nz, dtype=dtype)
        if type(other) is float or type(other) is int:
            out.array = self.array * other
        elif len(other.shape) > 1:
            for d in ["x", "y", "z"]:
                out.from_matrix(self.to_matrix(d) * other, d)
        else:
            out.array = self.array * other
        return out

    def __add__(self, other, dtype=None):
        # field.py:423-458
        if dtype is None:
            dtype = self.dtype
        out = OField(self.Nx, self.Ny, self.Nz, dtype=dtype)
        if type(other) is OField:
            out.field_x = self.field_x + other.field_x
            out.field_y = self.field_y + other.field_y
            out.field_z = self.field_z + other.field_z
        elif len(other.shape) > 1:
            for d in ["x", "y", "z"]:
                out.from_matrix(self.to_matrix(d) + other, d)
        else:
            out.array = self.array + other
        return out

    def copy(self):
        return copy.deepcopy(self)


# --------------------------------------------------------------------------
# Grid  (wakis/gridFIT3D.py)
# --------------------------------------------------------------------------
class OGrid:
    """Axes and FIT metric Fields.

    gridFIT3D.py:143-185 (axes from arrays or linspace extents) and
    gridFIT3D.py:263-305 ``_compute_grid``.
    """

    def __init__(self, xmin=None, xmax=None, ymin=None, ymax=None, zmin=None,
                 zmax=None, Nx=None, Ny=None, Nz=None, x=None, y=None, z=None):
        if x is not None and y is not None and z is not None:
            self.x, self.y, self.z = x, y, z
            self.Nx, self.Ny, self.Nz = len(x) - 1, len(y) - 1, len(z) - 1
        else:
            self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
            self.x = np.linspace(xmin, xmax, Nx + 1)
            self.y = np.linspace(ymin, ymax, Ny + 1)
            self.z = np.linspace(zmin, zmax, Nz + 1)
        self.dx = np.diff(self.x)
        self.dy = np.diff(self.y)
        self.dz = np.diff(self.z)
        self._compute_grid()

    def _compute_grid(self):
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        X, Y, Z = np.meshgrid(self.x, self.y, self.z, indexing="ij")
        self.L = OField(Nx, Ny, Nz)
        self.L.field_x = X[1:, 1:, 1:] - X[:-1, :-1, :-1]
        self.L.field_y = Y[1:, 1:, 1:] - Y[:-1, :-1, :-1]
        self.L.field_z = Z[1:, 1:, 1:] - Z[:-1, :-1, :-1]

        self.iA = OField(Nx, Ny, Nz)
        self.iA.field_x = np.divide(1.0, self.L.field_y * self.L.field_z)
        self.iA.field_y = np.divide(1.0, self.L.field_x * self.L.field_z)
        self.iA.field_z = np.divide(1.0, self.L.field_x * self.L.field_y)

        tx = (self.x[1:] + self.x[:-1]) / 2
        ty = (self.y[1:] + self.y[:-1]) / 2
        tz = (self.z[1:] + self.z[:-1]) / 2
        tx = np.append(tx, tx[-1])
        ty = np.append(ty, ty[-1])
        tz = np.append(tz, tz[-1])
        tX, tY, tZ = np.meshgrid(tx, ty, tz, indexing="ij")

        self.tL = OField(Nx, Ny, Nz)
        self.tL.field_x = tX[1:, 1:, 1:] - tX[:-1, :-1, :-1]
        self.tL.field_y = tY[1:, 1:, 1:] - tY[:-1, :-1, :-1]
        self.tL.field_z = tZ[1:, 1:, 1:] - tZ[:-1, :-1, :-1]

        self.itA = OField(Nx, Ny, Nz)
        aux = self.tL.field_y * self.tL.field_z
        self.itA.field_x = np.divide(1.0, aux, out=np.zeros_like(aux), where=aux != 0)
        aux = self.tL.field_x * self.tL.field_z
        self.itA.field_y = np.divide(1.0, aux, out=np.zeros_like(aux), where=aux != 0)
        aux = self.tL.field_x * self.tL.field_y
        self.itA.field_z = np.divide(1.0, aux, out=np.zeros_like(aux), where=aux != 0)


# --------------------------------------------------------------------------
# Solver  (wakis/solverFIT3D.py, wakis/boundaries.py, wakis/routines.py)
# --------------------------------------------------------------------------
_PEC_LIKE = ("electric", "pec", "pml")


class OSolver:
    """Restatement of ``SolverFIT3D`` (CPU, no MKL, no MPI).

    ``solids`` replaces the pyvista STL masks: a list of
    ``(mask[Nx,Ny,Nz] bool, material)`` applied in order, where material is
    a library key or ``[eps_r, mu_r(, sigma)]`` (solverFIT3D.py:438-502).
    """

    def __init__(self, grid, wake=None, cfln=0.5, dt=None,
                 bc_low=("Periodic",) * 3, bc_high=("Periodic",) * 3,
                 solids=None, use_sibc=True, fmax=1e9, dtype=np.float64,
                 n_pml=10, bg=(1.0, 1.0)):
        bc_low, bc_high = list(bc_low), list(bc_high)
        if isinstance(bg, tuple):
            bg = list(bg)
        self.step_0 = True
        self.use_sibc = use_sibc
        self.fmax = fmax
        self.activate_abc = False
        self.activate_pml = False
        self.use_conductivity = False
        self.grid = grid
        self.Nx, self.Ny, self.Nz = grid.Nx, grid.Ny, grid.Nz
        self.N = self.Nx * self.Ny * self.Nz
        self.dx, self.dy, self.dz = grid.dx, grid.dy, grid.dz
        # cell centres (solverFIT3D.py:161-163)
        self.x = grid.x[:-1] + self.dx / 2
        self.y = grid.y[:-1] + self.dy / 2
        self.z = grid.z[:-1] + self.dz / 2
        # metric Fields are copied: _apply_bc_to_C edits tL/itA in place
        self.L = grid.L.copy()
        self.iA = grid.iA.copy()
        self.tL = grid.tL.copy()
        self.itA = grid.itA.copy()
        self.wake = wake
        if wake is not None and fmax == 1e9:
            self.fmax = wake.fmax
        self.dtype = dtype
        Nx, Ny, Nz, N = self.Nx, self.Ny, self.Nz, self.N
        self.E = OField(Nx, Ny, Nz, dtype=dtype)
        self.H = OField(Nx, Ny, Nz, dtype=dtype)
        self.J = OField(Nx, Ny, Nz, dtype=dtype)

        # shift and curl matrices (solverFIT3D.py:207-227)
        self.Px = diags([-1, 1], [0, 1], shape=(N, N), dtype=np.int8)
        self.Py = diags([-1, 1], [0, Nx], shape=(N, N), dtype=np.int8)
        self.Pz = diags([-1, 1], [0, Nx * Ny], shape=(N, N), dtype=np.int8)
        self.Ds = diags(self.L.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        self.iDa = diags(self.iA.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        self.tDs = diags(self.tL.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        self.itDa = diags(self.itA.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        Z0 = csc_matrix((N, N))
        self.C = vstack(
            [
                hstack([Z0, -self.Pz, self.Py]),
                hstack([self.Pz, Z0, -self.Px]),
                hstack([-self.Py, self.Px, Z0]),
            ],
            dtype=np.int8,
        )
        self.bc_low, self.bc_high = bc_low, bc_high
        self._apply_bc_to_C()

        # materials (solverFIT3D.py:240-279)
        if type(bg) is str:
            bg = MATERIAL_LIB[bg.lower()]
        if len(bg) == 3:
            self.eps_bg, self.mu_bg, self.sigma_bg = bg[0] * eps_0, bg[1] * mu_0, bg[2]
            self.use_conductivity = True
        else:
            self.eps_bg, self.mu_bg, self.sigma_bg = bg[0] * eps_0, bg[1] * mu_0, 0.0
        self.ieps = OField(Nx, Ny, Nz, use_ones=True, dtype=dtype) * (1.0 / self.eps_bg)
        self.imu = OField(Nx, Ny, Nz, use_ones=True, dtype=dtype) * (1.0 / self.mu_bg)
        self.sigma = OField(Nx, Ny, Nz, use_ones=True, dtype=dtype) * self.sigma_bg
        if solids:
            self._apply_solid_materials(solids)

        if self.activate_pml:
            self.n_pml = n_pml
            self._initialize_PML()

        # dt (solverFIT3D.py:291-318)
        self.cfln = cfln
        if dt is None:
            self.dt = cfln / (
                c_light
                * np.sqrt(
                    1 / np.min(grid.dx) ** 2
                    + 1 / np.min(grid.dy) ** 2
                    + 1 / np.min(grid.dz) ** 2
                )
            )
        else:
            self.dt = dt
        self.dt = self.dtype(self.dt)
        if self.use_conductivity:
            mask = np.logical_and(self.sigma.toarray() != 0, self.ieps.toarray() != 0)
            self.tau = (1 / self.ieps.toarray()[mask]) / self.sigma.toarray()[mask]
            if self.dt > self.tau.min():
                self.dt = self.tau.min()

        # operator products (solverFIT3D.py:337-344)
        self.iDeps = diags(self.ieps.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        self.iDmu = diags(self.imu.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        self.Dsigma = diags(self.sigma.toarray(), shape=(3 * N, 3 * N), dtype=dtype)
        self.tDsiDmuiDaC = self.iDa * self.iDmu * self.C * self.Ds
        self.itDaiDepsDstC = self.iDeps * self.itDa * self.C.transpose() * self.tDs

    # -- boundaries.py:14-188 ------------------------------------------------
    def _apply_bc_to_C(self):
        xlo = ylo = zlo = 1.0
        xhi = yhi = zhi = 1.0
        bl, bh = self.bc_low, self.bc_high
        if any(True for x in bl if x.lower() == "periodic"):
            if bl[0].lower() == "periodic" and bh[0].lower() == "periodic":
                self.tL[-1, :, :, "x"] = self.L[0, :, :, "x"]
                self.itA[-1, :, :, "y"] = self.iA[0, :, :, "y"]
                self.itA[-1, :, :, "z"] = self.iA[0, :, :, "z"]
            if bl[1].lower() == "periodic" and bh[1].lower() == "periodic":
                self.tL[:, -1, :, "y"] = self.L[:, 0, :, "y"]
                self.itA[:, -1, :, "x"] = self.iA[:, 0, :, "x"]
                self.itA[:, -1, :, "z"] = self.iA[:, 0, :, "z"]
            if bl[2].lower() == "periodic" and bh[2].lower() == "periodic":
                self.tL[:, :, -1, "z"] = self.L[:, :, 0, "z"]
                self.itA[:, :, -1, "x"] = self.iA[:, :, 0, "x"]
                self.itA[:, :, -1, "y"] = self.iA[:, :, 0, "y"]
            self.tDs = diags(self.tL.toarray(), shape=(3 * self.N, 3 * self.N), dtype=self.dtype)
            self.itDa = diags(self.itA.toarray(), shape=(3 * self.N, 3 * self.N), dtype=self.dtype)

        if any(True for x in bl if x.lower() in _PEC_LIKE) or any(
            True for x in bh if x.lower() in _PEC_LIKE
        ):
            if bl[0].lower() in _PEC_LIKE:
                xlo = 0
            if bl[1].lower() in _PEC_LIKE:
                ylo = 0
            if bl[2].lower() in _PEC_LIKE:
                zlo = 0
            if bh[0].lower() in _PEC_LIKE:
                xhi = 0
            if bh[1].lower() in _PEC_LIKE:
                yhi = 0
            if bh[2].lower() in _PEC_LIKE:
                zhi = 0
            self.BC = self._face_mask(xlo, xhi, ylo, yhi, zlo, zhi)
            self.Dbc = diags(self.BC.toarray(), shape=(3 * self.N, 3 * self.N), dtype=np.int8)
            self.C = self.C * self.Dbc

        if any(True for x in bl if x.lower() in ("magnetic", "pmc")) or any(
            True for x in bh if x.lower() in ("magnetic", "pmc")
        ):
            # boundaries.py:116-127: case-sensitive "pmc", flags carried over
            if bl[0].lower() == "magnetic" or bl[0] == "pmc":
                xlo = 0
            if bl[1].lower() == "magnetic" or bl[1] == "pmc":
                ylo = 0
            if bl[2].lower() == "magnetic" or bl[2] == "pmc":
                zlo = 0
            if bh[0].lower() == "magnetic" or bh[0] == "pmc":
                xhi = 0
            if bh[1].lower() == "magnetic" or bh[1] == "pmc":
                yhi = 0
            if bh[2].lower() == "magnetic" or bh[2] == "pmc":
                zhi = 0
            self.BC = self._face_mask(xlo, xhi, ylo, yhi, zlo, zhi)
            self.Dbc = diags(self.BC.toarray(), shape=(3 * self.N, 3 * self.N), dtype=np.int8)
            self.C = self.Dbc * self.C

        if any(True for x in bl if x.lower() == "abc") or any(
            True for x in bh if x.lower() == "abc"
        ):
            if bh[0].lower() == "abc":
                self.tL[-1, :, :, "x"] = self.L[0, :, :, "x"]
                self.itA[-1, :, :, "y"] = self.iA[0, :, :, "y"]
                self.itA[-1, :, :, "z"] = self.iA[0, :, :, "z"]
            if bh[1].lower() == "abc":
                self.tL[:, -1, :, "y"] = self.L[:, 0, :, "y"]
                self.itA[:, -1, :, "x"] = self.iA[:, 0, :, "x"]
                self.itA[:, -1, :, "z"] = self.iA[:, 0, :, "z"]
            if bh[2].lower() == "abc":
                self.tL[:, :, -1, "z"] = self.L[:, :, 0, "z"]
                self.itA[:, :, -1, "x"] = self.iA[:, :, 0, "x"]
                self.itA[:, :, -1, "y"] = self.iA[:, :, 0, "y"]
            self.tDs = diags(self.tL.toarray(), shape=(3 * self.N, 3 * self.N), dtype=self.dtype)
            self.itDa = diags(self.itA.toarray(), shape=(3 * self.N, 3 * self.N), dtype=self.dtype)
            self.activate_abc = True

        if any(True for x in bl if x.lower() == "pml") or any(
            True for x in bh if x.lower() == "pml"
        ):
            self.activate_pml = True
            self.use_conductivity = True

    def _face_mask(self, xlo, xhi, ylo, yhi, zlo, zhi):
        BC = OField(self.Nx, self.Ny, self.Nz, dtype=np.int8, use_ones=True)
        for d in ["x", "y", "z"]:
            if d != "x":
                BC[0, :, :, d] = xlo
                BC[-1, :, :, d] = xhi
            if d != "y":
                BC[:, 0, :, d] = ylo
                BC[:, -1, :, d] = yhi
            if d != "z":
                BC[:, :, 0, d] = zlo
                BC[:, :, -1, d] = zhi
        return BC

    # -- boundaries.py:190-305 ----------------------------------------------
    def _initialize_PML(self):
        sx, sy, sz = np.zeros(self.Nx), np.zeros(self.Ny), np.zeros(self.Nz)
        lo, hi, n = 5.0e-3, 1.0, self.n_pml
        ieps0 = 1 / eps_0
        if self.bc_low[0].lower() == "pml":
            sx[0:n] = np.geomspace(hi, lo, n)
            for d in "xyz":
                for i in range(n):
                    self.ieps[i, :, :, d] = ieps0
                    self.sigma[i, :, :, d] = 0.0 + 1 * sx[i]
        if self.bc_low[1].lower() == "pml":
            sy[0:n] = np.geomspace(hi, lo, n)
            for d in "xyz":
                for j in range(n):
                    self.ieps[:, j, :, d] = ieps0
                    self.sigma[:, j, :, d] = 0.0 + 1 * sy[j]
        if self.bc_low[2].lower() == "pml":
            sz[0:n] = np.geomspace(hi, lo, n)
            for d in "xyz":
                for k in range(n):
                    self.ieps[:, :, k, d] = ieps0
                    self.sigma[:, :, k, d] = 0.0 + 1 * sz[k]
        if self.bc_high[0].lower() == "pml":
            sx[-n:] = np.geomspace(lo, hi, n)
            for d in "xyz":
                for i in range(1, n + 1):
                    self.ieps[-i, :, :, d] = ieps0
                    self.sigma[-i, :, :, d] = 0.0 + 1 * sx[-i]
        if self.bc_high[1].lower() == "pml":
            sy[-n:] = np.geomspace(lo, hi, n)
            for d in "xyz":
                for j in range(1, n + 1):
                    self.ieps[:, -j, :, d] = ieps0
                    self.sigma[:, -j, :, d] = 0.0 + 1 * sy[-j]
        if self.bc_high[2].lower() == "pml":
            sz[-n:] = np.geomspace(lo, hi, n)
            for d in "xyz":
                for k in range(1, n + 1):
                    self.ieps[:, :, -k, d] = ieps0
                    self.sigma[:, :, -k, d] = 0.0 + 1 * sz[-k]

    # -- solverFIT3D.py:438-502 (masks given directly) ----------------------
    def _apply_solid_materials(self, solids):
        for mask_b, mat in solids:
            mask = np.asarray(mask_b).astype(int)
            if type(mat) is str:
                mat = MATERIAL_LIB[mat.lower()]
            mat = list(mat)
            eps = mat[0] * eps_0
            mu = mat[1] * mu_0
            if len(mat) == 3:
                sigma = mat[2]
                if self.use_sibc and sigma > 0.0:
                    eps = sigma * eps_0
                self.sigma += self.sigma * (-1.0 * mask)
                self.sigma += mask * sigma
                self.use_conductivity = True
            elif self.sigma_bg > 0.0:
                self.sigma += self.sigma * (-1.0 * mask)
            self.ieps += self.ieps * (-1.0 * mask)
            self.imu += self.imu * (-1.0 * mask)
            self.ieps += mask * 1.0 / eps
            self.imu += mask * 1.0 / mu

    # -- solverFIT3D.py:378-436 ---------------------------------------------
    def update_tensors(self, tensor="all"):
        N3 = 3 * self.N
        if tensor in ("ieps", "all"):
            self.iDeps = diags(self.ieps.toarray(), shape=(N3, N3), dtype=self.dtype)
        if tensor in ("imu", "all"):
            self.iDmu = diags(self.imu.toarray(), shape=(N3, N3), dtype=self.dtype)
        if tensor in ("sigma", "all"):
            self.Dsigma = diags(self.sigma.toarray(), shape=(N3, N3), dtype=self.dtype)
        self.tDsiDmuiDaC = self.iDa * self.iDmu * self.C * self.Ds
        self.itDaiDepsDstC = self.iDeps * self.itDa * self.C.transpose() * self.tDs
        self.step_0 = False

    # -- solverFIT3D.py:504-525 ---------------------------------------------
    def one_step(self):
        if self.step_0:
            self._set_ghosts_to_0()
            self.step_0 = False
        self.H.fromarray(self.H.toarray() - self.dt * self.tDsiDmuiDaC * self.E.toarray())
        self.E.fromarray(
            self.E.toarray()
            + self.dt
            * (self.itDaiDepsDstC * self.H.toarray() - self.ieps.toarray() * self.J.toarray())
        )
        if self.use_conductivity:
            self.J.fromarray(self.sigma.toarray() * self.E.toarray())

    # -- solverFIT3D.py:842-862 ---------------------------------------------
    def _set_ghosts_to_0(self):
        for d in ["x", "y", "z"]:
            if d != "x":
                self.H[-1, :, :, d] = 0.0
            if d != "y":
                self.H[:, -1, :, d] = 0.0
            if d != "z":
                self.H[:, :, -1, d] = 0.0
        self.E[-1, :, :, "x"] = 0.0
        self.E[:, -1, :, "y"] = 0.0
        self.E[:, :, -1, "z"] = 0.0

    # -- boundaries.py:307-404 (CPU: get_abc returns live views) -------------
    def get_abc(self):
        E_abc, H_abc = {}, {}
        bl, bh = self.bc_low, self.bc_high
        planes = [
            (bl[0], 0, "lo", (1, slice(None), slice(None))),
            (bl[1], 1, "lo", (slice(None), 1, slice(None))),
            (bl[2], 2, "lo", (slice(None), slice(None), 1)),
            (bh[0], 0, "hi", (-1, slice(None), slice(None))),
            (bh[1], 1, "hi", (slice(None), -1, slice(None))),
            (bh[2], 2, "hi", (slice(None), slice(None), -1)),
        ]
        for bc, ax, side, idx in planes:
            if bc.lower() == "abc":
                E_abc[ax] = {}
                H_abc[ax] = {}
                for d in "xyz":
                    E_abc[ax][d + side] = self.E[idx + (d,)]
                    H_abc[ax][d + side] = self.H[idx + (d,)]
        return E_abc, H_abc

    def update_abc(self, E_abc, H_abc):
        bl, bh = self.bc_low, self.bc_high
        planes = [
            (bl[0], 0, "lo", (0, slice(None), slice(None))),
            (bl[1], 1, "lo", (slice(None), 0, slice(None))),
            (bl[2], 2, "lo", (slice(None), slice(None), 0)),
            (bh[0], 0, "hi", (-1, slice(None), slice(None))),
            (bh[1], 1, "hi", (slice(None), -1, slice(None))),
            (bh[2], 2, "hi", (slice(None), slice(None), -1)),
        ]
        for bc, ax, side, idx in planes:
            if bc.lower() == "abc":
                for d in "xyz":
                    self.E[idx + (d,)] = E_abc[ax][d + side]
                    self.H[idx + (d,)] = H_abc[ax][d + side]

    # -- routines.py:17-229 (no save/plot) -----------------------------------
    def emsolve(self, Nt, source=None, callback=None):
        if self.activate_abc:
            E_abc_2, H_abc_2 = self.get_abc()
            E_abc_1, H_abc_1 = self.get_abc()
        for n in range(Nt):
            if source is not None:
                source.update(self, n * self.dt)
            self.one_step()
            if self.activate_abc:
                self.update_abc(E_abc_2, H_abc_2)
                E_abc_2, H_abc_2 = E_abc_1, H_abc_1
                E_abc_1, H_abc_1 = self.get_abc()
            if callback is not None:
                callback(self, n * self.dt)

    # -- routines.py:231-485 (HDF5 replaced by a dict) -----------------------
    def wakesolve(self, wakelength=None, wake=None, compute_plane="both",
                  callback=None, solve=True):
        if wake is not None:
            self.wake = wake
        if self.wake is None:
            raise AttributeError("Wake solver information not passed to the solver instantiation")
        if wakelength is None:
            wakelength = self.wake.wakelength
        self.ixt = np.abs(self.x - self.wake.xtest).argmin()
        self.iyt = np.abs(self.y - self.wake.ytest).argmin()
        if compute_plane.lower() == "longitudinal":
            xx, yy = self.ixt, self.iyt
        else:
            xx = slice(self.ixt - 1, self.ixt + 2)
            yy = slice(self.iyt - 1, self.iyt + 2)
        self.wake.wakelength = wakelength
        self.ti = self.wake.ti
        self.v = self.wake.v
        z = self.z
        zz = slice(0, self.Nz)
        tmax = (wakelength + self.ti * self.v + (z.max() - z.min())) / self.v
        Nt = int(tmax / self.dt)
        self.tmax, self.Nt = tmax, Nt
        beam = OBeam(q=self.wake.q, sigmaz=self.wake.sigmaz, beta=self.wake.beta,
                     xsource=self.wake.xsource, ysource=self.wake.ysource)
        hf = {}
        hf["x"], hf["y"], hf["z"] = self.x[xx], self.y[yy], z[zz]
        hf["dx"], hf["dy"], hf["dz"] = self.grid.dx, self.grid.dy, self.grid.dz
        hf["t"] = np.arange(0, Nt * self.dt, self.dt)
        for n in range(Nt):
            beam.update(self, n * self.dt)
            hf["#" + str(n).zfill(5)] = np.array(self.E[xx, yy, zz, "z"])
            self.one_step()
            if callback is not None:
                callback(self, n * self.dt)
        self.wake.Ez_hf = hf
        self.Ez_hf = hf
        if solve:
            self.wake.solve(compute_plane=compute_plane)
        return hf


# --------------------------------------------------------------------------
# Field monitor  (wakis/field_monitors.py)
# --------------------------------------------------------------------------
class OFieldMonitor:
    """wakis/field_monitors.py:14-131: complex128 accumulators (nf, Nx, Ny,
    Nz) per E component; update adds E_c * exp(-2j*pi*f*t), t = time_index *
    dt with dt fixed at the first update (96-104)."""

    def __init__(self, frequencies):
        self.frequencies = np.array(frequencies)
        self.time_index = 0
        self.dt = None
        self.Ex_acc = self.Ey_acc = self.Ez_acc = None
        self.shape = None

    def update(self, E, dt):
        if self.dt is None:
            self.dt = dt
            self.shape = (E.Nx, E.Ny, E.Nz)
            shape = (len(self.frequencies), *self.shape)
            self.Ex_acc = np.zeros(shape, dtype=np.complex128)
            self.Ey_acc = np.zeros(shape, dtype=np.complex128)
            self.Ez_acc = np.zeros(shape, dtype=np.complex128)
        t = self.time_index * self.dt
        Ex, Ey, Ez = E.to_matrix("x"), E.to_matrix("y"), E.to_matrix("z")
        for i, f in enumerate(self.frequencies):
            phase = np.exp(-2j * np.pi * f * t)
            self.Ex_acc[i] += Ex * phase
            self.Ey_acc[i] += Ey * phase
            self.Ez_acc[i] += Ez * phase
        self.time_index += 1

    def get_components(self):
        return {"Ex": self.Ex_acc, "Ey": self.Ey_acc, "Ez": self.Ez_acc}


# --------------------------------------------------------------------------
# Sources  (wakis/sources.py)
# --------------------------------------------------------------------------
class OBeam:
    """wakis/sources.py:22-110 (Gaussian bunch written into J_z each step)."""

    def __init__(self, xsource=0.0, ysource=0.0, beta=1.0, q=1e-9, sigmaz=None, ti=None):
        self.xsource, self.ysource = xsource, ysource
        self.sigmaz = sigmaz
        self.q = q
        self.beta = beta
        self.v = c_light * beta
        self.ti = ti if ti is not None else 8.548921333333334 * self.sigmaz / self.v
        self.is_first_update = True

    def line(self, solver, t):
        if self.is_first_update:
            self.ixs = np.abs(solver.x - self.xsource).argmin()
            self.iys = np.abs(solver.y - self.ysource).argmin()
            self.is_first_update = False
            self.zmin = solver.z.min()
        s0 = self.zmin - self.v * self.ti
        s = solver.z - self.v * t
        profile = 1 / np.sqrt(2 * np.pi * self.sigmaz**2) * np.exp(
            -((s - s0) ** 2) / (2 * self.sigmaz**2)
        )
        return self.q * self.v * profile / solver.dx[self.ixs] / solver.dy[self.iys]

    def update(self, solver, t):
        vals = self.line(solver, t)
        solver.J[self.ixs, self.iys, :, "z"] = vals


class OPlaneWave:
    """wakis/sources.py:142-248 (E_x/H_y written on a z slice each step)."""

    def __init__(self, xs=None, ys=None, zs=0, nodes=None, f=None, amplitude=1.0,
                 beta=1.0, phase=0):
        self.nodes, self.beta = nodes, beta
        self.xs, self.ys, self.zs = xs, ys, zs
        self.f, self.amplitude, self.phase = f, amplitude, phase
        self.is_first_update = True
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
                self.amplitude * np.cos(self.w * t + self.phase) / (self.kz / (mu_0 * self.vp))
            )
        else:
            solver.H[self.xs, self.ys, self.zs, "y"] = 0.0
            solver.E[self.xs, self.ys, self.zs, "x"] = 0.0


# --------------------------------------------------------------------------
# Wake  (wakis/wakeSolver.py)
# --------------------------------------------------------------------------
class OWake:
    """WakeSolver: beam parameters, WP(s) and Z(f) post-processing."""

    def __init__(self, wakelength=None, q=1e-9, sigmaz=1e-3, beta=1.0, xsource=0.0,
                 ysource=0.0, xtest=0.0, ytest=0.0, ti=None, fmax=None, skip_cells=0,
                 counter_moving=False):
        # wakeSolver.py:133-173
        self.q, self.sigmaz, self.beta = q, sigmaz, beta
        self.v = self.beta * c_light
        self.xsource, self.ysource = xsource, ysource
        self.xtest, self.ytest = xtest, ytest
        self.fmax = self.v / (3.0 * self.sigmaz) if fmax is None else fmax
        self.skip_cells = skip_cells
        if ti is not None:
            self.ti = ti
        else:
            self.ti = 8.548921333333334 * self.sigmaz / (np.sqrt(self.beta) * self.v)
        self.wakelength = wakelength
        self.n_transverse_cells = 1
        self.lambdas = None
        self.chargedist = None
        self.counter_moving = counter_moving
        self.Ez_hf = None

    # read_Ez (wakeSolver.py:1317-1364) from the in-memory dict
    def _read(self):
        hf = self.Ez_hf
        self.xf, self.yf, self.zf = np.array(hf["x"]), np.array(hf["y"]), np.array(hf["z"])
        self.dx, self.dy, self.dz = np.array(hf["dx"]), np.array(hf["dy"]), np.array(hf["dz"])
        self.t = np.array(hf["t"])
        keys = sorted(k for k in hf.keys() if isinstance(k, str) and k.startswith("#"))
        return keys

    def _setup(self):
        keys = self._read()
        nt = len(self.t)
        dt = self.t[2] - self.t[1]
        ti = self.ti
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
            wakelength = nt * dt * self.v - (zmax - zmin) - ti * self.v
            self.wakelength = wakelength
        s = np.arange(-self.ti * self.v, wakelength, dt * self.v)
        nt = min(nt, len(keys))
        return keys, nt, dt, ti, zmin, zmax, zz, z, s

    def bucket_table(self):
        """it[n, k] = int(ts/dt) exactly as wakeSolver.py:517-529 (vectorised
        over n; every element op has the dtype of the scalar loop)."""
        keys, nt, dt, ti, zmin, zmax, zz, z, s = self._setup()
        it = np.empty((len(s), len(z)), dtype=np.int64)
        for k in range(len(z)):
            ts = (z[k] + s) / self.v - zmin / self.v - self.t[0] + ti
            it[:, k] = np.trunc(ts / dt).astype(np.int64)
        return it, nt, s

    def calc_long_WP_3d(self, loops=False):
        # wakeSolver.py:423-551
        keys, nt, dt, ti, zmin, zmax, zz, z, s = self._setup()
        nz = len(z)
        i0 = j0 = self.n_transverse_cells
        WP = np.zeros_like(s)
        WP_3d = np.zeros((i0 * 2 + 1, j0 * 2 + 1, len(s)))
        Ezt = np.zeros((nz, len(self.t)))
        for i in range(-i0, i0 + 1):
            for j in range(-j0, j0 + 1):
                for n in range(nt):
                    Ez = self.Ez_hf[keys[n]]
                    Ezt[:, n] = Ez[Ez.shape[0] // 2 + i, Ez.shape[1] // 2 + j, zz]
                if self.counter_moving and loops:
                    # wakeSolver.py:500-513
                    for n in range(len(s)):
                        for k in range(nz):
                            ts = ((z[-k - 1] - s[n]) / (-1 * self.v) - zmax / (-1 * self.v)
                                  - self.t[0] + ti)
                            it = int(ts / dt)
                            if it < nt:
                                WP[n] = WP[n] + (Ezt[-k - 1, it]) * (-1 * self.dz[k])
                elif self.counter_moving:
                    WP = _wp_accumulate_counter(WP, Ezt, z, s, self.v, zmax, self.t[0], ti,
                                                dt, nt, self.dz)
                elif loops:
                    for n in range(len(s)):
                        for k in range(nz):
                            ts = (z[k] + s[n]) / self.v - zmin / self.v - self.t[0] + ti
                            it = int(ts / dt)
                            if it < nt:
                                WP[n] = WP[n] + (Ezt[k, it]) * self.dz[k]
                else:
                    WP = _wp_accumulate(WP, Ezt, z, s, self.v, zmin, self.t[0], ti, dt,
                                        nt, self.dz)
                WP = WP / (self.q * 1e12)
                WP_3d[i0 + i, j0 + j, :] = WP
        self.s = s
        self.WP = WP_3d[i0, j0, :]
        self.WP_3d = WP_3d

    def calc_long_WP(self, loops=False):
        # wakeSolver.py:311-421 (window centre only, fresh WP)
        keys, nt, dt, ti, zmin, zmax, zz, z, s = self._setup()
        nz = len(z)
        WP = np.zeros_like(s)
        Ezt = np.zeros((nz, nt))
        Ez = self.Ez_hf[keys[0]]
        if len(Ez.shape) == 3:
            for n in range(nt):
                Ez = self.Ez_hf[keys[n]]
                Ezt[:, n] = Ez[Ez.shape[0] // 2 + 1, Ez.shape[1] // 2 + 1, zz]
        elif len(Ez.shape) == 1:
            for n in range(nt):
                Ezt[:, n] = self.Ez_hf[keys[n]]
        self.Ezt = Ezt
        if loops:
            for n in range(len(s)):
                for k in range(nz):
                    ts = (z[k] + s[n]) / self.v - zmin / self.v - self.t[0] + ti
                    it = int(ts / dt)
                    if it < nt:
                        WP[n] = WP[n] + (Ezt[k, it]) * self.dz[k]
        else:
            WP = _wp_accumulate(WP, Ezt, z, s, self.v, zmin, self.t[0], ti, dt, nt, self.dz)
        WP = WP / (self.q * 1e12)
        self.s = s
        self.WP = WP

    def calc_trans_WP(self):
        # wakeSolver.py:553-626 (cumulative sums restated with the same
        # left-to-right order as np.sum over WP_3d[..., 0:n])
        ds = self.s[2] - self.s[1]
        i0 = j0 = self.n_transverse_cells
        WPx = np.zeros_like(self.s)
        WPy = np.zeros_like(self.s)
        int_WP = np.zeros_like(self.WP_3d)
        dxf = np.diff(self.xf)
        dyf = np.diff(self.yf)
        for n in range(len(self.s)):
            for i in range(-i0, i0 + 1):
                for j in range(-j0, j0 + 1):
                    int_WP[i0 + i, j0 + j, n] = np.sum(self.WP_3d[i0 + i, j0 + j, 0:n]) * ds
            WPx[n] = -(int_WP[i0 + 1, j0, n] - int_WP[i0 - 1, j0, n]) / (dxf[i0 - 1] + dxf[i0])
            WPy[n] = -(int_WP[i0, j0 + 1, n] - int_WP[i0, j0 - 1, n]) / (dyf[j0 - 1] + dyf[j0])
        self.WPx, self.WPy = WPx, WPy

    def calc_lambdas_analytic(self):
        # wakeSolver.py:842-874
        self.lambdas = 1 / (self.sigmaz * np.sqrt(2 * np.pi)) * np.exp(
            -(self.s**2) / (2 * self.sigmaz**2)
        )

    def calc_long_Z(self, samples=1001, fmax=None):
        # wakeSolver.py:628-720
        if self.lambdas is None:
            self.calc_lambdas_analytic()
        ds = np.mean(self.s[1:] - self.s[:-1])
        fmax = self.fmax if fmax is None else fmax
        N = int((self.v / ds) // fmax * samples)
        lambdafft = np.fft.fft(self.lambdas * self.v, n=N)
        WPfft = np.fft.fft(self.WP * 1e12, n=N)
        ffft = np.fft.fftfreq(len(WPfft), ds / self.v)
        mask = np.logical_and(ffft >= 0, ffft < fmax)
        WPf = WPfft[mask] * ds
        lambdaf = lambdafft[mask] * ds
        self.f = ffft[mask]
        self.Z = -WPf / lambdaf
        self.lambdaf = lambdaf
        self.N_dft = N

    def calc_trans_Z(self, samples=1001, fmax=None):
        # wakeSolver.py:722-785
        ds = np.mean(self.s[1:] - self.s[:-1])
        fmax = self.fmax if fmax is None else fmax
        N = int((self.v / ds) // fmax * samples)
        lambdafft = np.fft.fft(self.lambdas * self.v, n=N)
        ffft = np.fft.fftfreq(len(lambdafft), ds / self.v)
        mask = np.logical_and(ffft >= 0, ffft < fmax)
        lambdaf = lambdafft[mask] * ds
        self.Zx = 1j * (np.fft.fft(self.WPx * 1e12, n=N)[mask] * ds) / lambdaf
        self.Zy = 1j * (np.fft.fft(self.WPy * 1e12, n=N)[mask] * ds) / lambdaf
        self.fx = self.fy = ffft[mask]

    def solve(self, compute_plane="both", trans=True):
        # wakeSolver.py:211-255
        if compute_plane.lower() in ["both", "transverse"]:
            self.calc_long_WP_3d()
            if trans:
                self.calc_trans_WP()
            self.calc_long_Z()
            if trans:
                self.calc_trans_Z()
        elif compute_plane.lower() == "longitudinal":
            self.calc_long_WP()
            self.calc_long_Z()


def _wp_accumulate_counter(WP, Ezt, z, s, v, zmax, t0, ti, dt, nt, dz):
    """The counter-moving branch of wakeSolver.py:500-513, vectorised over n
    with the loop's per-n order over k: term k samples z[-k-1] at the time
    measured from zmax with the velocity negated, weighted by -dz[k] (dz
    indexed from the bottom while z is walked from the top)."""
    WP = WP.copy()
    for k in range(len(z)):
        ts = (z[-k - 1] - s) / (-1 * v) - zmax / (-1 * v) - t0 + ti
        it = np.trunc(ts / dt).astype(np.int64)
        ok = it < nt
        idx = np.where(it < 0, it + Ezt.shape[1], it)
        WP[ok] = WP[ok] + (Ezt[-k - 1, idx[ok]]) * (-1 * dz[k])
    return WP


def _wp_accumulate(WP, Ezt, z, s, v, zmin, t0, ti, dt, nt, dz):
    """Vectorised over n, sequential over k: the same per-n addition order as
    the double loop at wakeSolver.py:517-529, so results are bit-identical."""
    WP = WP.copy()
    for k in range(len(z)):
        ts = (z[k] + s) / v - zmin / v - t0 + ti
        it = np.trunc(ts / dt).astype(np.int64)
        ok = it < nt
        # int() of a value in (-1, 0) is 0; a value <= -1 would index from the
        # end in the reference: reproduce Python's negative indexing.
        idx = np.where(it < 0, it + Ezt.shape[1], it)
        WP[ok] = WP[ok] + Ezt[k, idx[ok]] * dz[k]
    return WP
