"""``GridFIT3D``: axes, FIT metrics and solid masks.

Axes and spacings follow wakis/gridFIT3D.py:143-185 (custom ``x, y, z``
arrays, or ``linspace`` from extents and cell counts).  The FIT metrics of
wakis/gridFIT3D.py:263-305 are separable on a rectilinear grid, so the
solver keeps them as 1-D and 2-D tables (``metric_tables``) instead of four
3N Fields; the Fields ``L, iA, tL, itA`` are still available, built on
first access.

Solids: ``stl_solids``/``stl_materials`` as in the reference (an own STL
reader and ray-casting voxeliser replace pyvista; a cell belongs to a solid
if at least one of its 8 corner nodes is inside, as the reference's
``point_data_to_cell_data() > tol`` test, wakis/gridFIT3D.py:502-510), or
``solids=[(mask, material), ...]`` with ready (Nx, Ny, Nz) boolean masks.
"""

from __future__ import annotations

import numpy as np

from .field import Field
from .geometry import read_stl, transform_triangles, voxelize_nodes
from .materials import material_lib


class GridFIT3D:
    def __init__(self, xmin=None, xmax=None, ymin=None, ymax=None, zmin=None, zmax=None,
                 Nx=None, Ny=None, Nz=None, x=None, y=None, z=None, use_mpi=False,
                 stl_solids=None, stl_materials=None, stl_rotate=(0.0, 0.0, 0.0),
                 stl_translate=(0.0, 0.0, 0.0), stl_scale=1.0, stl_colors=None,
                 stl_tol=1e-3, solids=None, verbose=1, **kwargs):
        if use_mpi:
            raise NotImplementedError(
                "use_mpi: multi-GPU z-slab decomposition is not part of this build")
        self.verbose = verbose
        self.use_mpi = False
        if x is not None and y is not None and z is not None:
            self.x, self.y, self.z = np.asarray(x), np.asarray(y), np.asarray(z)
            self.Nx, self.Ny, self.Nz = len(x) - 1, len(y) - 1, len(z) - 1
            self.xmin, self.xmax = self.x[0], self.x[-1]
            self.ymin, self.ymax = self.y[0], self.y[-1]
            self.zmin, self.zmax = self.z[0], self.z[-1]
        elif all(v is not None for v in [xmin, xmax, ymin, ymax, zmin, zmax]):
            self.xmin, self.xmax = xmin, xmax
            self.ymin, self.ymax = ymin, ymax
            self.zmin, self.zmax = zmin, zmax
            self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
            self.x = np.linspace(self.xmin, self.xmax, self.Nx + 1)
            self.y = np.linspace(self.ymin, self.ymax, self.Ny + 1)
            self.z = np.linspace(self.zmin, self.zmax, self.Nz + 1)
        else:
            raise ValueError(
                "[!] Error initializing GridFIT3D: provide x, y, z arrays or "
                "xmin, xmax, ymin, ymax, zmin, zmax, Nx, Ny, Nz")
        self.dx = np.diff(self.x)
        self.dy = np.diff(self.y)
        self.dz = np.diff(self.z)

        self.stl_solids = stl_solids
        self.stl_materials = stl_materials
        self.stl_colors = stl_colors
        self.stl_tol = stl_tol
        # list of (name, mask, material) in application order
        self.solids = []
        if stl_solids is not None:
            self._prepare_stl_dicts(stl_rotate, stl_translate, stl_scale)
            for key in self.stl_solids:
                self.solids.append((key, self._mark_cells_in_stl(key), self.stl_materials[key]))
        if solids is not None:
            for n, (mask, mat) in enumerate(solids):
                if isinstance(mask, tuple) and len(mask) == 3 and all(isinstance(m, slice) for m in mask):
                    # a box of cells given as index slices (no (Nx, Ny, Nz) mask
                    # array: grids of 1e9 cells)
                    self.solids.append((f"solid{n}", mask, mat))
                    continue
                mask = np.asarray(mask, dtype=bool)
                if mask.shape != (self.Nx, self.Ny, self.Nz):
                    raise ValueError("solid mask must have shape (Nx, Ny, Nz)")
                self.solids.append((f"solid{n}", mask, mat))
        self._fields = None

    # -- STL (wakis/gridFIT3D.py:397-652) --------------------------------------
    def _prepare_stl_dicts(self, rotate, translate, scale):
        if type(self.stl_solids) is not dict:
            if type(self.stl_solids) is str:
                self.stl_solids = {"Solid 1": self.stl_solids}
            else:
                raise Exception("[!] Attribute `stl_solids` must contain a string or a dictionary")
        if type(self.stl_materials) is not dict:
            if type(self.stl_materials) is str:
                self.stl_materials = {"Solid 1": self.stl_materials}
            else:
                raise Exception("[!] Attribute `stl_materials` must contain a string or a dictionary")

        def per_key(v):
            if isinstance(v, dict):
                return v
            return {k: v for k in self.stl_solids}

        self.stl_rotate = per_key(rotate)
        self.stl_translate = per_key(translate)
        self.stl_scale = per_key(scale)
        for key in self.stl_solids:
            mat = self.stl_materials[key]
            if type(mat) is str:
                m = material_lib[mat.lower()]
                self.stl_materials[key] = list(m)

    def read_stl(self, key):
        tri = read_stl(self.stl_solids[key])
        return transform_triangles(tri, self.stl_rotate[key], self.stl_translate[key],
                                   self.stl_scale[key])

    def _mark_cells_in_stl(self, key):
        inside = voxelize_nodes(self.read_stl(key), self.x, self.y, self.z)
        # cell (i,j,k) is in the solid if any of its 8 corner nodes is inside
        c = np.zeros((self.Nx, self.Ny, self.Nz), dtype=bool)
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    c |= inside[di:di + self.Nx, dj:dj + self.Ny, dk:dk + self.Nz]
        return c

    # -- metric Fields, built on demand (wakis/gridFIT3D.py:263-305) ----------
    def _metric_fields(self):
        if self._fields is None:
            Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
            L1, tL1, iA2, itA2 = metric_tables_raw(self.x, self.y, self.z)
            out = {}
            for name, tab in (("L", L1), ("tL", tL1)):
                f = Field(Nx, Ny, Nz)
                f.field_x = np.broadcast_to(tab[0][:, None, None], (Nx, Ny, Nz))
                f.field_y = np.broadcast_to(tab[1][None, :, None], (Nx, Ny, Nz))
                f.field_z = np.broadcast_to(tab[2][None, None, :], (Nx, Ny, Nz))
                out[name] = f
            for name, tab in (("iA", iA2), ("itA", itA2)):
                f = Field(Nx, Ny, Nz)
                f.field_x = np.broadcast_to(tab[0][None, :, :], (Nx, Ny, Nz))
                f.field_y = np.broadcast_to(tab[1][:, None, :], (Nx, Ny, Nz))
                f.field_z = np.broadcast_to(tab[2][:, :, None], (Nx, Ny, Nz))
                out[name] = f
            self._fields = out
        return self._fields

    L = property(lambda s: s._metric_fields()["L"])
    tL = property(lambda s: s._metric_fields()["tL"])
    iA = property(lambda s: s._metric_fields()["iA"])
    itA = property(lambda s: s._metric_fields()["itA"])


def metric_tables_raw(x, y, z):
    """1-D lengths and 2-D inverse areas of wakis/gridFIT3D.py:263-305.

    Returns (L, tL, iA, itA): L and tL are (Lx[i], Ly[j], Lz[k]); iA and itA
    are (A_x[j,k], A_y[i,k], A_z[i,j]) with the same element arithmetic as
    the reference (products of the 1-D lengths, then 1.0/...; 0 where the
    dual area is 0, i.e. on the duplicated last dual node).
    """
    L = [np.diff(x), np.diff(y), np.diff(z)]
    tl = []
    for a in (x, y, z):
        t = (a[1:] + a[:-1]) / 2
        t = np.append(t, t[-1])
        tl.append(np.diff(t))
    Lx, Ly, Lz = L
    tLx, tLy, tLz = tl
    iA = [np.divide(1.0, Ly[:, None] * Lz[None, :]),
          np.divide(1.0, Lx[:, None] * Lz[None, :]),
          np.divide(1.0, Lx[:, None] * Ly[None, :])]
    itA = []
    for a, b in ((tLy, tLz), (tLx, tLz), (tLx, tLy)):
        aux = a[:, None] * b[None, :]
        itA.append(np.divide(1.0, aux, out=np.zeros_like(aux), where=aux != 0))
    return L, tl, iA, itA


def metric_tables(x, y, z, bc_low, bc_high):
    """Metric tables after the boundary fix-ups of wakis/boundaries.py:36-70
    (periodic on both faces of an axis) and 153-181 (ABC on a high face):
    tL_d[-1] := L_d[0] and, for the two other components e, the last plane of
    itA_e along d := iA_e of plane 0.  Applied in the reference's order
    (periodic x, y, z, then ABC x, y, z).  Returns the packed FP64 array of
    include/wakis_fit.h (fit_set_geometry)."""
    L, tL, iA, itA = metric_tables_raw(x, y, z)
    tL = [t.copy() for t in tL]
    itA = [t.copy() for t in itA]
    bl = [b.lower() for b in bc_low]
    bh = [b.lower() for b in bc_high]

    def fix(d):
        tL[d][-1] = L[d][0]
        # component e's 2-D table axes: x -> (j,k), y -> (i,k), z -> (i,j)
        for e in range(3):
            if e == d:
                continue
            axes = [a for a in range(3) if a != e]
            pos = axes.index(d)
            if pos == 0:
                itA[e][-1, :] = iA[e][0, :]
            else:
                itA[e][:, -1] = iA[e][:, 0]

    if any(b == "periodic" for b in bl):
        for d in range(3):
            if bl[d] == "periodic" and bh[d] == "periodic":
                fix(d)
    if any(b == "abc" for b in bl) or any(b == "abc" for b in bh):
        for d in range(3):
            if bh[d] == "abc":
                fix(d)
    parts = list(L) + list(tL) + [np.reshape(t, -1, order="F") for t in iA] + \
        [np.reshape(t, -1, order="F") for t in itA]
    return np.concatenate(parts).astype(np.float64)
