"""``GridFIT3D``: axes, FIT metrics and solid masks.

Axes and spacings follow wakis/gridFIT3D.py:143-185 (custom ``x, y, z``
arrays, or ``linspace`` from extents and cell counts).  The FIT metrics of
wakis/gridFIT3D.py:263-305 are separable on a rectilinear grid, so the
solver keeps them as 1-D and 2-D tables (``metric_tables``) instead of four
3N Fields; the Fields ``L, iA, tL, itA`` are still available, built on
first access.

Solids: ``stl_solids``/``stl_materials`` as in the reference (an own STL
reader and ray-casting voxeliser replace pyvista), or
``solids=[(mask, material), ...]`` with ready (Nx, Ny, Nz) boolean masks.
``stl_method`` picks the cell rule of wakis/gridFIT3D.py:475-593 from the
node classification: "legacy" (default) marks a cell if at least one of its
8 corner nodes is inside (``point_data_to_cell_data() > tol``), and
"interior_points" if the corner mean exceeds 0.5 (5 or more corners);
"implicit_distance" marks a cell whose mean corner signed distance to the
surface is at most half the mean cell size; "voxelize_rectilinear" tests one
point per cell, (xmin + i*dx, ...) on the uniform spacing of the extents.

Mesh refinement (``use_mesh_refinement``, wakis/gridFIT3D.py:202-208,
654-870): snap points are the feature-edge points of the STL solids that
lie within ``snap_tol`` of the planes x=0, y=0 or z=0; each axis is then
rebuilt around its snap coordinates and relaxed by scipy's ``least_squares``
on the reference's three penalties (snap displacement, gaps below the
uniform gap, spread of the gaps).
"""

from __future__ import annotations

import numpy as np

import time

from .field import Field
from .geometry import (feature_edge_points, node_distance, read_stl, transform_triangles,
                       voxelize_nodes)
from .logger import Logger
from .materials import material_lib


class GridFIT3D:
    def __init__(self, xmin=None, xmax=None, ymin=None, ymax=None, zmin=None, zmax=None,
                 Nx=None, Ny=None, Nz=None, x=None, y=None, z=None, use_mpi=False,
                 stl_solids=None, stl_materials=None, stl_rotate=(0.0, 0.0, 0.0),
                 stl_translate=(0.0, 0.0, 0.0), stl_scale=1.0, stl_colors=None,
                 stl_tol=1e-3, stl_method="legacy", use_mesh_refinement=False,
                 refinement_method="insert", refinement_tol=1e-8, snap_points=None,
                 snap_tol=1e-5, snap_solids=None, solids=None, verbose=1, **kwargs):
        if use_mpi:
            raise NotImplementedError(
                "use_mpi: multi-GPU z-slab decomposition is not part of this build")
        t0 = time.time()
        self.logger = Logger()
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
        self.update_logger(["Nx", "Ny", "Nz", "dx", "dy", "dz"])
        self.update_logger(["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])

        self.stl_solids = stl_solids
        self.stl_materials = stl_materials
        self.stl_colors = stl_colors
        self.stl_tol = stl_tol
        self.stl_method = stl_method
        self.update_logger(["stl_solids", "stl_materials"])
        for name, val, default in (("stl_rotate", stl_rotate, (0.0, 0.0, 0.0)),
                                   ("stl_translate", stl_translate, (0.0, 0.0, 0.0)),
                                   ("stl_scale", stl_scale, 1.0)):
            setattr(self, name, val)
            if isinstance(val, dict) or np.any(np.asarray(val) != np.asarray(default)):
                self.update_logger([name])
        if stl_solids is not None:
            self._prepare_stl_dicts(stl_rotate, stl_translate, stl_scale)

        # axes refined around the solids' snap points (gridFIT3D.py:202-208)
        self.use_mesh_refinement = use_mesh_refinement
        self.refinement_method = refinement_method
        self.snap_points = snap_points
        self.snap_tol = snap_tol
        self.snap_solids = snap_solids
        self.update_logger(["use_mesh_refinement"])
        if use_mesh_refinement:
            if snap_points is None:
                if stl_solids is None:
                    raise ValueError("use_mesh_refinement needs stl_solids or snap_points")
                self._compute_snap_points(snap_solids=snap_solids, snap_tol=snap_tol)
            else:
                self._snaps_from_points(np.asarray(snap_points, dtype=np.float64).reshape(-1, 3))
            self._refine_xyz_axis(method=refinement_method, tol=refinement_tol)

        # list of (name, mask, material) in application order
        self.solids = []
        if stl_solids is not None:
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
        self.gridInitializationTime = time.time() - t0
        self.update_logger(["gridInitializationTime"])

    def update_logger(self, attrs):
        """Copy grid attributes into ``self.logger.grid`` (wakis/gridFIT3D.py:1088-1098)."""
        for a in attrs:
            self.logger.grid[a] = getattr(self, a)

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

    def _corner_sum(self, node):
        """Sum over each cell's 8 corner nodes of a (Nx+1, Ny+1, Nz+1) array."""
        c = np.zeros((self.Nx, self.Ny, self.Nz), dtype=np.result_type(node, np.int32))
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    c = c + node[di:di + self.Nx, dj:dj + self.Ny, dk:dk + self.Nz]
        return c

    def _mark_cells_in_stl(self, key):
        """Cell mask of one solid by the rule ``self.stl_method`` names."""
        tri = self.read_stl(key)
        method = self.stl_method.lower()
        if method == "voxelize_rectilinear":
            xs = self.xmin + np.arange(self.Nx) * ((self.xmax - self.xmin) / self.Nx)
            ys = self.ymin + np.arange(self.Ny) * ((self.ymax - self.ymin) / self.Ny)
            zs = self.zmin + np.arange(self.Nz) * ((self.zmax - self.zmin) / self.Nz)
            return voxelize_nodes(tri, xs, ys, zs)
        inside = voxelize_nodes(tri, self.x, self.y, self.z)
        if method == "legacy":
            return self._corner_sum(inside.astype(np.int32)) > 0
        if method == "interior_points":
            return self._corner_sum(inside.astype(np.int32)) > 4
        if method == "implicit_distance":
            thr = 0.5 * np.mean([np.mean(self.dx), np.mean(self.dy), np.mean(self.dz)])
            diag = np.sqrt(np.max(self.dx) ** 2 + np.max(self.dy) ** 2 + np.max(self.dz) ** 2)
            # exact distances within the band decide every cell: a corner
            # farther than thr + diag from the surface fixes its cell's answer
            d = node_distance(tri, self.x, self.y, self.z, band=thr + 1.01 * diag)
            d = np.where(inside, -d, d)
            return self._corner_sum(d) / 8.0 <= thr
        raise ValueError(f"[!] Error: stl_method {self.stl_method} not recognized. Use 'legacy', "
                         "'interior_points', 'implicit_distance', or 'voxelize_rectilinear'.")

    def _mark_cells_in_surface(self, key):
        """Keep only the surface layer of solid `key`: the cells with a face
        neighbour of the other mask value (the nonzero-gradient cells of
        wakis/gridFIT3D.py:606-622, which the reference's SIBC branch
        would call; that branch is unreachable there, solverFIT3D.py:476)."""
        for n, (name, mask, mat) in enumerate(self.solids):
            if name != key:
                continue
            m = np.asarray(self.cell_mask(mask), dtype=bool)
            edge = np.zeros_like(m)
            for ax in range(3):
                d = np.diff(m.astype(np.int8), axis=ax) != 0
                lo = [slice(None)] * 3
                hi = [slice(None)] * 3
                lo[ax] = slice(0, -1)
                hi[ax] = slice(1, None)
                edge[tuple(lo)] |= d
                edge[tuple(hi)] |= d
            self.solids[n] = (name, edge, mat)
            return edge
        raise KeyError(key)

    def cell_mask(self, mask):
        """(Nx, Ny, Nz) boolean array of a solid's mask (slice boxes expanded)."""
        if isinstance(mask, tuple):
            m = np.zeros((self.Nx, self.Ny, self.Nz), dtype=bool)
            m[mask] = True
            return m
        return mask

    # -- mesh refinement (wakis/gridFIT3D.py:654-870) --------------------------
    def _compute_snap_points(self, snap_solids=None, snap_tol=1e-8):
        """Feature-edge points of the union of the solids lying within
        `snap_tol` of the planes y=0, x=0 and z=0 (gridFIT3D.py:654-698)."""
        keys = list(self.stl_solids.keys()) if snap_solids is None else list(snap_solids)
        tri = np.concatenate([self.read_stl(k) for k in keys])
        p = feature_edge_points(tri)
        self.snap_points = np.r_[p[np.abs(p[:, 1]) < snap_tol], p[np.abs(p[:, 0]) < snap_tol],
                                 p[np.abs(p[:, 2]) < snap_tol]]
        self._snaps_from_points(self.snap_points)

    def _snaps_from_points(self, pts):
        """Per-axis snap coordinates: unique values rounded to 5 decimals, with
        the domain bounds added."""
        self.snap_points = pts
        for ax, (lo, hi) in enumerate(((self.xmin, self.xmax), (self.ymin, self.ymax),
                                       (self.zmin, self.zmax))):
            v = np.unique(np.round(pts[:, ax], 5))
            setattr(self, "xyz"[ax] + "_snaps", np.unique(np.concatenate(([lo], v, [hi]))))

    def refine_axis(self, xmin, xmax, Nx, x_snaps, method="insert", tol=1e-12):
        """An axis of about `Nx` nodes on [xmin, xmax] through the snap points
        (gridFIT3D.py:771-842).  Start: "insert" adds the snaps to a uniform
        axis of Nx - len(snaps) nodes, "neighbor" moves the nearest uniform
        node (within half a gap) onto each snap, "subdivision" starts from
        the snaps and halves the longest gap until there are Nx nodes (the
        reference's branch raises UnboundLocalError before it starts).  Then
        ``least_squares`` (dogbox, arctan loss, 3-point Jacobian) on the
        normalised nodes in [0, 1] minimises the three penalties."""
        from scipy.optimize import least_squares

        def penalties(u, u0, fixed):
            moved = np.sum((u[fixed] - u0[fixed]) ** 2) * 1000
            gap = np.diff(u)
            h = 1 / (len(u) - 1)
            small = np.sum((h - gap[gap < h]) ** 2)
            spread = np.std(np.diff(u)) * 10
            return np.hstack([moved, small, spread])

        s = (np.asarray(x_snaps, dtype=np.float64) - xmin) / (xmax - xmin)
        if method == "insert":
            u0 = np.unique(np.append(s, np.linspace(0, 1, Nx - len(s))))
        elif method == "neighbor":
            u0 = np.linspace(0, 1, Nx)
            h = np.diff(u0)[0]
            for v in s:
                near = np.isclose(u0, v, rtol=0.0, atol=h / 2)
                if np.sum(near) > 0:
                    u0[np.argmax(near)] = v
        elif method == "subdivision":
            u0 = np.unique(s)
            while len(u0) < Nx:
                i = np.argsort(np.diff(u0))[-1]
                u0 = np.unique(np.insert(u0, i + 1, u0[i] + (u0[i + 1] - u0[i]) / 2))
        else:
            raise ValueError(f"Method {method} not supported. Use 'insert', 'neighbor' or "
                             "'subdivision'.")
        fixed = np.isin(u0, s)
        res = least_squares(penalties, x0=u0.copy(), bounds=(0, 1), jac="3-point",
                            method="dogbox", loss="arctan", gtol=tol, ftol=tol, xtol=tol,
                            verbose=0, args=(u0.copy(), fixed.copy()))
        return res.x * (xmax - xmin) + xmin

    def _refine_xyz_axis(self, method="insert", tol=1e-6):
        """gridFIT3D.py:844-906: every axis refined, then counts and spacings."""
        self.x = self.refine_axis(self.xmin, self.xmax, self.Nx + 1, self.x_snaps, method, tol)
        self.y = self.refine_axis(self.ymin, self.ymax, self.Ny + 1, self.y_snaps, method, tol)
        self.z = self.refine_axis(self.zmin, self.zmax, self.Nz + 1, self.z_snaps, method, tol)
        self.Nx, self.Ny, self.Nz = len(self.x) - 1, len(self.y) - 1, len(self.z) - 1
        self.dx, self.dy, self.dz = np.diff(self.x), np.diff(self.y), np.diff(self.z)

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
