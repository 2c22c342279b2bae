"""STL reading and node voxelisation (replaces the pyvista calls of
wakis/gridFIT3D.py:475-652 for the solids that feed the material masks).

``voxelize_nodes`` classifies every grid node as inside/outside a closed
triangle surface by ray parity along +z: for each (x_i, y_j) node column the
crossings of the vertical line with the triangles are collected, and a node
at height z is inside when an odd number of crossings lies above it.  The
line is nudged by a tiny irrational offset so that it never passes exactly
through a triangle edge or vertex.
"""

from __future__ import annotations

import struct

import numpy as np


def read_stl(path: str) -> np.ndarray:
    """Triangles of an ASCII or binary STL file as an (n, 3, 3) float64 array."""
    with open(path, "rb") as f:
        data = f.read()
    head = data[:80]
    if len(data) >= 84:
        n = struct.unpack("<I", data[80:84])[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                      ("a", "<u2")]), count=n, offset=84)
            return rec["v"].astype(np.float64)
    if not head.lstrip().lower().startswith(b"solid"):
        raise ValueError(f"{path}: not an STL file")
    verts = []
    for line in data.decode(errors="replace").splitlines():
        t = line.split()
        if len(t) == 4 and t[0] == "vertex":
            verts.append([float(t[1]), float(t[2]), float(t[3])])
    v = np.asarray(verts, dtype=np.float64)
    if len(v) % 3:
        raise ValueError(f"{path}: vertex count is not a multiple of 3")
    return v.reshape(-1, 3, 3)


def _rot(axis: int, deg: float) -> np.ndarray:
    a = np.deg2rad(deg)
    c, s = np.cos(a), np.sin(a)
    if axis == 0:
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == 1:
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def transform_triangles(tri, rotate=(0, 0, 0), translate=(0, 0, 0), scale=1.0):
    """Rotate about x, y, z (degrees, about the origin), translate, then scale
    about the origin, in the order of wakis/gridFIT3D.py:641-650."""
    t = np.asarray(tri, dtype=np.float64).reshape(-1, 3)
    for ax in range(3):
        if rotate[ax]:
            t = t @ _rot(ax, rotate[ax]).T
    t = t + np.asarray(translate, dtype=np.float64)
    t = t * np.asarray(scale, dtype=np.float64)
    return t.reshape(-1, 3, 3)


def voxelize_nodes(tri, x, y, z) -> np.ndarray:
    """Boolean (len(x), len(y), len(z)) array: node inside the closed surface."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    span = max(x[-1] - x[0], y[-1] - y[0], z[-1] - z[0], 1e-30)
    ex, ey = 1.234567e-9 * span, 2.718281e-9 * span
    xs, ys = x + ex, y + ey
    crossings = [[[] for _ in range(len(y))] for _ in range(len(x))]
    p0, p1, p2 = tri[:, 0], tri[:, 1], tri[:, 2]
    for t in range(len(tri)):
        a, b, c = p0[t], p1[t], p2[t]
        det = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1])
        if det == 0.0:
            continue  # triangle parallel to the ray
        lo_x, hi_x = min(a[0], b[0], c[0]), max(a[0], b[0], c[0])
        lo_y, hi_y = min(a[1], b[1], c[1]), max(a[1], b[1], c[1])
        i0, i1 = np.searchsorted(xs, lo_x), np.searchsorted(xs, hi_x, side="right")
        j0, j1 = np.searchsorted(ys, lo_y), np.searchsorted(ys, hi_y, side="right")
        if i0 >= i1 or j0 >= j1:
            continue
        X, Y = np.meshgrid(xs[i0:i1], ys[j0:j1], indexing="ij")
        l1 = ((X - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (Y - a[1])) / det
        l2 = ((b[0] - a[0]) * (Y - a[1]) - (X - a[0]) * (b[1] - a[1])) / det
        l0 = 1.0 - l1 - l2
        hit = (l0 >= 0) & (l1 >= 0) & (l2 >= 0)
        if not hit.any():
            continue
        zc = l0 * a[2] + l1 * b[2] + l2 * c[2]
        for ii, jj in zip(*np.nonzero(hit)):
            crossings[i0 + ii][j0 + jj].append(zc[ii, jj])
    inside = np.zeros((len(x), len(y), len(z)), dtype=bool)
    for i in range(len(x)):
        for j in range(len(y)):
            cz = crossings[i][j]
            if not cz:
                continue
            cz = np.sort(np.asarray(cz))
            above = len(cz) - np.searchsorted(cz, z, side="right")
            inside[i, j, :] = (above % 2) == 1
    return inside
