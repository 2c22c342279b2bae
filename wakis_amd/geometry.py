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


def merge_points(tri):
    """Exactly coincident vertices merged: (points (m, 3), triangles (n, 3)
    of point ids), triangles with a repeated id dropped (the STL reader's
    point merging that the edge topology below relies on)."""
    v = np.asarray(tri, dtype=np.float64).reshape(-1, 3)
    pts, ids = np.unique(v, axis=0, return_inverse=True)
    ids = ids.reshape(-1, 3)
    ok = (ids[:, 0] != ids[:, 1]) & (ids[:, 1] != ids[:, 2]) & (ids[:, 0] != ids[:, 2])
    return pts, ids[ok]


def feature_edge_points(tri, feature_angle=30.0):
    """Points of the surface's boundary edges (one triangle), non-manifold
    edges (more than two) and feature edges (two triangles whose unit normals
    n1, n2 have n1.n2 <= cos(feature_angle)); manifold edges are left out.
    This is what wakis/gridFIT3D.py:681 asks of pyvista's
    ``extract_feature_edges(boundary_edges=True, manifold_edges=False)``."""
    pts, ids = merge_points(tri)
    if len(ids) == 0:
        return np.zeros((0, 3))
    p0, p1, p2 = pts[ids[:, 0]], pts[ids[:, 1]], pts[ids[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    ln = np.linalg.norm(nrm, axis=1)
    nrm = np.divide(nrm, ln[:, None], out=np.zeros_like(nrm), where=ln[:, None] > 0)
    e = np.concatenate([ids[:, [0, 1]], ids[:, [1, 2]], ids[:, [2, 0]]])
    owner = np.tile(np.arange(len(ids)), 3)
    e = np.sort(e, axis=1)
    key, inv, cnt = np.unique(e, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    keep = (cnt == 1) | (cnt > 2)
    two = np.nonzero(cnt == 2)[0]
    if len(two):
        order = np.argsort(inv, kind="stable")
        first = np.searchsorted(inv[order], two)
        ta, tb = owner[order[first]], owner[order[first + 1]]
        dot = np.einsum("ij,ij->i", nrm[ta], nrm[tb])
        keep[two] = dot <= np.cos(np.deg2rad(feature_angle))
    used = np.unique(key[keep].reshape(-1))
    return pts[used]


def _closest_dist2(P, a, b, c):
    """Squared distance of points P (m, 3) to the triangle (a, b, c), by the
    Voronoi-region method for the closest point on a triangle (Ericson,
    Real-Time Collision Detection, 5.1.5)."""
    ab, ac = b - a, c - a
    ap, bp, cp = P - a, P - b, P - c
    d1, d2 = ap @ ab, ap @ ac
    d3, d4 = bp @ ab, bp @ ac
    d5, d6 = cp @ ab, cp @ ac
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2
    with np.errstate(divide="ignore", invalid="ignore"):
        den = va + vb + vc
        v, w = vb / den, vc / den
        q = a + v[:, None] * ab + w[:, None] * ac
        # regions in increasing priority: edge BC, edge AC, vertex C, edge AB,
        # vertex B, vertex A
        m = (va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0)
        t = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        q = np.where(m[:, None], b + t[:, None] * (c - b), q)
        m = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
        t = d2 / (d2 - d6)
        q = np.where(m[:, None], a + t[:, None] * ac, q)
        m = (d6 >= 0) & (d5 <= d6)
        q = np.where(m[:, None], c, q)
        m = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
        t = d1 / (d1 - d3)
        q = np.where(m[:, None], a + t[:, None] * ab, q)
        m = (d3 >= 0) & (d4 <= d3)
        q = np.where(m[:, None], b, q)
        m = (d1 <= 0) & (d2 <= 0)
        q = np.where(m[:, None], a, q)
    r = P - q
    return np.einsum("ij,ij->i", r, r)


def node_distance(tri, x, y, z, band):
    """Unsigned distance of every node of the (x, y, z) tensor grid to the
    triangle surface, exact where it is below `band` and +inf elsewhere."""
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    d2 = np.full((len(x), len(y), len(z)), np.inf)
    for a, b, c in np.asarray(tri, dtype=np.float64):
        lo = np.minimum(np.minimum(a, b), c) - band
        hi = np.maximum(np.maximum(a, b), c) + band
        i0, i1 = np.searchsorted(x, lo[0]), np.searchsorted(x, hi[0], side="right")
        j0, j1 = np.searchsorted(y, lo[1]), np.searchsorted(y, hi[1], side="right")
        k0, k1 = np.searchsorted(z, lo[2]), np.searchsorted(z, hi[2], side="right")
        if i0 >= i1 or j0 >= j1 or k0 >= k1:
            continue
        X, Y, Z = np.meshgrid(x[i0:i1], y[j0:j1], z[k0:k1], indexing="ij")
        P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        dd = _closest_dist2(P, a, b, c).reshape(X.shape)
        blk = d2[i0:i1, j0:j1, k0:k1]
        np.minimum(blk, dd, out=blk)
    d = np.sqrt(d2)
    d[d > band] = np.inf
    return d
