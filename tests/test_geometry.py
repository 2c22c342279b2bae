"""Geometry builder on the CPU: STL feature edges, point-triangle distance,
the STL cell rules, mesh refinement (the grid assertions of the reference's
tests/test_011_smart_mesh.py:105-147) and the run log."""

import json
import os

import numpy as np
import pytest

from helpers import write_ascii_stl

HERE = os.path.dirname(os.path.abspath(__file__))


def _box_triangles(lo, hi):
    """A closed axis-aligned box, two triangles per face, outward normals."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1],
                   lo[2] if i & 4 == 0 else hi[2]] for i in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tri = []
    for a, b, cc, d in quads:
        tri += [[c[a], c[b], c[cc]], [c[a], c[cc], c[d]]]
    return np.array(tri)


def _cavity_stl(tmp_path):
    geo = json.load(open(os.path.join(HERE, "golden", "001_cubic_cavity_box.json")))
    path = str(tmp_path / "001_cubic_cavity.stl")
    write_ascii_stl(path, geo["triangles"])
    return path, geo


def test_feature_edges_box_and_open_surfaces():
    from wakis_amd.geometry import feature_edge_points

    box = _box_triangles((0, 0, 0), (1, 2, 3))
    p = feature_edge_points(box)
    # the face diagonals are coplanar (manifold, 0 degrees): only the corners
    assert len(p) == 8
    assert set(map(tuple, p)) == {(x, y, z) for x in (0, 1) for y in (0, 2) for z in (0, 3)}
    # one open triangle: three boundary edges
    one = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], float)
    assert len(feature_edge_points(one)) == 3
    # a flat fan of 4 triangles around (0,0,0): only the rim is kept
    rim = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0]], float)
    fan = np.array([[[0, 0, 0], rim[i], rim[(i + 1) % 4]] for i in range(4)])
    q = feature_edge_points(fan)
    assert len(q) == 4 and not np.any(np.all(q == 0, axis=1))
    # a fold of 20 degrees is below the 30-degree feature angle, 40 is above
    for deg, n in ((20.0, 4), (40.0, 4 + 0)):
        a = np.deg2rad(deg)
        t = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]],
                      [[1, 0, 0], [0, 0, 0], [0, -np.cos(a), np.sin(a)]]])
        pts = feature_edge_points(t)
        shared = np.any(np.all(pts == [0, 0, 0], axis=1)) and np.any(np.all(pts == [1, 0, 0], axis=1))
        assert shared  # boundary edges always contain both ends of the fold
        assert len(pts) == n
    # three triangles on one edge: non-manifold, its ends are kept with the rest
    t3 = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 0], [1, 0, 0], [0, -1, 0]],
                   [[0, 0, 0], [1, 0, 0], [0, 0, 1]]])
    assert len(feature_edge_points(t3)) == 5


def test_cavity_feature_points_are_its_corners(tmp_path):
    from wakis_amd.geometry import feature_edge_points, read_stl

    path, geo = _cavity_stl(tmp_path)
    p = feature_edge_points(read_stl(path))
    # box (+-25 mm, z +-15 mm) and pipe (+-7.5 mm to z +-50 mm): 8 + 16 corners
    assert len(p) == 24
    assert sorted(np.unique(np.round(p[:, 2], 6))) == [-0.05, -0.015, 0.015, 0.05]


def test_point_triangle_distance_against_sampling():
    from wakis_amd.geometry import _closest_dist2

    rng = np.random.default_rng(3)
    n = 300
    u, v = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    keep = u + v <= 1
    for _ in range(4):
        a, b, c = rng.normal(size=(3, 3))
        P = rng.normal(size=(500, 3)) * 2
        d = np.sqrt(_closest_dist2(P, a, b, c))
        S = a + u[keep][:, None] * (b - a) + v[keep][:, None] * (c - a)
        bf = np.sqrt(np.min(((P[:, None, :] - S[None]) ** 2).sum(-1), axis=1))
        h = max(np.linalg.norm(b - a), np.linalg.norm(c - a)) / (n - 1)
        assert np.all(d <= bf + 1e-12)  # never above a point of the triangle
        assert np.all(bf - d <= h)      # and within the sampling step of the minimum


def test_stl_cell_rules_on_a_box(tmp_path):
    """Every rule against closed forms for an axis-aligned box: corner counts
    from the node coordinates, and the signed distance of a box."""
    import wakis_amd as wa

    lo, hi = np.array([-0.31, -0.22, -0.405]), np.array([0.27, 0.33, 0.18])
    path = str(tmp_path / "box.stl")
    write_ascii_stl(path, _box_triangles(lo, hi))
    N = (13, 11, 17)
    out = {}
    for method in ("legacy", "interior_points", "implicit_distance", "voxelize_rectilinear"):
        g = wa.GridFIT3D(-0.5, 0.5, -0.5, 0.5, -0.6, 0.6, *N, stl_solids={"b": path},
                         stl_materials={"b": "vacuum"}, stl_method=method)
        out[method] = g.solids[0][1]
        assert g.logger.grid["Nx"] == N[0]
    X, Y, Z = np.meshgrid(g.x, g.y, g.z, indexing="ij")
    P = np.stack([X, Y, Z], -1)
    inside = np.all((P > lo) & (P < hi), axis=-1)
    cnt = sum(inside[a:a + N[0], b:b + N[1], c:c + N[2]].astype(int)
              for a in (0, 1) for b in (0, 1) for c in (0, 1))
    assert np.array_equal(out["legacy"], cnt > 0)
    assert np.array_equal(out["interior_points"], cnt > 4)
    q = np.abs(P - (lo + hi) / 2) - (hi - lo) / 2
    sd = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)
    msd = sum(sd[a:a + N[0], b:b + N[1], c:c + N[2]]
              for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8
    thr = 0.5 * np.mean([np.mean(g.dx), np.mean(g.dy), np.mean(g.dz)])
    assert np.array_equal(out["implicit_distance"], msd <= thr)
    lowc = P[:-1, :-1, :-1]
    assert np.array_equal(out["voxelize_rectilinear"],
                          np.all((lowc > lo) & (lowc < hi), axis=-1))
    with pytest.raises(ValueError):
        wa.GridFIT3D(-0.5, 0.5, -0.5, 0.5, -0.6, 0.6, *N, stl_solids={"b": path},
                     stl_materials={"b": "vacuum"}, stl_method="nope")


def test_surface_layer_mask():
    import wakis_amd as wa

    m = np.zeros((8, 7, 9), bool)
    m[2:6, 1:5, 3:8] = True
    g = wa.GridFIT3D(0, 1, 0, 1, 0, 1, 8, 7, 9, solids=[(m, "vacuum")])
    edge = g._mark_cells_in_surface("solid0")
    ref = np.zeros_like(m)
    for i in range(8):
        for j in range(7):
            for k in range(9):
                for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    a, b, c = i + d[0], j + d[1], k + d[2]
                    if 0 <= a < 8 and 0 <= b < 7 and 0 <= c < 9 and m[a, b, c] != m[i, j, k]:
                        ref[i, j, k] = True
    assert np.array_equal(edge, ref)
    assert g.solids[0][1] is edge


@pytest.fixture(scope="module")
def smart_grid(tmp_path_factory):
    import wakis_amd as wa

    path, geo = _cavity_stl(tmp_path_factory.mktemp("stl"))
    lo, hi = geo["bounds_lo"], geo["bounds_hi"]
    return wa.GridFIT3D(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 50, 50, 80,
                        stl_solids={"cavity": path}, stl_materials={"cavity": "vacuum"},
                        use_mesh_refinement=True, snap_tol=1e-2, refinement_tol=1e-8)


def test_smart_mesh_grid_generation(smart_grid):
    """tests/test_011_smart_mesh.py:105-147 on the refined cubic-cavity grid."""
    grid = smart_grid
    edges = {"z": [-0.05, -0.0075, 0.0075, 0.05], "y": [-0.025, -0.005, 0.005, 0.025],
             "x": [-0.025, -0.005, 0.005, 0.025]}
    for ax, ed in edges.items():
        a = getattr(grid, ax)
        for e in ed:
            assert np.min(np.abs(a - e)) <= 1e-8 + 1e-2
    for ax in "xyz":
        d = getattr(grid, "d" + ax)
        assert np.all(np.abs(np.diff(d)) <= 0.1 * d[:-1])
        n = getattr(grid, "N" + ax)
        assert np.min(d) > 0.5 * (getattr(grid, ax + "max") - getattr(grid, ax + "min")) / n
    # the snap planes of the solid (pipe walls, box walls in z) are grid nodes
    # up to the optimiser's soft constraint
    for ax, snaps in (("x", [-0.0075, 0.0075]), ("y", [-0.0075, 0.0075]),
                      ("z", [-0.05, -0.015, 0.015, 0.05])):
        a = getattr(grid, ax)
        for s in snaps:
            assert np.min(np.abs(a - s)) < 1e-4
    # "insert" adds the 2 interior snaps to Nx + 1 - 4 uniform nodes
    assert (grid.Nx, grid.Ny, grid.Nz) == (48, 48, 78)
    assert grid.logger.grid["use_mesh_refinement"] is True
    assert grid.solids[0][1].shape == (48, 48, 78)


@pytest.mark.parametrize("method", ["neighbor", "subdivision"])
def test_refine_axis_methods(method):
    import wakis_amd as wa

    g = wa.GridFIT3D(0, 1, 0, 1, 0, 1, 4, 4, 4)
    snaps = np.array([-0.2, -0.0333, 0.051, 0.3])
    a = g.refine_axis(-0.2, 0.3, 41, snaps, method=method, tol=1e-8)
    assert len(a) == 41 and np.all(np.diff(a) > 0)
    for s in snaps:
        assert np.min(np.abs(a - s)) < 2e-3
    with pytest.raises(ValueError):
        g.refine_axis(-0.2, 0.3, 41, snaps, method="other")


def test_run_log_layout(tmp_path):
    """wakis.log: title, then WakeSolver, Solver and Grid sections as JSON
    (wakis/logger.py:41-97), numpy values converted."""
    from wakis_amd.logger import Logger

    lg = Logger()
    lg.grid = {"Nx": 4, "dx": np.array([0.5, 0.25]), "n": np.float32(1.5)}
    lg.wakeSolver = {"q": 1e-9}
    lg.save_logs(str(tmp_path / "res"))
    text = open(tmp_path / "res" / "wakis.log", encoding="utf-8").read()
    assert text.startswith("Simulation Parameters\n=====================\n\n")
    parts = text.split("\n## ")
    assert [p.split(" ##")[0] for p in parts[1:]] == ["WakeSolver Logs", "Solver Logs", "Grid Logs"]
    assert parts[2].split("\n", 1)[1] == "(empty)\n"
    wk = json.loads(parts[1].split("\n", 1)[1])
    assert wk == {"q": 1e-9, "results_folder": str(tmp_path / "res")}
    gr = json.loads(parts[3].split("\n", 1)[1])
    assert gr == {"Nx": 4, "dx": [0.5, 0.25], "n": "1.5"}


def test_solver_and_wake_logs_host_setup():
    import wakis_amd as wa

    g = wa.GridFIT3D(0, 1, 0, 1, 0, 1, 6, 5, 7)
    w = wa.WakeSolver(q=2e-9, sigmaz=0.1, save=False)
    s = wa.SolverFIT3D(g, w, bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"],
                       device=None, n_pml=2)
    assert s.logger.grid is g.logger.grid
    assert s.logger.wakeSolver is w.logger.wakeSolver
    assert w.logger.wakeSolver["q"] == 2e-9 and w.logger.wakeSolver["results_folder"] == "results/"
    for k in ("use_gpu", "use_mpi", "background", "bc_low", "bc_high", "n_pml", "dt",
              "solverInitializationTime"):
        assert k in s.logger.solver
    # logged as the CFL step, before the relaxation clamp (solverFIT3D.py:306-318)
    from scipy.constants import c as c_light
    cfl = 0.5 / (c_light * np.sqrt(1 / np.min(g.dx) ** 2 + 1 / np.min(g.dy) ** 2
                                   + 1 / np.min(g.dz) ** 2))
    assert s.logger.solver["dt"] == pytest.approx(cfl, rel=1e-12)
    assert s.dt < s.logger.solver["dt"]
