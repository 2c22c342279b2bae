"""Synthetic cavity workloads (geometry + materials + beam) used by the
benchmark and the tests.

Cells are marked with the reference's STL rule: a cell belongs to a solid if
at least one of its 8 corner nodes lies inside it (wakis/gridFIT3D.py:
502-510).  Shapes are analytic, so no STL file is needed.
"""

from __future__ import annotations

import numpy as np


def cells_from_nodes(inside: np.ndarray) -> np.ndarray:
    """(Nx+1, Ny+1, Nz+1) node flags -> (Nx, Ny, Nz) cells with >= 1 node in."""
    nx, ny, nz = (s - 1 for s in inside.shape)
    c = np.zeros((nx, ny, nz), dtype=bool)
    for a in (0, 1):
        for b in (0, 1):
            for d in (0, 1):
                c |= inside[a:a + nx, b:b + ny, d:d + nz]
    return c


def rect_cavity_mask(x, y, z, half_box, half_len, half_pipe):
    """Square box |x|,|y| <= half_box, |z| <= half_len, plus a square pipe
    |x|,|y| <= half_pipe along the whole z range."""
    bx = (np.abs(x) <= half_box)[:, None, None]
    by = (np.abs(y) <= half_box)[None, :, None]
    bz = (np.abs(z) <= half_len)[None, None, :]
    px = (np.abs(x) <= half_pipe)[:, None, None]
    py = (np.abs(y) <= half_pipe)[None, :, None]
    return cells_from_nodes((bx & by & bz) | (px & py & np.ones_like(bz)))


def pillbox_mask(x, y, z, radius, half_len, pipe_radius):
    """Cylinder r <= radius, |z| <= half_len, plus a round pipe r <= pipe_radius."""
    r2 = x[:, None] ** 2 + y[None, :] ** 2
    cav = (r2 <= radius**2)[:, :, None] & (np.abs(z) <= half_len)[None, None, :]
    pipe = (r2 <= pipe_radius**2)[:, :, None] & np.ones((1, 1, len(z)), dtype=bool)
    return cells_from_nodes(cav | pipe)


def pec_cubic_cavity(nx=64, ny=64, nz=192):
    """PEC box 100x100x100 mm with a 30x30 mm square pipe to z = +-200 mm,
    vacuum inside, PEC outside; bunch q = 1 nC, sigma_z = 25 mm, beta = 1;
    wakelength 3 m; cfl 1.0."""
    x = np.linspace(-0.05, 0.05, nx + 1)
    y = np.linspace(-0.05, 0.05, ny + 1)
    z = np.linspace(-0.2, 0.2, nz + 1)
    mask = rect_cavity_mask(x, y, z, 0.05, 0.05, 0.015)
    return dict(
        name=f"pec_cubic_cavity_{nx}x{ny}x{nz}",
        grid=dict(x=x, y=y, z=z, solids=[(mask, "vacuum")]),
        solver=dict(bc_low=["pec"] * 3, bc_high=["pec"] * 3, bg="pec", use_stl=True, cfln=1.0),
        wake=dict(q=1e-9, sigmaz=25e-3, beta=1.0, wakelength=3.0),
    )


def lossy_pillbox(nx=256, ny=256, nz=768, pml_profile="cubic"):
    """Pillbox cavity r = 60 mm, length 80 mm, beam pipe r = 20 mm, staircased
    in a 120 x 120 mm transverse box with dz = dx (z = +-180 mm at 768 cells).
    Walls are a conductor sigma = 1e5 S/m given as a solid (so, as in the
    reference's STL path with use_sibc, eps = sigma*eps_0 there); 10-cell PML
    on both z ends with a cubic sigma profile peaking at 1 S/m (or the
    reference's geomspace grading, pml_profile="geomspace"), PEC on x and y;
    bunch sigma_z = 10 mm, beta = 1, wakelength 5 m.  CFL factor 0.5 (the
    reference's default): the conductivity term J = sigma*E is explicit
    (wakis/solverFIT3D.py:514-525), and at the CFL limit the leapfrog mode of
    highest frequency then grows (its amplification solves
    l^2 + (2 + a) l + 1 - a = 0 with a = dt*sigma/eps > 0, so |l| > 1)."""
    half_t = 0.06
    dx = 2 * half_t / nx
    half_z = dx * nz / 2
    x = np.linspace(-half_t, half_t, nx + 1)
    y = np.linspace(-half_t, half_t, ny + 1)
    z = np.linspace(-half_z, half_z, nz + 1)
    vac = pillbox_mask(x, y, z, 0.06, 0.04, 0.02)
    return dict(
        name=f"lossy_pillbox_{nx}x{ny}x{nz}",
        grid=dict(x=x, y=y, z=z, solids=[(~vac, [1.0, 1.0, 1e5])]),
        solver=dict(bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"], bg="vacuum",
                    use_stl=True, n_pml=10, cfln=0.5, pml_profile=pml_profile),
        wake=dict(q=1e-9, sigmaz=10e-3, beta=1.0, wakelength=5.0),
    )


def rect_cavity(side, half_pipe, length, dx=1.5e-3, beta=1.0, n_pml=10, sigmaz=10e-3,
                wakelength=1.0, name=None):
    """Rectangular PEC cavity: square box of the given side and length, a
    square beam pipe of half-width half_pipe running one cavity length out
    on each side (total z = 3*length), vacuum inside, PEC elsewhere, grid
    spacing dx in every direction, PEC x/y faces and a graded PML on both z
    ends (CFL factor 0.5: the PML conductivity is explicit, see
    lossy_pillbox); bunch q = 1 nC, sigma_z, beta, wakelength."""
    nxy = max(8, int(round(side / dx)))
    nz = max(16, int(round(3 * length / dx)))
    hx = nxy * dx / 2
    hz = nz * dx / 2
    x = np.linspace(-hx, hx, nxy + 1)
    y = np.linspace(-hx, hx, nxy + 1)
    z = np.linspace(-hz, hz, nz + 1)
    mask = rect_cavity_mask(x, y, z, side / 2, length / 2, half_pipe)
    return dict(
        name=name or f"rect_cavity_{nxy}x{nxy}x{nz}",
        grid=dict(x=x, y=y, z=z, solids=[(mask, "vacuum")]),
        solver=dict(bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"], bg="pec",
                    use_stl=True, n_pml=n_pml, cfln=0.5),
        wake=dict(q=1e-9, sigmaz=sigmaz, beta=float(beta), wakelength=wakelength),
    )


def cavity_batch(n=64, seed=3, dx=1.5e-3, wakelength=1.0, sigmaz=10e-3, n_pml=10):
    """A batch of n rectangular PEC cavities with box side U(60, 140) mm, pipe
    half-width U(10, 25) mm, length U(60, 160) mm and beta from
    {0.8, 0.9, 1.0}, all drawn from one seed; dx = 1.5 mm gives grids from
    about 40 x 40 x 120 to 93 x 93 x 320 cells."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = rng.uniform(0.060, 0.140)
        pipe = rng.uniform(0.010, 0.025)
        length = rng.uniform(0.060, 0.160)
        beta = float(rng.choice([0.8, 0.9, 1.0]))
        out.append(rect_cavity(side, min(pipe, 0.45 * side), length, dx=dx, beta=beta,
                               n_pml=n_pml, sigmaz=sigmaz, wakelength=wakelength,
                               name=f"cavity{i:02d}"))
    return out


def stratified_box(n=1024, seed=5, half=0.256, n_pml=12, sigmaz=5e-3, nlayers=16):
    """Large single solve: an n^3 box of side 2*half (n = 1024: dx = 0.5 mm,
    1.07e9 cells) filled with `nlayers` z-stratified dielectric layers
    (eps_r U(1, 12), sigma U(0, 10) S/m, drawn from `seed`) and a central
    cube of side n/4 whose permittivity is a diagonal tensor (eps_x, eps_y,
    eps_z each U(1, 12), the layer conductivity kept), graded PML of n_pml
    cells on all six faces, Gaussian bunch q = 1 nC, sigma_z, beta = 1 on
    the axis.  Solids are index-slice boxes, so no (n, n, n) mask is built.
    CFL factor 0.5: with a = dt*sigma/eps <= 0.6 here the explicit
    conductivity term stays stable (see lossy_pillbox)."""
    rng = np.random.default_rng(seed)
    x = np.linspace(-half, half, n + 1)
    solids = []
    bounds = np.linspace(0, n, nlayers + 1).round().astype(int)
    for z0, z1 in zip(bounds[:-1], bounds[1:]):
        er = float(rng.uniform(1.0, 12.0))
        sg = float(rng.uniform(0.0, 10.0))
        solids.append(((slice(None), slice(None), slice(int(z0), int(z1))), [er, 1.0, sg]))
    b0, b1 = n // 2 - n // 8, n // 2 + n // 8
    eps3 = tuple(float(v) for v in rng.uniform(1.0, 12.0, 3))
    solids.append(((slice(b0, b1),) * 3, {"eps": eps3, "mu": 1.0}))
    return dict(
        name=f"stratified_box_{n}x{n}x{n}",
        grid=dict(x=x, y=x.copy(), z=x.copy(), solids=solids),
        solver=dict(bc_low=["pml"] * 3, bc_high=["pml"] * 3, bg="vacuum", use_stl=True,
                    use_sibc=False, n_pml=n_pml, cfln=0.5),
        wake=dict(q=1e-9, sigmaz=sigmaz, beta=1.0, wakelength=1.0),
    )


def build(workload: dict, dtype=np.float64, device=0, kernel="auto", save=False,
          precision="reference"):
    """(grid, wake, solver) of the wakis_amd API for a workload dict."""
    from . import GridFIT3D, SolverFIT3D, WakeSolver

    g = GridFIT3D(**workload["grid"])
    w = WakeSolver(save=save, **workload["wake"])
    s = SolverFIT3D(g, w, dtype=dtype, device=device, kernel=kernel, precision=precision,
                    **workload["solver"])
    return g, w, s
