"""Wake path on the device vs the CPU oracle: in-kernel bunch current, E_z
probe, wake potential (bit-exact given the same history), impedance DFT,
transverse wake, and the reference's test_001 golden vectors."""

import json
import os

import numpy as np
import pytest

from helpers import OGrid, OSolver, OWake, block_mask, make_pair, rel, write_ascii_stl
from oracle.fit_oracle import _wp_accumulate

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _cavity(shape):
    Nx, Ny, Nz = shape
    cav = block_mask(shape, (2, 2, Nz // 3), (Nx - 2, Ny - 2, 2 * Nz // 3))
    pipe = block_mask(shape, (Nx // 2 - 2, Ny // 2 - 2, 0), (Nx // 2 + 2, Ny // 2 + 2, Nz))
    return [(cav | pipe, "vacuum")]


WAKE_KW = dict(wakelength=0.15, q=1e-9, sigmaz=0.012, beta=1.0, skip_cells=2)
# the reference's 1-window path reads the whole E_z line into a skip_cells
# sliced array and fails for skip_cells != 0 (wakeSolver.py:370-389)
WAKE_KW_LONG = dict(WAKE_KW, skip_cells=0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("plane", ["both", "longitudinal"])
def test_wakesolve(dtype, plane):
    shape = (14, 13, 36)
    print(f"[case] wakesolve {np.dtype(dtype).name} {plane} shape={shape}", flush=True)
    kw = WAKE_KW if plane == "both" else WAKE_KW_LONG
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, dtype=dtype, bg="pec",
                         solids=_cavity(shape), init=False, cfln=1.0, wake_kw=kw)
    osv.wakesolve(compute_plane=plane)
    psv.wakesolve(compute_plane=plane)
    ow, pw = osv.wake, psv.wake
    assert psv.Nt == osv.Nt and len(pw.s) == len(ow.s) and len(pw.Z) == len(ow.Z)
    tol = 1e-10 if dtype == np.float64 else 2e-4
    errs = {"E": rel(psv.E.toarray(), osv.E.toarray()), "H": rel(psv.H.toarray(), osv.H.toarray()),
            "WP": rel(pw.WP, ow.WP), "Z": rel(pw.Z, ow.Z)}
    if plane == "both":
        errs["WP3d"] = rel(pw.WP_3d, ow.WP_3d)
        errs["WPx"] = rel(pw.WPx, ow.WPx)
        errs["WPy"] = rel(pw.WPy, ow.WPy)
        errs["Zx"] = rel(pw.Zx, ow.Zx)
    print(errs)
    assert max(errs.values()) <= tol, errs
    # J after the run holds the last beam line (non-conductive solver)
    assert rel(psv.J.toarray(), osv.J.toarray()) <= tol


def test_wake_potential_bit_exact():
    """Given the same E_z history, the device WP equals the oracle's
    accumulation bit for bit (same order of additions, no contraction)."""
    shape = (12, 12, 30)
    print("[case] WP bit-exact", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg="pec", solids=_cavity(shape),
                         init=False, cfln=1.0, wake_kw=WAKE_KW)
    psv.wakesolve(compute_plane="both")
    pw = psv.wake
    hist = psv.ctx.wake_history()  # (9, nz, nt)
    # feed the device history to the oracle's restated WP loop
    ow = OWake(**WAKE_KW)
    hf = {"x": pw.xf, "y": pw.yf, "z": pw.zf, "dx": pw.dx, "dy": pw.dy, "dz": pw.dz, "t": pw.t}
    for n in range(psv.Nt):
        hf["#" + str(n).zfill(5)] = hist[:, :, n].reshape(3, 3, -1)
    ow.Ez_hf = hf
    ow.calc_long_WP_3d()
    assert np.array_equal(ow.WP_3d, pw.WP_3d)


def test_longitudinal_skip_cells_fails_like_reference():
    shape = (10, 10, 24)
    print("[case] longitudinal skip_cells", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg="pec", solids=_cavity(shape),
                         init=False, cfln=1.0, wake_kw=WAKE_KW)
    with pytest.raises(ValueError):
        osv.wakesolve(compute_plane="longitudinal")
    with pytest.raises(ValueError):
        psv.wakesolve(compute_plane="longitudinal")


def test_wakesolve_callback_path_matches_fast_path():
    shape = (10, 10, 24)
    print("[case] callback vs fast path", flush=True)
    _, a = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg="pec", solids=_cavity(shape),
                     init=False, cfln=1.0, wake_kw=WAKE_KW_LONG)
    _, b = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg="pec", solids=_cavity(shape),
                     init=False, cfln=1.0, wake_kw=WAKE_KW_LONG)
    a.wakesolve(compute_plane="longitudinal")
    seen = []
    b.wakesolve(compute_plane="longitudinal", callback=lambda s, t: seen.append(t))
    assert len(seen) == b.Nt
    assert rel(b.wake.WP, a.wake.WP) <= 1e-12
    assert rel(b.E.toarray(), a.E.toarray()) <= 1e-12


def test_conductive_wakesolve_pml_fp64():
    """Lossy walls + z PML (config-2 class, small): sigma*E J mode, PML
    grading and the relaxation dt clamp."""
    shape = (12, 12, 40)
    print("[case] conductive pml wakesolve", flush=True)
    Nx, Ny, Nz = shape
    wall = ~block_mask(shape, (2, 2, 0), (Nx - 2, Ny - 2, Nz))
    solids = [(wall, [1.0, 1.0, 1e3])]
    osv, psv = make_pair(*shape, ["pec", "pec", "pml"], ["pec", "pec", "pml"], bg=[1.0, 1.0],
                         solids=solids, init=False, cfln=0.9, n_pml=5, wake_kw=WAKE_KW_LONG)
    assert float(psv.dt) == float(osv.dt)
    osv.wakesolve(compute_plane="longitudinal")
    psv.wakesolve(compute_plane="longitudinal")
    errs = {"E": rel(psv.E.toarray(), osv.E.toarray()), "WP": rel(psv.wake.WP, osv.wake.WP),
            "Z": rel(psv.wake.Z, osv.wake.Z), "J": rel(psv.J.toarray(), osv.J.toarray())}
    print(errs)
    assert max(errs.values()) <= 1e-10, errs


def test_golden_001_fp32():
    """tests/test_001_pec_cubic_cavity.py goldens on the device, FP32."""
    import wakis_amd as wa

    print("[case] golden 001 fp32", flush=True)
    g = json.load(open(os.path.join(HERE, "golden", "goldens.json")))["test_001"]
    stl = os.path.join(HERE, "golden", "001_cubic_cavity_box.json")
    geo = json.load(open(stl))
    lo, hi = geo["bounds_lo"], geo["bounds_hi"]
    grid = wa.GridFIT3D(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 50, 50, 150,
                        solids=[(_cubic_cavity_mask(geo), "vacuum")])
    wake = wa.WakeSolver(wakelength=1.0, q=1e-9, sigmaz=18.5e-3, beta=1.0, skip_cells=12,
                         save=False)
    solver = wa.SolverFIT3D(grid, wake, bc_low=["pec"] * 3, bc_high=["pec"] * 3, use_stl=True,
                            bg="pec", dtype=np.float32)
    solver.wakesolve(wakelength=1.0)
    WP, Z = wake.WP, wake.Z
    gz = np.array(g["Z_every20_re"]) + 1j * np.array(g["Z_every20_im"])
    assert np.allclose(WP[::50], g["WP_every50"], rtol=g["tol_wp"], atol=g["tol_wp"])
    assert np.cumsum(np.abs(WP))[-1] == pytest.approx(g["WP_abs_cumsum"], g["cumsum_rel"])
    for part in (np.abs, np.real, np.imag):
        assert np.allclose(part(Z)[::20], part(gz), rtol=g["tol_z"], atol=g["tol_z"])
    assert np.cumsum(np.abs(Z))[-1] == pytest.approx(g["Z_abs_cumsum"], g["cumsum_rel"])


def _cubic_cavity_mask(geo):
    from wakis_amd.geometry import voxelize_nodes

    x = np.linspace(geo["bounds_lo"][0], geo["bounds_hi"][0], 51)
    y = np.linspace(geo["bounds_lo"][1], geo["bounds_hi"][1], 51)
    z = np.linspace(geo["bounds_lo"][2], geo["bounds_hi"][2], 151)
    tri = np.asarray(geo["triangles"], dtype=np.float64)
    inside = voxelize_nodes(tri, x, y, z)
    c = np.zeros((50, 50, 150), bool)
    for a in (0, 1):
        for b in (0, 1):
            for d in (0, 1):
                c |= inside[a:a + 50, b:b + 50, d:d + 150]
    return c


def _smart_grid(tmp_path):
    """The reference's tests/test_011_smart_mesh.py grid: the cubic cavity
    STL on 50 x 50 x 80 cells refined around its snap points."""
    import wakis_amd as wa

    geo = json.load(open(os.path.join(HERE, "golden", "001_cubic_cavity_box.json")))
    path = str(tmp_path / "001_cubic_cavity.stl")
    write_ascii_stl(path, geo["triangles"])
    lo, hi = geo["bounds_lo"], geo["bounds_hi"]
    return wa.GridFIT3D(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 50, 50, 80,
                        stl_solids={"cavity": path}, stl_materials={"cavity": "vacuum"},
                        use_mesh_refinement=True, snap_tol=1e-2, refinement_tol=1e-8)


def test_smart_mesh_wakesolve_fp32(tmp_path):
    """tests/test_011_smart_mesh.py:21-103, 149-153: FP32 wakesolve on the
    refined non-uniform grid, no charge accumulation in Z (DC bin below 1 %
    of the peak); the run log lands in the results folder."""
    import wakis_amd as wa

    grid = _smart_grid(tmp_path)
    print(f"[case] smart mesh wakesolve fp32 shape={(grid.Nx, grid.Ny, grid.Nz)}", flush=True)
    folder = str(tmp_path / "results")
    wake = wa.WakeSolver(q=1e-9, sigmaz=18.5e-3, beta=1.0, xsource=0.0, ysource=0.0,
                         xtest=0.0, ytest=0.0, save=False, skip_cells=12, results_folder=folder)
    solver = wa.SolverFIT3D(grid, wake, bc_low=["pec"] * 3, bc_high=["pec"] * 3, use_stl=True,
                            bg="pec", dtype=np.float32)
    solver.wakesolve(wakelength=1.0, save_J=False)
    Z = wake.Z
    assert np.all(np.isfinite(Z)) and np.all(np.isfinite(wake.WP))
    assert np.abs(Z[0]) < 0.01 * np.abs(np.max(Z))
    log = open(os.path.join(folder, "wakis.log"), encoding="utf-8").read()
    assert '"wakelength": 1.0' in log and '"use_mesh_refinement": true' in log


def test_smart_mesh_parity_fp64(tmp_path):
    """Device vs oracle on the refined grid with its STL mask (FP64, short wake)."""
    grid = _smart_grid(tmp_path)
    import wakis_amd as wa

    shape = (grid.Nx, grid.Ny, grid.Nz)
    print(f"[case] smart mesh parity fp64 shape={shape}", flush=True)
    mask = grid.solids[0][1]
    kw = dict(wakelength=0.05, q=1e-9, sigmaz=18.5e-3, beta=1.0, skip_cells=12)
    og = OGrid(x=grid.x, y=grid.y, z=grid.z)
    osv = OSolver(og, OWake(**kw), bc_low=["pec"] * 3, bc_high=["pec"] * 3,
                  solids=[(mask, "vacuum")], bg="pec", dtype=np.float64)
    pg = wa.GridFIT3D(x=grid.x, y=grid.y, z=grid.z, solids=[(mask, "vacuum")])
    psv = wa.SolverFIT3D(pg, wa.WakeSolver(save=False, results_folder=str(tmp_path / "r"), **kw),
                         bc_low=["pec"] * 3, bc_high=["pec"] * 3, use_stl=True, bg="pec",
                         dtype=np.float64)
    osv.wakesolve(compute_plane="both")
    psv.wakesolve(compute_plane="both")
    errs = {"E": rel(psv.E.toarray(), osv.E.toarray()), "H": rel(psv.H.toarray(), osv.H.toarray()),
            "WP": rel(psv.wake.WP, osv.wake.WP), "Z": rel(psv.wake.Z, osv.wake.Z),
            "WPx": rel(psv.wake.WPx, osv.wake.WPx)}
    print(errs)
    assert max(errs.values()) <= 1e-10, errs


def test_config1_pec_cubic_cavity_fp64():
    """The reference's own CPU case in full on the device, FP64: the PEC cubic
    cavity (64 x 64 x 192, CFL 1.0, wakelength 3 m, both planes) against the
    oracle's full run on the same inputs (tests/golden/make_config1_golden.py
    wrote the fixture)."""
    from wakis_amd.workloads import build, pec_cubic_cavity

    g = np.load(os.path.join(HERE, "golden", "config1_pec_cubic_cavity_fp64.npz"))
    print("[case] config 1 cubic cavity fp64 full wakesolve", flush=True)
    _, wake, solver = build(pec_cubic_cavity(), dtype=np.float64)
    solver.wakesolve(compute_plane="both")
    assert solver.Nt == int(g["Nt"]) and float(solver.dt) == float(g["dt"])
    assert len(wake.s) == len(g["s"]) and len(wake.Z) == len(g["Z"])
    errs = {k: rel(getattr(wake, k), g[k]) for k in ("s", "WP", "WPx", "WPy", "f", "Z", "Zx", "Zy")}
    stride = int(g["sample"])
    E, H = solver.E.toarray(), solver.H.toarray()
    errs["E_sample"] = rel(E[::stride], g["E_sample"])
    errs["H_sample"] = rel(H[::stride], g["H_sample"])
    errs["E_norm"] = abs(np.linalg.norm(E) / float(g["E_norm"]) - 1)
    errs["H_norm"] = abs(np.linalg.norm(H) / float(g["H_norm"]) - 1)
    print(errs)
    assert max(errs.values()) <= 1e-10, errs


def _sphere_case(dtype):
    """tests/test_003_planewave_sphere.py:19-95: a dielectric sphere (eps_r
    10, the reference's tests/stl/003_sphere.stl scaled by 1e-3) on 60 x 60 x
    120 cells, periodic x/y, PEC z low, PML z high, a plane wave from z = 1."""
    import wakis_amd as wa
    from wakis_amd.geometry import read_stl, transform_triangles

    stl = os.path.join(HERE, "golden", "003_sphere.stl")
    tri = transform_triangles(read_stl(stl), (0, 0, 0), (0, 0, 0), 1e-3).reshape(-1, 3)
    lo, hi = tri.min(0), tri.max(0)
    pad = (hi - lo) * np.array([0.2, 0.2, 1.0])
    lo, hi = lo - pad, hi + pad
    Nx, Ny, Nz = 60, 60, 120
    grid = wa.GridFIT3D(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], Nx, Ny, Nz,
                        stl_solids={"Sphere": stl}, stl_rotate=[0, 0, 0], stl_scale=1e-3,
                        stl_materials={"Sphere": [10.0, 1.0]})
    bl, bh = ["periodic", "periodic", "pec"], ["periodic", "periodic", "pml"]
    psv = wa.SolverFIT3D(grid, use_stl=True, bc_low=bl, bc_high=bh, dtype=dtype)
    mask = grid.solids[0][1]
    osv = OSolver(OGrid(x=grid.x, y=grid.y, z=grid.z), bc_low=bl, bc_high=bh,
                  solids=[(mask, [10.0, 1.0])], dtype=dtype)
    return grid, psv, osv


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
def test_planewave_sphere_emsolve(dtype):
    """The reference's test_003 simulation on the device against the oracle
    (fields after the whole emsolve run)."""
    import wakis_amd as wa
    from helpers import OPlaneWave

    grid, psv, osv = _sphere_case(dtype)
    Nx, Ny = grid.Nx, grid.Ny
    assert 0 < grid.solids[0][1].sum() < grid.solids[0][1].size
    f = 15 / ((psv.z.max() - psv.z.min()) / 299792458.0)
    Nt = int(1.0 * (psv.z.max() - psv.z.min()) / 299792458.0 / psv.dt)
    print(f"[case] planewave sphere {np.dtype(dtype).name} Nt={Nt}", flush=True)
    assert float(psv.dt) == float(osv.dt)
    osv.emsolve(Nt, OPlaneWave(xs=slice(1, Nx - 1), ys=slice(1, Ny - 1), zs=1, f=f, beta=1.0))
    psv.emsolve(Nt, wa.PlaneWave(xs=slice(1, Nx - 1), ys=slice(1, Ny - 1), zs=1, f=f, beta=1.0))
    errs = {n: rel(getattr(psv, n).toarray(), getattr(osv, n).toarray()) for n in "EHJ"}
    print(errs)
    assert max(errs.values()) <= (1e-10 if dtype == np.float64 else 1e-4), errs


def test_te101_resonator_device():
    """tests/test_002_analytic_resonator.py:83-145 on the device: the TE101
    mode of a 1 m PEC cube on 30^3 cells (FP64), E_y at the centre read
    after every step for 5000 steps, rings at c/sqrt(2) = 2.121e8 Hz within
    5 %; the same series as the oracle's within 1e-10."""
    import wakis_amd as wa
    from test_oracle import te101_initial_h

    L, N, steps = 1.0, 30, 5000
    d = L / N
    lo, hi = -L / 2 + d / 2, L / 2 + d / 2
    print("[case] TE101 resonator on the device", flush=True)
    psv = wa.SolverFIT3D(wa.GridFIT3D(lo, hi, lo, hi, lo, hi, N, N, N),
                         bc_low=["pec"] * 3, bc_high=["pec"] * 3)
    osv = OSolver(OGrid(lo, hi, lo, hi, lo, hi, N, N, N), bc_low=["pec"] * 3,
                  bc_high=["pec"] * 3)
    hx, hy, hz = te101_initial_h(N, L)
    for s in (psv, osv):
        s.H.from_matrix(hx, "x")
        s.H.from_matrix(hy, "y")
        s.H.from_matrix(hz, "z")
    ep, eo = [], []
    for _ in range(steps):
        ep.append(psv.E[N // 2, N // 2, N // 2, "y"])
        eo.append(osv.E[N // 2, N // 2, N // 2, "y"])
        psv.one_step()
        osv.one_step()
    ep, eo = np.array(ep, dtype=np.float64), np.array(eo, dtype=np.float64)
    assert rel(ep, eo) <= 1e-10
    F = np.abs(np.fft.fft(ep))
    f = np.fft.fftfreq(len(ep), d=float(psv.dt))
    half = f >= 0
    assert f[half][np.argmax(F[half])] == pytest.approx(2.121e8, rel=0.05)


@pytest.mark.parametrize("profile", ["geomspace", "cubic"])
@pytest.mark.parametrize("eps_r,sigma", [(1.0, 0.0), (3.0, 1.0), (1.0, 1.0)],
                         ids=["vacuum", "resistive", "high-resistivity"])
def test_pml_reflection_device(eps_r, sigma, profile):
    """tests/test_005_pml_boundaries.py:19-336 on the device (FP32): a plane
    wave of amplitude 100 run into a 30-cell z PML through a background of
    (eps_r, sigma); after the round trip the field left on the axis is below
    10 % of the amplitude, for H and for E / (mu_0 c).  Both sigma gradings:
    the reference's geomspace and the cubic profile (1 S/m peak)."""
    import wakis_amd as wa
    from scipy.constants import c, mu_0

    Nx, Ny, Nz = 20, 20, 200
    print(f"[case] PML reflection eps_r={eps_r} sigma={sigma} {profile}", flush=True)
    grid = wa.GridFIT3D(-1.0, 1.0, -1.0, 1.0, 0.0, 1.0, Nx, Ny, Nz)
    solver = wa.SolverFIT3D(grid, use_stl=False, bg=[eps_r, 1.0, sigma],
                            bc_low=["periodic", "periodic", "pec"],
                            bc_high=["periodic", "periodic", "pml"], n_pml=30, dtype=np.float32,
                            pml_profile=profile)
    amplitude = 100.0
    f = 15 / ((solver.z.max() - solver.z.min()) / c)
    pw = wa.PlaneWave(xs=slice(0, Nx), ys=slice(0, Ny), zs=0, beta=1.0, amplitude=amplitude,
                      f=f, nodes=7, phase=np.pi / 2)
    solver.dt = 1 / f / 200
    z_pml = sum(solver.dz[Nz - 1 - k] for k in range(solver.n_pml))
    Nt = int(eps_r * 2.0 * (1.0 - 0.0 - z_pml) / c / solver.dt)
    for n in range(Nt):
        pw.update(solver, n * solver.dt)
        solver.one_step()
    refl_h = solver.H.get_abs()[Nx // 2, Ny // 2, :].max()
    refl_e = solver.E.get_abs()[Nx // 2, Ny // 2, :].max() / (mu_0 * c)
    print({"Nt": Nt, "H": float(refl_h), "E": float(refl_e)})
    assert np.isfinite(refl_h) and np.isfinite(refl_e)
    assert refl_h <= 10 and refl_e <= 10


def test_run_log_keys_like_reference(tmp_path):
    """tests/test_007_mpi_lossy_cavity.py:88-136, 453-536: after a wakesolve
    the grid, solver and wake logs hold exactly the reference's keys (a PEC
    background, PEC faces, STL solids) and wakis.log is in the results
    folder with the three sections."""
    import json as js

    import wakis_amd as wa

    geo = json.load(open(os.path.join(HERE, "golden", "001_cubic_cavity_box.json")))
    stl = str(tmp_path / "cavity.stl")
    write_ascii_stl(stl, geo["triangles"])
    lo, hi = geo["bounds_lo"], geo["bounds_hi"]
    print("[case] run log keys", flush=True)
    grid = wa.GridFIT3D(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 16, 16, 40,
                        stl_solids={"cavity": stl}, stl_materials={"cavity": [1.0, 1.0]})
    folder = str(tmp_path / "007_results") + "/"
    wake = wa.WakeSolver(q=1e-9, sigmaz=0.1, beta=1.0, xsource=0.0, ysource=0.0, xtest=0.0,
                         ytest=0.0, skip_cells=0, results_folder=folder, save=False)
    solver = wa.SolverFIT3D(grid, wake, bc_low=["pec"] * 3, bc_high=["pec"] * 3, use_stl=True,
                            bg="pec", dtype=np.float32)
    solver.wakesolve(wakelength=0.2, compute_plane="longitudinal")
    assert set(solver.logger.grid) == {
        "use_mesh_refinement", "Nx", "Ny", "Nz", "dx", "dy", "dz", "xmin", "xmax", "ymin",
        "ymax", "zmin", "zmax", "stl_solids", "stl_materials", "gridInitializationTime"}
    assert set(solver.logger.solver) == {
        "use_gpu", "use_mpi", "background", "bc_low", "bc_high", "dt", "solverInitializationTime"}
    assert set(solver.logger.wakeSolver) == {
        "ti", "q", "sigmaz", "beta", "xsource", "ysource", "xtest", "ytest", "chargedist",
        "skip_cells", "results_folder", "wakelength", "simulationTime"}
    assert solver.logger.wakeSolver["results_folder"] == folder
    assert solver.logger.wakeSolver["wakelength"] == 0.2
    assert solver.logger.solver["dt"] == solver.dt and solver.logger.grid["Nz"] == 40
    text = open(os.path.join(folder, "wakis.log"), encoding="utf-8").read()
    parts = text.split("\n## ")
    assert [p.split(" ##")[0] for p in parts[1:]] == ["WakeSolver Logs", "Solver Logs",
                                                      "Grid Logs"]
    grid_log = js.loads(parts[3].split("\n", 1)[1])
    assert grid_log["stl_materials"] == {"cavity": [1.0, 1.0]} and grid_log["Nx"] == 16


@pytest.mark.parametrize("skip", [0, 2])
def test_counter_moving_wakesolve(skip):
    """WakeSolver(counter_moving=True): the wake integral of
    wakis/wakeSolver.py:500-513 on the device after a wakesolve, against the
    oracle's (whose counter-moving branch is pinned by the z-mirror identity,
    tests/test_oracle.py), FP64 within 1e-10."""
    shape = (14, 13, 36)
    kw = dict(WAKE_KW, skip_cells=skip, counter_moving=True)
    print(f"[case] counter-moving wakesolve skip={skip} shape={shape}", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg="pec", solids=_cavity(shape),
                         init=False, cfln=1.0, wake_kw=kw)
    osv.wakesolve(compute_plane="both")
    psv.wakesolve(compute_plane="both")
    ow, pw = osv.wake, psv.wake
    assert np.abs(ow.WP_3d).max() > 0
    errs = {"WP3d": rel(pw.WP_3d, ow.WP_3d), "WP": rel(pw.WP, ow.WP), "Z": rel(pw.Z, ow.Z),
            "WPx": rel(pw.WPx, ow.WPx), "Zx": rel(pw.Zx, ow.Zx)}
    print(errs)
    assert max(errs.values()) <= 1e-10, errs


def _mirror_history(nz=20, nt=90, seed=9, dtype=np.float32):
    rng = np.random.default_rng(seed)
    zc = (np.arange(nz) - (nz - 1) / 2) * 0.004
    dz = rng.uniform(0.003, 0.005, nz)
    dt = np.float32(6.1e-12)
    hf = {"x": np.arange(3.0) * 1e-3, "y": np.arange(3.0) * 1e-3, "z": zc,
          "dx": np.full(3, 1e-3), "dy": np.full(3, 1e-3), "dz": dz,
          "t": np.arange(0, nt * dt, dt)}
    for n in range(nt):
        hf["#" + str(n).zfill(5)] = rng.standard_normal((3, 3, nz)).astype(dtype)
    return hf


@pytest.mark.parametrize("skip", [0, 3])
def test_counter_moving_mirror_identity_on_device(skip):
    """On a z axis symmetric about 0, the device's counter-moving WP of the
    z-reversed history is minus its co-moving WP of the history, bit for bit,
    and both equal the oracle's."""
    import wakis_amd as wa

    print(f"[case] counter-moving mirror identity skip={skip}", flush=True)
    hf = _mirror_history()
    rev = dict(hf)
    for key in hf:
        if key.startswith("#"):
            rev[key] = np.ascontiguousarray(hf[key][:, :, ::-1])
    kw = dict(wakelength=0.05, sigmaz=0.005, skip_cells=skip)
    co = wa.WakeSolver(save=False, **kw)
    cm = wa.WakeSolver(save=False, counter_moving=True, **kw)
    co.Ez_hf, cm.Ez_hf = hf, rev
    co.calc_long_WP_3d()
    cm.calc_long_WP_3d()
    assert np.abs(co.WP_3d).max() > 0
    assert np.array_equal(cm.WP_3d, -co.WP_3d)
    ocm = OWake(counter_moving=True, **kw)
    ocm.Ez_hf = rev
    ocm.calc_long_WP_3d()
    assert np.array_equal(cm.WP_3d, ocm.WP_3d)


@pytest.mark.parametrize("source", ["dict", "npz", "npz_f32"])
def test_standalone_solve_from_recorded_history(source, tmp_path):
    """WakeSolver.solve without a time loop (wakis/wakeSolver.py:211-255 on
    a history read by read_Ez, 1317-1364): the oracle's recorded E_z
    datasets ('#%05d' keys plus x, y, z, dx, dy, dz, t), given as a dict
    (Ez_hf) or an .npz archive (Ez_file), are moved to the device and give
    the oracle's WP_3d bit for bit and its WPx, WPy, Z, Zx, Zy within 1e-10."""
    import wakis_amd as wa

    shape = (12, 12, 30)
    print(f"[case] standalone solve from {source}", flush=True)
    dtype = np.float32 if source == "npz_f32" else np.float64
    osv, _ = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg="pec", solids=_cavity(shape),
                       init=False, cfln=1.0, wake_kw=WAKE_KW, dtype=dtype)
    osv.wakesolve(compute_plane="both")
    ow = osv.wake
    hf = dict(ow.Ez_hf) if ow.Ez_hf is not None else dict(osv.Ez_hf)
    assert any(k.startswith("#") for k in hf)
    if source == "dict":
        pw = wa.WakeSolver(save=False, **WAKE_KW)
        pw.Ez_hf = hf
    else:
        path = str(tmp_path / "Ez.npz")
        np.savez(path, **hf)
        pw = wa.WakeSolver(save=False, Ez_file=path, **WAKE_KW)
    pw.solve(compute_plane="both")
    assert np.array_equal(pw.s, ow.s)
    assert np.array_equal(pw.WP_3d, ow.WP_3d)
    errs = {"WPx": rel(pw.WPx, ow.WPx), "WPy": rel(pw.WPy, ow.WPy), "Z": rel(pw.Z, ow.Z),
            "Zx": rel(pw.Zx, ow.Zx), "Zy": rel(pw.Zy, ow.Zy)}
    print(errs)
    assert max(errs.values()) <= 1e-10, errs
    # the longitudinal-only path reads window (2, 2) of the 3x3 datasets
    pl = wa.WakeSolver(save=False, **WAKE_KW_LONG)
    pl.Ez_hf = hf
    pl.solve(compute_plane="longitudinal")
    ol = OWake(**WAKE_KW_LONG)
    ol.Ez_hf = hf
    ol.calc_long_WP()
    assert np.array_equal(pl.WP, ol.WP)
