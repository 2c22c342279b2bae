"""Host-side logic without a GPU: the C-ABI library exports every declared
symbol, and the solver's compact setup (metric tables, face masks, material
table, dt) reproduces the oracle's Fields and operators exactly."""

import ctypes
import os
import re

import numpy as np
import pytest

from helpers import axes, block_mask
from oracle.fit_oracle import OGrid, OSolver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "wakis_fit.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fit_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from wakis_amd import _lib

    names = _declared_symbols()
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_product_fails_loudly_without_device():
    import wakis_amd as wa
    from wakis_amd import _lib

    g = wa.GridFIT3D(0, 1, 0, 1, 0, 1, 3, 3, 3)
    s = wa.SolverFIT3D(g, bc_low=["pec"] * 3, bc_high=["pec"] * 3, device=None)
    with pytest.raises(_lib.FitError):
        s.one_step()


BCS = [
    (["pec"] * 3, ["pec"] * 3),
    (["periodic"] * 3, ["periodic"] * 3),
    (["periodic", "pec", "pec"], ["periodic", "pmc", "pml"]),
    (["pmc", "pec", "magnetic"], ["pec", "PMC", "pec"]),
    (["pec", "abc", "periodic"], ["abc", "pec", "periodic"]),
    (["abc", "periodic", "pml"], ["abc", "periodic", "pml"]),
]


def _pair(bl, bh, dtype=np.float64, nonuniform=True, solids=(), bg=(1.0, 1.0), n_pml=2):
    import wakis_amd as wa

    x, y, z = axes(5, 4, 7, nonuniform=nonuniform, seed=7)
    osv = OSolver(OGrid(x=x, y=y, z=z), bc_low=bl, bc_high=bh, dtype=dtype, solids=list(solids),
                  bg=bg, n_pml=n_pml)
    psv = wa.SolverFIT3D(wa.GridFIT3D(x=x, y=y, z=z, solids=list(solids)), bc_low=bl,
                         bc_high=bh, dtype=dtype, use_stl=bool(solids), bg=bg, n_pml=n_pml,
                         device=None)
    return osv, psv


def _expand(psv):
    """Packed metric tables -> 3N arrays (L, tL, iA, itA) like the Fields."""
    Nx, Ny, Nz = psv.Nx, psv.Ny, psv.Nz
    p = psv.geometry
    o = 0
    one = []
    for _ in range(2):
        one.append((p[o:o + Nx], p[o + Nx:o + Nx + Ny], p[o + Nx + Ny:o + Nx + Ny + Nz]))
        o += Nx + Ny + Nz
    two = []
    for _ in range(2):
        ax = p[o:o + Ny * Nz].reshape((Ny, Nz), order="F")
        o += Ny * Nz
        ay = p[o:o + Nx * Nz].reshape((Nx, Nz), order="F")
        o += Nx * Nz
        az = p[o:o + Nx * Ny].reshape((Nx, Ny), order="F")
        o += Nx * Ny
        two.append((ax, ay, az))
    assert o == p.size
    sh = (Nx, Ny, Nz)

    def f3(parts):
        return np.concatenate([np.reshape(np.broadcast_to(a, sh), -1, order="F") for a in parts])

    L = f3([one[0][0][:, None, None], one[0][1][None, :, None], one[0][2][None, None, :]])
    tL = f3([one[1][0][:, None, None], one[1][1][None, :, None], one[1][2][None, None, :]])
    iA = f3([two[0][0][None], two[0][1][:, None, :], two[0][2][:, :, None]])
    itA = f3([two[1][0][None], two[1][1][:, None, :], two[1][2][:, :, None]])
    return L, tL, iA, itA


@pytest.mark.parametrize("bl,bh", BCS)
def test_metric_tables_equal_oracle_fields(bl, bh):
    osv, psv = _pair(bl, bh)
    L, tL, iA, itA = _expand(psv)
    assert np.array_equal(L, osv.L.toarray())
    assert np.array_equal(tL, osv.tL.toarray())
    assert np.array_equal(iA, osv.iA.toarray())
    assert np.array_equal(itA, osv.itA.toarray())


@pytest.mark.parametrize("bl,bh", BCS)
def test_face_masks_reproduce_the_operator(bl, bh):
    """Rebuild C from the face bits and compare with the oracle's C after
    _apply_bc_to_C (column PEC mask, row PMC mask with carried flags)."""
    from scipy.sparse import diags

    osv, psv = _pair(bl, bh)
    Nx, Ny, Nz = psv.Nx, psv.Ny, psv.Nz

    def mask(bits):
        # the kernel's rule: faces in order xlo, xhi, ylo, yhi, zlo, zhi, the
        # last face containing the cell decides (masked iff its bit is set)
        m = np.ones((3, Nx, Ny, Nz), dtype=np.int8)
        for c in range(3):
            if c != 0:
                m[c, 0] = 0 if bits & 1 else 1
                m[c, -1] = 0 if bits & 2 else 1
            if c != 1:
                m[c, :, 0] = 0 if bits & 4 else 1
                m[c, :, -1] = 0 if bits & 8 else 1
            if c != 2:
                m[c, :, :, 0] = 0 if bits & 16 else 1
                m[c, :, :, -1] = 0 if bits & 32 else 1
        return np.concatenate([np.reshape(m[c], -1, order="F") for c in range(3)])

    g = OGrid(x=psv.grid.x, y=psv.grid.y, z=psv.grid.z)
    raw = OSolver(g, bc_low=["periodic"] * 3, bc_high=["periodic"] * 3)
    C = raw.C
    if psv.pec_bits:
        C = C @ diags(mask(psv.pec_bits))
    if psv.pmc_bits:
        C = diags(mask(psv.pmc_bits)) @ C
    assert abs(C - osv.C).sum() == 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_materials_and_dt_equal_oracle(dtype):
    shape = (5, 4, 7)
    solids = [
        (block_mask(shape, (0, 0, 1), (5, 4, 6)), "vacuum"),
        (block_mask(shape, (1, 1, 2), (3, 3, 4)), [4.0, 1.5]),
        (block_mask(shape, (2, 1, 3), (4, 4, 5)), [2.0, 1.0, 50.0]),
        (block_mask(shape, (0, 2, 3), (2, 4, 6)), "pec"),
        (block_mask(shape, (3, 0, 0), (5, 2, 3)), [3.0, 2.0]),
    ]
    for bl, bh, bg in [(["pec"] * 3, ["pec", "pec", "pml"], "pec"),
                       (["pml"] * 3, ["pml"] * 3, [1.0, 1.0, 2.0]),
                       (["pec"] * 3, ["pec"] * 3, [2.0, 1.0])]:
        osv, psv = _pair(bl, bh, dtype=dtype, solids=solids, bg=bg, n_pml=2)
        for name in ("ieps", "imu", "sigma"):
            assert np.array_equal(getattr(psv, name).toarray(),
                                  getattr(osv, name).toarray().astype(np.float64)), name
        assert type(psv.dt) is type(osv.dt) and psv.dt == osv.dt
        assert psv.use_conductivity == osv.use_conductivity


def test_empty_conductive_background_raises_like_reference():
    """bg with sigma=0 and no PML: the reference's relaxation clamp takes the
    min of an empty array (wakis/solverFIT3D.py:309-318)."""
    with pytest.raises(ValueError):
        _pair(["pec"] * 3, ["pec"] * 3, bg=[1.0, 1.0, 0.0])
    with pytest.raises(ValueError):
        OSolver(OGrid(0, 1, 0, 1, 0, 1, 3, 3, 3), bc_low=["pec"] * 3, bc_high=["pec"] * 3,
                bg=[1.0, 1.0, 0.0])


def test_bucket_columns_match_reference_expression():
    """The host bucket table is int(ts/dt) of wakeSolver.py:517-529."""
    import wakis_amd as wa

    w = wa.WakeSolver(wakelength=0.05, sigmaz=0.005, save=False, skip_cells=1)
    z = np.linspace(-0.04, 0.04, 17)
    zc = (z[1:] + z[:-1]) / 2
    dt = np.float32(2.7e-12)
    nt = 300
    w.attach_device_history(None, {"x": np.arange(3.0), "y": np.arange(3.0), "z": zc,
                                   "dx": 0, "dy": 0, "dz": np.diff(z),
                                   "t": np.arange(0, nt * dt, dt), "nsteps": nt, "half": 1})
    cols, s, zs = w.bucket_columns(nt)
    for n in range(0, len(s), 7):
        for k in range(len(zs)):
            ts = (zs[k] + s[n]) / w.v - zc.min() / w.v - w.t[0] + w.ti
            it = int(ts / (w.t[2] - w.t[1]))
            assert cols[k, n] == (it if it < nt else -1)


def test_counter_moving_bucket_columns_match_reference_expression():
    """counter=True: int(ts/dt) of the counter-moving branch,
    wakeSolver.py:500-508, term k at z[-k-1] with times from zmax."""
    import wakis_amd as wa

    w = wa.WakeSolver(wakelength=0.05, sigmaz=0.005, save=False, skip_cells=2,
                      counter_moving=True)
    z = np.sort(np.random.default_rng(3).uniform(-0.04, 0.04, 18))
    zc = (z[1:] + z[:-1]) / 2
    dt = np.float32(2.7e-12)
    nt = 300
    w.attach_device_history(None, {"x": np.arange(3.0), "y": np.arange(3.0), "z": zc,
                                   "dx": 0, "dy": 0, "dz": np.diff(z),
                                   "t": np.arange(0, nt * dt, dt), "nsteps": nt, "half": 1})
    cols, s, zs = w.bucket_columns(nt, counter=True)
    assert (cols >= 0).any()
    for n in range(0, len(s), 7):
        for k in range(len(zs)):
            ts = ((zs[-k - 1] - s[n]) / (-1 * w.v) - zc.max() / (-1 * w.v) - w.t[0] + w.ti)
            it = int(ts / (w.t[2] - w.t[1]))
            assert cols[k, n] == (it if it < nt else -1)


class _HistoryCapture:
    """Stands in for the device context: keeps what would be uploaded."""

    made = []

    def __init__(self, nx, ny, nz, dtype=np.float64, device=0):
        self.shape, self.dtype = (nx, ny, nz), np.dtype(dtype)
        _HistoryCapture.made.append(self)

    def wake_history_load(self, ixt, iyt, h, hist):
        self.probe, self.hist = (ixt, iyt, h), np.array(hist)


@pytest.mark.parametrize("win", [3, 5, 1])
def test_recorded_history_layout(win, monkeypatch, tmp_path):
    """read_Ez's datasets become the device history [w][k][t] with
    w = (di+1)*3 + (dj+1) over the centred 3x3 window (rows shape//2 + i, as
    wakeSolver.py:492-498 indexes a dataset), or the line of a 1-D dataset;
    the first min(len(t), number of datasets) steps are kept; an .npz
    archive gives the same history as the dict."""
    import wakis_amd as wa
    from wakis_amd import _lib

    monkeypatch.setattr(_lib, "Context", _HistoryCapture)
    rng = np.random.default_rng(2)
    nz, nkeys = 7, 11
    t = np.arange(9) * 1e-12  # fewer times than datasets: nt = 9
    hf = {"x": np.arange(3.0), "y": np.arange(3.0), "z": np.arange(nz) * 1e-3,
          "dx": np.ones(3), "dy": np.ones(3), "dz": np.full(nz, 1e-3), "t": t}
    shape = (win, win, nz) if win > 1 else (nz,)
    for n in range(nkeys):
        hf["#" + str(n).zfill(5)] = rng.standard_normal(shape).astype(np.float32)
    w = wa.WakeSolver(save=False)
    w.Ez_hf = hf
    w._load_host_history()
    cap = _HistoryCapture.made[-1]
    assert cap.dtype == np.float32 and w._meta["nsteps"] == 9
    if win > 1:
        c = win // 2
        assert cap.shape == (3, 3, nz) and cap.probe == (1, 1, 1)
        for n in range(9):
            e = hf["#" + str(n).zfill(5)]
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    assert np.array_equal(cap.hist[(di + 1) * 3 + (dj + 1), :, n],
                                          e[c + di, c + dj, :])
    else:
        assert cap.probe == (0, 0, 0) and cap.hist.shape == (1, nz, 9)
        for n in range(9):
            assert np.array_equal(cap.hist[0, :, n], hf["#" + str(n).zfill(5)])
    path = str(tmp_path / "Ez.npz")
    np.savez(path, **hf)
    w2 = wa.WakeSolver(save=False, Ez_file=path)
    w2._need()
    assert np.array_equal(_HistoryCapture.made[-1].hist, cap.hist)
    assert np.array_equal(w2.t, t) and np.array_equal(w2.dz, hf["dz"])


def test_missing_history_raises():
    import wakis_amd as wa

    w = wa.WakeSolver(save=False, Ez_file="/nonexistent/Ez.npz")
    with pytest.raises(RuntimeError, match="no E_z history"):
        w.calc_long_WP_3d()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_slice_solids_and_tensor_material_equal_oracle(dtype):
    """Index-slice solids and a diagonal-tensor material {"eps": (ex, ey, ez)}
    give the material Fields the oracle holds after per-component ieps edits
    (the stratified-box workload at 12^3, 2-cell PML on all faces)."""
    from helpers import stratified_pair

    osv, psv, wl = stratified_pair(n=12, nlayers=4, n_pml=2, dtype=dtype, device=None)
    for name in ("ieps", "imu", "sigma"):
        assert np.array_equal(getattr(psv, name).toarray(),
                              getattr(osv, name).toarray().astype(np.float64)), name
    assert psv.dt == osv.dt and psv.use_conductivity == osv.use_conductivity
    # the tensor block really is anisotropic
    b = psv.ieps.toarray().reshape(3, -1)[:, np.ravel_multi_index((6, 6, 6), (12, 12, 12), order="F")]
    assert len(set(b.tolist())) == 3


def test_stratified_box_host_setup_is_lean():
    """The 1024^3 box's setup holds no (n, n, n) solid masks: its solids are
    slice boxes and its material ids fit the 256-row device table."""
    from wakis_amd.workloads import stratified_box

    wl = stratified_box(n=1024)
    assert all(isinstance(r, tuple) for r, _ in wl["grid"]["solids"])
    assert len(wl["grid"]["solids"]) == 17
    assert wl["grid"]["x"].size == 1025


def test_result_files_roundtrip(tmp_path):
    """save_txt / read_txt / load_results with the reference's file layout
    (wakis/wakeSolver.py:1366-1511): header keys "name[unit]", complex
    columns read back as complex."""
    import wakis_amd as wa

    w = wa.WakeSolver(save=True, results_folder=str(tmp_path) + "/")
    rng = np.random.default_rng(3)
    s = np.linspace(-0.1, 1.0, 50)
    f = np.linspace(0, 5e9, 30)
    cols = {"WP": rng.standard_normal(50), "WPx": rng.standard_normal(50),
            "WPy": rng.standard_normal(50), "lambda": rng.standard_normal(50),
            "Z": rng.standard_normal(30) + 1j * rng.standard_normal(30),
            "Zx": rng.standard_normal(30) + 1j * rng.standard_normal(30),
            "Zy": rng.standard_normal(30) + 1j * rng.standard_normal(30),
            "spectrum": rng.standard_normal(30) + 1j * rng.standard_normal(30)}
    for name, v in cols.items():
        x = s if v.size == 50 else f
        w.save_txt(str(tmp_path / name), x, v, "s [m]" if v.size == 50 else "f [Hz]", name + " [-]")
    d = w.read_txt(str(tmp_path / "WP.txt"))
    assert list(d.keys()) == ["s[m]", "WP[-]"]
    w2 = wa.WakeSolver(save=False)
    w2.load_results(str(tmp_path))
    assert np.array_equal(w2.s, s) and np.array_equal(w2.WP, cols["WP"])
    assert np.array_equal(w2.Z, cols["Z"]) and np.array_equal(w2.f, f)
    assert np.array_equal(w2.lambdaf, cols["spectrum"]) and np.array_equal(w2.WPy, cols["WPy"])
    assert w2.wakelength == s[-1] and w2.Ez_file.endswith("Ez.h5")


def test_wake_impedance_transforms_sine():
    """tests/test_004_wakes_and_impedances.py:11-52: a sine wake of amplitude
    A at f_r gives |Z| peaking at A at f_r, and wake -> Z -> wake round trips
    keep the wake within 1 (twice)."""
    from wakis_amd import WakeSolver as wk

    fr, A = 0.5e9, 100
    t = np.linspace(0, 100 * 1e-9, 3000)
    wake = A * np.sin(2 * np.pi * fr * t)
    f, Z = wk.calc_impedance_from_wake([t, wake], verbose=False)
    tt, wwake = wk.calc_wake_from_impedance([f, Z], samples=3000, verbose=False)
    ff, Zz = wk.calc_impedance_from_wake([tt, wwake], verbose=False)
    ttt, wwwake = wk.calc_wake_from_impedance([ff, Zz], samples=3000, verbose=False)
    assert np.allclose(wake, wwake, atol=1)
    assert np.allclose(wake, wwwake, atol=1)
    assert np.allclose(np.max(np.abs(Z)), A, atol=1)
    assert np.allclose(np.max(np.abs(Zz)), A, atol=1)
    assert np.allclose(f[np.argmax(Z)], fr, atol=1e6)
    assert np.allclose(ff[np.argmax(Zz)], fr, atol=1e6)
    # the transform pair by its definition: a DFT bin of the wake, and a
    # length-preserving inverse of a one-sided spectrum
    n = len(wake)
    k = 17
    ref = 2 / n * np.sum(wake * np.exp(-2j * np.pi * k * np.arange(n) / n))
    assert abs(Z[k] - ref) <= 1e-9 * abs(ref) + 1e-9
    assert len(f) == (n + 1) // 2 and np.all(f >= 0) and np.all(np.diff(f) > 0)
    assert len(tt) == 3000 and tt[0] == 0.0 and tt[-1] == pytest.approx(1 / np.mean(np.diff(f)))
    # s in metres instead of t; fmax resampling; errors without a time axis
    f2, Z2 = wk.calc_impedance_from_wake(wake, s=t * 299792458.0, verbose=False)
    assert np.allclose(f2, f, rtol=1e-9) and np.allclose(Z2, Z)
    f3, _ = wk.calc_impedance_from_wake([t, wake], fmax=2e9, verbose=False)
    assert f3.max() < 2e9
    with pytest.raises(AttributeError):
        wk.calc_impedance_from_wake(wake, verbose=False)
    with pytest.raises(AttributeError):
        wk.calc_wake_from_impedance(Z, verbose=False)


@pytest.mark.parametrize("n_cells", [1, 2])
def test_trans_wp_host_closed_form(n_cells):
    """WakeSolver.calc_trans_WP (host numpy, wakis/wakeSolver.py:553-626) on a
    field with a closed-form Panofsky-Wenzel transverse wake, and equal to the
    oracle's per-n loop to rounding on random data."""
    from test_oracle import _linear_transverse_wake
    from oracle.fit_oracle import OWake
    from wakis_amd.wake import WakeSolver

    s, xf, yf, wp3d, wpx, wpy = _linear_transverse_wake(n_cells, 120, 21 + n_cells)
    w = WakeSolver(save=False, verbose=0)
    w.calc_trans_WP(s=s, xf=xf, yf=yf, WP_3d=wp3d, n_transverse_cells=n_cells)
    scale = np.abs(wpx).max()
    assert np.allclose(w.WPx, wpx, rtol=0, atol=1e-11 * scale)
    assert np.allclose(w.WPy, wpy, rtol=0, atol=1e-11 * scale)
    rnd = np.random.default_rng(5).standard_normal(wp3d.shape)
    o = OWake()
    o.s, o.xf, o.yf, o.WP_3d, o.n_transverse_cells = s, xf, yf, rnd, n_cells
    o.calc_trans_WP()
    w.calc_trans_WP(WP_3d=rnd)
    assert np.allclose(w.WPx, o.WPx, rtol=1e-12, atol=1e-12 * np.abs(o.WPx).max())
    assert np.allclose(w.WPy, o.WPy, rtol=1e-12, atol=1e-12 * np.abs(o.WPy).max())
