"""Pinning the CPU oracle (no GPU): reference golden vectors, closed forms,
invariants and brute force, each chosen so that a dropped term, a wrong
sign/index or a transposed operand fails."""

import json
import os

import numpy as np
import pytest
from scipy.constants import c as c_light
from scipy.sparse import hstack

from oracle.fit_oracle import OGrid, OPlaneWave, OSolver, OWake, _wp_accumulate

HERE = os.path.dirname(os.path.abspath(__file__))


def _cubic_cavity():
    from wakis_amd.geometry import voxelize_nodes

    geo = json.load(open(os.path.join(HERE, "golden", "001_cubic_cavity_box.json")))
    lo, hi = geo["bounds_lo"], geo["bounds_hi"]
    g = OGrid(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 50, 50, 150)
    inside = voxelize_nodes(np.asarray(geo["triangles"]), g.x, g.y, g.z)
    c = np.zeros((50, 50, 150), bool)
    for a in (0, 1):
        for b in (0, 1):
            for d in (0, 1):
                c |= inside[a:a + 50, b:b + 50, d:d + 150]
    return g, c


def test_cubic_cavity_mask_matches_box_and_pipe():
    """The voxelised STL equals the analytic box (50x50x30 mm) + pipe
    (15x15 mm) mask under the one-corner-inside rule."""
    g, c = _cubic_cavity()
    xc = lambda a: (a[:-1], a[1:])  # noqa: E731
    (x0, x1), (y0, y1), (z0, z1) = xc(g.x), xc(g.y), xc(g.z)
    eps = 1e-12

    def any_in(lo, hi, a, b):  # some node of [lo, hi] strictly inside (a, b)
        return ((lo > a + eps) & (lo < b - eps)) | ((hi > a + eps) & (hi < b - eps))

    X = any_in(x0, x1, -0.025 - 1, 0.025 + 1)[:, None, None]
    box = (any_in(x0, x1, -0.0250001, 0.0250001)[:, None, None]
           & any_in(y0, y1, -0.0250001, 0.0250001)[None, :, None]
           & any_in(z0, z1, -0.015, 0.015)[None, None, :])
    pipe = (any_in(x0, x1, -0.0075, 0.0075)[:, None, None]
            & any_in(y0, y1, -0.0075, 0.0075)[None, :, None]
            & any_in(z0, z1, -0.0500001, 0.0500001)[None, None, :])
    assert X.all()
    assert np.array_equal(c, box | pipe)


@pytest.mark.slow
def test_golden_001_pec_cubic_cavity_fp32():
    """tests/test_001_pec_cubic_cavity.py:18-67, 153-177 (FP32, 5184 steps)."""
    gold = json.load(open(os.path.join(HERE, "golden", "goldens.json")))["test_001"]
    g, c = _cubic_cavity()
    wake = OWake(wakelength=1.0, q=1e-9, sigmaz=18.5e-3, beta=1.0, skip_cells=12)
    s = OSolver(g, wake, bc_low=["pec"] * 3, bc_high=["pec"] * 3, solids=[(c, "vacuum")],
                bg="pec", dtype=np.float32)
    s.wakesolve(wakelength=1.0, solve=False)
    wake.calc_long_WP_3d()
    wake.calc_long_Z()
    WP, Z = wake.WP, wake.Z
    tw, tz = gold["tol_wp"], gold["tol_z"]
    assert len(WP) == 4776 and len(Z) == 998
    assert np.allclose(WP[::50], gold["WP_every50"], rtol=tw, atol=tw)
    assert np.cumsum(np.abs(WP))[-1] == pytest.approx(gold["WP_abs_cumsum"], gold["cumsum_rel"])
    gz = np.array(gold["Z_every20_re"]) + 1j * np.array(gold["Z_every20_im"])
    for part in (np.abs, np.real, np.imag):
        assert np.allclose(part(Z)[::20], part(gz), rtol=tz, atol=tz)
    assert np.cumsum(np.abs(Z))[-1] == pytest.approx(gold["Z_abs_cumsum"], gold["cumsum_rel"])


def te101_initial_h(N, L=1.0):
    """H of the TE101 mode at t = -dt/2 as tests/test_002_analytic_resonator.py:
    12-127 samples it (m=1, n=0, p=1, theta=0; x = i*dx + xmin); the time
    factor is a common scale and is dropped."""
    d = L / N
    xmin = -L / 2 + d / 2
    x = np.arange(N) * d + xmin
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    k = np.pi / L
    hz = np.cos(k * (X - L / 2)) * np.cos(0 * Y) * np.sin(k * (Z - L / 2))
    hx = -2 / (2 * k**2) * k * k * np.sin(k * (X - L / 2)) * np.cos(0 * Y) * np.cos(k * (Z - L / 2))
    return hx, np.zeros_like(hx), hz


def test_te101_resonator_frequency():
    """tests/test_002_analytic_resonator.py:12-144 on a 20^3 grid (3000
    steps): the TE101 mode of a 1 m PEC cube rings at c/sqrt(2) = 2.12e8 Hz."""
    L, N = 1.0, 20
    d = L / N
    g = OGrid(-L / 2 + d / 2, L / 2 + d / 2, -L / 2 + d / 2, L / 2 + d / 2,
              -L / 2 + d / 2, L / 2 + d / 2, N, N, N)
    s = OSolver(g, bc_low=["pec"] * 3, bc_high=["pec"] * 3)
    hx, hy, hz = te101_initial_h(N, L)
    s.H.from_matrix(hx, "x")
    s.H.from_matrix(hy, "y")
    s.H.from_matrix(hz, "z")
    ey = []
    for _ in range(3000):
        ey.append(s.E[N // 2, N // 2, N // 2, "y"])
        s.one_step()
    ey = np.array(ey)
    F = np.abs(np.fft.rfft(ey))
    f = np.fft.rfftfreq(len(ey), d=s.dt)
    assert f[np.argmax(F)] == pytest.approx(2.121e8, rel=0.05)


def test_discrete_divergence_of_curl_is_zero():
    """Yee/FIT identity S C = 0 (div curl = 0) for the assembled operator,
    where S = [Px Py Pz] is the primal divergence built from the same shift
    matrices; a wrong sign or a swapped block breaks it."""
    g = OGrid(0, 1, 0, 1, 0, 1, 5, 4, 6)
    s = OSolver(g, bc_low=["periodic"] * 3, bc_high=["periodic"] * 3)
    S = hstack([s.Px, s.Py, s.Pz]).astype(np.int64)
    C = s.C.astype(np.int64)
    assert abs(S @ C).sum() == 0


def test_pec_cavity_energy_bounded_and_div_b_conserved():
    g = OGrid(0, 0.1, 0, 0.12, 0, 0.14, 6, 7, 8)
    s = OSolver(g, bc_low=["pec"] * 3, bc_high=["pec"] * 3, cfln=0.9)
    rng = np.random.default_rng(0)
    s.E.fromarray(rng.standard_normal(3 * s.N))
    s.one_step()  # zero ghosts first
    S = hstack([s.Px, s.Py, s.Pz])
    divB0 = S @ (s.H.toarray() / s.iA.toarray())
    e0 = np.linalg.norm(s.E.toarray())
    for _ in range(200):
        s.one_step()
    divB1 = S @ (s.H.toarray() / s.iA.toarray())
    assert np.linalg.norm(divB1 - divB0) <= 1e-9 * max(np.linalg.norm(divB0), 1.0)
    assert np.linalg.norm(s.E.toarray()) < 10 * e0


def test_pml_plane_wave_reflection():
    """tests/test_005_pml_boundaries.py:19-101, on a 4x4x200 column: the
    reflected wave stays below 10 % of the amplitude."""
    Nx, Ny, Nz = 4, 4, 200
    g = OGrid(-1.0, 1.0, -1.0, 1.0, 0.0, 1.0, Nx, Ny, Nz)
    s = OSolver(g, bc_low=["periodic", "periodic", "pec"], bc_high=["periodic", "periodic", "pml"],
                bg=[1.0, 1.0, 0.0], n_pml=30, dtype=np.float32)
    amplitude, nodes = 100.0, 7
    f = 15 / ((s.z.max() - s.z.min()) / c_light)
    pw = OPlaneWave(xs=slice(0, Nx), ys=slice(0, Ny), zs=0, amplitude=amplitude, f=f,
                    nodes=nodes, phase=np.pi / 2)
    s.dt = 1 / f / 200
    z_pml = sum(s.dz[Nz - 1 - k] for k in range(s.n_pml))
    Nt = int(2.0 * (1.0 - z_pml) / c_light / s.dt)
    for n in range(Nt):
        pw.update(s, n * s.dt)
        s.one_step()
    absH = np.sqrt(sum(s.H.to_matrix(d) ** 2 for d in "xyz"))
    assert absH[Nx // 2, Ny // 2, :].max() <= 10


def test_wp_vectorised_equals_loops():
    """The vectorised WP accumulation is bit-identical to the reference's
    double loop (wakeSolver.py:517-529) on random inputs, including the
    carry-over of calc_long_WP_3d and skip_cells."""
    rng = np.random.default_rng(4)
    nz, nt = 9, 40
    z = np.sort(rng.uniform(-0.05, 0.05, nz + 1))
    zc = (z[1:] + z[:-1]) / 2
    dt = np.float32(3.1e-12)
    hf = {"x": np.arange(3.0), "y": np.arange(3.0), "z": zc, "dx": np.ones(3), "dy": np.ones(3),
          "dz": np.diff(z), "t": np.arange(0, nt * dt, dt)}
    for n in range(nt):
        hf["#" + str(n).zfill(5)] = rng.standard_normal((3, 3, nz)).astype(np.float32)
    for skip in (0, 2):
        a = OWake(wakelength=0.02, sigmaz=0.004, skip_cells=skip)
        b = OWake(wakelength=0.02, sigmaz=0.004, skip_cells=skip)
        a.Ez_hf = b.Ez_hf = hf
        a.calc_long_WP_3d(loops=True)
        b.calc_long_WP_3d(loops=False)
        assert np.array_equal(a.WP_3d, b.WP_3d)
        a.calc_long_WP(loops=True)
        b.calc_long_WP(loops=False)
        assert np.array_equal(a.WP, b.WP)


def _mirror_history(nz=10, nt=60, seed=9, dtype=np.float32):
    """A 3x3 history on a z axis symmetric about 0 (z[nz-1-k] == -z[k]
    exactly) with an arbitrary, non-symmetric dz array."""
    rng = np.random.default_rng(seed)
    zc = (np.arange(nz) - (nz - 1) / 2) * 0.004
    assert np.array_equal(zc[::-1], -zc)
    dz = rng.uniform(0.003, 0.005, nz)
    dt = np.float32(6.1e-12)
    hf = {"x": np.arange(3.0), "y": np.arange(3.0), "z": zc, "dx": np.ones(3), "dy": np.ones(3),
          "dz": dz, "t": np.arange(0, nt * dt, dt)}
    for n in range(nt):
        hf["#" + str(n).zfill(5)] = rng.standard_normal((3, 3, nz)).astype(dtype)
    return hf


def test_counter_moving_wp_vectorised_equals_loops():
    """The vectorised counter-moving accumulation equals the reference's
    double loop of wakeSolver.py:500-513 bit for bit, with skip_cells and
    the carry-over between windows."""
    hf = _mirror_history(nz=9, nt=40, seed=4)
    for skip in (0, 2):
        a = OWake(wakelength=0.02, sigmaz=0.004, skip_cells=skip, counter_moving=True)
        b = OWake(wakelength=0.02, sigmaz=0.004, skip_cells=skip, counter_moving=True)
        a.Ez_hf = b.Ez_hf = hf
        a.calc_long_WP_3d(loops=True)
        b.calc_long_WP_3d(loops=False)
        assert np.array_equal(a.WP_3d, b.WP_3d)
        assert np.abs(a.WP_3d).max() > 0


@pytest.mark.parametrize("skip", [0, 2])
def test_counter_moving_wp_mirror_identity(skip):
    """On a z axis symmetric about 0, the counter-moving integral of the
    z-reversed history is minus the co-moving integral of the history, bit
    for bit: the counter-moving time (z[-k-1] - s)/(-v) - zmax/(-v) equals
    the co-moving (z[k] + s)/v - zmin/v exactly there, term k reads row
    nz-1-k of the reversed history (row k of the original) and is weighted
    by -dz[k].  With a non-symmetric dz this fails for a dz indexed from the
    top, a dropped sign, a time measured from zmin, or an unreversed z walk."""
    hf = _mirror_history()
    rev = dict(hf)
    for key in hf:
        if key.startswith("#"):
            rev[key] = np.ascontiguousarray(hf[key][:, :, ::-1])
    co = OWake(wakelength=0.03, sigmaz=0.005, skip_cells=skip)
    cm = OWake(wakelength=0.03, sigmaz=0.005, skip_cells=skip, counter_moving=True)
    co.Ez_hf, cm.Ez_hf = hf, rev
    co.calc_long_WP_3d()
    cm.calc_long_WP_3d()
    assert np.abs(co.WP_3d).max() > 0
    assert np.array_equal(cm.WP_3d, -co.WP_3d)
    assert np.array_equal(cm.s, co.s)


def test_wp_closed_form_travelling_wave():
    """Ez(z, t) = g(z - v t) sampled on the time grid integrates to
    WP(s) = -sum_k g(-s) dz / (q 1e12) up to the time bucketing: with
    g = const the line integral is exactly the column sum of dz."""
    nz, nt = 12, 1000
    z = np.linspace(-0.05, 0.05, nz + 1)
    zc = (z[1:] + z[:-1]) / 2
    dt = 1e-12
    hf = {"x": np.arange(3.0), "y": np.arange(3.0), "z": zc, "dx": np.ones(3), "dy": np.ones(3),
          "dz": np.diff(z), "t": np.arange(0, nt * dt, dt)}
    for n in range(nt):
        hf["#" + str(n).zfill(5)] = np.full((3, 3, nz), 2.0)
    w = OWake(wakelength=0.01, sigmaz=0.003, q=1e-9)
    w.Ez_hf = hf
    w.calc_long_WP()
    assert np.allclose(w.WP, 2.0 * 0.1 / (1e-9 * 1e12), rtol=1e-13)


def test_long_z_matches_direct_dft():
    """calc_long_Z (numpy.fft, zero-padded) equals an independent O(K N)
    direct DFT of the same samples on the kept bins."""
    w = OWake(sigmaz=0.01)
    w.s = np.arange(-0.085, 0.3, 3e-4)
    w.WP = np.sin(w.s * 40) * np.exp(-w.s * 3)
    w.calc_long_Z()
    N = w.N_dft
    n = np.arange(len(w.s))
    K = len(w.f)
    ds = np.mean(np.diff(w.s))
    lam = w.lambdas * w.v
    x = w.WP * 1e12
    W = np.array([(x * np.exp(-2j * np.pi * j * n / N)).sum() for j in range(K)]) * ds
    Lf = np.array([(lam * np.exp(-2j * np.pi * j * n / N)).sum() for j in range(K)]) * ds
    assert np.allclose(w.Z, -W / Lf, rtol=1e-9, atol=1e-12 * np.abs(w.Z).max())


def _linear_transverse_wake(n_cells, nt, seed):
    """WP_3d(x, y, s) = a x + b y + f (x - x0)(y - y0) + w(s) on a
    non-uniform transverse grid.  Panofsky-Wenzel fixes the transverse wake of
    this field: the s-integral of the x-gradient along the centre row is
    a * (left Riemann sum of ds over s[0:n]) = a * n * ds, so
    WPx[n] = -a n ds and WPy[n] = -b n ds exactly.  The cross term vanishes on
    the centre row and column, and w(s) has no transverse gradient."""
    rng = np.random.default_rng(seed)
    m = 2 * n_cells + 1
    xf = np.cumsum(rng.uniform(0.5e-3, 2e-3, m))
    yf = np.cumsum(rng.uniform(0.5e-3, 2e-3, m))
    a, b, f = 3.0, -1.7, 250.0
    s = np.linspace(-0.02, 0.1, nt)
    w = rng.standard_normal(nt)
    X, Y = xf[:, None, None], yf[None, :, None]
    wp3d = (a * X + b * Y + f * (X - xf[n_cells]) * (Y - yf[n_cells]) + w[None, None, :])
    ds = s[2] - s[1]
    return s, xf, yf, wp3d, -a * ds * np.arange(nt), -b * ds * np.arange(nt)


@pytest.mark.parametrize("n_cells", [1, 2])
def test_trans_wp_panofsky_wenzel_closed_form(n_cells):
    """calc_trans_WP (wakeSolver.py:553-626) on a field whose transverse wake
    is known in closed form: pins the sign, the exclusive prefix s[0:n], the
    x/y assignment, the centre indices and the non-uniform denominator."""
    s, xf, yf, wp3d, wpx, wpy = _linear_transverse_wake(n_cells, 120, 11 + n_cells)
    w = OWake()
    w.s, w.xf, w.yf, w.WP_3d = s, xf, yf, wp3d
    w.n_transverse_cells = n_cells
    w.calc_trans_WP()
    scale = np.abs(wpx).max()
    assert np.allclose(w.WPx, wpx, rtol=0, atol=1e-11 * scale)
    assert np.allclose(w.WPy, wpy, rtol=0, atol=1e-11 * scale)


def test_trans_z_matches_direct_dft():
    """calc_trans_Z (wakeSolver.py:722-785): Zx = 1j DFT_N(WPx 1e12) ds /
    DFT_N(lambda v) ds on the kept bins, against an O(K N) direct DFT."""
    w = OWake(sigmaz=0.01)
    w.s = np.arange(-0.085, 0.3, 3e-4)
    w.calc_lambdas_analytic()
    w.WPx = np.sin(w.s * 30) * np.exp(-w.s * 2)
    w.WPy = np.cos(w.s * 55) * np.exp(-w.s * 4)
    w.calc_trans_Z()
    n = np.arange(len(w.s))
    ds = np.mean(np.diff(w.s))
    N = int((w.v / ds) // w.fmax * 1001)
    K = len(w.fx)
    E = np.exp(-2j * np.pi * np.outer(np.arange(K), n) / N)
    Lf = E @ (w.lambdas * w.v) * ds
    for got, wp in ((w.Zx, w.WPx), (w.Zy, w.WPy)):
        want = 1j * (E @ (wp * 1e12) * ds) / Lf
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_abc_aliasing_matches_reference_vectors():
    """The oracle's emsolve ABC lag (get_abc / update_abc on views of the
    live field) reproduces the reference's own Field + BCsMixin run on the
    same inputs bit for bit (tests/golden/make_abc_golden.py wrote the
    vectors), including low+high ABC on one axis raising KeyError('xlo')
    after the first step."""
    g = np.load(os.path.join(HERE, "golden", "abc_aliasing.npz"))
    nx, ny, nz, nt = (int(v) for v in g["dims"])
    a_e, a_h = (float(v) for v in g["a"])
    for name in g["cases"]:
        name = str(name)
        bl, bh = {
            "low_x": (("abc", "pec", "pec"), ("pec",) * 3),
            "low_y": (("pec", "abc", "pec"), ("pec",) * 3),
            "low_z": (("pec", "pec", "abc"), ("pec",) * 3),
            "high_z": (("pec",) * 3, ("pec", "pec", "abc")),
            "high_x_low_y": (("pec", "abc", "pec"), ("abc", "pec", "pec")),
            "low_xyz": (("abc",) * 3, ("pec",) * 3),
            "low_high_z": (("pec", "pec", "abc"), ("pec", "pec", "abc")),
            "low_x_low_high_z": (("abc", "pec", "abc"), ("pec", "pec", "abc")),
        }[name]
        s = OSolver(OGrid(0, 1, 0, 1, 0, 1, nx, ny, nz), bc_low=bl, bc_high=bh)
        assert s.activate_abc
        s.E.fromarray(g["E0"])
        s.H.fromarray(g["H0"])
        step = [0]

        def one_step(s=s):
            n = step[0]
            s.E.fromarray(a_e * s.E.toarray() + g["dE"][n])
            s.H.fromarray(a_h * s.H.toarray() + g["dH"][n])
            step[0] += 1

        s.one_step = one_step
        err = ""
        try:
            s.emsolve(nt)
        except KeyError as e:
            err = f"KeyError:{e.args[0]}:step{step[0] - 1}"
        assert err == str(g[name + "/error"]), name
        assert np.array_equal(s.E.toarray(), g[name + "/E"]), name
        assert np.array_equal(s.H.toarray(), g[name + "/H"]), name
