"""FieldMonitor (wakis/field_monitors.py:14-131): the closed-form checks of
the reference's tests/test_012_field_monitors.py on the oracle's
restatement and on this package's host path (CPU), and the device
accumulation against the oracle (GPU: user loop and inside wakesolve)."""

import numpy as np
import pytest

from helpers import OFieldMonitor, make_pair


def crel(a, b):
    """Relative L2 distance of complex arrays."""
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _constant_field(cls, Nx=2, Ny=2, Nz=2, values=(1.0, 2.0, 3.0)):
    f = cls(Nx, Ny, Nz, dtype=float)
    for v, c in zip(values, "xyz"):
        f.from_matrix(np.full((Nx, Ny, Nz), v, dtype=float), c)
    return f


def _monitors():
    import wakis_amd as wa
    from oracle.fit_oracle import OField

    return [(OFieldMonitor, OField), (wa.FieldMonitor, wa.Field)]


def test_initial_state_and_empty_components():
    import wakis_amd as wa

    m = wa.FieldMonitor([1.0e9, 2.0e9])
    assert isinstance(m.frequencies, np.ndarray) and np.allclose(m.frequencies, [1e9, 2e9])
    assert m.time_index == 0 and m.dt is None and m.shape is None and m.xp is None
    assert m.Ex_acc is None and m.Ey_acc is None and m.Ez_acc is None
    assert all(v is None for v in m.get_components().values())


@pytest.mark.parametrize("which", [0, 1], ids=["oracle", "host"])
def test_single_update_zero_frequency(which):
    Mon, Fld = _monitors()[which]
    field = _constant_field(Fld)
    m = Mon([0.0])
    m.update(field, 1.0e-9)
    comps = m.get_components()
    assert m.dt == 1.0e-9 and m.shape == (2, 2, 2) and m.time_index == 1
    for key, v in zip(("Ex", "Ey", "Ez"), (1.0, 2.0, 3.0)):
        assert comps[key].shape == (1, 2, 2, 2) and np.iscomplexobj(comps[key])
        assert np.allclose(comps[key][0], v)


@pytest.mark.parametrize("which", [0, 1], ids=["oracle", "host"])
def test_zero_frequency_sums_linearly(which):
    Mon, Fld = _monitors()[which]
    field = _constant_field(Fld, values=(1.0, 0.0, 0.0))
    m = Mon([0.0])
    for _ in range(5):
        m.update(field, 0.1)
    comps = m.get_components()
    assert m.time_index == 5 and m.dt == 0.1
    assert np.allclose(comps["Ex"][0], 5.0)
    assert np.allclose(comps["Ey"], 0.0) and np.allclose(comps["Ez"], 0.0)


@pytest.mark.parametrize("which", [0, 1], ids=["oracle", "host"])
def test_half_period_phase_sums_to_zero(which):
    """f*dt = 1/2: the phases alternate +1, -1, so four updates cancel."""
    Mon, Fld = _monitors()[which]
    field = _constant_field(Fld, 1, 1, 1, values=(1.0, 0.0, 0.0))
    m = Mon([0.5])
    for _ in range(4):
        m.update(field, 1.0)
    assert np.allclose(m.get_components()["Ex"][0], 0.0, atol=1e-12)


def test_host_path_equals_oracle_bitwise():
    """Random fields, several frequencies and steps: the host path is the
    reference's numpy code, so it matches the oracle exactly."""
    import wakis_amd as wa
    from oracle.fit_oracle import OField

    rng = np.random.default_rng(7)
    freqs = [0.0, 1.3e9, 2.7e9]
    om, pm = OFieldMonitor(freqs), wa.FieldMonitor(freqs)
    of, pf = OField(3, 4, 5), wa.Field(3, 4, 5)
    for _ in range(6):
        a = rng.standard_normal(3 * 60)
        of.fromarray(a.copy())
        pf.fromarray(a.copy())
        om.update(of, 3.1e-12)
        pm.update(pf, 3.1e-12)
    for k in ("Ex", "Ey", "Ez"):
        assert np.array_equal(om.get_components()[k], pm.get_components()[k])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
def test_device_monitor_user_loop(dtype):
    """monitor.update(solver.E, dt) after each step of a user loop: device
    accumulators against the oracle's (FP64 1e-10, FP32 1e-4)."""
    import wakis_amd as wa

    shape = (9, 7, 12)
    print(f"[case] device field monitor user loop {np.dtype(dtype).name}", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec", "pec", "pml"], dtype=dtype, seed=13)
    freqs = [0.0, 0.9e9, 2.1e9]
    om, pm = OFieldMonitor(freqs), wa.FieldMonitor(freqs)
    for _ in range(10):
        osv.one_step()
        psv.one_step()
        om.update(osv.E, osv.dt)
        pm.update(psv.E, psv.dt)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    for k in ("Ex", "Ey", "Ez"):
        a, b = pm.get_components()[k], om.get_components()[k]
        assert a.shape == b.shape == (3, *shape)
        err = crel(a, b)
        assert err <= tol, (k, err)


@pytest.mark.gpu
def test_device_monitor_in_wakesolve():
    """wakesolve(use_field_monitor=True) with this package's FieldMonitor
    accumulates on the device inside the launch sequence (phases staged per
    chunk); the result equals the oracle's wakesolve with the reference's
    monitor update after each step, and agrees with the host-driven per-step
    path on the device (whose beam is evaluated on the host) to 1e-12."""
    import wakis_amd as wa
    from helpers import block_mask

    shape = (9, 9, 24)
    print("[case] device field monitor in wakesolve", flush=True)
    kw = dict(q=1e-9, sigmaz=0.01, beta=1.0, wakelength=0.05)
    solids = [(block_mask(shape, (1, 1, 0), (8, 8, 24)), "vacuum")]
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, init=False, solids=solids, bg="pec",
                         wake_kw=kw)
    _, psv2 = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, init=False, solids=solids, bg="pec",
                        wake_kw=kw)
    freqs = [0.5e9, 1.5e9]
    om = OFieldMonitor(freqs)
    osv.wakesolve(compute_plane="longitudinal", callback=lambda s, t: om.update(s.E, s.dt))
    pm = wa.FieldMonitor(freqs)
    psv.wakesolve(compute_plane="longitudinal", use_field_monitor=True, field_monitor=pm, chunk=7)
    pm2 = wa.FieldMonitor(freqs)
    psv2.wakesolve(compute_plane="longitudinal", use_field_monitor=True, field_monitor=pm2,
                   callback=lambda s, t: None)
    assert pm.time_index == om.time_index == psv.Nt
    for k in ("Ex", "Ey", "Ez"):
        a, b, c = pm.get_components()[k], om.get_components()[k], pm2.get_components()[k]
        assert crel(a, c) <= 1e-12, k
        assert crel(a, b) <= 1e-10, k


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tile", [(np.float64, 1), (np.float64, 4), (np.float64, 8),
                                        (np.float32, 5), (np.float32, 2)],
                         ids=["fp64-t1", "fp64-t4", "fp64-row8", "fp32-t5", "fp32-t2"])
def test_fused_monitor_update(dtype, tile):
    """The monitor update fused into the one-step sweep (every chunk on the
    general tile path while a monitor update is staged) on a grid of several
    tiles and z chunks: accumulators against the oracle's reference update
    after each step (FP64 1e-10, FP32 1e-4) and against the separate
    accumulation pass of the split kernels (1e-12 FP64).  Tile 8 is a
    row-sweep tile, which keeps the separate pass."""
    import wakis_amd as wa
    from helpers import block_mask

    shape = (70, 20, 76)
    print(f"[case] fused monitor {np.dtype(dtype).name} tile={tile}", flush=True)
    kw = dict(q=1e-9, sigmaz=0.01, beta=1.0, wakelength=0.04)
    solids = [(block_mask(shape, (2, 2, 0), (68, 18, 76)), "vacuum"),
              (block_mask(shape, (20, 5, 30), (40, 12, 50)), [3.0, 1.0, 30.0])]
    bc = (["pec"] * 3, ["pec", "pec", "pml"])
    osv, pf = make_pair(*shape, *bc, dtype=dtype, init=False, solids=solids, bg="pec",
                        wake_kw=kw, kernel="fused", tile=tile, n_pml=4)
    _, ps = make_pair(*shape, *bc, dtype=dtype, init=False, solids=solids, bg="pec",
                      wake_kw=kw, kernel="split", n_pml=4)
    freqs = [0.7e9, 2.2e9]
    om = OFieldMonitor(freqs)
    osv.wakesolve(compute_plane="longitudinal", callback=lambda s, t: om.update(s.E, s.dt))
    mf, ms = wa.FieldMonitor(freqs), wa.FieldMonitor(freqs)
    pf.wakesolve(compute_plane="longitudinal", use_field_monitor=True, field_monitor=mf)
    ps.wakesolve(compute_plane="longitudinal", use_field_monitor=True, field_monitor=ms)
    assert pf.ctx.launch_counts()[1] == 0  # one-step sweeps while the monitor runs
    tol = 1e-10 if dtype == np.float64 else 1e-4
    for k in ("Ex", "Ey", "Ez"):
        a, b, c = mf.get_components()[k], om.get_components()[k], ms.get_components()[k]
        assert np.abs(b).max() > 0, k
        assert crel(a, b) <= tol, (k, crel(a, b))
        if dtype == np.float64:
            assert crel(a, c) <= 1e-12, (k, crel(a, c))


@pytest.mark.gpu
def test_monitor_staging_after_frequency_count_change():
    """Phases staged for one frequency, then, after fit_monitor_set with two,
    for two: the staged table is re-sized by bytes, and the fused updates
    equal a separate update after each step (regression: a table sized for
    one frequency's rows was reused for two)."""
    shape = (40, 12, 30)
    print("[case] monitor restaging", flush=True)
    _, pa = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, seed=5, kernel="fused")
    _, pb = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, seed=5, kernel="fused")
    for s in (pa, pb):
        s.ctx.zero_ghosts()
        s.step_0 = False
    n = 6
    for nf in (1, 2):
        ph = np.exp(-2j * np.pi * np.outer(np.arange(n) + 1.0, 0.05 * (np.arange(nf) + 1)))
        pa.ctx.monitor_set(nf)
        pb.ctx.monitor_set(nf)
        pa.ctx.monitor_phases(ph)
        pa.ctx.step(n, False)
        for q in range(n):
            pb.ctx.step(1, False)
            pb.ctx.monitor_accumulate(ph[q])
        for f in range(nf):
            for comp in range(3):
                a = pa.ctx.monitor_read(f, comp)
                b = pb.ctx.monitor_read(f, comp)
                assert np.abs(b).max() > 0
                assert crel(a, b) <= 1e-12, (nf, f, comp, crel(a, b))
