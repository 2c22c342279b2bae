"""The two-step fused sweep (two leapfrog steps per launch; used by fit_step
for pairs of steps on grids with nx >= 16 and ny >= 16, in tiles of 64 x 8,
32 x 16, 32 x 8, 16 x 16 or 60 x 12 columns in FP32, 32 x 8, 16 x 16 or
28 x 12 in FP64) against the oracle (relative L2 <= 1e-4 in FP32, <= 1e-10 in FP64:
the bars of the other parity tests) and against the one-step fused sweep on
the same context setup (<= 1e-5 in FP32, <= 1e-12 in FP64: the two kernels
evaluate the same per-cell expressions).  Cases:
every boundary kind, the helical x/y wrap across tile and grid edges, odd
step counts (a single step after the pairs), odd plane counts, non-uniform
spacing, conductive solids with PML (J = sigma*E in both steps), and the
whole FP32 wakesolve (in-kernel beam and E_z probe of both steps)."""

import numpy as np
import pytest

from helpers import block_mask, make_pair, rel

pytestmark = pytest.mark.gpu

TOL_ORACLE = 1e-4
TOL_KERNEL = 1e-5

CASES = [
    ("pec", (64, 16, 24), ["pec"] * 3, ["pec"] * 3, False, False),
    # tiles overhanging the high x and y edges (96 = 64 + 32, 16 = 2 x 7 + 2)
    ("periodic_partial", (96, 16, 20), ["periodic"] * 3, ["periodic"] * 3, False, False),
    ("mixed_odd_nz", (128, 16, 17), ["periodic", "pec", "pec"], ["periodic", "pmc", "pml"],
     False, False),
    ("pmc", (64, 24, 13), ["pmc", "pec", "magnetic"], ["pec", "pmc", "pec"], False, False),
    ("pml_conductive", (64, 16, 40), ["pec", "pec", "pml"], ["pec", "pec", "pml"], True, False),
    ("nonuniform_abc_open", (80, 17, 30), ["pec", "periodic", "pec"],
     ["periodic", "periodic", "pec"], False, True),
    # tall grids: lean mid chunks between 4-plane face chunks
    ("periodic_lean", (64, 21, 48), ["periodic"] * 3, ["periodic"] * 3, False, False),
    ("pmc_lean_nonuniform", (128, 16, 45), ["pmc", "pec", "magnetic"], ["pec", "pmc", "pml"],
     False, True),
]


def _solids(shape):
    Nx, Ny, Nz = shape
    return [
        (block_mask(shape, (2, 2, 3), (Nx - 2, Ny - 2, Nz - 3)), "vacuum"),
        (block_mask(shape, (10, 1, 4), (30, 6, 9)), [4.0, 1.5]),
        (block_mask(shape, (20, 3, 5), (50, 7, 12)), [2.0, 1.0, 50.0]),
        (block_mask(shape, (5, 4, 6), (9, 8, 10)), "pec"),
    ]


def _pair(shape, bl, bh, solids, nonuniform, kernel, dtype=np.float32):
    osv, psv = make_pair(*shape, bl, bh, dtype=dtype, solids=_solids(shape) if solids else (),
                         bg="pec" if solids else (1.0, 1.0), n_pml=3, nonuniform=nonuniform,
                         init=False, kernel=kernel, seed=3)
    rng = np.random.default_rng(17)
    for name in ("E", "H"):  # J stays 0 (J mode zero / sigma*E)
        a = rng.standard_normal(3 * osv.N).astype(dtype)
        getattr(osv, name).fromarray(a)
        getattr(psv, name).fromarray(a)
    return osv, psv


@pytest.mark.parametrize("name,shape,bl,bh,solids,nonuniform", CASES, ids=[c[0] for c in CASES])
def test_two_step_sweep_matches_oracle_and_one_step(name, shape, bl, bh, solids, nonuniform):
    nsteps = 7  # three two-step launches and one single step
    print(f"[case] two-step {name} {shape}", flush=True)
    osv, p2 = _pair(shape, bl, bh, solids, nonuniform, "fused2")
    _, p1 = _pair(shape, bl, bh, solids, nonuniform, "fused")
    from wakis_amd import _lib

    assert p2.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
    assert p1.ctx.kernel_in_use() == _lib.KERNEL_FUSED
    for _ in range(nsteps):
        osv.one_step()
    p2._advance(nsteps)
    p1._advance(nsteps)
    n1, n2 = p2.ctx.launch_counts()
    assert (n1, n2) == (1, 3), (n1, n2)
    assert p1.ctx.launch_counts() == (nsteps, 0)
    errs = {}
    for f in "EHJ":
        o = getattr(osv, f).toarray()
        a2 = getattr(p2, f).toarray()
        a1 = getattr(p1, f).toarray()
        errs[f] = (rel(a2, o), rel(a2, a1))
    print(errs, flush=True)
    for f, (eo, ek) in errs.items():
        assert eo <= TOL_ORACLE and ek <= TOL_KERNEL, (f, eo, ek)


def test_two_step_wakesolve_matches_one_step():
    """The FP32 wakesolve of a conductive cavity with z PML, beam on the axis
    and the 3 x 3 E_z probe: fields, WP and Z of the two-step path against
    the one-step path (and the fields against the FP32 oracle)."""
    import wakis_amd as wa
    from wakis_amd import _lib

    shape = (64, 16, 48)
    Nx, Ny, Nz = shape
    wall = ~block_mask(shape, (4, 3, 0), (Nx - 4, Ny - 3, Nz))
    kw = dict(wakelength=0.12, q=1e-9, sigmaz=0.01, beta=1.0)
    res = {}
    for kernel in ("fused2", "fused"):
        osv, psv = make_pair(*shape, ["pec", "pec", "pml"], ["pec", "pec", "pml"],
                             dtype=np.float32, solids=[(wall, [1.0, 1.0, 1e3])], n_pml=5,
                             init=False, kernel=kernel, wake_kw=kw)
        print(f"[case] two-step wakesolve kernel={kernel}", flush=True)
        psv.wakesolve(compute_plane="longitudinal")
        res[kernel] = (psv, osv)
    p2, osv = res["fused2"]
    p1, _ = res["fused"]
    assert p2.ctx.launch_counts()[1] > p2.Nt // 4
    osv.wakesolve(compute_plane="longitudinal")
    errs = {"E": rel(p2.E.toarray(), p1.E.toarray()), "H": rel(p2.H.toarray(), p1.H.toarray()),
            "WP": rel(p2.wake.WP, p1.wake.WP), "Z": rel(p2.wake.Z, p1.wake.Z),
            "E_oracle": rel(p2.E.toarray(), osv.E.toarray()),
            "WP_oracle": rel(p2.wake.WP, osv.wake.WP)}
    print(errs, flush=True)
    assert max(errs["E"], errs["H"], errs["WP"], errs["Z"]) <= TOL_KERNEL, errs
    assert errs["E_oracle"] <= TOL_ORACLE and errs["WP_oracle"] <= TOL_ORACLE, errs


SHAPE_CASES = [
    # tiles overhanging the high x and y edges for every shape
    ("periodic", (72, 40, 30), ["periodic"] * 3, ["periodic"] * 3, False, False),
    # narrower than 64 cells: 64 x 8 does not apply
    ("mixed_narrow_lean", (50, 24, 41), ["periodic", "pec", "pec"],
     ["periodic", "pmc", "pml"], True, True),
    ("pml_conductive", (64, 32, 44), ["pec", "pec", "pml"], ["pec", "pec", "pml"], True, False),
    # 61 wide: 60 x 12 overhangs by one column, 64 x 8 does not apply
    ("overhang_61", (61, 26, 38), ["pec", "periodic", "pec"], ["pmc", "periodic", "pml"], True,
     True),
]


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4],
                         ids=["64x8", "32x16", "32x8", "16x16", "60x12"])
@pytest.mark.parametrize("name,shape,bl,bh,solids,nonuniform", SHAPE_CASES,
                         ids=[c[0] for c in SHAPE_CASES])
def test_two_step_tile_shapes(cfg, name, shape, bl, bh, solids, nonuniform):
    """Every tile shape of the single-grid two-step sweep (fit_set_tile2)
    against the FP32 oracle and the one-step sweep; a shape the grid cannot
    take falls back to the default shape."""
    from wakis_amd import _lib

    nsteps = 9
    print(f"[case] two-step shape {cfg} {name} {shape}", flush=True)
    osv, p2 = _pair(shape, bl, bh, solids, nonuniform, "fused2")
    _, p1 = _pair(shape, bl, bh, solids, nonuniform, "fused")
    p2.ctx.set_tile2(cfg)
    tx = (64, 32, 32, 16, 60)[cfg]
    if shape[0] >= tx:
        assert p2.ctx.get_tile2()[0] == cfg
    else:  # falls back to the modelled default
        assert p2.ctx.get_tile2()[0] not in (-1, cfg)
    assert p2.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
    for _ in range(nsteps):
        osv.one_step()
    p2._advance(nsteps)
    p1._advance(nsteps)
    assert p2.ctx.launch_counts() == (1, 4)
    errs = {}
    for f in "EHJ":
        o = getattr(osv, f).toarray()
        a2 = getattr(p2, f).toarray()
        a1 = getattr(p1, f).toarray()
        errs[f] = (rel(a2, o), rel(a2, a1))
    print(errs, flush=True)
    for f, (eo, ek) in errs.items():
        assert eo <= TOL_ORACLE and ek <= TOL_KERNEL, (f, eo, ek)


def test_two_step_tuner_keeps_state():
    """fit_tune2 times the shapes without changing the fields; the tuned run
    agrees with an untuned one (the shapes evaluate the same per-cell
    expressions; chunk boundaries, and with them the planes the lean path
    takes, may differ)."""
    shape = (72, 40, 30)
    print("[case] two-step tuner", flush=True)
    _, pa = _pair(shape, ["pec"] * 3, ["pec"] * 3, True, False, "fused2")
    _, pb = _pair(shape, ["pec"] * 3, ["pec"] * 3, True, False, "fused2")
    e0 = pa.E.toarray().copy()
    ms = pa.ctx.tune2(reps=2)
    assert np.array_equal(pa.E.toarray(), e0)
    assert (ms[:5] > 0).all() and ms[5] == -1.0, ms  # 28 x 12 is built for FP64 only
    assert pa.ctx.get_tile2()[0] == int(np.argmin(np.where(ms > 0, ms, np.inf)))
    pa._advance(6)
    pb._advance(6)
    for f in "EH":
        assert rel(getattr(pa, f).toarray(), getattr(pb, f).toarray()) <= TOL_KERNEL, f


@pytest.mark.parametrize("name,shape,bl,bh,solids,nonuniform",
                         [c for c in CASES if c[0] in ("pml_conductive", "periodic_lean",
                                                       "pmc_lean_nonuniform", "mixed_odd_nz")],
                         ids=["pml_conductive", "periodic_lean", "pmc_lean_nonuniform",
                              "mixed_odd_nz"])
@pytest.mark.parametrize("cfg", [-1, 4], ids=["default", "set_60x12"])
def test_two_step_per_component_resources(cfg, name, shape, bl, bh, solids, nonuniform):
    """The per-component, chunk-based buffer resources of the two-step sweep
    (what fields of 2^32 bytes or more take), forced on small grids with
    fit_set_chunk_rsrc: against the FP32 oracle and the one-step sweep.  The
    64 x 8 shape is the one built with them; a shape set that is not falls
    back to it."""
    from wakis_amd import _lib

    nsteps = 7
    print(f"[case] two-step per-component resources {cfg} {name} {shape}", flush=True)
    osv, p2 = _pair(shape, bl, bh, solids, nonuniform, "fused2")
    _, p1 = _pair(shape, bl, bh, solids, nonuniform, "fused")
    p2.ctx.set_chunk_rsrc(1)
    p2.ctx.set_tile2(cfg)
    assert p2.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
    assert p2.ctx.get_tile2()[0] == 0
    for _ in range(nsteps):
        osv.one_step()
    p2._advance(nsteps)
    p1._advance(nsteps)
    assert p2.ctx.launch_counts() == (1, 3)
    for f in "EHJ":
        o = getattr(osv, f).toarray()
        a2 = getattr(p2, f).toarray()
        a1 = getattr(p1, f).toarray()
        assert rel(a2, o) <= TOL_ORACLE and rel(a2, a1) <= TOL_KERNEL, (f, rel(a2, o), rel(a2, a1))


# ---------------- FP64 ----------------

TOL_ORACLE64 = 1e-10
TOL_KERNEL64 = 1e-12


@pytest.mark.parametrize("name,shape,bl,bh,solids,nonuniform", CASES, ids=[c[0] for c in CASES])
def test_two_step_fp64_matches_oracle_and_one_step(name, shape, bl, bh, solids, nonuniform):
    """The FP64 two-step sweep (default shape) against the FP64 oracle and the
    FP64 one-step sweep: three two-step launches and one single step."""
    from wakis_amd import _lib

    nsteps = 7
    print(f"[case] two-step fp64 {name} {shape}", flush=True)
    osv, p2 = _pair(shape, bl, bh, solids, nonuniform, "fused2", np.float64)
    _, p1 = _pair(shape, bl, bh, solids, nonuniform, "fused", np.float64)
    assert p2.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
    assert p2.ctx.get_tile2()[0] in (2, 3, 5)
    for _ in range(nsteps):
        osv.one_step()
    p2._advance(nsteps)
    p1._advance(nsteps)
    assert p2.ctx.launch_counts() == (1, 3), p2.ctx.launch_counts()
    assert p1.ctx.launch_counts() == (nsteps, 0)
    errs = {f: (rel(getattr(p2, f).toarray(), getattr(osv, f).toarray()),
                rel(getattr(p2, f).toarray(), getattr(p1, f).toarray())) for f in "EHJ"}
    print(errs, flush=True)
    for f, (eo, ek) in errs.items():
        assert eo <= TOL_ORACLE64 and ek <= TOL_KERNEL64, (f, eo, ek)


@pytest.mark.parametrize("cfg", [0, 2, 3, 5], ids=["64x8_falls_back", "32x8", "16x16", "28x12"])
@pytest.mark.parametrize("name,shape,bl,bh,solids,nonuniform", SHAPE_CASES,
                         ids=[c[0] for c in SHAPE_CASES])
def test_two_step_fp64_tile_shapes(cfg, name, shape, bl, bh, solids, nonuniform):
    """The FP64 shapes (32 x 8, 16 x 16, 28 x 12) against the FP64 oracle
    and the one-step sweep; an FP32-only shape set on an FP64 context falls
    back to the modelled default."""
    from wakis_amd import _lib

    nsteps = 9
    print(f"[case] two-step fp64 shape {cfg} {name} {shape}", flush=True)
    osv, p2 = _pair(shape, bl, bh, solids, nonuniform, "fused2", np.float64)
    _, p1 = _pair(shape, bl, bh, solids, nonuniform, "fused", np.float64)
    p2.ctx.set_tile2(cfg)
    got = p2.ctx.get_tile2()[0]
    assert got == cfg if cfg in (2, 3, 5) else got in (2, 3, 5), got
    assert p2.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
    for _ in range(nsteps):
        osv.one_step()
    p2._advance(nsteps)
    p1._advance(nsteps)
    assert p2.ctx.launch_counts() == (1, 4)
    errs = {f: (rel(getattr(p2, f).toarray(), getattr(osv, f).toarray()),
                rel(getattr(p2, f).toarray(), getattr(p1, f).toarray())) for f in "EHJ"}
    print(errs, flush=True)
    for f, (eo, ek) in errs.items():
        assert eo <= TOL_ORACLE64 and ek <= TOL_KERNEL64, (f, eo, ek)


def test_two_step_fp64_wakesolve_matches_oracle():
    """The FP64 wakesolve of a conductive cavity with z PML, beam on the axis
    and the 3 x 3 E_z probe on the two-step path: fields, WP and Z against
    the FP64 oracle and the one-step path."""
    import wakis_amd as wa  # noqa: F401

    shape = (48, 20, 48)
    Nx, Ny, Nz = shape
    wall = ~block_mask(shape, (4, 3, 0), (Nx - 4, Ny - 3, Nz))
    kw = dict(wakelength=0.12, q=1e-9, sigmaz=0.01, beta=1.0)
    res = {}
    for kernel in ("fused2", "fused"):
        osv, psv = make_pair(*shape, ["pec", "pec", "pml"], ["pec", "pec", "pml"],
                             dtype=np.float64, solids=[(wall, [1.0, 1.0, 1e3])], n_pml=5,
                             init=False, kernel=kernel, wake_kw=kw)
        print(f"[case] two-step fp64 wakesolve kernel={kernel}", flush=True)
        psv.wakesolve(compute_plane="longitudinal")
        res[kernel] = (psv, osv)
    p2, osv = res["fused2"]
    p1, _ = res["fused"]
    assert p2.ctx.launch_counts()[1] > p2.Nt // 4
    osv.wakesolve(compute_plane="longitudinal")
    errs = {"E": rel(p2.E.toarray(), osv.E.toarray()), "H": rel(p2.H.toarray(), osv.H.toarray()),
            "WP": rel(p2.wake.WP, osv.wake.WP), "Z": rel(p2.wake.Z, osv.wake.Z),
            "E_one_step": rel(p2.E.toarray(), p1.E.toarray()),
            "WP_one_step": rel(p2.wake.WP, p1.wake.WP)}
    print(errs, flush=True)
    assert max(errs["E"], errs["H"], errs["WP"], errs["Z"]) <= TOL_ORACLE64, errs
    assert max(errs["E_one_step"], errs["WP_one_step"]) <= TOL_KERNEL64, errs


@pytest.mark.parametrize("name,shape,bl,bh,solids,nonuniform",
                         [c for c in CASES if c[0] in ("pml_conductive", "periodic_lean",
                                                       "pmc_lean_nonuniform", "mixed_odd_nz")],
                         ids=["pml_conductive", "periodic_lean", "pmc_lean_nonuniform",
                              "mixed_odd_nz"])
@pytest.mark.parametrize("cfg", [-1, 3, 5], ids=["default", "set_16x16", "set_28x12"])
def test_two_step_fp64_per_component_resources(cfg, name, shape, bl, bh, solids, nonuniform):
    """The FP64 two-step sweep with per-component, chunk-based buffer
    resources (what FP64 fields of 2^32 bytes or more take), forced on small
    grids: against the FP64 oracle and the one-step sweep.  32 x 8 and 28 x 12
    are the FP64 shapes built with them; a shape set that is not falls back
    to the default among those two."""
    from wakis_amd import _lib

    nsteps = 7
    print(f"[case] two-step fp64 per-component resources {cfg} {name} {shape}", flush=True)
    osv, p2 = _pair(shape, bl, bh, solids, nonuniform, "fused2", np.float64)
    _, p1 = _pair(shape, bl, bh, solids, nonuniform, "fused", np.float64)
    p2.ctx.set_chunk_rsrc(1)
    p2.ctx.set_tile2(cfg)
    assert p2.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
    got = p2.ctx.get_tile2()[0]
    assert got == 5 if cfg == 5 else got in (2, 5), got
    for _ in range(nsteps):
        osv.one_step()
    p2._advance(nsteps)
    p1._advance(nsteps)
    assert p2.ctx.launch_counts() == (1, 3)
    for f in "EHJ":
        o = getattr(osv, f).toarray()
        a2 = getattr(p2, f).toarray()
        a1 = getattr(p1, f).toarray()
        assert rel(a2, o) <= TOL_ORACLE64 and rel(a2, a1) <= TOL_KERNEL64, \
            (f, rel(a2, o), rel(a2, a1))


@pytest.mark.parametrize("dtype,tol", [(np.float64, TOL_KERNEL64), (np.float32, TOL_KERNEL)],
                         ids=["fp64", "fp32"])
def test_full_size_pillbox_two_step_matches_one_step(dtype, tol):
    """The lossy pillbox at its full size (256 x 256 x 768: lossy walls, 10-cell
    cubic z PML, J = sigma*E), 41 steps from the same random fields: the
    two-step sweep (20 launches of the default shape and one single step)
    against the one-step sweep, relative L2 over the whole grid (the kernels
    evaluate the same per-cell expressions; the oracle is checked on the
    smaller grids above)."""
    from wakis_amd.workloads import build, lossy_pillbox

    nsteps = 41
    wl = lossy_pillbox(256, 256, 768)
    out = {}
    for kernel in ("fused2", "fused"):
        print(f"[case] full-size pillbox {np.dtype(dtype).name} kernel={kernel}", flush=True)
        _, _, s = build(wl, dtype=dtype, kernel=kernel)
        rng = np.random.default_rng(11)
        for name in ("E", "H"):
            getattr(s, name).fromarray(rng.standard_normal(3 * s.N).astype(dtype))
        s._advance(nsteps)
        out[kernel] = (s.E.toarray(), s.H.toarray(), s.ctx.launch_counts())
        s.ctx.close()
    e2, h2, c2 = out["fused2"]
    e1, h1, c1 = out["fused"]
    assert c2 == (1, 20) and c1 == (nsteps, 0), (c2, c1)
    assert np.isfinite(e2).all() and np.isfinite(h2).all()
    errs = (rel(e2, e1), rel(h2, h1))
    print(errs, flush=True)
    assert max(errs) <= tol, errs
