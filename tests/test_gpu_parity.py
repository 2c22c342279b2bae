"""HIP step kernels vs the CPU oracle (FP64: relative L2 <= 1e-10; FP32:
<= 1e-4 against the FP32 oracle), on small grids with every boundary kind,
materials, non-uniform spacing, the flat-index wrap, J modes and the beam."""

import numpy as np
import pytest

from helpers import OBeam, OPlaneWave, block_mask, compare_fields, make_pair, rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-4

BC_CASES = [
    ("pec", ["pec"] * 3, ["pec"] * 3),
    ("periodic", ["periodic"] * 3, ["periodic"] * 3),
    ("mixed", ["periodic", "pec", "pec"], ["periodic", "pmc", "pml"]),
    ("pmc", ["pmc", "pec", "magnetic"], ["pec", "pmc", "pec"]),
    ("abc", ["pec", "abc", "pec"], ["abc", "pec", "abc"]),
    ("pml_z", ["pec", "pec", "pml"], ["pec", "pec", "pml"]),
    ("open_x", ["pec", "periodic", "pec"], ["periodic", "periodic", "pec"]),
]


def _solids(shape):
    Nx, Ny, Nz = shape
    return [
        (block_mask(shape, (1, 1, 2), (Nx - 1, Ny - 1, Nz - 2)), "vacuum"),
        (block_mask(shape, (2, 1, 3), (4, 3, 6)), [4.0, 1.5]),
        (block_mask(shape, (3, 2, 4), (5, 4, 7)), [2.0, 1.0, 50.0]),
        (block_mask(shape, (1, 3, 5), (3, 5, 7)), "pec"),
    ]


KERNELS = ["split", "fused"]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name,bl,bh", BC_CASES, ids=[c[0] for c in BC_CASES])
@pytest.mark.parametrize("nonuniform", [False, True], ids=["uniform", "nonuniform"])
def test_step_fp64(name, bl, bh, nonuniform, kernel):
    shape = (7, 6, 11)
    print(f"[case] step_fp64 {name} nonuniform={nonuniform} kernel={kernel} shape={shape}",
          flush=True)
    osv, psv = make_pair(*shape, bl, bh, solids=_solids(shape), nonuniform=nonuniform,
                         seed=3, n_pml=3, kernel=kernel)
    for _ in range(25):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(name, err)
    assert max(err.values()) <= TOL64, err


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name,bl,bh", BC_CASES[:4], ids=[c[0] for c in BC_CASES[:4]])
def test_step_fp32(name, bl, bh, kernel):
    shape = (8, 7, 10)
    print(f"[case] step_fp32 {name} kernel={kernel} shape={shape}", flush=True)
    osv, psv = make_pair(*shape, bl, bh, dtype=np.float32, solids=_solids(shape), seed=5,
                         kernel=kernel)
    for _ in range(25):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(name, err)
    assert max(err.values()) <= TOL32, err


def test_conductive_background_and_update_tensors():
    shape = (6, 7, 9)
    print("[case] conductive bg + update_tensors", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg=[2.0, 1.0, 3.0], seed=9)
    rng = np.random.default_rng(1)
    # a few distinct values per component: the device table holds 256 rows
    ieps = rng.choice([0.5, 0.7, 0.9, 1.0], 3 * osv.N) / 8.854e-12
    for s in (osv, psv):
        f = s.ieps
        f.fromarray(ieps)
        s.ieps = f
        s.update_tensors()
    for _ in range(15):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL64, err


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("tensor", ["all", "ieps"])
def test_graded_material_tensors_beyond_256_rows(dtype, tensor):
    """Per-cell material tensors of any content through update_tensors
    (wakis/solverFIT3D.py:378-436): a dielectric graded in z and x (eps_r
    1..3, a different value in nearly every cell, several thousand distinct
    rows) with a conductive slab, stepped with J = sigma*E and a beam line;
    the solver takes the wide material ids (split kernels) and matches the
    oracle.  tensor="ieps" rebuilds only the permittivity (sigma is left as
    the background gives it)."""
    import wakis_amd as wa

    shape = (14, 12, 20)
    print(f"[case] graded tensors {np.dtype(dtype).name} tensor={tensor}", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, bg=[1.0, 1.0, 10.0], seed=11,
                         dtype=dtype, init=False)
    Nx, Ny, Nz = shape
    i = np.arange(Nx)[:, None, None]
    k = np.arange(Nz)[None, None, :]
    j = np.arange(Ny)[None, :, None]
    epsr = 1.0 + 2.0 * (k + 0.37 * i + 0.011 * j) / (Nz + 0.37 * Nx)
    ieps_m = np.broadcast_to(1.0 / (epsr * 8.8541878128e-12), shape)
    sig_m = np.where((k >= 8) & (k < 11), 25.0, 10.0) * np.ones(shape)
    for s in (osv, psv):
        f = s.ieps
        for d in "xyz":
            f[:, :, :, d] = ieps_m
        s.ieps = f
        if tensor == "all":
            g = s.sigma
            for d in "xyz":
                g[:, :, :, d] = sig_m
            s.sigma = g
        s.update_tensors(tensor)
    assert psv.ctx.kernel_in_use() == wa._lib.KERNEL_SPLIT
    rows = len(psv._mt.table())
    assert rows > 256, rows
    ob = OBeam(q=1e-9, sigmaz=0.01, beta=1.0)
    pb = wa.Beam(q=1e-9, sigmaz=0.01, beta=1.0)
    for n in range(30):
        ob.update(osv, n * osv.dt)
        pb.update(psv, n * psv.dt)
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(rows, err)
    assert max(err.values()) <= (TOL64 if dtype == np.float64 else TOL32), err


def test_wide_materials_refuse_fused_kernels():
    """An explicit fused kernel on a context with more than 256 material
    rows is refused; batches refuse such members."""
    import wakis_amd as wa

    shape = (8, 8, 10)
    print("[case] wide materials refuse fused", flush=True)
    _, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, init=False, kernel="fused")
    f = psv.ieps
    f.fromarray(np.linspace(1.0, 2.0, 3 * psv.N) / 8.8541878128e-12)
    psv.ieps = f
    psv.update_tensors()
    with pytest.raises(wa._lib.FitError, match="256 material rows"):
        psv.one_step()


def test_beam_user_loop_fp64():
    """Beam.update + one_step in a user loop (wakis/sources.py:76-110)."""
    import wakis_amd as wa

    shape = (9, 9, 16)
    print("[case] beam user loop", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, init=False,
                         solids=[(block_mask(shape, (1, 1, 0), (8, 8, 16)), "vacuum")], bg="pec")
    ob = OBeam(q=1e-9, sigmaz=0.01, beta=1.0)
    pb = wa.Beam(q=1e-9, sigmaz=0.01, beta=1.0)
    for n in range(40):
        ob.update(osv, n * osv.dt)
        pb.update(psv, n * psv.dt)
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL64, err


def test_planewave_pml_fp32():
    """PlaneWave source writing E/H planes each step into a PML-terminated
    column (tests/test_005_pml_boundaries.py:19-137 shape, shorter)."""
    import wakis_amd as wa

    print("[case] planewave pml fp32", flush=True)
    Nx, Ny, Nz = 6, 6, 60
    osv, psv = make_pair(Nx, Ny, Nz, ["periodic", "periodic", "pec"],
                         ["periodic", "periodic", "pml"], dtype=np.float32, bg=[1.0, 1.0, 0.0],
                         n_pml=10, init=False)
    f = 15 / ((osv.z.max() - osv.z.min()) / 299792458.0)
    for s in (osv, psv):
        s.dt = 1 / f / 200
    ow = OPlaneWave(xs=slice(0, Nx), ys=slice(0, Ny), zs=0, amplitude=100.0, f=f, nodes=7,
                    phase=np.pi / 2)
    pw = wa.PlaneWave(xs=slice(0, Nx), ys=slice(0, Ny), zs=0, amplitude=100.0, f=f, nodes=7,
                      phase=np.pi / 2)
    for n in range(300):
        ow.update(osv, n * osv.dt)
        pw.update(psv, n * psv.dt)
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL32, err


def test_emsolve_abc_fp64():
    """emsolve with low-face ABC planes (CPU aliasing semantics)."""
    print("[case] emsolve abc", flush=True)
    shape = (7, 8, 12)
    osv, psv = make_pair(*shape, ["abc", "pec", "abc"], ["pec", "abc", "pec"], seed=11)
    osv.emsolve(20)
    psv.emsolve(20)
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL64, err


def test_emsolve_abc_both_faces_one_axis_raises_like_reference():
    """ABC on both z faces: the reference's update_abc writes the low x face,
    then fails on the z axis' missing 'xlo' key after the first step; the
    device run stops at the same point with the same fields."""
    print("[case] emsolve abc low+high z", flush=True)
    shape = (6, 5, 9)
    osv, psv = make_pair(*shape, ["abc", "pec", "abc"], ["pec", "pec", "abc"], seed=12)
    for sv in (osv, psv):
        with pytest.raises(KeyError, match="xlo"):
            sv.emsolve(5)
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL64, err


def test_wrap_large_nx_fp64():
    """A grid whose rows exceed one workgroup, periodic (helical wrap)."""
    print("[case] wrap large nx", flush=True)
    osv, psv = make_pair(300, 3, 4, ["periodic"] * 3, ["periodic"] * 3, seed=13)
    for _ in range(10):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL64, err


def test_field_indexing_roundtrip():
    """DeviceField tuple indexing agrees with the host Field."""
    from wakis_amd.field import Field

    print("[case] field indexing", flush=True)
    osv, psv = make_pair(5, 6, 7, ["pec"] * 3, ["pec"] * 3, seed=2)
    host = Field(5, 6, 7)
    host.fromarray(psv.E.toarray())
    keys = [(1, 2, 3, "x"), (slice(None), 2, slice(1, 5), "y"), (-1, slice(0, 3), 4, "z"),
            (slice(1, 4), slice(2, 6), slice(0, 7), "z"), (np.int64(2), 3, slice(None), "x")]
    for k in keys:
        assert np.array_equal(np.asarray(psv.E[k]), np.asarray(host[k])), k
    psv.E[1:3, :, 2, "y"] = 7.0
    host[1:3, :, 2, "y"] = 7.0
    psv.H[:, 0, :, "z"] = np.arange(35.0).reshape(5, 7)
    assert np.array_equal(psv.E.toarray(), host.toarray())
    assert np.array_equal(psv.H[:, 0, :, "z"], np.arange(35.0).reshape(5, 7))


@pytest.mark.parametrize("tile", list(range(12)))
@pytest.mark.parametrize("name,bl,bh", BC_CASES, ids=[c[0] for c in BC_CASES])
def test_fused_multi_tile_fp64(name, bl, bh, tile):
    """A grid spanning several fused tiles in x, y and z for every tile
    configuration: tile seams, z-chunk restarts, partial tiles, the XCD
    remap, and the flat-window wrap at every face."""
    shape = (131, 19, 70)
    print(f"[case] fused multi-tile {name} tile={tile} shape={shape}", flush=True)
    Nx, Ny, Nz = shape
    solids = [(block_mask(shape, (3, 2, 5), (Nx - 4, Ny - 3, Nz - 6)), "vacuum"),
              (block_mask(shape, (40, 5, 20), (90, 12, 40)), [3.0, 1.2, 20.0]),
              (block_mask(shape, (60, 8, 30), (70, 15, 50)), "pec")]
    # tiles 0 (320 x 3), 3 and 5 (256 x 4) do not fit the LDS in FP64: the
    # default tile runs
    osv, psv = make_pair(*shape, bl, bh, solids=solids, seed=21, n_pml=4, kernel="fused",
                         tile=None if tile in (0, 3, 5) else tile)
    for _ in range(6):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(name, err)
    assert max(err.values()) <= TOL64, err


@pytest.mark.parametrize("tile", list(range(6, 12)))
def test_row_sweep_j_buffer_fp64(tile):
    """The row-wave tiles (ids 6-11) in FP64 with J read from its buffer on
    every step (a non-conductive solver given a J field), on a grid whose
    last z chunk runs the tail loop for several planes: the variants that
    run (spill-free builds; a variant that spills is never launched) match
    the oracle within 1e-10 on E, H and J."""
    shape = (70, 13, 70)  # chunks of 32 planes: [0, 32), [32, 64), [64, 70)
    print(f"[case] row sweep J buffer fp64 tile={tile} shape={shape}", flush=True)
    Nx, Ny, Nz = shape
    solids = [(block_mask(shape, (3, 2, 5), (Nx - 4, Ny - 3, Nz - 6)), "vacuum"),
              (block_mask(shape, (20, 4, 30), (50, 10, 66)), [3.0, 1.2])]
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, solids=solids, seed=23,
                         kernel="fused", tile=tile)
    runs = psv.ctx.tune(reps=1)[tile] > 0
    psv.ctx.set_tile(tile)
    print("runs as itself:", runs, flush=True)
    for _ in range(8):
        osv.one_step()
        psv.one_step()
    assert psv.ctx.get_state()[0] == 2  # FIT_J_BUFFER throughout
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= TOL64, err


def test_fused_equals_split_bitwise_structure():
    """Fused and split kernels agree to rounding on a mid-size grid (FP64)."""
    shape = (70, 20, 40)
    print("[case] fused vs split", flush=True)
    _, a = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, seed=8, kernel="split")
    _, b = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, seed=8, kernel="fused")
    for _ in range(20):
        a.one_step()
        b.one_step()
    for n in "EHJ":
        assert rel(getattr(b, n).toarray(), getattr(a, n).toarray()) <= 1e-13


@pytest.mark.parametrize("tile", list(range(12)))
def test_fused_tiles_fp32(tile):
    """Every tile configuration in FP32 (256x4 fits the LDS only in FP32)."""
    shape = (140, 21, 40)
    print(f"[case] fused tiles fp32 tile={tile}", flush=True)
    osv, psv = make_pair(*shape, ["periodic", "pec", "pml"], ["pec", "periodic", "pml"],
                         dtype=np.float32, solids=_solids(shape), seed=31, kernel="fused",
                         tile=tile)
    for _ in range(8):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL32, err


@pytest.mark.parametrize("tile", list(range(12)))
def test_fused_inner_chunks_fp32(tile):
    """Chunks away from the z ends (the lean buffer-addressed path) in FP32
    for every tile configuration: 80 planes give chunks starting at plane 32
    (KC = 32) that touch neither z face; periodic x/y wrap, PMC/PEC faces,
    conductive solids."""
    shape = (137, 13, 80)
    print(f"[case] fused inner chunks fp32 tile={tile} shape={shape}", flush=True)
    Nx, Ny, Nz = shape
    solids = [(block_mask(shape, (2, 1, 4), (Nx - 3, Ny - 2, Nz - 5)), "vacuum"),
              (block_mask(shape, (30, 3, 35), (80, 9, 60)), [3.0, 1.5, 40.0]),
              (block_mask(shape, (90, 2, 33), (100, 11, 70)), "pec")]
    osv, psv = make_pair(*shape, ["periodic", "pmc", "pml"], ["pec", "periodic", "pml"],
                         dtype=np.float32, solids=solids, seed=41, kernel="fused", tile=tile)
    for _ in range(8):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= TOL32, err


@pytest.mark.parametrize("tile", list(range(6)))
def test_fused_chunk_resources_fp32(tile):
    """The lean path with a buffer resource per field component based at each
    chunk (the addressing of fields of 2^32 bytes or more), forced on a small
    grid: face chunks, interior chunks, periodic/PMC/PEC/PML faces,
    conductive and PEC solids, initial J; against the oracle."""
    shape = (300, 11, 150)
    print(f"[case] fused chunk resources fp32 tile={tile} shape={shape}", flush=True)
    Nx, Ny, Nz = shape
    solids = [(block_mask(shape, (2, 1, 4), (Nx - 3, Ny - 2, Nz - 5)), "vacuum"),
              (block_mask(shape, (30, 3, 20), (200, 9, 110)), [4.0, 1.3, 30.0]),
              (block_mask(shape, (220, 2, 60), (260, 10, 140)), "pec")]
    osv, psv = make_pair(*shape, ["periodic", "pmc", "pml"], ["pec", "periodic", "pml"],
                         dtype=np.float32, solids=solids, seed=43, kernel="fused", tile=tile)
    psv.ctx.set_chunk_rsrc(True)
    for _ in range(8):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= TOL32, err


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_fused_inner_chunks_fp64(tile):
    """The lean path in FP64 (80 planes in chunks of 32: the chunk at planes
    32..63 touches neither z end) on every tile that fits FP64: periodic x/y
    wrap, PMC/PEC/PML faces, conductive and PEC solids, initial J; against
    the FP64 oracle at 1e-10."""
    shape = (137, 13, 80)
    print(f"[case] fused inner chunks fp64 tile={tile} shape={shape}", flush=True)
    Nx, Ny, Nz = shape
    solids = [(block_mask(shape, (2, 1, 4), (Nx - 3, Ny - 2, Nz - 5)), "vacuum"),
              (block_mask(shape, (30, 3, 35), (80, 9, 60)), [3.0, 1.5, 40.0]),
              (block_mask(shape, (90, 2, 33), (100, 11, 70)), "pec")]
    osv, psv = make_pair(*shape, ["periodic", "pmc", "pml"], ["pec", "periodic", "pml"],
                         dtype=np.float64, solids=solids, seed=47, kernel="fused", tile=tile)
    for _ in range(8):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= TOL64, err


@pytest.mark.parametrize("tile", [0, 5])
def test_fused_inner_chunks_k64_fp32(tile):
    """The lean path of the 64-plane tiles (320 x 3, 256 x 4) in FP32: 150
    planes give a chunk at planes 64..127 away from both z faces; a grid
    narrower than the tile in x (partial rows) and several tiles in y."""
    shape = (300, 11, 150)
    print(f"[case] fused inner chunks K64 fp32 tile={tile} shape={shape}", flush=True)
    Nx, Ny, Nz = shape
    solids = [(block_mask(shape, (2, 1, 4), (Nx - 3, Ny - 2, Nz - 5)), "vacuum"),
              (block_mask(shape, (30, 3, 70), (200, 9, 110)), [4.0, 1.3, 30.0]),
              (block_mask(shape, (220, 2, 60), (260, 10, 120)), "pec")]
    osv, psv = make_pair(*shape, ["periodic", "pec", "pml"], ["pec", "pmc", "pml"],
                         dtype=np.float32, solids=solids, seed=47, kernel="fused", tile=tile)
    for _ in range(8):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= TOL32, err


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_fused_inner_chunks_jbuffer_beam(tile):
    """The lean path in the J buffer mode (a non-conductive solver whose J is
    written from the host: Beam.update in a user loop), FP64, 80 planes."""
    import wakis_amd as wa

    shape = (70, 9, 80)
    print(f"[case] fused inner chunks J buffer + beam tile={tile}", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, seed=43, kernel="fused", tile=tile,
                         solids=[(block_mask(shape, (1, 1, 0), (69, 8, 80)), "vacuum")],
                         bg="pec")
    ob = OBeam(q=1e-9, sigmaz=0.004, beta=1.0, ti=0.0)
    pb = wa.Beam(q=1e-9, sigmaz=0.004, beta=1.0, ti=0.0)
    for n in range(12):
        ob.update(osv, n * osv.dt)
        pb.update(psv, n * psv.dt)
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= TOL64, err


def test_tiny_grids_fused():
    """Grids smaller than one tile in every direction and with 2 cells on an
    axis (the window wraps several rows)."""
    for shape in [(2, 3, 4), (5, 2, 3), (3, 4, 2), (9, 7, 5)]:
        print(f"[case] tiny {shape}", flush=True)
        osv, psv = make_pair(*shape, ["periodic", "pec", "pec"], ["periodic", "pmc", "pec"],
                             seed=41, kernel="fused")
        for _ in range(6):
            osv.one_step()
            psv.one_step()
        err = compare_fields(osv, psv)
        assert max(err.values()) <= TOL64, (shape, err)


def test_tune_keeps_state():
    shape = (70, 20, 40)
    print("[case] tune", flush=True)
    osv, psv = make_pair(*shape, ["pec"] * 3, ["pec"] * 3, seed=8, kernel="fused")
    before = psv.E.toarray().copy()
    ms = psv.tune(reps=2)
    assert len(ms) >= 5 and all(v > 0 for v in ms.values())
    assert np.array_equal(psv.E.toarray(), before)
    for _ in range(5):
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    assert max(err.values()) <= TOL64, err


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
def test_stratified_box_tensor_block_beam(dtype):
    """The stratified-box workload at 16^3: index-slice dielectric layers
    with conductivity, a diagonal-tensor block, graded PML on all six faces
    and the bunch on the axis (entering within the run), 40 steps from
    random fields."""
    import wakis_amd as wa
    from helpers import stratified_pair

    print(f"[case] stratified box 16^3 {np.dtype(dtype).name}", flush=True)
    osv, psv, wl = stratified_pair(n=16, nlayers=4, n_pml=3, dtype=dtype)
    rng = np.random.default_rng(21)
    for name in ("E", "H"):
        a = rng.standard_normal(3 * osv.N).astype(dtype)
        getattr(osv, name).fromarray(a)
        getattr(psv, name).fromarray(a)
    w = wl["wake"]
    ti = 2 * w["sigmaz"] / 299792458.0
    ob = OBeam(q=w["q"], sigmaz=w["sigmaz"], beta=1.0, ti=ti)
    pb = wa.Beam(q=w["q"], sigmaz=w["sigmaz"], beta=1.0, ti=ti)
    for n in range(40):
        ob.update(osv, n * osv.dt)
        pb.update(psv, n * psv.dt)
        osv.one_step()
        psv.one_step()
    err = compare_fields(osv, psv)
    print(err)
    assert max(err.values()) <= (TOL64 if dtype == np.float64 else TOL32), err


def test_save_reset_load_state(tmp_path):
    """save_state / reset_fields / load_state (wakis/solverFIT3D.py:908-1035)
    round trip: the restored run continues exactly like an uninterrupted one."""
    shape = (8, 7, 10)
    print("[case] save/reset/load state", flush=True)
    _, a = make_pair(*shape, ["pec"] * 3, ["pec", "pec", "pml"], seed=17)
    _, b = make_pair(*shape, ["pec"] * 3, ["pec", "pec", "pml"], seed=17)
    for _ in range(5):
        a.one_step()
        b.one_step()
    fn = str(tmp_path / "state.h5")
    b.save_state(fn)
    b.reset_fields()
    assert not np.any(b.E.toarray()) and not np.any(b.H.toarray()) and not np.any(b.J.toarray())
    b.load_state(fn)
    for _ in range(5):
        a.one_step()
        b.one_step()
    for n in "EHJ":
        assert np.array_equal(getattr(a, n).toarray(), getattr(b, n).toarray()), n


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_tiled_copy_calibration_runs(dtype):
    """The sweep-traversal copy used to calibrate the roofline runs for every
    tile configuration, ragged grid edges included."""
    from wakis_amd import _lib

    print(f"[case] tiled copy calibration {np.dtype(dtype).name}", flush=True)
    for cfg in range(6):
        ms = _lib.copy_bench_tiled(70, 37, 45, cfg, dtype=dtype, reps=2)
        assert ms > 0
