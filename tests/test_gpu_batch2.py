"""The batched two-step sweep (FP32 batches whose members are all at least 32
cells wide and 16 deep, FP64 ones at least 28 wide: pairs of steps in two
launches over every member):
each member against the FP32 oracle (relative L2 <= 1e-4) and against the
one-step sweep alone (<= 1e-5: the same per-cell expressions), on uneven
members with lean and face chunks, a conductive member with PML and an odd
last step; graph replays equal eager launch groups; the wakesolve of a
beta-mixed batch against each member solved alone."""

import numpy as np
import pytest

from helpers import block_mask, make_pair, rel

pytestmark = pytest.mark.gpu

TOL_ORACLE = 1e-4
TOL_KERNEL = 1e-5

MEMBERS = [
    ((32, 16, 24), ["pec"] * 3, ["pec"] * 3, False, False),
    ((40, 33, 45), ["pec", "pec", "pml"], ["pec", "pec", "pml"], True, True),
    ((70, 20, 17), ["periodic", "pec", "pec"], ["periodic", "pmc", "pec"], False, False),
    ((33, 40, 61), ["pmc", "pec", "magnetic"], ["pec", "pmc", "pml"], False, True),
]


def _solids(shape):
    Nx, Ny, Nz = shape
    return [
        (block_mask(shape, (2, 2, 3), (Nx - 2, Ny - 2, Nz - 3)), "vacuum"),
        (block_mask(shape, (5, 1, 4), (20, 6, 9)), [4.0, 1.5]),
        (block_mask(shape, (12, 3, 5), (30, 9, 12)), [2.0, 1.0, 50.0]),
        (block_mask(shape, (5, 4, 6), (9, 8, 10)), "pec"),
    ]


def _members(kernel, seed=17, dtype=np.float32):
    out = []
    rng = np.random.default_rng(seed)
    for shape, bl, bh, solids, nonuniform in MEMBERS:
        osv, psv = make_pair(*shape, bl, bh, dtype=dtype,
                             solids=_solids(shape) if solids else (),
                             bg="pec" if solids else (1.0, 1.0), n_pml=3, nonuniform=nonuniform,
                             init=False, kernel=kernel, seed=3)
        for name in ("E", "H"):
            a = rng.standard_normal(3 * osv.N).astype(dtype)
            getattr(osv, name).fromarray(a)
            getattr(psv, name).fromarray(a)
        out.append((osv, psv))
    return out


def test_two_step_batch_matches_oracle_and_one_step():
    import wakis_amd as wa

    nsteps = 7  # the conductive member's first step, three pairs, no single
    print("[case] two-step batch vs oracle", flush=True)
    batch = _members("auto")
    alone = _members("fused")
    st = wa.BatchStepper([p for _, p in batch], graph=False)
    assert st.two_step()[0]
    st.step(nsteps)
    for (osv, pb), (_, pa) in zip(batch, alone):
        for _ in range(nsteps):
            osv.one_step()
        pa._advance(nsteps)
        assert pb.ctx.launch_counts()[1] == 3, pb.ctx.launch_counts()
        errs = {}
        for f in "EHJ":
            o = getattr(osv, f).toarray()
            b = getattr(pb, f).toarray()
            a = getattr(pa, f).toarray()
            errs[f] = (rel(b, o), rel(b, a))
        print(pb.Nx, pb.Ny, pb.Nz, errs, flush=True)
        for f, (eo, ek) in errs.items():
            assert eo <= TOL_ORACLE and ek <= TOL_KERNEL, (pb.Nx, f, eo, ek)
    assert st.two_step()[1] == 0  # a 16-deep member rules out 24 x 24
    assert st.two_step()[2] >= 32


def test_two_step_batch_odd_calls_and_graph_replays():
    """Calls of 5, 6, 1, 9 and 4 steps (odd calls end with a one-step
    launch; the pairs of the next call continue from its buffer parity),
    graph replays of 2-pair groups against eager groups, bit for bit, and
    the result against the oracle."""
    import wakis_amd as wa

    print("[case] two-step batch graph replays", flush=True)
    a = _members("auto", seed=5)
    b = _members("auto", seed=5)
    ga = wa.BatchStepper([p for _, p in a], graph=True, group=4)
    gb = wa.BatchStepper([p for _, p in b], graph=False, group=4)
    total = 0
    for n in (5, 6, 1, 9, 4):
        ga.step(n)
        gb.step(n)
        total += n
        for (_, pa), (_, pb) in zip(a, b):
            for f in "EH":
                assert np.array_equal(getattr(pa, f).toarray(), getattr(pb, f).toarray()), (n, f)
    for osv, pa in a:
        for _ in range(total):
            osv.one_step()
        for f in "EHJ":
            e = rel(getattr(pa, f).toarray(), getattr(osv, f).toarray())
            assert e <= TOL_ORACLE, (pa.Nx, f, e)


def test_two_step_batch_wakesolve_matches_alone():
    """Three cavities 40-47 cells wide with z PML and beta 0.8, 0.9, 1.0:
    WP, Z and fields of the batched wakesolve (two-step pairs, the beam from
    the device table) against each one solved alone."""
    import wakis_amd as wa
    from wakis_amd.workloads import build, rect_cavity

    print("[case] two-step batch wakesolve", flush=True)

    def make():
        rng = np.random.default_rng(8)
        out = []
        for i in range(3):
            wl = rect_cavity(side=rng.uniform(0.060, 0.070), half_pipe=rng.uniform(0.008, 0.012),
                             length=rng.uniform(0.02, 0.03), dx=1.5e-3,
                             beta=[0.8, 0.9, 1.0][i], n_pml=4, wakelength=0.06)
            out.append(build(wl, dtype=np.float32))
        return out

    alone = make()
    for _, _, s in alone:
        s.set_kernel("fused")
        s.wakesolve(compute_plane="longitudinal")
    mixed = make()
    assert min(s.Nx for _, _, s in mixed) >= 32
    wa.wakesolve_batch([s for _, _, s in mixed], compute_plane="longitudinal", graph=True,
                       group=8)
    for (_, wa_, sa), (_, wb, sb) in zip(alone, mixed):
        assert sa.Nt == sb.Nt
        assert sb.ctx.launch_counts()[1] > sb.Nt // 4
        errs = {"WP": rel(wb.WP, wa_.WP), "Z": rel(wb.Z, wa_.Z),
                "E": rel(sb.E.toarray(), sa.E.toarray()), "H": rel(sb.H.toarray(), sa.H.toarray())}
        print(sb.Nx, sb.Nz, sb.Nt, errs, flush=True)
        assert max(errs.values()) <= TOL_KERNEL, errs


def test_two_step_batch_fp64_matches_oracle_and_one_step():
    """The FP64 two-step batch (28 x 12 tiles; the FP64 two-step shape of at
    most 512 window cells the batch builds) on the same uneven members:
    against the FP64 oracle (1e-10) and the FP64 one-step sweep alone
    (1e-12); graph replays of 2-pair groups equal eager launch groups bit
    for bit."""
    import wakis_amd as wa

    nsteps = 7
    print("[case] two-step batch fp64 vs oracle", flush=True)
    batch = _members("auto", dtype=np.float64)
    alone = _members("fused", dtype=np.float64)
    st = wa.BatchStepper([p for _, p in batch], graph=False)
    on, cfg, kc = st.two_step()
    assert on and cfg == 2 and kc >= 16, (on, cfg, kc)
    st.step(nsteps)
    for (osv, pb), (_, pa) in zip(batch, alone):
        for _ in range(nsteps):
            osv.one_step()
        pa._advance(nsteps)
        assert pb.ctx.launch_counts()[1] == 3, pb.ctx.launch_counts()
        errs = {}
        for f in "EHJ":
            o = getattr(osv, f).toarray()
            b = getattr(pb, f).toarray()
            a = getattr(pa, f).toarray()
            errs[f] = (rel(b, o), rel(b, a))
        print(pb.Nx, pb.Ny, pb.Nz, errs, flush=True)
        for f, (eo, ek) in errs.items():
            assert eo <= 1e-10 and ek <= 1e-12, (pb.Nx, f, eo, ek)
    ga_m = _members("auto", seed=5, dtype=np.float64)
    gb_m = _members("auto", seed=5, dtype=np.float64)
    ga = wa.BatchStepper([p for _, p in ga_m], graph=True, group=4)
    gb = wa.BatchStepper([p for _, p in gb_m], graph=False, group=4)
    for n in (5, 6, 1):
        ga.step(n)
        gb.step(n)
        for (_, pa), (_, pb) in zip(ga_m, gb_m):
            for f in "EH":
                assert np.array_equal(getattr(pa, f).toarray(), getattr(pb, f).toarray()), (n, f)
