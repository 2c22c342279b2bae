"""Batched stepping (one persistent launch per step over many grids): each
member's results are the same alone and in the batch, bit for bit, with the
same tile configuration; graph replays equal eager launches."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cavities(dtype, seed=5, n=3, length=(0.02, 0.04)):
    import wakis_amd as wa
    from wakis_amd.workloads import build, rect_cavity

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        wl = rect_cavity(side=rng.uniform(0.024, 0.036), half_pipe=rng.uniform(0.004, 0.007),
                         length=rng.uniform(*length), dx=1.5e-3,
                         beta=[0.8, 0.9, 1.0][i % 3], n_pml=4, wakelength=0.08)
        out.append(build(wl, dtype=dtype))
    assert len({s.N for _, _, s in out}) == n  # uneven grids
    del wa
    return out


def _fields(s):
    return s.E.toarray().copy(), s.H.toarray().copy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_wakesolve_batch_equals_alone(dtype):
    import wakis_amd as wa

    print(f"[case] batch == alone {np.dtype(dtype).name}", flush=True)
    tile = 1  # 128 x 4 fits FP32 and FP64
    alone = _cavities(dtype)
    for _, w, s in alone:
        s.ctx.set_tile(tile)
        s.wakesolve(compute_plane="longitudinal")
    mixed = _cavities(dtype)
    nts = []
    wa.wakesolve_batch([s for _, _, s in mixed], compute_plane="longitudinal", graph=True,
                       group=4, tile=tile)
    for (_, wa_, sa), (_, wb, sb) in zip(alone, mixed):
        nts.append(sa.Nt)
        assert sa.Nt == sb.Nt
        assert np.array_equal(np.asarray(wa_.WP), np.asarray(wb.WP))
        assert np.array_equal(np.asarray(wa_.Z), np.asarray(wb.Z))
        ea, ha = _fields(sa)
        eb, hb = _fields(sb)
        assert np.array_equal(ea, eb) and np.array_equal(ha, hb)
    # members of different beta run different step counts: some left early
    assert len(set(nts)) > 1, nts


def test_wakesolve_batch_equals_alone_fp64_inner_chunks():
    """FP64 members of 155-167 planes, so chunks of 32 planes lie away from
    both z ends and the batched launch runs its lean path there: the same
    bits as each member solved alone."""
    import wakis_amd as wa

    print("[case] batch == alone float64 inner chunks", flush=True)
    tile = 4
    alone = _cavities(np.float64, seed=6, length=(0.07, 0.09))
    assert min(s.Nz for _, _, s in alone) >= 68
    for _, w, s in alone:
        s.ctx.set_tile(tile)
        s.wakesolve(compute_plane="longitudinal")
    mixed = _cavities(np.float64, seed=6, length=(0.07, 0.09))
    wa.wakesolve_batch([s for _, _, s in mixed], compute_plane="longitudinal", graph=True,
                       group=4, tile=tile)
    for (_, wa_, sa), (_, wb, sb) in zip(alone, mixed):
        assert sa.Nt == sb.Nt
        assert np.array_equal(np.asarray(wa_.WP), np.asarray(wb.WP))
        ea, ha = _fields(sa)
        eb, hb = _fields(sb)
        assert np.array_equal(ea, eb) and np.array_equal(ha, hb)


def test_batch_graph_replays_equal_eager():
    """Graph replays (group 3, with a remainder group) against eager launch
    groups, checked after every call with host work in between."""
    import wakis_amd as wa

    print("[case] batch graph replays", flush=True)
    a = _cavities(np.float32, seed=9, n=2)
    b = _cavities(np.float32, seed=9, n=2)
    rng = np.random.default_rng(0)
    for (_, _, sa), (_, _, sb) in zip(a, b):
        for nm in ("E", "H"):
            x = rng.standard_normal(3 * sa.N).astype(np.float32)
            getattr(sa, nm).fromarray(x)
            getattr(sb, nm).fromarray(x)
    ga = wa.BatchStepper([s for _, _, s in a], graph=True, group=3, tile=0)
    gb = wa.BatchStepper([s for _, _, s in b], graph=False, group=3, tile=0)
    for n in (7, 6, 1, 9):
        ga.step(n)
        gb.step(n)
        for (_, _, sa), (_, _, sb) in zip(a, b):
            ea, ha = _fields(sa)
            eb, hb = _fields(sb)
            assert np.array_equal(ea, eb) and np.array_equal(ha, hb), n


def test_batch_matches_single_steps_conductive():
    """A conductive member (J = sigma*E after its first step) and a lossless
    one, stepped in a batch and alone, FP64."""
    from wakis_amd.workloads import build, lossy_pillbox

    print("[case] batch conductive", flush=True)
    import wakis_amd as wa

    def make():
        out = []
        for nz in (40, 56):
            wl = lossy_pillbox(32, 32, nz)
            wl["solver"]["n_pml"] = 4
            _, _, s = build(wl, dtype=np.float64)
            out.append(s)
        _, _, s = build(_cavity_small(), dtype=np.float64)
        out.append(s)
        rng = np.random.default_rng(4)
        for s in out:
            s.E.fromarray(rng.standard_normal(3 * s.N))
            s.H.fromarray(rng.standard_normal(3 * s.N))
            s.ctx.set_tile(2)
        return out

    alone = make()
    mixed = make()
    assert alone[0].use_conductivity
    for s in alone:
        for _ in range(5):
            s.one_step()
    st = wa.BatchStepper(mixed, graph=False, tile=2)
    st.step(2)
    st.step(3)
    for sa, sb in zip(alone, mixed):
        ea, ha = _fields(sa)
        eb, hb = _fields(sb)
        assert np.array_equal(ea, eb) and np.array_equal(ha, hb)
        assert np.array_equal(sa.J.toarray(), sb.J.toarray())


def _cavity_small():
    from wakis_amd.workloads import rect_cavity

    return rect_cavity(side=0.03, half_pipe=0.006, length=0.03, dx=1.5e-3, n_pml=4)


def test_batch_tune_keeps_state():
    import wakis_amd as wa

    print("[case] batch tune", flush=True)
    mem = _cavities(np.float32, seed=2, n=2)
    rng = np.random.default_rng(1)
    for _, _, s in mem:
        s.E.fromarray(rng.standard_normal(3 * s.N).astype(np.float32))
    before = [_fields(s) for _, _, s in mem]
    st = wa.BatchStepper([s for _, _, s in mem], graph=False)
    ms = st.tune(reps=2)
    assert len(ms) >= 4 and all(v > 0 for v in ms.values())
    for (e0, h0), (_, _, s) in zip(before, mem):
        e1, h1 = _fields(s)
        assert np.array_equal(e0, e1) and np.array_equal(h0, h1)
    assert st.tile()[0] in ms
