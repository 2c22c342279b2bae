"""The north star's workloads on the device against the oracle, at reduced
sizes: config 2 (the lossy pillbox with cubic z PML) in FP64 against a
committed oracle run and in mixed precision (FP32 fields, FP64 coefficients,
wake and DFT) against the same run; config 3 (batched rectangular cavities
with beta 0.8, 0.9 and 1.0, 10-cell z PML) in FP64 against the oracle; and
config 4 (Poisson-arriving cavity requests with lossy dielectric slabs,
served by continuous batching in mixed precision) against the oracle's FP64
impedance.  FP64: relative L2 <= 1e-10.  Mixed: max |dZ| / max |Z| <= 1e-3
and |dZ| / |Z| <= 1e-3 on every kept bin (all below the bunch cutoff) where
|Z| >= 1 % of its maximum; WP relative L2 <= 1e-3."""

import os

import numpy as np
import pytest

from helpers import oracle_from_workload, rel, z_errors

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL64 = 1e-10
TOLZ = 1e-3


def _config2(dtype, precision):
    from wakis_amd.workloads import build, lossy_pillbox

    gold = np.load(os.path.join(HERE, "golden", "config2_lossy_pillbox_fp64.npz"),
                   allow_pickle=False)
    w = lossy_pillbox(*(int(v) for v in gold["shape"]))
    w["wake"]["wakelength"] = float(gold["wakelength"])
    _, wake, s = build(w, dtype=dtype, precision=precision)
    s.wakesolve(compute_plane="longitudinal")
    return gold, wake, s


def test_config2_pillbox_fp64_matches_oracle():
    print("[case] config 2 fp64 64x64x192", flush=True)
    gold, wake, s = _config2(np.float64, "reference")
    assert s.Nt == int(gold["Nt"]) and float(s.dt) == float(gold["dt"])
    k = int(gold["sample"])
    E, H = s.E.toarray(), s.H.toarray()
    errs = {"WP": rel(wake.WP, gold["WP"]), "Z": rel(wake.Z, gold["Z"]),
            "E": rel(E[::k], gold["E_sample"]), "H": rel(H[::k], gold["H_sample"]),
            "E_norm": abs(np.linalg.norm(E) / float(gold["E_norm"]) - 1),
            "H_norm": abs(np.linalg.norm(H) / float(gold["H_norm"]) - 1)}
    print(errs, flush=True)
    assert np.array_equal(np.asarray(wake.s), gold["s"]) and np.array_equal(wake.f, gold["f"])
    assert max(errs.values()) <= TOL64, errs


def test_config2_pillbox_mixed_fp32_within_1e3():
    from wakis_amd import _lib

    print("[case] config 2 mixed fp32 64x64x192", flush=True)
    gold, wake, s = _config2(np.float32, "mixed")
    assert s.ctx.kernel_in_use() == _lib.KERNEL_FUSED2  # the two-step sweep
    assert s.ctx.launch_counts()[1] >= s.Nt // 2 - 1
    zmax_err, zbin_err = z_errors(wake.Z, gold["Z"])
    errs = {"WP": rel(wake.WP, gold["WP"]), "Z_max_normalised": zmax_err,
            "Z_per_bin": zbin_err, "E": rel(s.E.toarray()[::int(gold["sample"])],
                                            gold["E_sample"])}
    print(errs, flush=True)
    assert errs["WP"] <= TOLZ and zmax_err <= TOLZ and zbin_err <= TOLZ, errs


def _cavities3():
    from wakis_amd.workloads import rect_cavity

    # three uneven cavities at 3 mm cells (config 3's shapes, coarser grid)
    specs = [(0.066, 0.012, 0.066, 0.8), (0.090, 0.018, 0.096, 0.9), (0.120, 0.024, 0.072, 1.0)]
    return [rect_cavity(side, pipe, length, dx=3e-3, beta=beta, n_pml=10, wakelength=0.25,
                        name=f"c{i}") for i, (side, pipe, length, beta) in enumerate(specs)]


def test_config3_batch_members_fp64_match_oracle():
    """wakesolve_batch of three uneven members with beta 0.8, 0.9, 1.0 (the
    beam's injection time without sqrt(beta), the wake's with it) against
    each member's oracle wakesolve."""
    import wakis_amd as wa
    from wakis_amd.workloads import build

    print("[case] config 3 batch fp64", flush=True)
    wls = _cavities3()
    solvers = [build(w, dtype=np.float64)[2] for w in wls]
    assert len({s.N for s in solvers}) == 3
    wa.wakesolve_batch(solvers, compute_plane="longitudinal", graph=True, group=8)
    for w, s in zip(wls, solvers):
        osv = oracle_from_workload(w)
        osv.wakesolve(compute_plane="longitudinal")
        errs = {"E": rel(s.E.toarray(), osv.E.toarray()), "H": rel(s.H.toarray(), osv.H.toarray()),
                "WP": rel(s.wake.WP, osv.wake.WP), "Z": rel(s.wake.Z, osv.wake.Z)}
        print(w["name"], w["wake"]["beta"], s.N, s.Nt, errs, flush=True)
        assert s.Nt == osv.Nt
        assert max(errs.values()) <= TOL64, (w["name"], errs)


def test_config4_served_requests_match_oracle():
    """Three Poisson requests (seed 4 draws, cell counts scaled down, lossy
    dielectric slabs where drawn) served by the continuous-batching server
    in mixed precision: each Z against the oracle's FP64 Z."""
    from wakis_amd.serve import Server, make_requests

    print("[case] config 4 served requests", flush=True)
    reqs = make_requests(3, seed=4, median_cells=4e4, sigma_ln=0.3, clip=(2e4, 6e4),
                         slab_prob=1.0, wakelength=0.06)
    t0 = reqs[0].arrival
    for r in reqs:
        r.workload["solver"]["n_pml"] = 5
        # the trace's spacing, from 0 and compressed tenfold on the virtual
        # clock (which advances only while the server steps: every request
        # must arrive before the earlier ones are done)
        r.arrival = (r.arrival - t0) * 0.1

    class VClock:  # virtual seconds = steps the server has taken x 1 ms
        server = None

        def __call__(self):
            return 0.0 if self.server is None else sum(e[2] for e in self.server.log) * 1e-3

    clock = VClock()
    srv = Server(dtype=np.float32, budget=288e9, chunk_steps=64, graph=True, group=8,
                 clock=clock)
    clock.server = srv
    done = srv.run(reqs)
    assert max(e[1] for e in srv.log) >= 2  # requests overlapped in the batch
    for r in done:
        osv = oracle_from_workload(r.workload)
        osv.wakesolve(compute_plane="longitudinal")
        zmax_err, zbin_err = z_errors(r.result["Z"], osv.wake.Z)
        errs = {"WP": rel(r.result["WP"], osv.wake.WP), "Z_max_normalised": zmax_err,
                "Z_per_bin": zbin_err}
        print(r.rid, r.cells, r.result["Nt"], errs, flush=True)
        assert r.result["Nt"] == osv.Nt
        assert max(errs.values()) <= TOLZ, (r.rid, errs)
