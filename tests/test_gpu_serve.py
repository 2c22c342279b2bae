"""Serving loop: requests that arrive while others run join the batch, each
result is the same as the request solved alone (bit for bit, same dtype and
tile), and admission respects the memory budget: the device bytes every
context reports (fit_device_bytes) stay within its admission model, and the
stepped members' bytes within the budget."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class VClock:
    """Virtual seconds = steps taken by the server so far x 1 ms."""

    def __init__(self):
        self.server = None

    def __call__(self):
        if self.server is None:
            return 0.0
        return sum(e[2] for e in self.server.log) * 1e-3


def _requests():
    from wakis_amd.serve import Request, serve_cavity

    specs = [("rect", 0.015, 0.020, 0.005, 5e4, None, 0.0),
             ("cyl", 0.018, 0.024, 0.006, 8e4, (4.0, 2e-3), 0.0),
             ("rect", 0.016, 0.018, 0.005, 6e4, (2.5, 5e-4), 0.06),
             ("cyl", 0.014, 0.022, 0.004, 4e4, None, 0.15)]
    out = []
    for i, (shape, r, L, p, cells, slab, t) in enumerate(specs):
        wl = serve_cavity(shape, r, L, p, cells, slab=slab, n_pml=4, wakelength=0.05,
                          name=f"t{i}")
        n = (len(wl["grid"]["x"]) - 1) * (len(wl["grid"]["y"]) - 1) * (len(wl["grid"]["z"]) - 1)
        out.append(Request(rid=i, arrival=t, workload=wl, cells=n))
    return out


@pytest.mark.parametrize("budget", [288e9, 9e6], ids=["hbm", "tight"])
def test_serve_results_equal_alone(budget):
    from wakis_amd.serve import Server, solve_alone

    print(f"[case] serve budget={budget:g}", flush=True)
    tile = 2
    clock = VClock()
    srv = Server(dtype=np.float32, budget=budget, chunk_steps=40, graph=True, group=8, tile=tile,
                 clock=clock)
    clock.server = srv
    done = srv.run(_requests())
    assert all(r.result is not None for r in done)
    # continuous batching: several members stepped together, and a request
    # arrived (virtually) while others were running
    assert max(e[1] for e in srv.log) >= 2
    assert any(r.arrival > 0 and r.admitted > 0 for r in done)
    # the device bytes of the stepped members and the batch's tables never
    # exceeded the bytes admitted, nor these the budget
    assert all(e[6] <= e[3] <= budget for e in srv.log)
    assert srv.peak_admitted <= budget
    for r in done:
        # what each context held at once, its finish scratch included
        assert 0 < r.result["device_peak"] <= r.result["device_model"], r.rid
        assert np.isfinite(r.result["WP"]).all() and np.isfinite(r.result["Z"]).all()
        _, alone = solve_alone(r, dtype=np.float32, tile=tile)
        assert r.result["Nt"] == alone["Nt"]
        assert np.array_equal(r.result["WP"], alone["WP"]), r.rid
        assert np.array_equal(r.result["Z"], alone["Z"]), r.rid


@pytest.mark.parametrize("tune2", [False, True], ids=["model", "tuned"])
def test_serve_two_step_members_equal_alone(tune2):
    """Requests wide enough for the two-step sweep (nx >= 64) are stepped by
    their own fit_step calls beside the batch; each is bit for bit the same
    as solved alone (with the two-step tile shape the server's tuner chose,
    when it tunes), and the context bytes stay within the model."""
    from wakis_amd import _lib
    from wakis_amd.serve import Request, Server, serve_cavity, solve_alone

    print(f"[case] serve two-step members tune2={tune2}", flush=True)
    specs = [("rect", 0.050, 0.060, 0.015, 8e5, None, 0.0),
             ("cyl", 0.050, 0.070, 0.020, 9e5, (3.0, 1e-3), 0.05)]
    reqs = []
    for i, (shape, r, L, p, cells, slab, t) in enumerate(specs):
        wl = serve_cavity(shape, r, L, p, cells, slab=slab, n_pml=4, wakelength=0.05,
                          name=f"w{i}")
        n = (len(wl["grid"]["x"]) - 1) * (len(wl["grid"]["y"]) - 1) * (len(wl["grid"]["z"]) - 1)
        reqs.append(Request(rid=i, arrival=t, workload=wl, cells=n))
    for r in _requests()[:2]:
        r.rid += 10
        r.arrival += 0.02
        reqs.append(r)
    tile = 2
    clock = VClock()
    srv = Server(dtype=np.float32, budget=288e9, chunk_steps=40, graph=True, group=8, tile=tile,
                 clock=clock, tune2=tune2)
    clock.server = srv
    done = srv.run(reqs)
    assert max(e[8] for e in srv.log) >= 1 and max(e[1] for e in srv.log) >= 2
    for r in done:
        assert 0 < r.result["device_peak"] <= r.result["device_model"], r.rid
        t, alone = solve_alone(r, dtype=np.float32, tile=tile, tile2=r.result["tile2"])
        assert r.result["Nt"] == alone["Nt"]
        assert np.array_equal(r.result["WP"], alone["WP"]), r.rid
        assert np.array_equal(r.result["Z"], alone["Z"]), r.rid
        print(r.rid, r.cells, r.steps, "peak", r.result["device_peak"], "model",
              r.result["device_model"], flush=True)


def test_watchdog_and_cancel_on_device():
    """A request whose fields hold a NaN (written after it is attached) is
    stopped by the device watchdog at the first round boundary, a request
    cancelled before it starts never steps, and the others give the same
    results as alone."""
    from wakis_amd.serve import Server, solve_alone

    print("[case] serve watchdog + cancel", flush=True)
    reqs = _requests()
    clock = VClock()
    srv = Server(dtype=np.float32, chunk_steps=40, graph=True, group=8, tile=2, clock=clock)
    clock.server = srv

    def poison(req, solver):
        if req.rid == 1:
            solver.E[3, 4, 5, "z"] = np.nan

    srv.after_attach = poison
    srv.cancel(3)
    clock.server = srv
    done = {r.rid: r for r in srv.run(reqs)}
    assert "non-finite" in done[1].result["error"] and done[1].steps == 40
    assert "cancelled" in done[3].result["error"] and done[3].steps == 0
    for rid in (0, 2):
        _, ref = solve_alone(done[rid], dtype=np.float32, tile=2)
        assert np.array_equal(done[rid].result["Z"], ref["Z"]), rid
