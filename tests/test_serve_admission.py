"""Serving admission without a device: a fake context that counts the bytes
the library would allocate (the sizes of fit_create, fit_set_geometry,
fit_set_materials, fit_wake_probe, the sweep's second buffers and the finish
scratch) shows that a burst of requests never holds more device memory than
the budget, that nothing is allocated before admission, that a request too
large for the budget is rejected, and that a virtual clock is advanced over
idle gaps."""

import threading

import numpy as np
import pytest

from wakis_amd import _lib, serve


class Ledger:
    def __init__(self):
        self.lock = threading.Lock()
        self.live = 0
        self.peak = 0
        self.created = []  # (request grid shape, live bytes after creation)

    def add(self, n):
        with self.lock:
            self.live += n
            self.peak = max(self.peak, self.live)

    def transient(self, n):
        with self.lock:
            self.peak = max(self.peak, self.live + n)


class FakeContext:
    ledger: Ledger = None

    def __init__(self, nx, ny, nz, dtype=np.float64, device=0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny * nz
        self.esz = np.dtype(dtype).itemsize
        self.np_dtype = np.dtype(dtype)
        self.held = {}
        self.peak = 0
        self.closed = False
        self._hold("fields", 9 * self.N * self.esz + self.N + nz * (self.esz + 8))
        self.ledger.created.append((nx, ny, nz))

    def _hold(self, key, n):
        self.ledger.add(n - self.held.get(key, 0))
        self.held[key] = n
        self.peak = max(self.peak, sum(self.held.values()))

    def _transient(self, n):
        self.ledger.transient(n)
        self.peak = max(self.peak, sum(self.held.values()) + n)

    def set_geometry(self, packed):
        self._hold("geo", np.asarray(packed).size * self.esz)

    def set_materials(self, ids, table):
        self._hold("mtab", np.asarray(table).shape[0] * 9 * self.esz)

    def wake_probe(self, ixt, iyt, h, nt):
        self.probe = (ixt, iyt, h, int(nt))
        self._hold("hist", (2 * h + 1) ** 2 * self.nz * int(nt) * self.esz)

    def second_buffers(self):
        self._hold("alt", 6 * self.N * self.esz)

    def kernel_in_use(self):
        # the two-step sweep applies from nx >= 64, ny >= 16 in FP32
        two = self.esz == 4 and self.nx >= 64 and self.ny >= 16
        return _lib.KERNEL_FUSED2 if two else _lib.KERNEL_FUSED

    # the two-step tile shape: the tuner (no allocation; the second buffers
    # are already held) picks shape 4 and is recorded with the steps taken
    # before it
    def tune2(self, reps):
        self.tuned = getattr(self, "tuned", []) + [getattr(self, "steps", 0)]
        self.tile2 = 4
        return np.array([1.0, 2.0, 2.0, 2.0, 0.5])

    def get_tile2(self):
        return (getattr(self, "tile2", 0), 64, 8)

    def step(self, n, conductive=False):
        assert not self.closed
        self.second_buffers()
        self._hold("btab", max(int(n), self.held.get("btab", 0) // (self.nz * self.esz))
                   * self.nz * self.esz)
        self.steps = getattr(self, "steps", 0) + int(n)
        self.calls = getattr(self, "calls", []) + [int(n)]

    def wake_potential(self, window_ids, carry, cols, k0, dz, qnorm):
        nzw, ns = np.asarray(cols).shape
        nwin = len(window_ids)
        self._transient(4 * nwin + 4 * max(nzw * ns, 1) + 8 * max(nzw, 1) + 8 * nwin * ns)
        return np.zeros((nwin, ns))

    def impedance(self, x, lam, N, K, ds, noalloc=False):
        self._transient(16 * min(len(x), N) + 48 * K)
        z = np.zeros(K, dtype=complex)
        return z, z.copy(), z.copy()

    def device_bytes(self):
        return sum(self.held.values()), self.peak

    # the watchdog: a context whose shape is in `diverge` reports non-finite
    # fields once it has stepped more than `diverge_after` steps
    diverge = ()
    diverge_after = 0

    def field_finite(self, which=0):
        self._hold("flag", 256)
        bad = (self.nx, self.ny, self.nz) in self.diverge
        return not (bad and getattr(self, "steps", 0) > self.diverge_after)

    def close(self):
        if not self.closed:
            self.ledger.add(-sum(self.held.values()))
            self.held = {}
            self.closed = True

    def set_bc(self, *a): pass
    def set_dt(self, *a): pass
    def set_kernel(self, *a): pass
    def set_tile(self, *a): pass
    def set_step_index(self, *a): pass
    def set_beam_gauss(self, *a): pass
    def set_beam_values(self, *a): pass
    def beam_off(self): pass
    def zero_ghosts(self): pass
    def sync(self): pass


class FakeBatch:
    """Steps nothing; members get their second field buffers on joining, as
    the batched sweep's member check allocates them."""

    def __init__(self, solvers, graph=True, group=16, tile=None):
        self.group = group
        self.set_members(solvers)

    def set_members(self, solvers):
        for s in solvers:
            assert not s.ctx.closed
            s.ctx.second_buffers()
        self.members = list(solvers)

    def step(self, n):
        for s in self.members:
            s.ctx.steps = getattr(s.ctx, "steps", 0) + int(n)
        return 0.1 * n

    def device_bytes(self):
        return sum(self.group * s.Nz * s.ctx.esz + 4096 for s in self.members)

    def close(self):
        pass


class VClock:
    """Virtual seconds: 1 ms per step the server takes, plus idle jumps."""

    def __init__(self):
        self.server = None
        self.offset = 0.0

    def __call__(self):
        steps = 0 if self.server is None else sum(e[2] for e in self.server.log)
        return steps * 1e-3 + self.offset

    def wait_until(self, t):
        self.offset += max(0.0, t - self())


@pytest.fixture
def fake_device(monkeypatch):
    led = Ledger()
    FakeContext.ledger = led
    monkeypatch.setattr(_lib, "Context", FakeContext)
    monkeypatch.setattr(serve, "BatchStepper", FakeBatch)
    return led


def _burst(n, arrivals=None):
    reqs = serve.make_requests(n, seed=7, median_cells=2e4, sigma_ln=0.4, clip=(1e4, 4e4),
                               slab_prob=0.5, wakelength=0.05)
    for i, r in enumerate(reqs):
        r.workload["solver"]["n_pml"] = 4
        r.arrival = 0.0 if arrivals is None else arrivals[i]
    return reqs


def _model(r, esz=4):
    from wakis_amd.workloads import build

    _, _, s = build(r.workload, dtype=np.float32, device=None, precision="mixed")
    run = s._wake_plan(None, None, "longitudinal")
    return serve.device_bytes(s.Nx, s.Ny, s.Nz, run["Nt"], esz, nmat=len(s._mt.table()),
                              half=run["h"], beam_rows=50)


@pytest.mark.timeout(300)
def test_burst_stays_within_budget(fake_device):
    reqs = _burst(10)
    models = {r.rid: _model(r) for r in reqs}
    budget = 2.5 * max(models.values())
    clock = VClock()
    srv = serve.Server(dtype=np.float32, budget=budget, chunk_steps=50, group=8, clock=clock)
    clock.server = srv
    done = srv.run(reqs)
    assert all(r.result is not None and "error" not in r.result for r in done)
    # the device never held more than the budget, the finish scratch included
    assert fake_device.peak <= budget, (fake_device.peak, budget)
    assert fake_device.live == 0
    assert len(fake_device.created) == len(reqs)
    # admission reserved at most the budget, and the burst had to queue
    assert srv.peak_admitted <= budget
    assert max(e[1] for e in srv.log) >= 2
    assert any(r.admitted > 0 for r in done)
    for r in done:
        # the model bounds what each request held at once
        assert r.result["device_peak"] <= r.result["device_model"] == models[r.rid]
        assert r.admitted >= r.arrival and r.finished >= r.admitted
    for e in srv.log:
        assert e[6] <= e[3] <= budget  # stepped members' bytes <= admitted bytes


@pytest.mark.timeout(300)
def test_oversized_request_is_rejected(fake_device):
    reqs = _burst(4)
    models = {r.rid: _model(r) for r in reqs}
    big = max(models, key=models.get)
    budget = sorted(models.values())[-2] * 1.01  # every request but the largest fits
    clock = VClock()
    srv = serve.Server(dtype=np.float32, budget=budget, chunk_steps=50, group=8, clock=clock)
    clock.server = srv
    done = srv.run(reqs)
    for r in done:
        if r.rid == big:
            assert "error" in r.result and "budget" in r.result["error"]
        else:
            assert "error" not in r.result
    assert fake_device.peak <= budget
    assert len(fake_device.created) == len(reqs) - 1


@pytest.mark.timeout(300)
def test_virtual_clock_crosses_idle_gaps(fake_device):
    # the second request arrives long after the first is done: the server
    # must move the virtual clock to it instead of waiting for ever
    reqs = _burst(2, arrivals=[0.0, 50.0])
    clock = VClock()
    srv = serve.Server(dtype=np.float32, budget=288e9, chunk_steps=50, group=8, clock=clock)
    clock.server = srv
    done = srv.run(reqs)
    assert all("error" not in r.result for r in done)
    assert done[1].admitted >= 50.0
    assert done[0].finished < 50.0


def test_model_terms():
    # the model is the library's allocation list, term by term
    nx, ny, nz, nt, esz = 20, 22, 60, 300, 4
    n = nx * ny * nz
    b = serve.device_bytes(nx, ny, nz, nt, esz, nmat=3, half=0, beam_rows=8, samples=11)
    fields = 15 * n * esz + n
    metric = (2 * (nx + ny + nz) + 2 * (ny * nz + nx * nz + nx * ny)) * esz
    tables = 27 * esz + nz * (esz + 8)
    hist = nz * nt * esz
    ns = nt + 2
    finish = 4 + 4 * nz * ns + 8 * nz + 8 * ns + 16 * ns + 48 * 11
    watchdog = 256
    assert b == fields + metric + tables + hist + 8 * nz * esz + 4096 + finish + watchdog
    # a 3 x 3 probe window holds nine histories
    b9 = serve.device_bytes(nx, ny, nz, nt, esz, nmat=3, half=1, beam_rows=8, samples=11)
    assert b9 - b == 8 * nz * nt * esz + 32 + 64 * ns


@pytest.mark.timeout(300)
@pytest.mark.parametrize("tune2", [False, True], ids=["model", "tuned"])
def test_two_step_members_step_alone_in_even_calls(fake_device, tune2):
    # requests wide enough for the two-step sweep are stepped by their own
    # fit_step calls of min(chunk, steps left): every call but the last is
    # even, so pairs of steps start at even indices as in a solve alone; with
    # tune2 each is tuned once, before its first step, and its result records
    # the shape
    reqs = serve.make_requests(3, seed=7, median_cells=8e5, sigma_ln=0.0, clip=(8e5, 8e5),
                               slab_prob=0.0, wakelength=0.05)
    small = _burst(2)
    for r in small:
        r.rid += 10
    reqs += small
    for r in reqs:
        r.workload["solver"]["n_pml"] = 4
        r.arrival = 0.0
    clock = VClock()
    srv = serve.Server(dtype=np.float32, budget=288e9, chunk_steps=50, group=8, clock=clock,
                       tune2=tune2)
    clock.server = srv
    made = []
    orig = FakeContext.__init__

    def track(self, *a, **k):
        orig(self, *a, **k)
        made.append(self)

    FakeContext.__init__ = track
    try:
        done = srv.run(reqs)
    finally:
        FakeContext.__init__ = orig
    assert all("error" not in r.result for r in done)
    solo = [c for c in made if c.kernel_in_use() == _lib.KERNEL_FUSED2]
    assert len(solo) == 3 and len(made) == 5
    for c in solo:
        assert all(n == 50 for n in c.calls[:-1]) and 0 < c.calls[-1] <= 50
        assert getattr(c, "tuned", []) == ([0] if tune2 else [])
    for c in made:
        if c not in solo:
            assert not hasattr(c, "calls")  # batched members are stepped by the batch
            assert not hasattr(c, "tuned")
    wide = [r for r in done if r.rid < 10]
    assert all(r.result["tile2"] == (4 if tune2 else 0) for r in wide)
    assert all(r.result["tile2"] is None for r in done if r.rid >= 10)
    assert max(e[8] for e in srv.log) >= 1
    assert sum(e[7] for e in srv.log) == sum(r.cells * r.steps for r in done)
    assert fake_device.peak <= srv.peak_admitted


@pytest.mark.timeout(300)
def test_watchdog_stops_a_diverging_solve(fake_device, monkeypatch):
    """A solve whose fields turn non-finite is stopped at the next round
    boundary with an error result and its memory released; the other
    requests complete."""
    reqs = _burst(4)
    from wakis_amd.workloads import build

    _, _, s = build(reqs[1].workload, dtype=np.float32, device=None, precision="mixed")
    monkeypatch.setattr(FakeContext, "diverge", ((s.Nx, s.Ny, s.Nz),))
    monkeypatch.setattr(FakeContext, "diverge_after", 60)
    clock = VClock()
    srv = serve.Server(dtype=np.float32, chunk_steps=50, group=8, clock=clock)
    clock.server = srv
    done = srv.run(reqs)
    bad = [r for r in done if "error" in r.result]
    assert [r.rid for r in bad] == [reqs[1].rid], [(r.rid, r.result.get("error")) for r in done]
    err = bad[0].result["error"]
    assert "non-finite" in err and "watchdog" in err
    assert 60 < bad[0].steps <= 110  # caught at the first round boundary past step 60
    assert all("error" not in r.result for r in done if r.rid != reqs[1].rid)
    assert fake_device.live == 0


@pytest.mark.timeout(300)
def test_cancel_before_and_during_the_run(fake_device):
    """cancel() drops a request before preparation, one waiting for
    admission, and one stopped between rounds; the rest complete and every
    byte is released."""
    reqs = _burst(6, arrivals=[0.0, 0.0, 0.0, 0.0, 0.0, 5.0])
    models = {r.rid: _model(r) for r in reqs}
    budget = 1.05 * max(models.values())  # one request at a time: the others queue
    clock = VClock()
    srv = serve.Server(dtype=np.float32, budget=budget, chunk_steps=50, group=8, clock=clock)
    clock.server = srv
    srv.cancel(reqs[5].rid)  # arrives later: never prepared
    first = []

    # the first request to reach the device is cancelled before its first
    # step; reqs[3] is cancelled as soon as it is prepared
    def after_attach(req, solver):
        if not first:
            first.append(req.rid)
            srv.cancel(req.rid)

    prepare = srv._prepare

    def prepare_then_cancel(req):
        it = prepare(req)
        if req.rid == reqs[3].rid:
            srv.cancel(req.rid)
        return it

    srv.after_attach = after_attach
    srv._prepare = prepare_then_cancel
    done = srv.run(reqs)
    res = {r.rid: r.result for r in done}
    assert "cancelled before preparation" in res[reqs[5].rid]["error"]
    assert "cancelled before admission" in res[reqs[3].rid]["error"]
    if first[0] != reqs[3].rid:
        assert "cancelled before its first step" in res[first[0]]["error"]
    for r in done:
        if r.rid not in (reqs[5].rid, reqs[3].rid, first[0]):
            assert "error" not in r.result, r.result
    assert fake_device.live == 0
    assert all(r.finished >= r.arrival for r in done)


@pytest.mark.timeout(300)
def test_cancel_a_running_solve(fake_device, monkeypatch):
    """A request cancelled while it is stepping stops at the next round
    boundary with the steps it took; its memory is released."""
    reqs = _burst(3)
    from wakis_amd.workloads import build

    _, _, s = build(reqs[2].workload, dtype=np.float32, device=None, precision="mixed")
    shape = (s.Nx, s.Ny, s.Nz)
    clock = VClock()
    srv = serve.Server(dtype=np.float32, chunk_steps=50, group=8, clock=clock)
    clock.server = srv
    step = FakeBatch.step

    def step_then_cancel(self, n):
        ms = step(self, n)
        if any((m.Nx, m.Ny, m.Nz) == shape for m in self.members):
            srv.cancel(reqs[2].rid)
        return ms

    monkeypatch.setattr(FakeBatch, "step", step_then_cancel)
    done = srv.run(reqs)
    res = {r.rid: r for r in done}
    err = res[reqs[2].rid].result["error"]
    assert err == "cancelled after 50 steps", err
    assert res[reqs[2].rid].steps == 50
    assert all("error" not in res[r.rid].result for r in reqs[:2])
    assert fake_device.live == 0
