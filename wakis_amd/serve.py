"""Serving loop: impedance requests arriving over time, solved with
continuous batching on one device.

Requests (cavity wake + impedance solves of different sizes) arrive as a
Poisson process.  A request is admitted when the device memory it needs fits
the HBM budget next to the solves already resident; admitted solves keep
their fields on the device and are stepped together by the batched launch
(``wakis_amd.batch``), in chunks, so requests that arrive while others run
join the batch at the next chunk boundary and their work overlaps on the
device.  A solve leaves the batch after its own Nt steps; its wake
potential and impedance are then computed on the device and its memory is
released.  Each member's arithmetic is the same as a solve alone, so the
results do not depend on the mix.

The host side of a new request (geometry, materials, metric tables) is
prepared by worker threads without touching the device.  Admission reserves
the request's device bytes (device_bytes, an upper bound of what the library
allocates for it) against the budget; only then does a device worker
allocate and upload it, and, after its last step, compute its wake and
impedance and free it, while the main thread steps the batch (ctypes
releases the GIL during library calls).

There is no counterpart in the reference (it runs one solve per process,
wakis/routines.py:231-485).
"""

from __future__ import annotations

import dataclasses
import queue
import threading
import time
from typing import Callable

import numpy as np

from . import _lib
from .batch import BatchStepper
from .workloads import build, cells_from_nodes

EPS0 = 8.8541878128e-12


@dataclasses.dataclass
class Request:
    """One impedance solve: a cavity workload and its arrival time (s)."""
    rid: int
    arrival: float
    workload: dict
    cells: int
    # filled by the server
    admitted: float = float("nan")
    finished: float = float("nan")
    steps: int = 0
    result: dict | None = None
    prep_s: float = float("nan")     # host preparation (worker thread)
    finish_s: float = float("nan")   # wake potential + impedance after the last step

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def serve_cavity(shape, radius, length, pipe, cells, slab=None, n_pml=10, sigmaz=10e-3,
                 beta=1.0, wakelength=1.0, name="req"):
    """A rectangular (side 2*radius) or staircased cylindrical cavity of the
    given length with a beam pipe of half-width/radius `pipe` running one
    cavity length out on each side, meshed with cubic cells so that the grid
    has about `cells` cells; PEC walls, graded PML on both z ends.  `slab`
    (eps_r, sigma) adds a dielectric insert filling y < -radius/2 over the
    middle half of the cavity length."""
    wt = 2.0 * radius
    lz = 3.0 * length
    dx = (wt * wt * lz / float(cells)) ** (1.0 / 3.0)
    nxy = max(8, int(round(wt / dx)))
    nz = max(16, int(round(lz / dx)))
    x = np.linspace(-radius, radius, nxy + 1)
    y = np.linspace(-radius, radius, nxy + 1)
    z = np.linspace(-lz / 2, lz / 2, nz + 1)
    if shape == "rect":
        bx = (np.abs(x) <= radius)[:, None, None]
        by = (np.abs(y) <= radius)[None, :, None]
        px = (np.abs(x) <= pipe)[:, None, None]
        py = (np.abs(y) <= pipe)[None, :, None]
        bz = (np.abs(z) <= length / 2)[None, None, :]
        vac = cells_from_nodes((bx & by & bz) | (px & py & np.ones_like(bz)))
    else:
        r2 = x[:, None] ** 2 + y[None, :] ** 2
        cav = (r2 <= radius**2)[:, :, None] & (np.abs(z) <= length / 2)[None, None, :]
        pp = (r2 <= pipe**2)[:, :, None] & np.ones((1, 1, len(z)), dtype=bool)
        vac = cells_from_nodes(cav | pp)
    solids = [(vac, "vacuum")]
    if slab is not None:
        eps_r, sigma = slab
        sy = (y[:-1] + y[1:]) / 2 < -radius / 2
        sz = np.abs((z[:-1] + z[1:]) / 2) <= length / 4
        m = vac & sy[None, :, None] & sz[None, None, :]
        solids.append((m, [float(eps_r), 1.0, float(sigma)]))
    return dict(
        name=name,
        grid=dict(x=x, y=y, z=z, solids=solids),
        # use_sibc=False: the slab is a lossy dielectric (eps_r kept), not a
        # conductor whose eps the reference's surface-impedance rule replaces
        # by sigma*eps_0 (wakis/solverFIT3D.py:471-472)
        solver=dict(bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"], bg="pec",
                    use_stl=True, use_sibc=False, n_pml=n_pml, cfln=0.5),
        wake=dict(q=1e-9, sigmaz=sigmaz, beta=float(beta), wakelength=wakelength),
    )


def make_requests(n, seed=4, rate=2.0, median_cells=8e6, sigma_ln=1.0,
                  clip=(1e6, 1.2e8), slab_prob=0.5, f_ref=1e9, wakelength=1.0):
    """n requests with Poisson arrivals of the given rate (1/s): rectangular
    or cylindrical cavities with radius U(45, 75) mm, length U(60, 100) mm,
    pipe U(15, 25) mm, a lognormal cell count (median, sigma_ln) clipped to
    `clip`, and with probability slab_prob a dielectric insert with
    eps_r U(2, 10) and loss tangent U(1e-4, 1e-2), i.e. a conductivity
    sigma = 2*pi*f_ref*eps_0*eps_r*tan(delta)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        shape = "rect" if rng.random() < 0.5 else "cyl"
        radius = rng.uniform(0.045, 0.075)
        length = rng.uniform(0.060, 0.100)
        pipe = rng.uniform(0.015, 0.025)
        cells = float(np.clip(median_cells * np.exp(sigma_ln * rng.standard_normal()), *clip))
        slab = None
        if rng.random() < slab_prob:
            eps_r = rng.uniform(2.0, 10.0)
            tand = rng.uniform(1e-4, 1e-2)
            slab = (eps_r, 2 * np.pi * f_ref * EPS0 * eps_r * tand)
        wl = serve_cavity(shape, radius, length, pipe, cells, slab=slab, wakelength=wakelength,
                          name=f"req{i:03d}")
        n_cells = (len(wl["grid"]["x"]) - 1) * (len(wl["grid"]["y"]) - 1) * (len(wl["grid"]["z"]) - 1)
        out.append(Request(rid=i, arrival=float(t[i]), workload=wl, cells=n_cells))
    return out


def device_bytes(nx, ny, nz, nt, esz, nmat=256, half=0, beam_rows=256, samples=1001):
    """Most device bytes one served solve holds at once, term by term as the
    library allocates them (fit_device_bytes reports the real figure):

    fields      E, H, J and the sweep's second E, H buffers: 15 N esz
    ids         the uint8 material id of every cell: N
    metric      the packed metric tables: (2 (nx+ny+nz) + 2 (ny nz + nx nz + nx ny)) esz
    materials   nmat rows of dt/mu, dt/eps, sigma: 9 nmat esz
    lines       the beam line and cell centres: nz (esz + 8)
    history     the E_z probe: (2 half + 1)^2 nz nt esz
    beam        the beam table: beam_rows x nz values (a solve stepped alone holds
                one row per step of a fit_step call, a batched one its rows of
                the batch's table, group per member), and the member's entries
                in the batch's tables (4 KiB)
    finish      the wake potential's scratch (window ids, nz x ns time columns,
                dz, the WP rows) and the impedance DFT's (2 ns + 6 K doubles),
                with ns <= nt + 2 wake samples and K <= samples DFT bins; both
                are counted although they are held one after the other
    watchdog    the 256-byte flag of fit_field_finite."""
    n = nx * ny * nz
    nwin = (2 * half + 1) ** 2
    ns = nt + 2
    fields = 15 * n * esz + n
    metric = (2 * (nx + ny + nz) + 2 * (ny * nz + nx * nz + nx * ny)) * esz
    tables = 9 * nmat * esz + nz * (esz + 8)
    history = nwin * nz * max(nt, 1) * esz
    beam = beam_rows * nz * esz + 4096
    finish = (4 * nwin + 4 * nz * ns + 8 * nz + 8 * nwin * ns) + (16 * ns + 48 * samples)
    watchdog = 256
    return int(fields + metric + tables + history + beam + finish + watchdog)


class Server:
    """Continuous-batching server on one device.

    budget: device bytes the resident solves may use (default 288 GB); a
    request is admitted only when its device_bytes fit next to those already
    admitted, and nothing of it is allocated on the device before that, so
    the device memory of the served solves (fit_device_bytes, plus the
    batch's own tables) stays within the budget under any burst.  A request
    whose device_bytes alone exceed the budget is rejected: its result is
    {"error": ...}.  chunk_steps: steps per batch call (a request attached
    while a chunk runs joins at the next call).  clock: seconds since start
    (wall time by default); a caller's clock that advances only with the
    server's own work (a virtual clock) should have wait_until(t), which the
    server calls when it is idle until the next arrival at t.
    device: the device the solves run on.

    two_step: a solve on which fit_step runs the two-step sweep (FP32, nx >=
    64, ny >= 16) is stepped by its own fit_step calls on its own stream,
    min(chunk_steps, steps left) per round, so its pairs of steps fall as
    they do when it is solved alone (chunk_steps must be even) and its
    launches overlap on the device with the other solves' and with the
    batched launch of the remaining solves.  on_done(request) is called in the
    main loop as each request completes (or is rejected, cancelled or
    stopped by the watchdog).

    watchdog: after every round each stepped solve's E is checked on the
    device (fit_field_finite); a solve with a NaN or infinite value is
    stopped, its memory released, and its result is {"error": ...} with the
    step count.  cancel(rid) (any thread) drops a request before it is
    prepared or admitted, or stops it at the next round boundary; its result
    is {"error": "cancelled ..."}.  after_attach(request, solver) runs on the
    device worker once a request's fields are on the device, before its first
    step (to set initial fields, for instance).

    tune2: a solve stepped alone first times the two-step tile shapes on its
    own grid and data (fit_tune2: a few launches into its second field
    buffers, no allocation, fields unchanged) and keeps the fastest; its
    result records the shape as "tile2", so solve_alone(request,
    tile2=result["tile2"]) repeats it bit for bit.  Off (the default): the
    launch model's shape, the one solve_alone takes by default.  In a busy mix
    the tuner times the shapes while the other solves' launches share the
    device, and its picks scatter: on the 50-request trace it served 3074
    solves/h against 3562 with the model's shapes."""

    def __init__(self, dtype=np.float32, budget=288e9, chunk_steps=256, graph=True, group=16,
                 tile=None, compute_plane="longitudinal", clock: Callable[[], float] | None = None,
                 prep_workers=2, precision="mixed", device=0, two_step=True,
                 on_done: Callable[[Request], None] | None = None, watchdog=True,
                 after_attach: Callable[[Request, object], None] | None = None, tune2=False):
        if int(chunk_steps) < 2 or int(chunk_steps) % 2:
            raise ValueError("chunk_steps must be even (pairs of steps start at even indices)")
        self.two_step = bool(two_step)
        self.tune2 = bool(tune2)
        self.on_done = on_done
        self.watchdog = bool(watchdog)
        self.after_attach = after_attach
        self._cancel = set()
        self._cancel_lock = threading.Lock()
        self.dtype = np.dtype(dtype)
        self.precision = precision
        self.budget = float(budget)
        self.chunk_steps = int(chunk_steps)
        self.graph, self.group, self.tile = graph, group, tile
        self.compute_plane = compute_plane
        self.device = device
        self._t0 = None
        self._clock = clock
        self.prep_workers = prep_workers
        # per round: (time, members, steps (most of any member), admitted
        #             bytes, ms, cells, device bytes of the stepped members and
        #             the batch, cell-steps, members stepped alone)
        self.log = []
        self.peak_admitted = 0.0

    def cancel(self, rid: int) -> None:
        """Cancel request rid (safe from any thread); see the class notes."""
        with self._cancel_lock:
            self._cancel.add(int(rid))

    def _cancelled(self, req: Request) -> bool:
        with self._cancel_lock:
            return req.rid in self._cancel

    def now(self) -> float:
        if self._clock is not None:
            return self._clock()
        return time.monotonic() - self._t0

    # -- host preparation of a request (worker thread, no device calls) -------
    def _prepare(self, req: Request):
        t0 = time.monotonic()
        grid, wake, solver = build(req.workload, dtype=self.dtype, device=None,
                                   precision=self.precision)
        run = solver._wake_plan(None, None, self.compute_plane)
        mem = device_bytes(solver.Nx, solver.Ny, solver.Nz, run["Nt"], self.dtype.itemsize,
                           nmat=len(solver._mt.table()), half=run["h"],
                           beam_rows=max(self.group, self.chunk_steps))
        req.prep_s = time.monotonic() - t0
        return {"req": req, "solver": solver, "wake": wake, "run": run, "done": 0, "mem": mem}

    # -- device work of a request (the device worker thread) ------------------
    def _attach(self, it):
        """Allocate and upload an admitted request and arm its beam and probe."""
        t0 = time.monotonic()
        s, run = it["solver"], it["run"]
        s.attach_device(self.device, tile=self.tile)
        s._wake_attach(run)
        s._wake_device_beam(run)
        if self.after_attach is not None:
            self.after_attach(it["req"], s)
        s.ctx.sync()
        it["solo"] = self.two_step and s.ctx.kernel_in_use() == _lib.KERNEL_FUSED2
        if it["solo"] and self.tune2:
            s.tune2(reps=2)
        it["attach_s"] = time.monotonic() - t0
        return it

    def _finish(self, it):
        """Wake potential and impedance of a request after its last step;
        frees its device memory."""
        tf = time.monotonic()
        s = it["solver"]
        try:
            s._wake_device_end(it["run"])
            s._wake_finish(it["run"], self.compute_plane, save_logs=False)
            w = it["wake"]
            it["req"].result = {"WP": np.asarray(w.WP).copy(), "Z": np.asarray(w.Z).copy(),
                                "f": np.asarray(w.f).copy(), "s": np.asarray(w.s).copy(),
                                "Nt": it["run"]["Nt"], "N": s.N,
                                "device_peak": s.ctx.device_bytes()[1], "device_model": it["mem"],
                                "tile2": s.ctx.get_tile2()[0] if it.get("solo") else None}
        finally:
            s.ctx.close()
        it["req"].finish_s = time.monotonic() - tf
        return it

    def _abort(self, it):
        """Release a stopped request's device memory (device worker)."""
        it["solver"].ctx.close()
        return it

    def run(self, requests):
        """Serve the requests (sorted by arrival); returns them with latency,
        steps and results filled in."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self._t0 = time.monotonic()
        ready: queue.Queue = queue.Queue()    # host-prepared, not admitted
        todo: queue.Queue = queue.Queue()     # arrived, not prepared
        dev_in: queue.Queue = queue.Queue()   # ("attach" | "finish", item)
        dev_out: queue.Queue = queue.Queue()  # (kind, item or error)
        stop = threading.Event()

        def prep_worker():
            while not stop.is_set():
                try:
                    r = todo.get(timeout=0.05)
                except queue.Empty:
                    continue
                try:
                    ready.put(self._prepare(r))
                except BaseException as e:  # surfaced by the main loop
                    ready.put({"req": r, "error": e})

        def device_worker():
            while not stop.is_set():
                try:
                    kind, it = dev_in.get(timeout=0.05)
                except queue.Empty:
                    continue
                try:
                    fn = {"attach": self._attach, "finish": self._finish,
                          "abort": self._abort}[kind]
                    dev_out.put((kind, fn(it)))
                except BaseException as e:
                    dev_out.put(("error", e))

        threads = [threading.Thread(target=prep_worker, daemon=True)
                   for _ in range(self.prep_workers)]
        threads.append(threading.Thread(target=device_worker, daemon=True))
        for th in threads:
            th.start()
        nxt = 0          # next request not yet handed to the workers
        received = 0     # requests back from the preparation workers
        waiting = []     # prepared, not admitted
        active = []      # resident and stepping
        in_flight = 0    # admitted items at the device worker (attach or finish)
        used = 0.0       # device_bytes of every admitted, unfinished request
        stepper = None
        finished = 0

        def done(req, error=None):
            nonlocal finished
            if error is not None:
                req.result = {"error": error}
            req.finished = self.now()
            finished += 1
            if self.on_done is not None:
                self.on_done(req)

        def stop_active(it, error):
            # a resident request stopped before its last step
            nonlocal in_flight
            it["error"] = error
            dev_in.put(("abort", it))
            in_flight += 1

        try:
            while finished < len(reqs):
                now = self.now()
                while nxt < len(reqs) and reqs[nxt].arrival <= now:
                    if self._cancelled(reqs[nxt]):
                        received += 1  # never prepared
                        done(reqs[nxt], "cancelled before preparation")
                    else:
                        todo.put(reqs[nxt])
                    nxt += 1
                while True:
                    try:
                        item = ready.get_nowait()
                    except queue.Empty:
                        break
                    if "error" in item:
                        raise item["error"]
                    received += 1
                    if self._cancelled(item["req"]):
                        done(item["req"], "cancelled before admission")
                    else:
                        waiting.append(item)
                while True:
                    try:
                        kind, it = dev_out.get_nowait()
                    except queue.Empty:
                        break
                    if kind == "error":
                        raise it
                    in_flight -= 1
                    if kind == "attach":
                        if self._cancelled(it["req"]):
                            stop_active(it, "cancelled before its first step")
                        else:
                            active.append(it)
                    elif kind == "abort":
                        used -= it["mem"]
                        done(it["req"], it["error"])
                    else:
                        used -= it["mem"]
                        done(it["req"])
                # admission against the budget, in arrival order; a request
                # that does not fit waits, later smaller ones may pass it
                waiting.sort(key=lambda it: it["req"].arrival)
                keep = []
                for it in waiting:
                    if self._cancelled(it["req"]):
                        done(it["req"], "cancelled before admission")
                    elif it["mem"] > self.budget:
                        done(it["req"], f"needs {it['mem']} device bytes, more than the "
                                        f"{self.budget:.0f} byte budget")
                    elif used + it["mem"] <= self.budget:
                        it["req"].admitted = self.now()
                        used += it["mem"]
                        self.peak_admitted = max(self.peak_admitted, used)
                        dev_in.put(("attach", it))
                        in_flight += 1
                    else:
                        keep.append(it)
                waiting = keep
                if not active:
                    # idle: wait for the next arrival, preparation or attach
                    if (nxt < len(reqs) and received == nxt and not waiting and in_flight == 0
                            and hasattr(self._clock, "wait_until")):
                        self._clock.wait_until(reqs[nxt].arrival)
                    elif nxt < len(reqs) and self._clock is None:
                        time.sleep(max(0.0005, min(0.01, reqs[nxt].arrival - self.now())))
                    else:
                        time.sleep(0.002)
                    continue
                solo = [it for it in active if it["solo"]]
                batched = [it for it in active if not it["solo"]]
                t_round = time.perf_counter()
                for it in solo:  # asynchronous, each on its solver's stream
                    it["n"] = min(self.chunk_steps, it["run"]["Nt"] - it["done"])
                    it["solver"]._advance(it["n"])
                ms = 0.0
                if batched:
                    n = min(self.chunk_steps, min(it["run"]["Nt"] - it["done"] for it in batched))
                    members = [it["solver"] for it in batched]
                    if stepper is None:
                        stepper = BatchStepper(members, graph=self.graph, group=self.group,
                                               tile=self.tile)
                    else:
                        stepper.set_members(members)
                    ms = stepper.step(n) if n > 0 else 0.0
                    for it in batched:
                        it["n"] = n
                for it in solo:
                    it["solver"].ctx.sync()
                if solo:
                    ms = (time.perf_counter() - t_round) * 1e3
                for it in active:
                    it["done"] += it["n"]
                    it["req"].steps = it["done"]
                real = sum(it["solver"].ctx.device_bytes()[0] for it in active)
                if stepper is not None:
                    real += stepper.device_bytes()
                self.log.append((self.now(), len(active), max(it["n"] for it in active), used, ms,
                                 sum(it["solver"].N for it in active), real,
                                 sum(it["solver"].N * it["n"] for it in active), len(solo)))
                still = []
                for it in active:
                    if self._cancelled(it["req"]):
                        stop_active(it, f"cancelled after {it['done']} steps")
                    elif self.watchdog and not it["solver"].ctx.field_finite(_lib.FIELD_E):
                        stop_active(it, f"non-finite fields after {it['done']} steps "
                                        "(the solve diverged); stopped by the watchdog")
                    elif it["done"] >= it["run"]["Nt"]:
                        dev_in.put(("finish", it))
                        in_flight += 1
                    else:
                        still.append(it)
                active = still
        finally:
            stop.set()
            for th in threads:
                th.join(timeout=5)
            if stepper is not None:
                stepper.close()
        return reqs


def solve_alone(req: Request, dtype=np.float64, tile=None, compute_plane="longitudinal",
                precision="mixed", tile2=None):
    """One request by itself with the plain per-solver wakesolve (the
    one-at-a-time path); returns (seconds, result dict).  tile, tile2: the
    one-step and two-step tile shapes (None: the defaults)."""
    grid, wake, solver = build(req.workload, dtype=dtype, precision=precision)
    if tile is not None:
        solver.ctx.set_tile(int(tile))
    if tile2 is not None:
        solver.ctx.set_tile2(int(tile2))
    t0 = time.monotonic()
    solver.wakesolve(compute_plane=compute_plane)
    solver.ctx.sync()
    dt = time.monotonic() - t0
    res = {"WP": np.asarray(wake.WP).copy(), "Z": np.asarray(wake.Z).copy(),
           "f": np.asarray(wake.f).copy(), "Nt": solver.Nt, "N": solver.N,
           "stepping_s": solver.simulationTime}
    solver.ctx.close()
    return dt, res


def fifo_latencies(arrivals, durations):
    """Latencies of a one-at-a-time FIFO server with the given service times."""
    t = 0.0
    out = []
    for a, d in zip(arrivals, durations):
        start = max(t, a)
        t = start + d
        out.append(t - a)
    return np.asarray(out)


def z_rel_error(z, z_ref, f, fcut):
    """max |Z - Z_ref| / max |Z_ref| over f < fcut."""
    m = np.asarray(f) < fcut
    zr = np.asarray(z_ref)[m]
    return float(np.max(np.abs(np.asarray(z)[m] - zr)) / max(np.max(np.abs(zr)), 1e-300))
