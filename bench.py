"""Benchmark of the FIT leapfrog hot path on one MI355X.

Workload: the lossy pillbox cavity, 256 x 256 x 768 cells (5.03e7), FP32
fields, conductive walls (J = sigma*E), 10-cell z PML, PEC x/y faces; every
step is the wakesolve inner loop on the device: in-kernel Gaussian bunch
current, E_z probe into the HBM history, H and E updates.

Metric: Gcell-steps/s = Nx*Ny*Nz*steps / stepping time (HIP events on the
solver stream around exactly `--steps` steps, after `--warmup` untimed steps,
with device synchronisation on both sides).  The JSON line also carries the
HBM roofline of the dominant kernel and the CPU oracle timed on a bounded
sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_info():
    """Host CPU model, the logical CPUs the OS reports (nproc) and those this
    process may run on."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "usable_cpus": usable}


def _oracle_for(spec):
    """Oracle solver (FP64 scipy.sparse) of a CPU-baseline sample: a z-slab
    of the pillbox grid, or one cavity of the batch workload."""
    from oracle.fit_oracle import OGrid, OSolver
    from wakis_amd.workloads import cavity_batch, lossy_pillbox

    if spec["kind"] == "pillbox_slab":
        # the reference's geomspace PML grading (the oracle restates the
        # reference; the grading changes sigma values only, not the work)
        w = lossy_pillbox(256, 256, spec["nz"], pml_profile="geomspace")
    else:
        w = cavity_batch(n=64, seed=3)[spec["index"]]
    g, sv = w["grid"], w["solver"]
    og = OGrid(x=g["x"], y=g["y"], z=g["z"])
    return OSolver(og, cfln=sv["cfln"], bc_low=sv["bc_low"], bc_high=sv["bc_high"],
                   solids=g["solids"], bg=sv["bg"], n_pml=sv["n_pml"], dtype=np.float64)


def cpu_worker(spec):
    """One CPU-baseline process: build the sample's oracle, report ready,
    wait for the start line on stdin, time spec["steps"] steps."""
    osv = _oracle_for(spec)
    osv.one_step()
    print(json.dumps({"ready": True}), flush=True)
    sys.stdin.readline()
    t0 = time.perf_counter()
    for _ in range(spec["steps"]):
        osv.one_step()
    dt = time.perf_counter() - t0
    print(json.dumps({"cells": osv.N, "steps": spec["steps"], "seconds": dt}), flush=True)


def cpu_parallel(specs):
    """len(specs) oracle processes started together (each builds its sample,
    then all start stepping at once); returns the sum of their cell-step
    rates and the per-worker results.  The workers touch no GPU."""
    import subprocess

    procs = []
    for spec in specs:
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", "cpu-worker",
               "--cpu-spec", json.dumps(spec)]
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1",
                   MKL_NUM_THREADS="1")
        procs.append(subprocess.Popen(cmd, stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                                      env=env, text=True))
    try:
        for p in procs:
            line = p.stdout.readline()
            if not line or not json.loads(line).get("ready"):
                raise RuntimeError("cpu worker failed to start")
        for p in procs:
            p.stdin.write("go\n")
            p.stdin.flush()
        res = [json.loads(p.stdout.readline()) for p in procs]
        for p in procs:
            p.wait(timeout=120)
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
    rate = sum(r["cells"] * r["steps"] / r["seconds"] for r in res)
    return rate, res


def cpu_procs():
    """Processes for the multi-process CPU figure: the usable CPUs, at most 16
    (the host share of one GPU)."""
    return max(1, min(cpu_info()["usable_cpus"], 16))


def traffic_from_profiles(kernel, dtype, grid, tile):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (each scaled by what the counter reads for a copy kernel of
    known size), from the committed profile table; None if not profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.tsv")
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None, None
    for line in lines[1:]:
        f = line.split("\t")
        if len(f) >= 11 and f[0] == kernel and f[1] == dtype and f[2] == grid and f[3] == str(tile):
            return int(float(f[8])), f[10]
    return None, None


def cpu_baseline_large(n_sample=128, steps=40):
    """The CPU oracle on an n_sample^3 sub-box of the stratified box with the
    same layering (16 z-layers, the central diagonal-tensor block, 12-cell
    PML on six faces): the full 1024^3 grid is out of reach on the host (its
    two CSC operators alone would take ~200 GB), so the rate per cell-step is
    measured on the sub-box and reported as extrapolated."""
    from scipy.constants import epsilon_0 as eps_0

    from oracle.fit_oracle import OGrid, OSolver
    from wakis_amd.workloads import stratified_box

    n = n_sample
    wl = stratified_box(n=n, half=0.256 * n / 1024)
    g, sv = wl["grid"], wl["solver"]
    iso = [(r, m) for r, m in g["solids"] if not isinstance(m, dict)]
    tensors = [(r, m) for r, m in g["solids"] if isinstance(m, dict)]

    def mask(region):
        m = np.zeros((n, n, n), bool)
        m[region] = True
        return m

    osv = OSolver(OGrid(x=g["x"], y=g["y"], z=g["z"]), cfln=sv["cfln"], bc_low=sv["bc_low"],
                  bc_high=sv["bc_high"], solids=[(mask(r), m) for r, m in iso], bg=sv["bg"],
                  dtype=np.float64, n_pml=sv["n_pml"], use_sibc=sv["use_sibc"])
    for region, mat in tensors:  # the reference's route: ieps edits + update_tensors
        for d, e in zip("xyz", mat["eps"]):
            osv.ieps[region + (d,)] = np.float64(1.0) / (e * eps_0)
    osv.update_tensors("ieps")
    osv.step_0 = True
    osv.one_step()
    t0 = time.perf_counter()
    for _ in range(steps):
        osv.one_step()
    dt = time.perf_counter() - t0
    return {
        "value": osv.N * steps / dt / 1e9,
        "unit": "Gcell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"extrapolated: oracle FP64 scipy.sparse one_step on a {n}^3 sub-box of the "
                  f"stratified box (same 16 layers, tensor block and 12-cell PML), {steps} steps, "
                  f"{dt:.2f} s on 1 core of {cpu_model()}; the rate per cell-step applies to the "
                  "1024^3 grid, whose operators do not fit the host",
        **cpu_info(),
    }


def cpu_baseline(nz_sample=32, steps=150):
    """The CPU oracle (scipy.sparse FP64, single-threaded csc/csr mat-vec, the
    reference's arithmetic path) on a z-slab of the same pillbox grid."""
    from oracle.fit_oracle import OGrid, OSolver
    from wakis_amd.workloads import lossy_pillbox

    # (the reference's geomspace PML grading: the oracle restates the
    # reference; the grading changes sigma values only, not the work per step)
    w = lossy_pillbox(256, 256, nz_sample, pml_profile="geomspace")
    g = w["grid"]
    og = OGrid(x=g["x"], y=g["y"], z=g["z"])
    sv = w["solver"]
    osv = OSolver(og, cfln=sv["cfln"], bc_low=sv["bc_low"], bc_high=sv["bc_high"],
                  solids=g["solids"], bg=sv["bg"], n_pml=sv["n_pml"], dtype=np.float64)
    osv.one_step()
    t0 = time.perf_counter()
    for _ in range(steps):
        osv.one_step()
    dt = time.perf_counter() - t0
    cells = osv.N * steps
    return {
        "value": cells / dt / 1e9,
        "unit": "Gcell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle FP64 scipy.sparse one_step, 256x256x{nz_sample} z-slab of the pillbox "
                  f"grid, {steps} steps, {dt:.2f} s on 1 core of {cpu_model()} (scipy's "
                  "sparse products and numpy's vector operations are single-threaded)",
        **cpu_info(),
    }


class PlainSparseStepper:
    """The reference's own GPU time step, for the plain-method baseline
    (wakis/solverFIT3D.py:337-367 and 504-525 as run with cupyx): the
    operators M_H = iDa iDmu C Ds and M_E = iDeps itDa C^T tDs as FP64
    sparse matrices on the device (the CSR structure of the reference, zero
    entries dropped as scipy's products drop them), stepped in the
    reference's expression order:

        H[:] = H - (dt * M_H) @ E        dt * M_H: a new matrix each step
                                         (values scaled)
        E[:] = E + dt * (M_E @ H - ieps * J)
        J[:] = sigma * E                 (conductive solvers)

    The products run as coalesced COO matrix-vector products through torch:
    its CSR product on this ROCm build moves about 12 GB/s (measured), the
    COO one about 30 times more.  The index arrays are shared by dt * M_H
    rather than copied as cupyx's scalar product copies them, which favours
    this baseline.

    `wake_step(n)` adds the rest of the reference's wakesolve iteration
    (wakis/routines.py:438-448): the beam line written into J_z from host
    values and the 3x3 E_z window read back to the host.  Built from a
    wakis_amd solver's host data (metric tables, material ids and table,
    face masks, dt); the fields start at zero."""

    def __init__(self, solver, device="cuda"):
        import warnings

        import torch

        warnings.filterwarnings("ignore", message="Sparse CSR tensor support is in beta")
        self.torch = torch
        dev = torch.device(device)
        self.dev = dev
        nx, ny, nz = solver.Nx, solver.Ny, solver.Nz
        N = nx * ny * nz
        self.solver, self.N, self.shape = solver, N, (nx, ny, nz)
        geo = np.asarray(solver.geometry, dtype=np.float64)
        sizes = [nx, ny, nz, nx, ny, nz, ny * nz, nx * nz, nx * ny, ny * nz, nx * nz, nx * ny]
        parts, o = [], 0
        for m in sizes:
            parts.append(torch.from_numpy(geo[o:o + m].copy()).to(dev))
            o += m
        Lt, tLt, iAt, itAt = parts[0:3], parts[3:6], parts[6:9], parts[9:12]
        n = torch.arange(N, device=dev, dtype=torch.int64)
        i, j, k = n % nx, (n // nx) % ny, n // (nx * ny)
        self._ijk = (i, j, k)
        L = [Lt[0][i], Lt[1][j], Lt[2][k]]
        tL = [tLt[0][i], tLt[1][j], tLt[2][k]]
        iA = [iAt[0][j + ny * k], iAt[1][i + nx * k], iAt[2][i + nx * j]]
        itA = [itAt[0][j + ny * k], itAt[1][i + nx * k], itAt[2][i + nx * j]]
        tab = torch.from_numpy(solver._mt.table()).to(dev)
        ids = torch.from_numpy(np.reshape(solver._mt.ids, -1, order="F").astype(np.int64)).to(dev)
        mat = tab[ids]  # (N, 9): 1/eps, 1/mu, sigma per component
        keep = lambda bits, c: (~self._masked(bits, c)).to(torch.float64)
        pec = [keep(solver.pec_bits, c) for c in range(3)]
        pmc = [keep(solver.pmc_bits, c) for c in range(3)]
        sx, sy, sz = 1, nx, nx * ny
        # curl rows: (column component, flat shift, sign) of each of the 4 terms
        h_terms = {0: [(1, 0, 1.0), (1, sz, -1.0), (2, 0, -1.0), (2, sy, 1.0)],
                   1: [(2, 0, 1.0), (2, sx, -1.0), (0, 0, -1.0), (0, sz, 1.0)],
                   2: [(0, 0, 1.0), (0, sy, -1.0), (1, 0, -1.0), (1, sx, 1.0)]}
        e_terms = {0: [(2, 0, 1.0), (2, -sy, -1.0), (1, 0, -1.0), (1, -sz, 1.0)],
                   1: [(0, 0, 1.0), (0, -sz, -1.0), (2, 0, -1.0), (2, -sx, 1.0)],
                   2: [(1, 0, 1.0), (1, -sx, -1.0), (0, 0, -1.0), (0, -sy, 1.0)]}
        # M_H row (c, n): (iA_c imu_c pmc_c)[n] * sign * (L_p pec_p)[m]
        rh = [iA[c] * mat[:, 3 + c] * pmc[c] for c in range(3)]
        self.MH = self._coo(self._csr(h_terms, rh, [L[p] * pec[p] for p in range(3)]))
        # M_E row (c, n): (ieps_c itA_c pec_c)[n] * sign * (tL_p pmc_p)[m]
        re = [mat[:, c] * itA[c] * pec[c] for c in range(3)]
        self.ME = self._coo(self._csr(e_terms, re, [tL[p] * pmc[p] for p in range(3)]))
        self.ieps = torch.cat([mat[:, c] for c in range(3)]).contiguous()
        self.sigma = torch.cat([mat[:, 6 + c] for c in range(3)]).contiguous()
        self.dt = float(solver.dt)
        self.cond = bool(solver.use_conductivity)
        self.E = torch.zeros(3 * N, dtype=torch.float64, device=dev)
        self.H = torch.zeros_like(self.E)
        self.J = torch.zeros_like(self.E)
        self.step_0 = True
        del L, tL, iA, itA, mat, ids, rh, re, pec, pmc

    def _masked(self, bits, c):
        """Face mask of component c (last face written wins, in the order
        xlo, xhi, ylo, yhi, zlo, zhi: wakis/boundaries.py:90-100, 130-140)."""
        torch = self.torch
        i, j, k = self._ijk
        nx, ny, nz = self.shape
        r = torch.zeros(self.N, dtype=torch.bool, device=self.dev)
        for axis, idx, last in ((0, i, nx - 1), (1, j, ny - 1), (2, k, nz - 1)):
            if c == axis:
                continue
            r = torch.where(idx == 0, bool(bits & (1 << (2 * axis))), r)
            r = torch.where(idx == last, bool(bits & (1 << (2 * axis + 1))), r)
        return r

    def _csr(self, terms, rowc, colv):
        torch = self.torch
        N = self.N
        n = torch.arange(N, device=self.dev, dtype=torch.int64)
        crow_parts, col_parts, val_parts = [], [], []
        for c in range(3):
            cols, vals = [], []
            for p, shift, sign in terms[c]:
                m = n + shift
                ok = (m >= 0) & (m < N)
                mc = m.clamp(0, N - 1)
                v = (rowc[c] * sign) * colv[p][mc]
                vals.append(torch.where(ok, v, torch.zeros_like(v)))
                cols.append(p * N + mc)
            col = torch.stack(cols, 1)
            val = torch.stack(vals, 1)
            del cols, vals
            nz = val != 0
            col = torch.where(nz, col, torch.full_like(col, 3 * N))
            col, order = torch.sort(col, dim=1)
            val = torch.gather(val, 1, order)
            nz = col < 3 * N
            crow_parts.append(nz.sum(1))
            col_parts.append(col[nz].to(torch.int32))
            val_parts.append(val[nz])
            del col, val, nz, order
        counts = torch.cat(crow_parts)
        crow = torch.zeros(3 * N + 1, dtype=torch.int64, device=self.dev)
        crow[1:] = torch.cumsum(counts, 0)
        return (crow.to(torch.int32), torch.cat(col_parts), torch.cat(val_parts))

    def _coo(self, csr):
        """(indices (2, nnz) int64, values) of a CSR triple, rows ascending."""
        torch = self.torch
        crow, col, val = csr
        counts = (crow[1:] - crow[:-1]).to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(len(counts), device=self.dev), counts)
        return torch.stack([rows, col.to(torch.int64)]), val

    def _zero_ghosts(self):
        """wakis/solverFIT3D.py:842-862."""
        i, j, k = self._ijk
        nx, ny, nz = self.shape
        N = self.N
        li, lj, lk = i == nx - 1, j == ny - 1, k == nz - 1
        for c, m in ((0, lj | lk), (1, li | lk), (2, li | lj)):
            self.H[c * N:(c + 1) * N][m] = 0.0
        for c, m in ((0, li), (1, lj), (2, lk)):
            self.E[c * N:(c + 1) * N][m] = 0.0

    def one_step(self):
        torch = self.torch
        if self.step_0:
            self._zero_ghosts()
            self.step_0 = False
        size = (3 * self.N, 3 * self.N)
        idx, val = self.MH
        mh_dt = torch.sparse_coo_tensor(idx, self.dt * val, size=size, is_coalesced=True)
        self.H.copy_(self.H - torch.mv(mh_dt, self.E))
        me = torch.sparse_coo_tensor(*self.ME, size=size, is_coalesced=True)
        self.E.copy_(self.E + self.dt * (torch.mv(me, self.H) - self.ieps * self.J))
        if self.cond:
            self.J.copy_(self.sigma * self.E)

    def setup_wake(self, beam, ixt, iyt):
        """Beam line and E_z window positions of a wakesolve loop."""
        torch = self.torch
        nx, ny, nz = self.shape
        N = self.N
        self.beam = beam
        beam.line(self.solver, 0.0)  # first update: source indices and zmin
        self._jline = 2 * N + int(beam.ixs) + nx * int(beam.iys) + nx * ny * torch.arange(
            nz, device=self.dev)
        ii = torch.arange(ixt - 1, ixt + 2, device=self.dev)
        jj = torch.arange(iyt - 1, iyt + 2, device=self.dev)
        kk = torch.arange(nz, device=self.dev)
        self._win = (2 * N + ii[:, None, None] + nx * jj[None, :, None]
                     + nx * ny * kk[None, None, :]).reshape(-1)

    def wake_step(self, n):
        """One wakesolve iteration: beam.update, the E_z window to the host,
        one_step (wakis/routines.py:438-448)."""
        line = self.beam.line(self.solver, n * self.dt)
        self.J[self._jline] = self.torch.from_numpy(np.asarray(line, dtype=np.float64)).to(self.dev)
        ez = self.E[self._win].cpu()
        self.one_step()
        return ez


def plain_ms_per_step(solver, steps=60, warmup=5, wake=True):
    """Wall milliseconds per wakesolve iteration of the plain FP64 GPU method
    on `solver`'s grid (the E_z read-back synchronises every iteration)."""
    import torch
    from wakis_amd.sources import Beam

    st = PlainSparseStepper(solver)
    w = solver.wake
    beam = Beam(q=w.q, sigmaz=w.sigmaz, beta=w.beta, xsource=w.xsource, ysource=w.ysource)
    ixt = int(np.abs(solver.x - w.xtest).argmin())
    iyt = int(np.abs(solver.y - w.ytest).argmin())
    st.setup_wake(beam, ixt, iyt)
    for n in range(warmup):
        st.wake_step(n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(warmup, warmup + steps):
        st.wake_step(n)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    nnz = int(st.MH[1].numel() + st.ME[1].numel())
    del st
    torch.cuda.empty_cache()
    return ms, nnz


def plain_child(nreq, steps, target="serve", grid="256,256,768"):
    """ms per wakesolve iteration of the plain FP64 GPU method for every
    request of the serving workload (or the pillbox grid), measured in a
    child process."""
    import subprocess

    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", "plain", "--nreq",
           str(nreq), "--plain-steps", str(steps), "--plain-target", target, "--grid", grid]
    out = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=900, check=True).stdout
    rows = json.loads(out.decode().strip().splitlines()[-1])["plain"]
    return {r["rid"]: (r["ms"], r["nnz"]) for r in rows}


def plain_floor_bytes(n_cells, nnz, conductive=True):
    """Fewest HBM bytes one step of the reference's sparse method can move
    (FP64): both matrices read once in CSR with 32-bit column indices
    (12 B per entry + 4 B row pointer per row), dt * M_H written as a new
    value array (16 B per entry of M_H, half of nnz), and the 27 vector
    passes of 3N doubles of the step's expression (H: product output, H - y,
    the copy back; E: product, ieps * J, the difference, dt * (), E + (),
    the copy back; J: sigma * E and the copy back), each read and write
    counted once."""
    words = 7 + 15 + (5 if conductive else 0)
    return 12 * nnz + 2 * 3 * n_cells * 4 + 16 * (nnz // 2) + words * 3 * n_cells * 8


def bench_plain(args):
    """Child of the serving benchmark: the plain method on each request's
    grid (host setup only in this process, no wakis_amd device context)."""
    import torch  # first: its HIP runtime owns the device in this process

    from wakis_amd.serve import make_requests
    from wakis_amd.workloads import build

    torch.cuda.init()
    rows = []
    if args.plain_target == "pillbox":
        from wakis_amd.workloads import lossy_pillbox

        nx, ny, nz = (int(v) for v in args.grid.split(","))
        _, _, hs = build(lossy_pillbox(nx, ny, nz), dtype=np.float64, device=None)
        ms, nnz = plain_ms_per_step(hs, steps=args.plain_steps)
        print(f"[plain] pillbox N {hs.N} nnz {nnz} {ms:.3f} ms/step", file=sys.stderr, flush=True)
        return {"plain": [{"rid": 0, "ms": ms, "nnz": nnz, "cells": hs.N}]}
    for r in make_requests(args.nreq, seed=4, rate=2.0):
        _, _, hs = build(r.workload, dtype=np.float64, device=None, precision="reference")
        ms, nnz = plain_ms_per_step(hs, steps=args.plain_steps)
        print(f"[plain] req {r.rid:3d} N {r.cells:9d} nnz {nnz} {ms:.3f} ms/step",
              file=sys.stderr, flush=True)
        rows.append({"rid": r.rid, "ms": ms, "nnz": nnz, "cells": r.cells})
    return {"plain": rows}


def bench_cubic(args):
    """The reference's own CPU case: the PEC cubic cavity (64 x 64 x 192,
    FP64, CFL 1.0, wakelength 3 m) solved in full on the device (wakesolve:
    steps, WP and Z, longitudinal and transverse), against the CPU oracle on
    the same grid (steps timed on a bounded sample, scaled to the run)."""
    from wakis_amd.workloads import build, pec_cubic_cavity

    wl = pec_cubic_cavity()
    grid, wake, solver = build(wl, dtype=np.float64)
    solver.wakesolve(compute_plane="both")  # warm-up run (builds, tunes, caches)
    grid, wake, solver = build(wl, dtype=np.float64)
    t0 = time.perf_counter()
    solver.wakesolve(compute_plane="both")
    wall = time.perf_counter() - t0
    step_s = solver.simulationTime
    N, Nt = solver.N, solver.Nt
    print(f"[bench] cubic cavity FP64 on the device: Nt {Nt}, wakesolve {wall:.3f} s "
          f"(stepping {step_s:.3f} s)", file=sys.stderr, flush=True)
    out = {
        "metric": "Gcell-steps/s",
        "value": round(N * Nt / step_s / 1e9, 3),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": int(Nt),
        "warmup": 0,
        "ms_per_step": round(step_s / Nt * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"PEC cubic cavity {wl['name']} ({N} cells), CFL 1.0, q 1 nC, "
                        "sigma_z 25 mm, wakelength 3 m, full wakesolve (both planes)",
            "wakesolve_s": round(wall, 4),
            "stepping_s": round(step_s, 4),
        },
    }
    if not args.no_cpu:
        from oracle.fit_oracle import OGrid, OSolver, OWake

        g, sv = wl["grid"], wl["solver"]
        osv = OSolver(OGrid(x=g["x"], y=g["y"], z=g["z"]), OWake(**wl["wake"]), cfln=sv["cfln"],
                      bc_low=sv["bc_low"], bc_high=sv["bc_high"], solids=g["solids"],
                      bg=sv["bg"], dtype=np.float64)
        osv.one_step()
        steps = 200
        t0 = time.perf_counter()
        for _ in range(steps):
            osv.one_step()
        cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": N * steps / cpu / 1e9,
            "unit": "Gcell-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle FP64 scipy.sparse one_step on the full grid, {steps} of the "
                      f"{Nt} steps, {cpu:.2f} s on 1 core of {cpu_model()}; the full stepping "
                      f"would take {cpu / steps * Nt:.0f} s",
        }
        out["config"]["speedup_vs_cpu_stepping"] = round(cpu / steps * Nt / step_s, 1)
    return out


def bench_batch(args):
    """64 uneven rectangular PEC cavities (box 60-140 mm, pipe 10-25 mm,
    length 60-160 mm, dx = 1.5 mm, z-end PML, beta in {0.8, 0.9, 1.0}) in
    FP32 (or FP64 with --dtype f64), stepped by one launch per pair of steps
    (two-step sweep) or per step, a block per tile, inside hipGraph launch
    groups (the wakesolve inner loop of every member: device beam + E_z
    probe + step), against the same cavities stepped one context at a time."""
    import wakis_amd as wa
    from wakis_amd.workloads import build, cavity_batch

    wls = cavity_batch(n=args.ncav, seed=3)
    t0 = time.perf_counter()
    dtype = np.float32 if args.dtype == "f32" else np.float64
    esz = np.dtype(dtype).itemsize
    solvers = [build(w, dtype=dtype)[2] for w in wls]
    setup_s = time.perf_counter() - t0
    runs = [s._wake_setup(None, None, "longitudinal") for s in solvers]
    for s, r in zip(solvers, runs):
        s._wake_device_beam(r)
    cells = sum(s.N for s in solvers)
    st = wa.BatchStepper(solvers, graph=True, group=16, tile=args.tile,
                         two_step=not args.one_step, kc2=args.kc2, cfg2=args.cfg2)
    pairs = st.two_step()[0]
    # the tuners pick the launch's tile shape (and, two-step, the chunk depth)
    tune = None
    if not args.no_tune and args.tile is None:
        tune = st.tune_two_step(reps=3) if pairs else st.tune(reps=3)
    cfg, nblk, ntiles = st.tile()
    st.step(args.warmup)
    _, cfg2, kc2 = st.two_step()
    print(f"[bench] batch {len(solvers)} cavities, {cells} cells, setup {setup_s:.1f}s, "
          f"two-step {pairs} (tile cfg {cfg2}, kc {kc2}), one-step tile {cfg} blocks {nblk} "
          f"tiles {ntiles} tune {tune}", file=sys.stderr, flush=True)
    ms = st.step(args.steps)
    value = cells * args.steps / (ms * 1e-3) / 1e9
    # the same steps one context at a time (the plain per-solver path)
    seq_steps = max(16, args.steps // 8)
    for s in solvers:
        s.ctx.set_tile(cfg)
        s.ctx.step(2, conductive=s.use_conductivity)
    for s in solvers:
        s.ctx.sync()
    t0 = time.perf_counter()
    for s in solvers:
        s.ctx.step(seq_steps, conductive=s.use_conductivity)
    for s in solvers:
        s.ctx.sync()
    seq_s = time.perf_counter() - t0
    seq_value = cells * seq_steps / seq_s / 1e9
    ms_step = ms / args.steps
    traffic, tsrc = (None, None)
    if pairs:
        # one two-step launch pair (lean chunks, then face chunks of every
        # member) per two steps: 12 field words per cell read and written
        kname = "fused2_batch"
        alg = 2 * 12 * esz * cells
        launch_ms = 2 * ms_step
        if len(solvers) == 64:
            tx2, ty2 = {0: (32, 16), 1: (24, 24), 2: (28, 12)}[cfg2]
            ctype = "float" if args.dtype == "f32" else "double"
            traffic, tsrc = traffic_from_profiles(f"fused2_batch<{ctype}, {tx2}, {ty2}>",
                                                  args.dtype, "batch64", f"f2b-kc{kc2}")
    else:
        kname = "fused_batch"
        alg = 12 * esz * cells
        launch_ms = ms_step
        dims = {0: (320, 3, 64), 1: (128, 4, 32), 2: (64, 8, 32), 3: (256, 4, 32), 4: (64, 4, 32),
                5: (256, 4, 64)}.get(cfg)
        if dims and len(solvers) == 64 and args.dtype == "f32":
            traffic, tsrc = traffic_from_profiles(
                f"fused_batch<float, {dims[0]}, {dims[1]}, {dims[2]}>", "f32", "batch64", cfg)
    achieved = alg / (launch_ms * 1e-3) / 1e9
    sizes = sorted(s.N for s in solvers)
    out_cpu = None if args.no_cpu else cpu_baseline_batch(solvers)
    return {
        "metric": "Gcell-steps/s",
        "value": round(value, 3),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"batch of {len(solvers)} rectangular PEC cavities (seed 3), "
                        f"{sizes[0]}..{sizes[-1]} cells each, {cells} cells in all; "
                        + ("two-step sweeps: one launch pair per two steps"
                           if pairs else "one launch per step")
                        + " in hipGraph groups of 16 steps; wakesolve inner loop",
            "two_step": pairs, "two_step_tile": ({0: "32x16", 1: "24x24", 2: "28x12"}.get(cfg2)
                                                 if pairs else None),
            "kc": kc2 if pairs else None,
            "tile": None if pairs else cfg, "blocks": None if pairs else nblk,
            "tiles": None if pairs else ntiles,
            "one_at_a_time_gcells": round(seq_value, 3),
            "speedup_vs_one_at_a_time": round(value / seq_value, 3),
        },
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": f"profiles/{tsrc}" if tsrc else None,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": round(launch_ms, 5)},
        **({"cpu_baseline": out_cpu} if out_cpu else {}),
    }


def cpu_baseline_batch(solvers, target=2.5e8):
    """The batch's cavities on the CPU oracle (FP64 scipy.sparse): one
    process alone on the median cavity, and P = cpu_procs() processes at once
    on P cavities spread over the size range, each timing about `target`
    cell-steps; value = the P processes' summed cell-step rate."""
    order = np.argsort([s.N for s in solvers])
    med = int(order[len(order) // 2])
    nsteps = lambda i: max(5, int(round(target / solvers[i].N)))
    r1, res1 = cpu_parallel([{"kind": "cavity", "index": med, "steps": nsteps(med)}])
    P = cpu_procs()
    idx = [int(order[(q * len(order)) // P]) for q in range(P)]
    rp, resp = cpu_parallel([{"kind": "cavity", "index": i, "steps": nsteps(i)} for i in idx])
    return {
        "value": rp / 1e9,
        "unit": "Gcell-steps/s",
        "cores": P,
        "kind": "port",
        "single_process_gcells": r1 / 1e9,
        "sample": f"oracle FP64 scipy.sparse one_step on the batch's cavities: {P} processes at "
                  f"once on cavities {idx} (about {target:.1e} cell-steps each, "
                  f"{max(r['seconds'] for r in resp):.1f} s), summed rate; one process alone "
                  f"on the median cavity {med}: {r1 / 1e9:.4f} Gcell-steps/s",
        **cpu_info(),
    }


def fifo_servers(arrivals, service, servers):
    """Completion times of first-come-first-served requests on `servers`
    identical servers."""
    import heapq

    free = [0.0] * servers
    heapq.heapify(free)
    fin = np.empty(len(arrivals))
    for q in np.argsort(arrivals, kind="stable"):
        t = max(float(arrivals[q]), heapq.heappop(free))
        fin[q] = t + float(service[q])
        heapq.heappush(free, fin[q])
    return fin


def cpu_baseline_serve(arrivals, cells, nts, nz_sample=32, target=2.5e8):
    """Solves/hour of the CPU oracle on the same arrival trace, extrapolated
    from measured stepping rates: one process (one request at a time) and
    P = cpu_procs() processes (P requests at a time, first come first
    served), each request's service time cells x Nt / (per-process rate).
    The rates are measured on 256 x 256 x nz_sample slabs of the pillbox
    grid (FP64 oracle, about `target` cell-steps per process); whole
    requests (1e6-1.2e8 cells, hours each on one core) are not run."""
    spec = {"kind": "pillbox_slab", "nz": nz_sample,
            "steps": max(5, int(round(target / (256 * 256 * nz_sample))))}
    r1, _ = cpu_parallel([spec])
    P = cpu_procs()
    rp, resp = cpu_parallel([spec] * P)
    work = np.asarray(cells, dtype=np.float64) * np.asarray(nts, dtype=np.float64)
    arr = np.asarray(arrivals, dtype=np.float64)
    out = {}
    for name, servers, rate in (("one_process", 1, r1), ("p_processes", P, rp / P)):
        fin = fifo_servers(arr, work / rate, servers)
        lat = fin - arr
        out[name] = {"solves_per_h": 3600.0 * len(arr) / (fin.max() - arr.min()),
                     "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)),
                     "gcell_steps_per_s": (rate * servers) / 1e9}
    return {
        "value": out["p_processes"]["solves_per_h"],
        "unit": "solves/h",
        "cores": P,
        "kind": "port",
        "one_process": out["one_process"],
        "p_processes": out["p_processes"],
        "sample": f"extrapolated: oracle FP64 scipy.sparse stepping rates measured on "
                  f"256x256x{nz_sample} pillbox slabs ({spec['steps']} steps; 1 process "
                  f"{r1 / 1e9:.4f}, {P} processes at once {rp / 1e9:.4f} Gcell-steps/s summed), "
                  f"applied to each request's cells x Nt on the same arrivals, FIFO",
        **cpu_info(),
    }


def bench_serve(args):
    """Serving loop: Poisson arrivals (2/s) of cavity impedance requests
    (lognormal cell count, median 8M, clipped 1M-120M; optional lossy
    dielectric slab), continuous batching in FP32 against a 288 GB budget,
    versus the same requests solved one at a time in FP64."""
    from wakis_amd.serve import Server, fifo_latencies, make_requests, solve_alone, z_rel_error

    reqs = make_requests(args.nreq, seed=4, rate=2.0)
    def done_line(r):
        print(f"[serve] req {r.rid:3d} N {r.cells:9d} Nt {r.steps:6d} done at {r.finished:7.2f}s "
              f"(arrived {r.arrival:6.2f}s)", file=sys.stderr, flush=True)

    srv = Server(dtype=np.float32, budget=288e9, chunk_steps=args.chunk, graph=True, group=16,
                 on_done=done_line, two_step=not args.all_batched)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - t0
    lat = np.array([r.latency for r in done])
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    served_per_h = 3600.0 * len(done) / span
    print(f"[bench] served {len(done)} requests in {wall:.1f}s: p50 {np.percentile(lat, 50):.2f}s "
          f"p99 {np.percentile(lat, 99):.2f}s", file=sys.stderr, flush=True)
    busy_ms = sum(e[4] for e in srv.log)
    cell_steps = sum(e[7] for e in srv.log)
    print(f"[bench] device stepping {busy_ms / 1e3:.1f}s of {wall:.1f}s wall, "
          f"{cell_steps / (busy_ms * 1e-3) / 1e9:.1f} Gcell-steps/s while stepping, "
          f"{len(srv.log)} chunks", file=sys.stderr, flush=True)
    if args.served_only:
        return {"metric": "solves/h", "value": round(served_per_h, 1), "unit": "solves/h",
                "n_gpus": 1, "steps": int(sum(r.steps for r in done)), "warmup": 0,
                "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": f"{len(done)} Poisson requests (served only, no baseline)",
                           "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                           "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
                           "stepping_gcell_steps_per_s":
                               round(cell_steps / (busy_ms * 1e-3) / 1e9, 2),
                           "all_batched": bool(args.all_batched)}}
    # one at a time in FP64, same arrivals, FIFO: (a) this library's FP64
    # kernel; (b) the reference's plain GPU method (sparse operators), its
    # stepping timed on each request's grid and scaled to the request's Nt,
    # plus (a)'s post-processing time
    durs, errs, durs_plain, durs_floor, nts, post = [], [], [], [], [], []
    for r in done:
        d, ref = solve_alone(r, dtype=np.float64)
        durs.append(d)
        errs.append(z_rel_error(r.result["Z"], ref["Z"], ref["f"], np.inf))
        nts.append(ref["Nt"])
        post.append(d - ref["stepping_s"])
        print(f"[bench] req {r.rid:3d} N {r.cells:9d} Nt {r.steps:6d} arr {r.arrival:6.2f} "
              f"adm {r.admitted:6.2f} fin {r.finished:6.2f} prep {r.prep_s:5.2f}s "
              f"post {r.finish_s:5.2f}s fp64 {d:5.2f}s dZ {errs[-1]:.2e} "
              f"finite {bool(np.isfinite(r.result['Z']).all())}/{bool(np.isfinite(ref['Z']).all())} "
              f"tile2 {r.result.get('tile2')}",
              file=sys.stderr, flush=True)
    if not args.no_plain:
        # torch carries its own HIP runtime, which cannot start on a device
        # this process's runtime already uses: the plain leg runs in a child
        from wakis_amd import _lib

        ms = plain_child(args.nreq, args.plain_steps)
        copy_gbs = 2 * (2 << 30) / (_lib.copy_bench(2 << 30, reps=10) * 1e-3) / 1e9
        for r, nt, p in zip(done, nts, post):
            pms, nnz = ms[r.rid]
            durs_plain.append(nt * pms * 1e-3 + p)
            floor_ms = plain_floor_bytes(r.cells, nnz) / (copy_gbs * 1e9) * 1e3
            durs_floor.append(nt * floor_ms * 1e-3 + p)
            print(f"[bench] req {r.rid:3d} plain FP64 GPU {pms:8.2f} ms/step (floor "
                  f"{floor_ms:7.2f} at {copy_gbs:.0f} GB/s) -> {durs_plain[-1]:7.1f}s "
                  f"(floor {durs_floor[-1]:7.1f}s)", file=sys.stderr, flush=True)
    arr = np.array([r.arrival for r in done])
    lat64 = fifo_latencies(arr, durs)
    span64 = max(arr + lat64) - arr.min()
    per_h64 = 3600.0 * len(done) / span64
    plain = {}
    if durs_plain:
        latp = fifo_latencies(arr, durs_plain)
        per_hp = 3600.0 * len(done) / (max(arr + latp) - arr.min())
        latf = fifo_latencies(arr, durs_floor)
        per_hf = 3600.0 * len(done) / (max(arr + latf) - arr.min())
        plain = {
            "plain_fp64_gpu_one_at_a_time_solves_per_h": round(per_hp, 2),
            "plain_fp64_gpu_p50_s": round(float(np.percentile(latp, 50)), 2),
            "plain_fp64_gpu_p99_s": round(float(np.percentile(latp, 99)), 2),
            "speedup_vs_plain_fp64_gpu": round(served_per_h / per_hp, 2),
            "plain_fp64_gpu_floor_solves_per_h": round(per_hf, 2),
            "speedup_vs_plain_fp64_gpu_floor": round(served_per_h / per_hf, 2),
            "plain_floor": "each plain step at its fewest bytes (plain_floor_bytes) moved at "
                           "the measured device copy rate",
            "plain_method": "reference's GPU step: FP64 sparse M_H, M_E (torch COO products), "
                            "dt*M_H rebuilt per step, vector temporaries, beam line H2D and "
                            f"E_z window D2H per step; {args.plain_steps} steps timed per request",
        }
    cells = np.array([r.cells for r in done])
    out_cpu = None if args.no_cpu else cpu_baseline_serve(arr, cells, nts)
    return {
        "metric": "solves/h",
        "value": round(served_per_h, 1),
        "unit": "solves/h",
        "n_gpus": 1,
        "steps": int(sum(r.steps for r in done)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(done)} Poisson (2/s, seed 4) cavity impedance requests, cells "
                        f"{int(cells.min())}..{int(cells.max())} (median {int(np.median(cells))}), "
                        "wakelength 1 m, continuous batching FP32 (FP64 wake/DFT), 288 GB budget",
            "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
            "fp64_one_at_a_time_solves_per_h": round(per_h64, 1),
            "fp64_one_at_a_time_p50_s": round(float(np.percentile(lat64, 50)), 3),
            "fp64_one_at_a_time_p99_s": round(float(np.percentile(lat64, 99)), 3),
            "fp64_mean_service_s": round(float(np.mean(durs)), 3),
            "speedup_solves_per_h": round(served_per_h / per_h64, 3),
            **plain,
            "max_rel_dZ_vs_fp64": float(np.max(errs)),
            "device_busy_frac": round(busy_ms / 1e3 / wall, 3),
            "chunk_steps": args.chunk,
        },
        **({"cpu_baseline": out_cpu} if out_cpu else {}),
    }


def bench_precision(args):
    """FP32 against FP64 on the lossy pillbox (the full wakesolve: wakelength
    5 m): relative L2 error of E and H after the run and of the wake
    potential, and max |Z - Z_64| / max |Z_64| over the kept bins (all below
    the bunch cutoff fmax = v/(3 sigma_z)), for FP32 with the reference's
    dtype rounding of dt and material values ("reference") and for FP32
    fields with an FP64 dt, time axis and coefficient setup ("mixed")."""
    from wakis_amd.workloads import build, lossy_pillbox

    nx, ny, nz = (int(v) for v in args.grid.split(","))
    wl = lossy_pillbox(nx, ny, nz)
    if args.wakelength:
        wl["wake"]["wakelength"] = args.wakelength
    out = {}
    res = {}
    for name, dtype, prec in (("fp64", np.float64, "reference"), ("fp32_mixed", np.float32, "mixed"),
                              ("fp32_reference", np.float32, "reference")):
        _, w, s = build(wl, dtype=dtype, precision=prec)
        t0 = time.perf_counter()
        s.wakesolve(compute_plane="longitudinal")
        s.ctx.sync()
        el = time.perf_counter() - t0
        res[name] = {"E": s.E.toarray().astype(np.float64), "H": s.H.toarray().astype(np.float64),
                     "WP": np.asarray(w.WP, dtype=np.float64), "Z": np.asarray(w.Z),
                     "f": np.asarray(w.f), "Nt": s.Nt, "s": np.asarray(w.s)}
        out[name] = {"seconds": round(el, 2), "Nt": int(s.Nt),
                     "gcell_steps_per_s": round(s.N * s.Nt / el / 1e9, 2)}
        print(f"[bench] precision {name}: Nt {s.Nt} {el:.1f}s", file=sys.stderr, flush=True)
        s.ctx.close()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    ref = res["fp64"]
    for name in ("fp32_mixed", "fp32_reference"):
        r = res[name]
        same = r["Nt"] == ref["Nt"] and r["WP"].shape == ref["WP"].shape
        out[name].update({
            "same_time_axis": bool(same),
            "E_rel_l2": rel(r["E"], ref["E"]) if same else None,
            "H_rel_l2": rel(r["H"], ref["H"]) if same else None,
            "WP_rel_l2": rel(r["WP"], ref["WP"]) if same else None,
            "Z_max_rel": (float(np.max(np.abs(r["Z"] - ref["Z"])) / np.max(np.abs(ref["Z"])))
                          if r["Z"].shape == ref["Z"].shape else None),
        })
    return {
        "metric": "FP32 vs FP64 relative error",
        "value": out["fp32_mixed"]["Z_max_rel"],
        "unit": "max |dZ| / max |Z_fp64|",
        "n_gpus": 1,
        "steps": int(ref["Nt"]),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"lossy pillbox {nx}x{ny}x{nz}, wakelength "
                               f"{wl['wake']['wakelength']} m, full wakesolve (longitudinal)",
                   "runs": out, "target": "Z within 1e-3 relative below the bunch cutoff"},
    }


def bench_large(args):
    """One large solve, FP64 against FP32: the stratified box (n^3 cells,
    n = 1024 by default: 1.07e9 cells; 16 dielectric layers eps_r U(1, 12),
    sigma U(0, 10) S/m, a diagonal-tensor central block, 12-cell PML on all
    six faces), the Gaussian bunch (sigma_z = 5 mm, beta = 1) evaluated on
    the device and the E_z probe, `--steps` steps (2000 by default) from
    zero fields.  Both solves are resident at once (FP64 E, H, their second
    buffers and J: 15 x 8 B per cell; FP32: 15 x 4 B), stepped one after the
    other; then the relative L2 error of FP32 E, H and the E_z probe history
    against FP64, read back slab by slab."""
    from wakis_amd import _lib
    from wakis_amd.sources import Beam
    from wakis_amd.workloads import build, stratified_box

    n = args.n_large
    wl = stratified_box(n=n)
    runs, out = {}, {}
    for name, dtype, prec in (("fp64", np.float64, "reference"), ("fp32_mixed", np.float32, "mixed")):
        t0 = time.perf_counter()
        _, wake, s = build(wl, dtype=dtype, precision=prec)
        setup_s = time.perf_counter() - t0
        ctx = s.ctx
        cond = s.use_conductivity
        # untimed warmup and tile search on the zero fields (they stay zero:
        # no source, J = sigma*E = 0), then the bunch and the probe from step 0
        tune_ms = None
        kin = ctx.kernel_in_use()
        if args.tile is not None and dtype == np.float32:
            ctx.set_tile(args.tile)  # the FP32 run's tile (profiling runs)
        elif args.tile2 is not None and kin == _lib.KERNEL_FUSED2:
            ctx.set_tile2(args.tile2)
        elif not args.no_tune and kin == _lib.KERNEL_FUSED:
            tune_ms = s.tune(reps=2)
        elif not args.no_tune and kin == _lib.KERNEL_FUSED2:
            tune_ms = [round(v, 4) for v in ctx.tune2(reps=2)]
        ctx.step(4, cond)
        ctx.sync()
        beam = Beam(q=wake.q, sigmaz=wake.sigmaz, beta=wake.beta)
        ixs, iys = int(np.abs(s.x).argmin()), int(np.abs(s.y).argmin())
        ctx.wake_probe(ixs, iys, 0, args.steps)
        ctx.set_step_index(0)
        ctx.set_beam_gauss(ixs, iys, beam.device_params(s), s.z)
        ctx.zero_ghosts()
        s.step_0 = False
        tile = ctx.get_tile()
        tile2 = ctx.get_tile2() if kin == _lib.KERNEL_FUSED2 else None
        print(f"[bench] large {name}: {s.N} cells, setup {setup_s:.1f}s, dt {float(s.dt):.6e}, "
              f"kernel {kin}, tile {tile}, tile2 {tile2}, tune {tune_ms}", file=sys.stderr,
              flush=True)
        total_ms, done = 0.0, 0
        chunk = 250
        while done < args.steps:
            k = min(chunk, args.steps - done)
            ms, _, _ = ctx.step_timed(k, cond, per_kernel=False)
            total_ms += ms
            done += k
            print(f"[bench] large {name}: step {done}/{args.steps}, {ms / k:.3f} ms/step",
                  file=sys.stderr, flush=True)
        runs[name] = s
        out[name] = {"setup_s": round(setup_s, 1), "ms_per_step": round(total_ms / args.steps, 4),
                     "gcell_steps_per_s": round(s.N * args.steps / (total_ms * 1e-3) / 1e9, 3),
                     "tile": ({"two_step_cfg": tile2[0], "tx": tile2[1], "ty": tile2[2]} if tile2
                              else {"cfg": tile[0], "tx": tile[1], "ty": tile[2], "kc": tile[3]}),
                     "esz": np.dtype(dtype).itemsize}

    # FP32 against FP64, z-slab by z-slab (full x-y planes are contiguous)
    r64, r32 = runs["fp64"], runs["fp32_mixed"]
    slab = max(1, (1 << 26) // (n * n))
    errs = {}
    for fname, which in (("E", _lib.FIELD_E), ("H", _lib.FIELD_H)):
        num = den = 0.0
        for c in range(3):
            for k0 in range(0, n, slab):
                box = [0, n, 0, n, k0, min(n, k0 + slab)]
                a = r64.ctx.read_box(which, c, box)
                b = r32.ctx.read_box(which, c, box).astype(np.float64)
                num += float(np.sum((b - a) ** 2))
                den += float(np.sum(a * a))
        errs[f"{fname}_rel_l2"] = float(np.sqrt(num / max(den, 1e-300)))
        print(f"[bench] large {fname} rel L2 {errs[fname + '_rel_l2']:.3e}", file=sys.stderr,
              flush=True)
    h64 = r64.ctx.wake_history().astype(np.float64)
    h32 = r32.ctx.wake_history().astype(np.float64)
    errs["Ez_probe_rel_l2"] = float(np.linalg.norm(h32 - h64) / max(np.linalg.norm(h64), 1e-300))
    finite = bool(np.isfinite(h64).all() and np.isfinite(h32).all())

    # average fused-kernel launch over a few more steps (after the comparison)
    # (two-step launches: two cell-steps of 12 field words per cell each)
    for name, s in runs.items():
        nk = 20
        c0 = s.ctx.launch_counts()
        _, k1, _ = s.ctx.step_timed(nk, s.use_conductivity, per_kernel=True)
        c1 = s.ctx.launch_counts()
        n_one, n_two = c1[0] - c0[0], c1[1] - c0[1]
        two = n_two > 0 and n_one == 0
        esz = out[name]["esz"]
        alg = (2 if two else 1) * 12 * esz * s.N
        kms = k1 / (n_two if two else nk)
        out[name]["roofline"] = {"bound": "hbm", "kernel": "fused2_sweep" if two else "fused_sweep",
                                 "achieved": round(alg / (kms * 1e-3) / 1e9, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "alg_bytes_per_launch": alg, "avg_launch_ms": round(kms, 4)}
        if two:
            out[name]["roofline"]["steps_per_launch"] = 2
    N = r32.N
    f32 = out["fp32_mixed"]
    tile = f32["tile"]
    if "two_step_cfg" in tile:
        traffic, tsrc = traffic_from_profiles(
            f"fused2_sweep<float, {tile['tx']}, {tile['ty']}>", "f32", f"{n}x{n}x{n}", "f2")
    else:
        traffic, tsrc = traffic_from_profiles(
            f"fused_sweep<float, {tile['tx']}, {tile['ty']}, {tile['kc']}>", "f32",
            f"{n}x{n}x{n}", tile["cfg"])
    roof = dict(f32["roofline"], traffic=traffic,
                traffic_source=f"profiles/{tsrc}" if tsrc else None)
    for s in runs.values():
        s.ctx.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_large()
        print(f"[bench] large cpu baseline {cpu['value']:.4f} Gcell-steps/s ({cpu['sample']})",
              file=sys.stderr, flush=True)
    return {
        "metric": "Gcell-steps/s",
        "value": f32["gcell_steps_per_s"],
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 4,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"stratified box {n}x{n}x{n} ({N} cells): 16 z-layers eps_r "
                               "U(1,12) sigma U(0,10) S/m, diagonal-tensor central block, "
                               "12-cell PML on six faces, bunch sigma_z 5 mm beta 1, "
                               f"{args.steps} steps, FP32 (mixed) vs FP64",
                   "runs": out, "fp32_vs_fp64": errs, "finite": finite},
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 1000; large: 2000 steps of the solve)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--grid", default="256,256,768")
    ap.add_argument("--kernel", default="auto", choices=["auto", "split", "fused", "fused2"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--tile", type=int, default=None, help="fused tile configuration id")
    ap.add_argument("--tile2", type=int, default=None,
                    help="two-step tile shape (0: 64x8, 1: 32x16, 2: 32x8, 3: 16x16, 4: 60x12, "
                         "5: 28x12; FP32 0-4, FP64 2, 3, 5)")
    ap.add_argument("--wakelength", type=float, default=None, help="precision: wakelength (m)")
    ap.add_argument("--workload", default="pillbox",
                    choices=["pillbox", "batch", "serve", "precision", "large", "plain", "cubic",
                             "cpu-worker"],
                    help="pillbox: one 256x256x768 solve (the headline); batch: 64 cavities in "
                         "one launch per step; serve: Poisson impedance requests; large: "
                         "the 1024^3 stratified box, FP32 vs FP64; cubic: the PEC cubic "
                         "cavity 64x64x192 FP64 wakesolve against the CPU oracle")
    ap.add_argument("--ncav", type=int, default=64, help="batch: number of cavities")
    ap.add_argument("--nreq", type=int, default=24, help="serve: number of requests")
    ap.add_argument("--chunk", type=int, default=256, help="serve: steps per batch call")
    ap.add_argument("--all-batched", action="store_true",
                    help="serve: every request in the batched launch (no solves stepped alone)")
    ap.add_argument("--served-only", action="store_true",
                    help="serve: skip the one-at-a-time FP64 baseline (no speedup reported)")
    ap.add_argument("--one-step", action="store_true",
                    help="batch: one-step launches only (no two-step pairs)")
    ap.add_argument("--kc2", type=int, default=0,
                    help="batch: planes per z chunk of the two-step sweep (0: modelled)")
    ap.add_argument("--cfg2", type=int, default=-1,
                    help="batch: two-step tile configuration (0: 32x16, 1: 24x24, -1: modelled)")
    ap.add_argument("--n-large", type=int, default=1024, help="large: cells per side")
    ap.add_argument("--no-plain", action="store_true",
                    help="serve: skip the plain (sparse-operator) FP64 GPU baseline")
    ap.add_argument("--plain-steps", type=int, default=60,
                    help="serve: timed iterations of the plain baseline per request")
    ap.add_argument("--plain", action="store_true",
                    help="pillbox: also time the plain (sparse-operator) FP64 GPU method")
    ap.add_argument("--plain-target", default="serve", choices=["serve", "pillbox"],
                    help=argparse.SUPPRESS)
    ap.add_argument("--cpu-spec", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.workload == "cpu-worker":
        cpu_worker(json.loads(args.cpu_spec))
        return
    if args.steps is None:
        args.steps = 2000 if args.workload == "large" else 1000
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.workload == "precision":
        print(json.dumps(bench_precision(args)), flush=True)
        return
    if args.workload == "large":
        print(json.dumps(bench_large(args)), flush=True)
        return
    if args.workload == "plain":
        print(json.dumps(bench_plain(args)), flush=True)
        return
    if args.workload == "cubic":
        print(json.dumps(bench_cubic(args)), flush=True)
        return

    from wakis_amd.workloads import build, lossy_pillbox

    nx, ny, nz = (int(v) for v in args.grid.split(","))
    dtype = np.float32 if args.dtype == "f32" else np.float64
    esz = np.dtype(dtype).itemsize
    wl = lossy_pillbox(nx, ny, nz)
    t0 = time.perf_counter()
    grid, wake, solver = build(wl, dtype=dtype, kernel=args.kernel)
    if args.tile is not None:
        solver.ctx.set_tile(args.tile)
    if args.tile2 is not None:
        solver.ctx.set_tile2(args.tile2)
    setup_s = time.perf_counter() - t0
    ctx = solver.ctx
    N = solver.N

    # the wakesolve inner loop: device bunch current + E_z probe + step
    from wakis_amd.sources import Beam

    beam = Beam(q=wake.q, sigmaz=wake.sigmaz, beta=wake.beta)
    ixt = int(np.abs(solver.x).argmin())
    iyt = int(np.abs(solver.y).argmin())
    ctx.wake_probe(ixt, iyt, 1, args.warmup + args.steps + 200)
    ctx.set_step_index(0)
    ctx.set_beam_gauss(*[int(v) for v in (np.abs(solver.x).argmin(), np.abs(solver.y).argmin())],
                       beam.device_params(solver), solver.z)
    ctx.zero_ghosts()
    solver.step_0 = False
    cond = solver.use_conductivity

    tune_ms = tune2_ms = None
    if not args.no_tune and args.tile is None and solver.ctx.kernel_in_use() == 2:
        tune_ms = solver.tune(reps=3)
    if not args.no_tune and args.tile2 is None and solver.ctx.kernel_in_use() == 3:
        tune2_ms = [round(v, 4) for v in ctx.tune2(reps=3)]
    tile = ctx.get_tile()
    tile2 = ctx.get_tile2()
    print(f"[bench] {wl['name']} {args.dtype} N={N} setup {setup_s:.1f}s warmup {args.warmup} "
          f"tile {tile} tune {tune_ms} tile2 {tile2} tune2 {tune2_ms}", file=sys.stderr,
          flush=True)
    ctx.step(args.warmup, cond)
    ctx.sync()
    # one timed window: HIP events around every launch of exactly --steps
    # steps; ms_per_step from the whole window, the dominant kernel's average
    # launch from the same launches
    from wakis_amd import _lib

    kin = ctx.kernel_in_use()
    c0 = ctx.launch_counts()
    wall0 = time.perf_counter()
    total_ms, k1, k2 = ctx.step_timed(args.steps, cond, per_kernel=True)
    wall = time.perf_counter() - wall0
    c1 = ctx.launch_counts()
    n_one, n_two = c1[0] - c0[0], c1[1] - c0[1]
    ms_per_step = total_ms / args.steps
    value = N * args.steps / (total_ms * 1e-3) / 1e9
    print(f"[bench] timed {args.steps} steps: {total_ms:.1f} ms (wall {wall * 1e3:.1f} ms), "
          f"launches: {n_two} two-step, {n_one} one-step", file=sys.stderr, flush=True)

    fused = kin in (_lib.KERNEL_FUSED, _lib.KERNEL_FUSED2)
    two = kin == _lib.KERNEL_FUSED2 and n_two > 0 and n_one == 0
    ctype = "float" if args.dtype == "f32" else "double"
    if two:
        # one launch = two steps of every cell: 2 x 12 field words per cell
        # algorithmically (the per-cell-step figure), while the launch itself
        # reads and writes E and H once (12 words per cell)
        kname, kms, alg_bytes = "fused2_sweep", k1 / n_two, 2 * 12 * esz * N
        traffic, tsrc = traffic_from_profiles(f"fused2_sweep<{ctype}, {tile2[1]}, {tile2[2]}>",
                                              args.dtype, f"{nx}x{ny}x{nz}", "f2")
        extra = {"traffic_source": f"profiles/{tsrc}" if tsrc else None,
                 "steps_per_launch": 2,
                 "launch_min_hbm_bytes": 12 * esz * N,
                 "launch_min_hbm_gbs": round(12 * esz * N / (kms * 1e-3) / 1e9, 1)}
    elif fused:
        # one launch per step: E and H read once and written once per cell
        kname, kms, alg_bytes = "fused_sweep", k1 / args.steps, 12 * esz * N
        traffic, tsrc = traffic_from_profiles(
            f"fused_sweep<{ctype}, {tile[1]}, {tile[2]}, {tile[3]}>", args.dtype,
            f"{nx}x{ny}x{nz}", tile[0])
        extra = {"traffic_source": f"profiles/{tsrc}" if tsrc else None}
    elif k2 >= k1:
        # E update: reads H and E, writes E (9 field words per cell)
        kname, kms, alg_bytes = "e_update_split", k2 / args.steps, 9 * esz * N
        extra = {"h_update_avg_ms": round(k1 / args.steps, 5)}
    else:
        kname, kms, alg_bytes = "h_update_split", k1 / args.steps, 9 * esz * N
        extra = {"e_update_avg_ms": round(k2 / args.steps, 5)}
    if not fused:
        traffic = None
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    step_alg = 12 * esz * N / (ms_per_step * 1e-3) / 1e9
    copy_bytes = 2 * 12 * esz * N  # same footprint as the fields (read + write)
    copy_ms = _lib.copy_bench(copy_bytes // 2, reps=10)
    copy_gbs = copy_bytes / (copy_ms * 1e-3) / 1e9
    if fused and not two:
        # a plain copy of the twelve field streams with the sweep's traversal
        try:
            tiled_ms = _lib.copy_bench_tiled(nx, ny, nz, tile[0], dtype=dtype, reps=10)
            tiled_gbs = 12 * esz * N / (tiled_ms * 1e-3) / 1e9
            extra["copy_same_traversal_gbs"] = round(tiled_gbs, 1)
            extra["frac_of_copy_same_traversal"] = round(achieved / tiled_gbs, 4)
        except _lib.FitError as e:
            print(f"[bench] tiled copy calibration failed: {e}", file=sys.stderr, flush=True)

    out = {
        "metric": "Gcell-steps/s",
        "value": round(value, 3),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"lossy pillbox cavity {nx}x{ny}x{nz} ({N} cells), sigma=1e5 walls, "
                        "10-cell z PML with a cubic sigma profile, wakesolve inner loop "
                        "(beam + E_z probe + step)",
            "grid": [nx, ny, nz],
            "kernel": "fused2" if two else ("fused" if fused else "split"),
            "tile": ({"two_step_cfg": tile2[0], "tx": tile2[1], "ty": tile2[2],
                      "tune2_ms": tune2_ms} if two else
                     {"cfg": tile[0], "tx": tile[1], "ty": tile[2], "kc": tile[3]} if fused
                     else None),
            "step_alg_gbs": round(step_alg, 1),
            "step_hbm_frac": round(step_alg / HBM_PEAK_GBS, 4),
            "copy_kernel_gbs": round(copy_gbs, 1),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": alg_bytes,
            "avg_launch_ms": round(kms, 5),
            **extra,
        },
    }
    if args.plain:
        pms, nnz = plain_child(0, args.plain_steps, "pillbox", args.grid)[0]
        out["config"]["plain_fp64_gpu"] = {
            "gcell_steps_per_s": round(N / (pms * 1e-3) / 1e9, 3),
            "ms_per_step": round(pms, 3),
            "nnz": nnz,
            "method": "the reference's GPU step (sparse M_H, M_E in FP64, torch COO products) "
                      "with its wakesolve iteration (beam line H2D, E_z window D2H)",
        }
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
