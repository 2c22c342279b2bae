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


def cpu_baseline(nz_sample=32, steps=150):
    """The CPU oracle (scipy.sparse FP64, single-threaded csc/csr mat-vec, the
    reference's arithmetic path) on a z-slab of the same pillbox grid."""
    from oracle.fit_oracle import OGrid, OSolver
    from wakis_amd.workloads import lossy_pillbox

    w = lossy_pillbox(256, 256, nz_sample)
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
                  f"grid, {steps} steps, {dt:.2f} s on 1 core of {cpu_model()}",
    }


def bench_batch(args):
    """64 uneven rectangular PEC cavities (box 60-140 mm, pipe 10-25 mm,
    length 60-160 mm, dx = 1.5 mm, z-end PML, beta in {0.8, 0.9, 1.0}) in
    FP32, stepped by one persistent launch per step inside hipGraph launch
    groups (the wakesolve inner loop of every member: device beam + E_z
    probe + step), against the same cavities stepped one context at a time."""
    import wakis_amd as wa
    from wakis_amd.workloads import build, cavity_batch

    wls = cavity_batch(n=args.ncav, seed=3)
    t0 = time.perf_counter()
    solvers = [build(w, dtype=np.float32)[2] for w in wls]
    setup_s = time.perf_counter() - t0
    runs = [s._wake_setup(None, None, "longitudinal") for s in solvers]
    for s, r in zip(solvers, runs):
        s._wake_device_beam(r)
    cells = sum(s.N for s in solvers)
    st = wa.BatchStepper(solvers, graph=True, group=16, tile=args.tile)
    tune = st.tune(reps=3) if not args.no_tune and args.tile is None else None
    cfg, nblk, ntiles = st.tile()
    print(f"[bench] batch {len(solvers)} cavities, {cells} cells, setup {setup_s:.1f}s, "
          f"tile {cfg} blocks {nblk} tiles {ntiles} tune {tune}", file=sys.stderr, flush=True)
    st.step(args.warmup)
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
    alg = 12 * 4 * cells
    dims = {0: (256, 2, 32), 1: (128, 4, 32), 2: (64, 8, 32), 3: (256, 4, 32), 4: (64, 4, 32),
            5: (256, 4, 64)}.get(cfg)
    traffic, tsrc = (None, None)
    if dims and len(solvers) == 64:
        traffic, tsrc = traffic_from_profiles(f"fused_batch<float, {dims[0]}, {dims[1]}, {dims[2]}>",
                                              "f32", "batch64", cfg)
    achieved = alg / (ms_step * 1e-3) / 1e9
    sizes = sorted(s.N for s in solvers)
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
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batch of {len(solvers)} rectangular PEC cavities (seed 3), "
                        f"{sizes[0]}..{sizes[-1]} cells each, {cells} cells in all; one persistent "
                        "launch per step in hipGraph groups of 16; wakesolve inner loop",
            "tile": cfg, "blocks": nblk, "tiles": ntiles,
            "one_at_a_time_gcells": round(seq_value, 3),
            "speedup_vs_one_at_a_time": round(value / seq_value, 3),
        },
        "roofline": {"bound": "hbm", "kernel": "fused_batch", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": f"profiles/{tsrc}" if tsrc else None,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": round(ms_step, 5)},
    }


def bench_serve(args):
    """Serving loop: Poisson arrivals (2/s) of cavity impedance requests
    (lognormal cell count, median 8M, clipped 1M-120M; optional lossy
    dielectric slab), continuous batching in FP32 against a 288 GB budget,
    versus the same requests solved one at a time in FP64."""
    from wakis_amd.serve import Server, fifo_latencies, make_requests, solve_alone, z_rel_error

    reqs = make_requests(args.nreq, seed=4, rate=2.0)
    srv = Server(dtype=np.float32, budget=288e9, chunk_steps=args.chunk, graph=True, group=16)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - t0
    lat = np.array([r.latency for r in done])
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    served_per_h = 3600.0 * len(done) / span
    print(f"[bench] served {len(done)} requests in {wall:.1f}s: p50 {np.percentile(lat, 50):.2f}s "
          f"p99 {np.percentile(lat, 99):.2f}s", file=sys.stderr, flush=True)
    busy_ms = sum(e[4] for e in srv.log)
    cell_steps = sum(e[5] * e[2] for e in srv.log)
    print(f"[bench] device stepping {busy_ms / 1e3:.1f}s of {wall:.1f}s wall, "
          f"{cell_steps / (busy_ms * 1e-3) / 1e9:.1f} Gcell-steps/s while stepping, "
          f"{len(srv.log)} chunks", file=sys.stderr, flush=True)
    # one at a time in FP64 (the plain path), same arrivals, FIFO
    durs, errs = [], []
    for r in done:
        d, ref = solve_alone(r, dtype=np.float64)
        durs.append(d)
        errs.append(z_rel_error(r.result["Z"], ref["Z"], ref["f"], np.inf))
        print(f"[bench] req {r.rid:3d} N {r.cells:9d} Nt {r.steps:6d} arr {r.arrival:6.2f} "
              f"adm {r.admitted:6.2f} fin {r.finished:6.2f} prep {r.prep_s:5.2f}s "
              f"post {r.finish_s:5.2f}s fp64 {d:5.2f}s dZ {errs[-1]:.2e} "
              f"finite {bool(np.isfinite(r.result['Z']).all())}/{bool(np.isfinite(ref['Z']).all())}",
              file=sys.stderr, flush=True)
    arr = np.array([r.arrival for r in done])
    lat64 = fifo_latencies(arr, durs)
    span64 = max(arr + lat64) - arr.min()
    per_h64 = 3600.0 * len(done) / span64
    cells = np.array([r.cells for r in done])
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
            "max_rel_dZ_vs_fp64": float(np.max(errs)),
            "device_busy_frac": round(busy_ms / 1e3 / wall, 3),
            "chunk_steps": args.chunk,
        },
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
        if not args.no_tune and ctx.kernel_in_use() == _lib.KERNEL_FUSED:
            tune_ms = s.tune(reps=2)
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
        print(f"[bench] large {name}: {s.N} cells, setup {setup_s:.1f}s, dt {float(s.dt):.6e}, "
              f"tile {tile}, tune {tune_ms}", file=sys.stderr, flush=True)
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
                     "tile": {"cfg": tile[0], "tx": tile[1], "ty": tile[2], "kc": tile[3]},
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
    for name, s in runs.items():
        nk = 20
        _, k1, _ = s.ctx.step_timed(nk, s.use_conductivity, per_kernel=True)
        esz = out[name]["esz"]
        alg = 12 * esz * s.N
        kms = k1 / nk
        out[name]["roofline"] = {"bound": "hbm", "kernel": "fused_sweep",
                                 "achieved": round(alg / (kms * 1e-3) / 1e9, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "alg_bytes_per_launch": alg, "avg_launch_ms": round(kms, 4)}
    N = r32.N
    f32 = out["fp32_mixed"]
    tile = f32["tile"]
    traffic, tsrc = traffic_from_profiles(
        f"fused_sweep<float, {tile['tx']}, {tile['ty']}, {tile['kc']}>", "f32", f"{n}x{n}x{n}",
        tile["cfg"])
    roof = dict(f32["roofline"], traffic=traffic,
                traffic_source=f"profiles/{tsrc}" if tsrc else None)
    for s in runs.values():
        s.ctx.close()
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
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 1000; large: 2000 steps of the solve)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--grid", default="256,256,768")
    ap.add_argument("--kernel", default="auto", choices=["auto", "split", "fused"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--tile", type=int, default=None, help="fused tile configuration id")
    ap.add_argument("--wakelength", type=float, default=None, help="precision: wakelength (m)")
    ap.add_argument("--workload", default="pillbox",
                    choices=["pillbox", "batch", "serve", "precision", "large"],
                    help="pillbox: one 256x256x768 solve (the headline); batch: 64 cavities in "
                         "one launch per step; serve: Poisson impedance requests; large: "
                         "the 1024^3 stratified box, FP32 vs FP64")
    ap.add_argument("--ncav", type=int, default=64, help="batch: number of cavities")
    ap.add_argument("--nreq", type=int, default=24, help="serve: number of requests")
    ap.add_argument("--chunk", type=int, default=256, help="serve: steps per batch call")
    ap.add_argument("--n-large", type=int, default=1024, help="large: cells per side")
    args = ap.parse_args()
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

    from wakis_amd.workloads import build, lossy_pillbox

    nx, ny, nz = (int(v) for v in args.grid.split(","))
    dtype = np.float32 if args.dtype == "f32" else np.float64
    esz = np.dtype(dtype).itemsize
    wl = lossy_pillbox(nx, ny, nz)
    t0 = time.perf_counter()
    grid, wake, solver = build(wl, dtype=dtype, kernel=args.kernel)
    if args.tile is not None:
        solver.ctx.set_tile(args.tile)
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

    tune_ms = None
    if not args.no_tune and args.tile is None and solver.ctx.kernel_in_use() == 2:
        tune_ms = solver.tune(reps=3)
    tile = ctx.get_tile()
    print(f"[bench] {wl['name']} {args.dtype} N={N} setup {setup_s:.1f}s warmup {args.warmup} "
          f"tile {tile} tune {tune_ms}", file=sys.stderr, flush=True)
    ctx.step(args.warmup, cond)
    ctx.sync()
    wall0 = time.perf_counter()
    total_ms, _, _ = ctx.step_timed(args.steps, cond, per_kernel=False)
    wall = time.perf_counter() - wall0
    ms_per_step = total_ms / args.steps
    value = N * args.steps / (total_ms * 1e-3) / 1e9
    print(f"[bench] timed {args.steps} steps: {total_ms:.1f} ms (wall {wall * 1e3:.1f} ms)",
          file=sys.stderr, flush=True)

    # per-kernel durations for the roofline of the dominant kernel
    from wakis_amd import _lib

    fused = ctx.kernel_in_use() == _lib.KERNEL_FUSED
    nk = min(args.steps, 100)
    _, k1, k2 = ctx.step_timed(nk, cond, per_kernel=True)
    if fused:
        # one launch per step: E and H read once and written once per cell
        kname, kms, alg_bytes = "fused_sweep", k1 / nk, 12 * esz * N
        ctype = "float" if args.dtype == "f32" else "double"
        traffic, tsrc = traffic_from_profiles(
            f"fused_sweep<{ctype}, {tile[1]}, {tile[2]}, {tile[3]}>", args.dtype,
            f"{nx}x{ny}x{nz}", tile[0])
        extra = {"traffic_source": f"profiles/{tsrc}" if tsrc else None}
    elif k2 >= k1:
        # E update: reads H and E, writes E (9 field words per cell)
        kname, kms, alg_bytes = "e_update_split", k2 / nk, 9 * esz * N
        extra = {"h_update_avg_ms": round(k1 / nk, 5)}
    else:
        kname, kms, alg_bytes = "h_update_split", k1 / nk, 9 * esz * N
        extra = {"e_update_avg_ms": round(k2 / nk, 5)}
    if not fused:
        traffic = None
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    step_alg = 12 * esz * N / (ms_per_step * 1e-3) / 1e9
    copy_bytes = 2 * 12 * esz * N  # same footprint as the fields (read + write)
    copy_ms = _lib.copy_bench(copy_bytes // 2, reps=10)
    copy_gbs = copy_bytes / (copy_ms * 1e-3) / 1e9

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
                        "10-cell z PML, wakesolve inner loop (beam + E_z probe + step)",
            "grid": [nx, ny, nz],
            "kernel": "fused" if fused else "split",
            "tile": {"cfg": tile[0], "tx": tile[1], "ty": tile[2], "kc": tile[3]} if fused else None,
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
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
