"""The plain-method baseline of the benchmarks (the reference's sparse
operator step, bench.PlainSparseStepper) against the oracle: on the CPU
here, and on the device (hipSPARSE) in the gpu-marked test."""

import numpy as np
import pytest

from helpers import OGrid, OSolver, axes, block_mask, rel

CASES = [
    (["pec"] * 3, ["pec"] * 3, "pec", True),
    (["pec", "pec", "pml"], ["pec", "pec", "pml"], (1.0, 1.0), False),
    (["pmc", "pec", "periodic"], ["pec", "pmc", "periodic"], (1.0, 1.0), False),
]


def _pair(bl, bh, bg, solids, shape=(9, 8, 12)):
    import wakis_amd as wa

    x, y, z = axes(*shape, nonuniform=True, seed=1)
    sol = []
    if solids:
        m = block_mask(shape, (2, 2, 3), (7, 6, 9)) | block_mask(shape, (3, 3, 0), (6, 5, 12))
        sol = [(m, "vacuum"), (block_mask(shape, (4, 3, 5), (6, 5, 7)), [3.0, 1.0, 10.0])]
    osv = OSolver(OGrid(x=x, y=y, z=z), cfln=0.5, bc_low=bl, bc_high=bh, solids=list(sol),
                  bg=bg, dtype=np.float64, n_pml=3)
    hs = wa.SolverFIT3D(wa.GridFIT3D(x=x, y=y, z=z, solids=list(sol)), cfln=0.5, bc_low=bl,
                        bc_high=bh, use_stl=bool(sol), bg=bg, dtype=np.float64, n_pml=3,
                        device=None)
    return osv, hs


def _run(device, case, nsteps=20):
    import torch

    from bench import PlainSparseStepper

    osv, hs = _pair(*case)
    st = PlainSparseStepper(hs, device=device)
    rng = np.random.default_rng(7)
    for name in "EHJ":
        a = rng.standard_normal(3 * osv.N)
        getattr(osv, name).fromarray(a)
        getattr(st, name).copy_(torch.from_numpy(a).to(st.dev))
    for _ in range(nsteps):
        osv.one_step()
        st.one_step()
    return {n: rel(getattr(st, n).cpu().numpy(), getattr(osv, n).toarray()) for n in "EHJ"}


@pytest.mark.parametrize("case", CASES, ids=["pec-solids", "pml-conductive", "pmc-periodic"])
def test_plain_stepper_cpu_matches_oracle(case):
    errs = _run("cpu", case)
    assert max(errs.values()) <= 1e-13, errs


@pytest.mark.gpu
def test_plain_stepper_device_matches_oracle():
    """On the device through hipSPARSE.  torch brings its own HIP runtime,
    which cannot start on a device that this process's runtime (the
    wakis_amd library, used by the other GPU tests) already holds, so the
    cases run in a child process."""
    import json
    import os
    import subprocess
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    code = ("import sys, json; sys.path[:0] = [%r, %r]; import torch; torch.cuda.init(); "
            "import test_plain_baseline as t; "
            "print(json.dumps([t._run('cuda', c) for c in t.CASES]))"
            % (here, os.path.dirname(here)))
    print("[case] plain stepper on the device (child process)", flush=True)
    out = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, timeout=600,
                         check=True, cwd=os.path.dirname(here)).stdout
    errs = json.loads(out.decode().strip().splitlines()[-1])
    assert len(errs) == len(CASES)
    for e in errs:
        assert max(e.values()) <= 1e-12, errs
