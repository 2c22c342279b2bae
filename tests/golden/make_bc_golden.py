"""Writes tests/golden/bc_pml_beam.npz: input/output vectors of three
reference functions that no closed form or reference test pins:

* ``BCsMixin._apply_bc_to_C`` (wakis/boundaries.py:14-188): periodic and ABC
  metric fix-ups (tL, itA), the PEC/PML column mask and the PMC row mask of
  the curl C (with the PMC block reusing the PEC block's face flags), and
  the activate_abc / activate_pml / use_conductivity flags;
* ``BCsMixin._initialize_PML`` (wakis/boundaries.py:190-305): ieps and sigma
  after the geomspace grading, including the corner overwrite order
  xlo, ylo, zlo, xhi, yhi, zhi;
* ``Beam.update`` (wakis/sources.py:65-110): the J_z line for beta 0.8, 0.9
  and 1.0 (the injection time without sqrt(beta)), with and without a given
  ``ti``, on a non-uniform z axis.

The reference's ``field.py``, ``boundaries.py`` and ``sources.py`` import with
numpy, scipy and matplotlib only, so they are imported here directly (the
package ``__init__`` is not run: it pulls in pyvista and h5py), each case is
run on a small duck-typed solver, and only the numbers are stored.  The curl
C given to ``_apply_bc_to_C`` is the reference's assembly
(wakis/solverFIT3D.py:207-227, restated below: that module imports h5py);
the metric Fields are seeded random positive values so that every copy is
visible.  Run from the repository root:

    python tests/golden/make_bc_golden.py
"""

import importlib
import os
import sys
import types
from types import SimpleNamespace

sys.dont_write_bytecode = True  # nothing is written under the reference tree

import numpy as np  # noqa: E402
from scipy.constants import epsilon_0  # noqa: E402
from scipy.sparse import csc_matrix, diags, hstack, vstack  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/wakis"

NX, NY, NZ = 4, 5, 6
# (bc_low, bc_high): every BC kind, mixes, case variants of pmc/magnetic
BC_CASES = {
    "pec": (["pec"] * 3, ["pec"] * 3),
    "periodic": (["periodic"] * 3, ["periodic"] * 3),
    "mixed": (["periodic", "pec", "pec"], ["periodic", "pmc", "pml"]),
    "pmc": (["pmc", "pec", "magnetic"], ["pec", "pmc", "pec"]),
    "pmc_upper": (["PMC", "pec", "pec"], ["pec", "Magnetic", "pec"]),
    "pmc_only": (["pmc"] * 3, ["pmc"] * 3),
    "pec_pmc_edges": (["pec", "pmc", "pec"], ["pmc", "pec", "pmc"]),
    "abc": (["pec", "abc", "pec"], ["abc", "pec", "abc"]),
    "abc_high_all": (["pec"] * 3, ["abc"] * 3),
    "abc_periodic": (["periodic", "abc", "pec"], ["periodic", "abc", "abc"]),
    "pml_z": (["pec", "pec", "pml"], ["pec", "pec", "pml"]),
    "pml_all": (["pml"] * 3, ["pml"] * 3),
    "pml_pmc": (["pml", "pmc", "pec"], ["pmc", "pml", "pml"]),
    "open_x": (["pec", "periodic", "pec"], ["periodic", "periodic", "pec"]),
    "electric": (["electric", "periodic", "pec"], ["electric", "periodic", "pml"]),
}
N_PML = 2
# PML cases for _initialize_PML (grid large enough for n_pml on both sides)
PML_CASES = {
    "pml_z": (["pec", "pec", "pml"], ["pec", "pec", "pml"]),
    "pml_all": (["pml"] * 3, ["pml"] * 3),
    "pml_low_x_high_y": (["pml", "pec", "pec"], ["pec", "pml", "pec"]),
    "pml_mixed": (["pml", "pmc", "pec"], ["pmc", "pml", "pml"]),
}
PNX, PNY, PNZ, PN = 7, 8, 9, 3
BETAS = (0.8, 0.9, 1.0)


def _reference_modules():
    pkg = types.ModuleType("_wakis_src")
    pkg.__path__ = [REF]
    sys.modules["_wakis_src"] = pkg
    field = importlib.import_module("_wakis_src.field")
    bcs = importlib.import_module("_wakis_src.boundaries")
    sources = importlib.import_module("_wakis_src.sources")
    return field.Field, bcs.BCsMixin, sources.Beam


def curl(nx, ny, nz):
    """The reference's curl assembly (wakis/solverFIT3D.py:207-227)."""
    N = nx * ny * nz
    Px = diags([-1, 1], [0, 1], shape=(N, N), dtype=np.int8)
    Py = diags([-1, 1], [0, nx], shape=(N, N), dtype=np.int8)
    Pz = diags([-1, 1], [0, nx * ny], shape=(N, N), dtype=np.int8)
    Z0 = csc_matrix((N, N))
    return vstack([hstack([Z0, -Pz, Py]), hstack([Pz, Z0, -Px]), hstack([-Py, Px, Z0])],
                  dtype=np.int8)


def coo(m):
    m = m.tocoo()
    order = np.lexsort((m.col, m.row))
    return np.stack([m.row[order], m.col[order]]).astype(np.int32), m.data[order]


def main():
    Field, BCs, Beam = _reference_modules()
    rng = np.random.default_rng(7)
    out = {"dims": np.array([NX, NY, NZ]), "cases": np.array(list(BC_CASES))}
    n3 = 3 * NX * NY * NZ
    metrics = {k: rng.uniform(0.5, 2.0, n3) for k in ("L", "iA", "tL", "itA")}
    for k, v in metrics.items():
        out["in/" + k] = v
    for name, (bl, bh) in BC_CASES.items():
        s = SimpleNamespace(use_mpi=False, grid=SimpleNamespace(use_mpi=False),
                            bc_low=list(bl), bc_high=list(bh), Nx=NX, Ny=NY, Nz=NZ,
                            N=NX * NY * NZ, dtype=np.float64, C=curl(NX, NY, NZ),
                            activate_abc=False, activate_pml=False, use_conductivity=False)
        for k, v in metrics.items():
            f = Field(NX, NY, NZ)
            f.fromarray(v.copy())
            setattr(s, k, f)
        BCs._apply_bc_to_C(s)
        rc, val = coo(s.C)
        out[name + "/C_idx"] = rc
        out[name + "/C_val"] = val.astype(np.int8)
        out[name + "/tL"] = s.tL.toarray().copy()
        out[name + "/itA"] = s.itA.toarray().copy()
        out[name + "/BC"] = (s.BC.toarray().copy() if hasattr(s, "BC")
                             else np.zeros(0, dtype=np.int8))
        out[name + "/flags"] = np.array([s.activate_abc, s.activate_pml, s.use_conductivity])
        print(name, "C nnz", val.size, flush=True)

    out["pml/dims"] = np.array([PNX, PNY, PNZ, PN])
    out["pml/cases"] = np.array(list(PML_CASES))
    for name, (bl, bh) in PML_CASES.items():
        # background eps_r = 2 (ieps != 1/eps_0 outside the PML), sigma 0.25
        s = SimpleNamespace(bc_low=list(bl), bc_high=list(bh), Nx=PNX, Ny=PNY, Nz=PNZ,
                            n_pml=PN)
        s.ieps = Field(PNX, PNY, PNZ, use_ones=True) * (1.0 / (2.0 * epsilon_0))
        s.sigma = Field(PNX, PNY, PNZ, use_ones=True) * 0.25
        BCs._initialize_PML(s)
        out["pml/" + name + "/ieps"] = s.ieps.toarray().copy()
        out["pml/" + name + "/sigma"] = s.sigma.toarray().copy()
        print("pml", name, flush=True)

    # Beam: non-uniform z, off-centre source, several times
    bx = np.linspace(-0.02, 0.02, 6)
    by = np.linspace(-0.015, 0.015, 5)
    dz = rng.uniform(0.6, 1.4, 11)
    bz = np.concatenate([[0.0], np.cumsum(dz)]) / dz.sum() * 0.1 - 0.05
    x, y, z = bx[:-1] + np.diff(bx) / 2, by[:-1] + np.diff(by) / 2, bz[:-1] + np.diff(bz) / 2
    times = np.array([0.0, 1.3e-11, 7.7e-11, 2.5e-10])
    out["beam/axes_x"], out["beam/axes_y"], out["beam/axes_z"] = bx, by, bz
    out["beam/times"] = times
    out["beam/betas"] = np.array(BETAS)
    for beta in BETAS:
        for ti in (None, 4.0e-11):
            key = f"beam/b{beta}_ti{'default' if ti is None else 'given'}"
            b = Beam(xsource=0.004, ysource=-0.003, beta=beta, q=2e-9, sigmaz=0.012, ti=ti)
            sol = SimpleNamespace(x=x, y=y, z=z, dx=np.diff(bx), dy=np.diff(by),
                                  dz=np.diff(bz), J=Field(len(x), len(y), len(z)))
            rows = []
            for t in times:
                b.update(sol, t)
                rows.append(sol.J.toarray().copy())
            out[key + "/J"] = np.array(rows)
            out[key + "/meta"] = np.array([b.ixs, b.iys, b.zmin, b.ti, b.v])
            print(key, flush=True)
    np.savez_compressed(os.path.join(HERE, "bc_pml_beam.npz"), **out)


if __name__ == "__main__":
    main()
