"""Writes tests/golden/planewave.npz: input/output vectors of the reference's
``PlaneWave.update`` (wakis/sources.py:142-248), which no closed form or
reference test pins on its own: the E_x and H_y slices it writes on the
reference's own ``Field`` (wakis/field.py), for default and given
transverse slices, an integer and a slice ``zs``, float32 and float64
fields, times before and after ``tmax = nodes / f``.

The reference's ``field.py`` and ``sources.py`` import with numpy, scipy
and matplotlib only, so they are imported here directly (the package
``__init__`` is not run: it pulls in pyvista and h5py); each case runs on a
duck-typed solver holding ``Nx, Ny, E, H`` and only the numbers are stored.
Run from the repository root:

    python tests/golden/make_planewave_golden.py
"""

import importlib
import os
import sys
import types
from types import SimpleNamespace

sys.dont_write_bytecode = True  # nothing is written under the reference tree

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/wakis"

NX, NY, NZ = 5, 4, 7
# name: (PlaneWave kwargs, dtype)
CASES = {
    "default_f64": (dict(nodes=3, f=1.5e9, amplitude=2.5, phase=0.3), np.float64),
    "slices_f32": (dict(xs=slice(1, 4), ys=slice(0, 3), zs=2, nodes=2, f=0.8e9, amplitude=100.0,
                        phase=np.pi / 2), np.float32),
    "zslice_beta": (dict(zs=slice(1, 3), f=2.0e9, amplitude=1.0, beta=0.7), np.float64),
}
TIMES = np.array([0.0, 0.3e-9, 1.1e-9, 2.4e-9, 3.9e-9])


def _reference_modules():
    pkg = types.ModuleType("_wakis_src")
    pkg.__path__ = [REF]
    sys.modules["_wakis_src"] = pkg
    field = importlib.import_module("_wakis_src.field")
    sources = importlib.import_module("_wakis_src.sources")
    return field.Field, sources.PlaneWave


def main():
    Field, PlaneWave = _reference_modules()
    out = {"dims": np.array([NX, NY, NZ]), "cases": np.array(list(CASES)), "times": TIMES}
    for name, (kw, dtype) in CASES.items():
        s = SimpleNamespace(Nx=NX, Ny=NY, Nz=NZ, E=Field(NX, NY, NZ, dtype=dtype),
                            H=Field(NX, NY, NZ, dtype=dtype))
        pw = PlaneWave(**kw)
        for q, t in enumerate(TIMES):
            pw.update(s, t)
            out[f"{name}/E/{q}"] = np.array(s.E.toarray())
            out[f"{name}/H/{q}"] = np.array(s.H.toarray())
    np.savez_compressed(os.path.join(HERE, "planewave.npz"), **out)
    print("wrote", os.path.join(HERE, "planewave.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
