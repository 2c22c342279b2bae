"""Writes tests/golden/abc_aliasing.npz: input/output vectors of the
reference's absorbing boundary (wakis/boundaries.py:307-404 ``get_abc`` /
``update_abc``) acting on the reference's own ``Field`` (wakis/field.py),
driven through the two-snapshot lag of ``emsolve`` (wakis/routines.py:182-216).

What these vectors fix is how the saved planes alias the live field: on the
CPU a tuple read of a ``Field`` returns a view, so the "n-2" snapshot is the
current plane 1 when ``update_abc`` runs.  No closed form or reference test
pins that, so the reference's two modules are imported here (numpy only; the
package ``__init__`` is not run, since it pulls in pyvista and h5py), each
case is run, and only the numbers are stored.

The time step is replaced by a deterministic in-place update
``X.fromarray(a X + dX[n])``, like ``SolverFIT3D._one_step``'s in-place
``fromarray`` (wakis/solverFIT3D.py:510-522).  Cases: each face kind alone,
a mix, all three low faces, and low+high ABC on one axis (which raises
``KeyError`` in ``update_abc``, after the low faces of earlier axes were
written: the high-face branch of ``get_abc`` rebuilds that axis'
dictionary).  Run from the repository root:

    python tests/golden/make_abc_golden.py
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

NX, NY, NZ, NT = 4, 5, 6, 4
A_E, A_H = 0.5, 0.25
CASES = {
    "low_x": (("abc", "pec", "pec"), ("pec", "pec", "pec")),
    "low_y": (("pec", "abc", "pec"), ("pec", "pec", "pec")),
    "low_z": (("pec", "pec", "abc"), ("pec", "pec", "pec")),
    "high_z": (("pec", "pec", "pec"), ("pec", "pec", "abc")),
    "high_x_low_y": (("pec", "abc", "pec"), ("abc", "pec", "pec")),
    "low_xyz": (("abc", "abc", "abc"), ("pec", "pec", "pec")),
    "low_high_z": (("pec", "pec", "abc"), ("pec", "pec", "abc")),
    "low_x_low_high_z": (("abc", "pec", "abc"), ("pec", "pec", "abc")),
}


def _reference_modules():
    pkg = types.ModuleType("_wakis_src")
    pkg.__path__ = [REF]
    sys.modules["_wakis_src"] = pkg
    field = importlib.import_module("_wakis_src.field")
    bcs = importlib.import_module("_wakis_src.boundaries")
    return field.Field, bcs.BCsMixin


def main():
    Field, BCs = _reference_modules()
    rng = np.random.default_rng(2024)
    n3 = 3 * NX * NY * NZ
    E0, H0 = rng.standard_normal(n3), rng.standard_normal(n3)
    dE, dH = rng.standard_normal((NT, n3)), rng.standard_normal((NT, n3))
    out = {"dims": np.array([NX, NY, NZ, NT]), "a": np.array([A_E, A_H]),
           "E0": E0, "H0": H0, "dE": dE, "dH": dH,
           "cases": np.array(list(CASES))}
    for name, (bl, bh) in CASES.items():
        s = SimpleNamespace(bc_low=list(bl), bc_high=list(bh),
                            E=Field(NX, NY, NZ), H=Field(NX, NY, NZ))
        s.E.fromarray(E0)
        s.H.fromarray(H0)
        err = ""
        # wakis/routines.py:182-216, one_step replaced as described above
        E_abc_2, H_abc_2 = BCs.get_abc(s)
        E_abc_1, H_abc_1 = BCs.get_abc(s)
        try:
            for n in range(NT):
                s.E.fromarray(A_E * s.E.toarray() + dE[n])
                s.H.fromarray(A_H * s.H.toarray() + dH[n])
                BCs.update_abc(s, E_abc_2, H_abc_2)
                E_abc_2, H_abc_2 = E_abc_1, H_abc_1
                E_abc_1, H_abc_1 = BCs.get_abc(s)
        except KeyError as e:
            err = f"KeyError:{e.args[0]}:step{n}"
        out[name + "/E"] = s.E.toarray().copy()
        out[name + "/H"] = s.H.toarray().copy()
        out[name + "/error"] = np.array(err)
        print(name, err or "ok", flush=True)
    np.savez_compressed(os.path.join(HERE, "abc_aliasing.npz"), **out)


if __name__ == "__main__":
    main()
