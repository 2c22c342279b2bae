"""Extract the reference's golden vectors (numbers only) into JSON fixtures.

Reads the reference test files as text (no import, no execution) and writes
tests/golden/goldens.json with:
  test_001: WP[::50] (96 values), Z[::20] (50 complex), cumsums
            (tests/test_001_pec_cubic_cavity.py:18-67, 159-177)
  test_007: shapes/cumsums/dt (tests/test_007_mpi_lossy_cavity.py:119, 394-451)
Run: python tests/golden/make_goldens.py /root/reference
"""
import ast
import json
import os
import re
import sys


def _array_literal(src, name):
    m = re.search(name + r"\s*=\s*np\.array\(\s*(\[.*?\])\s*,?\s*\)", src, re.S)
    txt = m.group(1)
    txt = re.sub(r"(\d)\s*([+-])\s*([\d.]+(?:e[+-]?\d+)?)j", r"\1\2\3j", txt)
    return ast.literal_eval(txt)


def _read_ascii_stl(path):
    verts = []
    for line in open(path):
        t = line.split()
        if len(t) == 4 and t[0] == "vertex":
            verts.append([float(t[1]), float(t[2]), float(t[3])])
    return [verts[i:i + 3] for i in range(0, len(verts), 3)]


def main(ref):
    out = {}
    src = open(os.path.join(ref, "tests/test_001_pec_cubic_cavity.py")).read()
    wp = _array_literal(src, "WP")
    z = _array_literal(src, "Z")
    out["test_001"] = {
        "WP_every50": wp,
        "Z_every20_re": [complex(v).real for v in z],
        "Z_every20_im": [complex(v).imag for v in z],
        "WP_abs_cumsum": 1325.6968037037557,
        "Z_abs_cumsum": 372019.59123029554,
        "tol_wp": 5e-5,
        "tol_z": 5e-4,
        "cumsum_rel": 0.1,
    }
    out["test_002"] = {"TE101_f": 2.121e8, "rel": 0.05}
    out["test_005"] = {"max_reflection_fraction": 0.1}
    here = os.path.dirname(os.path.abspath(__file__))
    # geometry of test_001 (tests/stl/001_cubic_cavity.stl, a data file of the
    # reference's tests): its triangles and bounds, in metres
    tri = _read_ascii_stl(os.path.join(ref, "tests/stl/001_cubic_cavity.stl"))
    flat = [v for t in tri for v in t]
    geo = {
        "source": "tests/stl/001_cubic_cavity.stl",
        "bounds_lo": [min(v[a] for v in flat) for a in range(3)],
        "bounds_hi": [max(v[a] for v in flat) for a in range(3)],
        "triangles": tri,
    }
    with open(os.path.join(here, "001_cubic_cavity_box.json"), "w") as f:
        json.dump(geo, f)
    with open(os.path.join(here, "goldens.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(here, "goldens.json"), len(wp), len(z))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
