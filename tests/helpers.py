"""Shared builders: the same configuration as an oracle solver (CPU, the
checker) and as a wakis_amd solver (HIP), with identical initial fields."""

import numpy as np

from oracle.fit_oracle import OBeam, OFieldMonitor, OGrid, OPlaneWave, OSolver, OWake  # noqa: F401


def axes(Nx, Ny, Nz, nonuniform=False, seed=0, ext=(0.1, 0.08, 0.2)):
    if not nonuniform:
        return (np.linspace(-ext[0] / 2, ext[0] / 2, Nx + 1),
                np.linspace(-ext[1] / 2, ext[1] / 2, Ny + 1),
                np.linspace(-ext[2] / 2, ext[2] / 2, Nz + 1))
    rng = np.random.default_rng(seed)
    out = []
    for n, e in zip((Nx, Ny, Nz), ext):
        d = rng.uniform(0.6, 1.4, n)
        a = np.concatenate([[0.0], np.cumsum(d)])
        out.append(a / a[-1] * e - e / 2)
    return tuple(out)


def write_ascii_stl(path, tri):
    """Triangles (n, 3, 3) as an ASCII STL file."""
    with open(path, "w") as f:
        f.write("solid t\n")
        for a, b, c in np.asarray(tri, dtype=np.float64):
            f.write(" facet normal 0 0 0\n  outer loop\n")
            for v in (a, b, c):
                f.write("   vertex %r %r %r\n" % (float(v[0]), float(v[1]), float(v[2])))
            f.write("  endloop\n endfacet\n")
        f.write("endsolid t\n")


def block_mask(shape, lo, hi):
    m = np.zeros(shape, bool)
    m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = True
    return m


def make_pair(Nx, Ny, Nz, bc_low, bc_high, dtype=np.float64, bg=(1.0, 1.0), solids=(),
              n_pml=3, cfln=0.5, nonuniform=False, seed=0, init=True, wake_kw=None,
              kernel="auto", tile=None):
    import wakis_amd as wa

    x, y, z = axes(Nx, Ny, Nz, nonuniform, seed)
    owake = OWake(**wake_kw) if wake_kw else None
    pwake = wa.WakeSolver(save=False, **wake_kw) if wake_kw else None
    og = OGrid(x=x, y=y, z=z)
    osv = OSolver(og, owake, cfln=cfln, bc_low=bc_low, bc_high=bc_high,
                  solids=list(solids), bg=bg, dtype=dtype, n_pml=n_pml)
    g = wa.GridFIT3D(x=x, y=y, z=z, solids=list(solids))
    psv = wa.SolverFIT3D(g, pwake, cfln=cfln, bc_low=bc_low, bc_high=bc_high,
                         use_stl=bool(solids), bg=bg, dtype=dtype, n_pml=n_pml, kernel=kernel,
                         tile=tile)
    if init:
        rng = np.random.default_rng(seed + 100)
        for name in ("E", "H", "J"):
            a = rng.standard_normal(3 * osv.N).astype(dtype)
            if name == "J":
                a *= 1e-3
            getattr(osv, name).fromarray(a)
            getattr(psv, name).fromarray(a)
    return osv, psv


def rel(a, b):
    """Relative L2 distance ||a - b|| / ||b|| (complex arrays compared as
    complex, real ones in float64)."""
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def compare_fields(osv, psv):
    return {n: rel(getattr(psv, n).toarray(), getattr(osv, n).toarray()) for n in "EHJ"}


def region_mask(shape, region):
    m = np.zeros(shape, bool)
    m[region] = True
    return m


def stratified_pair(n=12, nlayers=4, n_pml=2, dtype=np.float64, device=0, seed=5, kernel="auto"):
    """The stratified-box workload (index-slice solids, a diagonal-tensor
    block) as a wakis_amd solver and as an oracle solver; the oracle gets the
    slice boxes as masks and the tensor block as per-component edits of its
    ieps Field followed by update_tensors (the reference's route to an
    anisotropic material)."""
    import wakis_amd as wa
    from wakis_amd.workloads import stratified_box

    from scipy.constants import epsilon_0 as eps_0
    wl = stratified_box(n=n, seed=seed, half=0.256 * n / 1024, n_pml=n_pml, nlayers=nlayers)
    g = wl["grid"]
    sv = wl["solver"]
    shape = (n, n, n)
    iso, tensors = [], []
    for region, mat in g["solids"]:
        (tensors if isinstance(mat, dict) else iso).append((region, mat))
    og = OGrid(x=g["x"], y=g["y"], z=g["z"])
    osv = OSolver(og, cfln=sv["cfln"], bc_low=sv["bc_low"], bc_high=sv["bc_high"],
                  solids=[(region_mask(shape, r), m) for r, m in iso], bg=sv["bg"], dtype=dtype,
                  n_pml=sv["n_pml"], use_sibc=sv["use_sibc"])
    for region, mat in tensors:
        for d, e in zip("xyz", mat["eps"]):
            osv.ieps[region + (d,)] = np.float64(1.0) / (e * eps_0)
    if tensors:
        osv.update_tensors("ieps")
        osv.step_0 = True
    psv = wa.SolverFIT3D(wa.GridFIT3D(x=g["x"], y=g["y"], z=g["z"], solids=g["solids"]),
                         dtype=dtype, device=device, kernel=kernel, **sv)
    return osv, psv, wl


def oracle_from_workload(w, dtype=np.float64):
    """The oracle solver (with its wake) of a wakis_amd.workloads dict whose
    solids are (mask, material) pairs; the reference's geomspace PML grading
    (a workload with pml_profile "cubic" gets the cubic sigma through the
    sigma Field and update_tensors, the reference's route to per-cell
    conductivities)."""
    g, sv = w["grid"], w["solver"]
    og = OGrid(x=g["x"], y=g["y"], z=g["z"])
    osv = OSolver(og, OWake(**w["wake"]), cfln=sv.get("cfln", 0.5), bc_low=sv["bc_low"],
                  bc_high=sv["bc_high"], solids=list(g["solids"]), bg=sv["bg"], dtype=dtype,
                  n_pml=sv.get("n_pml", 10), use_sibc=sv.get("use_sibc", True))
    if sv.get("pml_profile", "geomspace") == "cubic":
        n, nz = sv["n_pml"], osv.Nz
        assert [b.lower() for b in sv["bc_low"]][:2] != ["pml", "pml"]  # z faces only
        for p in range(n):
            for d in "xyz":
                osv.sigma[:, :, p, d] = (float(n - p) / n) ** 3
                osv.sigma[:, :, nz - n + p, d] = (float(p + 1) / n) ** 3
        osv.update_tensors("sigma")
        osv.step_0 = True
    return osv


def z_errors(z, z_ref, rel_floor=1e-2):
    """(max |Z - Z_ref| / max |Z_ref|, max over bins with |Z_ref| >= rel_floor
    max |Z_ref| of |Z - Z_ref| / |Z_ref|) over the kept bins (all below the
    bunch cutoff fmax)."""
    z, z_ref = np.asarray(z), np.asarray(z_ref)
    zmax = np.max(np.abs(z_ref))
    d = np.abs(z - z_ref)
    big = np.abs(z_ref) >= rel_floor * zmax
    return float(np.max(d) / zmax), float(np.max(d[big] / np.abs(z_ref[big])))
