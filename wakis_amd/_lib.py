"""ctypes binding of the HIP FIT library (``libwakis_fit.so``).

The C ABI is declared in ``include/wakis_fit.h``.  There is no fallback: if
the library is missing or a call fails, a ``FitError`` is raised.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwakis_fit.so")

FIT_F64, FIT_F32 = 0, 1
FIELD_E, FIELD_H, FIELD_J = 0, 1, 2
J_ZERO, J_SIGMA_E, J_BUFFER = 0, 1, 2
BEAM_OFF, BEAM_VALUES, BEAM_GAUSS = 0, 1, 2
KERNEL_AUTO, KERNEL_SPLIT, KERNEL_FUSED, KERNEL_FUSED2 = 0, 1, 2, 3
NTILE = 12
NTILE2 = 6  # FIT_NTILE2: two-step tile shapes 64x8, 32x16, 32x8, 16x16, 60x12, 28x12


class FitError(RuntimeError):
    """A call into the HIP FIT library failed."""


class fit_desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("dtype", C.c_int32),
        ("device", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_U32 = C.c_uint32
_D = C.c_double
_PD = C.POINTER(C.c_double)
_PI32 = C.POINTER(C.c_int32)

# name -> argtypes (restype is int unless listed in _STR)
SIGNATURES = {
    "fit_last_error": [],
    "fit_version": [],
    "fit_create": [C.POINTER(fit_desc), C.POINTER(_P)],
    "fit_destroy": [_P],
    "fit_set_geometry": [_P, _PD, _I64],
    "fit_set_materials": [_P, _P, _I32, _PD],
    "fit_set_materials_wide": [_P, _PI32, _I32, _PD],
    "fit_set_bc": [_P, _U32, _U32],
    "fit_set_dt": [_P, _D],
    "fit_set_jmode": [_P, _I32],
    "fit_set_beam": [_P, _I32, _I32, _I32, _PD, _PD, _PD],
    "fit_field_read": [_P, _I32, _I32, _PI32, _P],
    "fit_field_write": [_P, _I32, _I32, _PI32, _P],
    "fit_field_zero": [_P, _I32],
    "fit_field_finite": [_P, _I32, _PI32],
    "fit_materialize_j": [_P],
    "fit_get_state": [_P, _PI32, _PI32],
    "fit_zero_ghosts": [_P],
    "fit_apply_abc": [_P, _U32],
    "fit_wake_probe": [_P, _I32, _I32, _I32, _I64],
    "fit_wake_history": [_P, _P, _I64],
    "fit_wake_history_load": [_P, _I32, _I32, _I32, _I64, _P],
    "fit_monitor_set": [_P, _I32],
    "fit_monitor_accumulate": [_P, _PD],
    "fit_monitor_phases": [_P, _PD, _I64],
    "fit_monitor_read": [_P, _I32, _I32, _PD],
    "fit_step": [_P, _I32, _I32],
    "fit_step_timed": [_P, _I32, _I32, _I32, _PD],
    "fit_copy_bench": [_I32, _I64, _I32, _PD],
    "fit_copy_bench_tiled": [_I32, _I32, _I32, _I32, _I32, _I32, _I32, _PD],
    "fit_set_step_index": [_P, _I64],
    "fit_get_step_index": [_P, C.POINTER(_I64)],
    "fit_set_kernel": [_P, _I32],
    "fit_kernel_in_use": [_P, _PI32],
    "fit_launch_counts": [_P, C.POINTER(_I64)],
    "fit_device_bytes": [_P, C.POINTER(_I64), C.POINTER(_I64)],
    "fit_set_tile": [_P, _I32],
    "fit_set_chunk_rsrc": [_P, _I32],
    "fit_get_tile": [_P, _PI32, _PI32, _PI32, _PI32],
    "fit_tune": [_P, _I32, _PD],
    "fit_set_tile2": [_P, _I32],
    "fit_get_tile2": [_P, _PI32, _PI32, _PI32],
    "fit_tune2": [_P, _I32, _PD],
    "fit_sync": [_P],
    "fit_wake_potential": [_P, _I32, _PI32, _I32, _I32, _I32, _I32, _I32, _PI32, _PD, _D,
                           _PD],
    "fit_impedance": [_P, _PD, _PD, _I64, _I64, _I32, _D, _PD, _PD, _PD, _PD, _PD, _PD],
    "fit_impedance_noalloc": [_P, _PD, _PD, _I64, _I64, _I32, _D, _PD, _PD, _PD, _PD, _PD,
                              _PD],
    "fit_batch_create": [C.POINTER(_P), _I32, _PI32, C.POINTER(_P)],
    "fit_batch_set_members": [_P, C.POINTER(_P), _I32, _PI32],
    "fit_batch_destroy": [_P],
    "fit_batch_step": [_P, _I64, _PD],
    "fit_batch_device_bytes": [_P, C.POINTER(_I64)],
    "fit_batch_set_two_step": [_P, _I32, _I32, _I32],
    "fit_batch_two_step": [_P, _PI32, _PI32, _PI32],
    "fit_batch_tune_two_step": [_P, _I32, _PI32, _PI32, _PD],
    "fit_batch_set_tile": [_P, _I32],
    "fit_batch_get_tile": [_P, _PI32, _PI32, _PI32],
    "fit_batch_tune": [_P, _I32, _PD],
    "fit_batch_set_graph": [_P, _I32, _I32],
}
_STR = {"fit_last_error", "fit_version"}

_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes library; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FitError(
            f"HIP FIT library not built: {path} is missing "
            "(run __graft_entry__.build() or `make -C wakis_amd/csrc`)"
        )
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name in _STR else C.c_int
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().fit_last_error().decode(errors="replace")
        raise FitError(f"{what} failed ({rc}): {msg}")


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_PD)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_PI32)


def vptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


def copy_bench(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """Average ms of a plain nbytes device copy (reads + writes nbytes each)."""
    out = np.zeros(1, dtype=np.float64)
    check(load().fit_copy_bench(int(device), int(nbytes), int(reps), dptr(out)), "fit_copy_bench")
    return float(out[0])


def copy_bench_tiled(nx: int, ny: int, nz: int, cfg: int, dtype=np.float32, reps: int = 10,
                     device: int = 0) -> float:
    """Average ms of copying six (nx, ny, nz) arrays into six others with the
    traversal of fused tile configuration cfg."""
    out = np.zeros(1, dtype=np.float64)
    dt = FIT_F64 if np.dtype(dtype) == np.float64 else FIT_F32
    check(load().fit_copy_bench_tiled(int(device), int(nx), int(ny), int(nz), dt, int(cfg),
                                      int(reps), dptr(out)), "fit_copy_bench_tiled")
    return float(out[0])


class Context:
    """Owns one device context (fields, materials, metrics) of one solve."""

    def __init__(self, nx: int, ny: int, nz: int, dtype=np.float64, device: int = 0):
        lib = load()
        self.lib = lib
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.N = self.nx * self.ny * self.nz
        self.np_dtype = np.dtype(dtype)
        if self.np_dtype == np.float64:
            code = FIT_F64
        elif self.np_dtype == np.float32:
            code = FIT_F32
        else:
            raise ValueError(f"unsupported dtype {dtype}: float64 or float32")
        d = fit_desc(self.nx, self.ny, self.nz, code, int(device))
        h = C.c_void_p()
        check(lib.fit_create(C.byref(d), C.byref(h)), "fit_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -------------------------------------------------------------
    def set_geometry(self, packed: np.ndarray):
        packed = np.ascontiguousarray(packed, dtype=np.float64)
        check(self.lib.fit_set_geometry(self.h, dptr(packed), packed.size), "fit_set_geometry")

    def set_materials(self, ids, table: np.ndarray):
        table = np.ascontiguousarray(table, dtype=np.float64)
        nmat = table.shape[0]
        if ids is None:
            p = None
        else:
            ids = np.ascontiguousarray(ids, dtype=np.uint8)
            assert ids.size == self.N
            p = C.c_void_p(ids.ctypes.data)
        check(self.lib.fit_set_materials(self.h, p, nmat, dptr(table)), "fit_set_materials")

    def set_materials_wide(self, ids, table: np.ndarray):
        """More than 256 material rows: int32 ids per cell (Fortran order)."""
        table = np.ascontiguousarray(table, dtype=np.float64)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        assert ids.size == self.N
        check(self.lib.fit_set_materials_wide(self.h, iptr(ids), table.shape[0], dptr(table)),
              "fit_set_materials_wide")

    def set_bc(self, pec: int, pmc: int):
        check(self.lib.fit_set_bc(self.h, pec, pmc), "fit_set_bc")

    def set_dt(self, dt: float):
        check(self.lib.fit_set_dt(self.h, float(dt)), "fit_set_dt")

    def set_jmode(self, mode: int):
        check(self.lib.fit_set_jmode(self.h, mode), "fit_set_jmode")

    def get_state(self):
        j, b = C.c_int32(), C.c_int32()
        check(self.lib.fit_get_state(self.h, C.byref(j), C.byref(b)), "fit_get_state")
        return j.value, b.value

    def set_beam_values(self, ixs: int, iys: int, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.nz
        check(self.lib.fit_set_beam(self.h, BEAM_VALUES, ixs, iys, dptr(v), None, None),
              "fit_set_beam")

    def set_beam_gauss(self, ixs: int, iys: int, params, zc: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float64)
        z = np.ascontiguousarray(zc, dtype=np.float64)
        assert p.size == 8 and z.size == self.nz
        check(self.lib.fit_set_beam(self.h, BEAM_GAUSS, ixs, iys, None, dptr(p), dptr(z)),
              "fit_set_beam")

    def beam_off(self):
        check(self.lib.fit_set_beam(self.h, BEAM_OFF, 0, 0, None, None, None), "fit_set_beam")

    # -- field transfer ----------------------------------------------------
    def read_box(self, which: int, comp: int, box) -> np.ndarray:
        b = np.ascontiguousarray(box, dtype=np.int32)
        shape = (b[1] - b[0], b[3] - b[2], b[5] - b[4])
        out = np.empty(int(np.prod(shape)), dtype=self.np_dtype)
        check(self.lib.fit_field_read(self.h, which, comp, iptr(b), vptr(out)), "fit_field_read")
        return out.reshape(shape, order="F")

    def write_box(self, which: int, comp: int, box, data: np.ndarray):
        b = np.ascontiguousarray(box, dtype=np.int32)
        shape = (int(b[1] - b[0]), int(b[3] - b[2]), int(b[5] - b[4]))
        d = np.asarray(data, dtype=self.np_dtype)
        d = np.broadcast_to(d, shape)
        flat = np.ascontiguousarray(np.reshape(d, -1, order="F"))
        check(self.lib.fit_field_write(self.h, which, comp, iptr(b), vptr(flat)), "fit_field_write")

    def read_field(self, which: int) -> np.ndarray:
        out = np.empty(3 * self.N, dtype=self.np_dtype)
        full = [0, self.nx, 0, self.ny, 0, self.nz]
        for c in range(3):
            out[c * self.N:(c + 1) * self.N] = np.reshape(
                self.read_box(which, c, full), -1, order="F")
        return out

    def write_field(self, which: int, array: np.ndarray):
        a = np.asarray(array, dtype=self.np_dtype)
        assert a.size == 3 * self.N
        full = [0, self.nx, 0, self.ny, 0, self.nz]
        for c in range(3):
            self.write_box(which, c, full, np.reshape(a[c * self.N:(c + 1) * self.N],
                                                       (self.nx, self.ny, self.nz), order="F"))

    def zero_field(self, which: int):
        check(self.lib.fit_field_zero(self.h, which), "fit_field_zero")

    def materialize_j(self):
        check(self.lib.fit_materialize_j(self.h), "fit_materialize_j")

    def field_finite(self, which: int = FIELD_E) -> bool:
        """Whether every value of a field is finite (fit_field_finite)."""
        out = C.c_int32()
        check(self.lib.fit_field_finite(self.h, int(which), C.byref(out)), "fit_field_finite")
        return bool(out.value)

    # -- stepping ----------------------------------------------------------
    def zero_ghosts(self):
        check(self.lib.fit_zero_ghosts(self.h), "fit_zero_ghosts")

    def apply_abc(self, mask: int):
        check(self.lib.fit_apply_abc(self.h, mask), "fit_apply_abc")

    def step(self, nsteps: int = 1, conductive: bool = False):
        check(self.lib.fit_step(self.h, int(nsteps), int(bool(conductive))), "fit_step")

    def step_timed(self, nsteps: int, conductive: bool, per_kernel: bool):
        """(total_ms, first_kernel_ms, second_kernel_ms) over nsteps steps."""
        out = np.zeros(3, dtype=np.float64)
        check(self.lib.fit_step_timed(self.h, int(nsteps), int(bool(conductive)),
                                      int(bool(per_kernel)), dptr(out)), "fit_step_timed")
        return float(out[0]), float(out[1]), float(out[2])

    def set_step_index(self, n: int):
        check(self.lib.fit_set_step_index(self.h, int(n)), "fit_set_step_index")

    def get_step_index(self) -> int:
        v = C.c_int64()
        check(self.lib.fit_get_step_index(self.h, C.byref(v)), "fit_get_step_index")
        return v.value

    def set_kernel(self, k: int):
        check(self.lib.fit_set_kernel(self.h, k), "fit_set_kernel")

    def kernel_in_use(self) -> int:
        v = C.c_int32()
        check(self.lib.fit_kernel_in_use(self.h, C.byref(v)), "fit_kernel_in_use")
        return v.value

    def launch_counts(self):
        """(one-step launches, two-step launches) since the context was created."""
        v = (C.c_int64 * 2)()
        check(self.lib.fit_launch_counts(self.h, v), "fit_launch_counts")
        return int(v[0]), int(v[1])

    def device_bytes(self):
        """(bytes held now, most bytes held at once) on the device."""
        now, peak = C.c_int64(0), C.c_int64(0)
        check(self.lib.fit_device_bytes(self.h, C.byref(now), C.byref(peak)), "fit_device_bytes")
        return int(now.value), int(peak.value)

    def set_tile(self, cfg: int):
        check(self.lib.fit_set_tile(self.h, int(cfg)), "fit_set_tile")

    def set_chunk_rsrc(self, on: bool):
        check(self.lib.fit_set_chunk_rsrc(self.h, int(bool(on))), "fit_set_chunk_rsrc")

    def get_tile(self):
        v = [C.c_int32() for _ in range(4)]
        check(self.lib.fit_get_tile(self.h, *[C.byref(x) for x in v]), "fit_get_tile")
        return tuple(x.value for x in v)

    def tune(self, reps: int = 5):
        """Time every fused tile configuration, select the fastest; returns
        the per-configuration average ms."""
        out = np.zeros(NTILE, dtype=np.float64)
        check(self.lib.fit_tune(self.h, int(reps), dptr(out)), "fit_tune")
        return out

    def set_tile2(self, cfg: int):
        """Tile shape of the two-step sweep (0: 64x8, 1: 32x16, 2: 32x8,
        3: 16x16, 4: 60x12, 5: 28x12, -1: the modelled default; FP32 builds
        0-4, FP64 2, 3 and 5)."""
        check(self.lib.fit_set_tile2(self.h, int(cfg)), "fit_set_tile2")

    def get_tile2(self):
        """(cfg, tx, ty) of the next two-step sweep; cfg -1 when none applies."""
        c, x, y = C.c_int32(), C.c_int32(), C.c_int32()
        check(self.lib.fit_get_tile2(self.h, C.byref(c), C.byref(x), C.byref(y)),
              "fit_get_tile2")
        return c.value, x.value, y.value

    def tune2(self, reps: int = 3):
        """Time every two-step tile shape that applies, select the fastest;
        returns the per-shape average ms (-1 where a shape does not apply)."""
        out = np.zeros(NTILE2, dtype=np.float64)
        check(self.lib.fit_tune2(self.h, int(reps), dptr(out)), "fit_tune2")
        return out

    def sync(self):
        check(self.lib.fit_sync(self.h), "fit_sync")

    # -- wake ----------------------------------------------------------------
    def wake_probe(self, ixt: int, iyt: int, h: int, nt_cap: int):
        check(self.lib.fit_wake_probe(self.h, ixt, iyt, h, int(nt_cap)), "fit_wake_probe")
        self.probe = (ixt, iyt, h, int(nt_cap))

    # -- field monitor ---------------------------------------------------------
    def monitor_set(self, nf: int):
        check(self.lib.fit_monitor_set(self.h, int(nf)), "fit_monitor_set")

    def monitor_accumulate(self, phase: np.ndarray):
        p = np.ascontiguousarray(np.asarray(phase, dtype=np.complex128)).view(np.float64)
        check(self.lib.fit_monitor_accumulate(self.h, dptr(p)), "fit_monitor_accumulate")

    def monitor_phases(self, phases: np.ndarray):
        """phases: (nsteps, nf) complex128, one row per step of the next step() call."""
        p = np.ascontiguousarray(np.asarray(phases, dtype=np.complex128))
        n = p.shape[0]
        check(self.lib.fit_monitor_phases(self.h, dptr(p.view(np.float64).reshape(-1)), n),
              "fit_monitor_phases")

    def monitor_read(self, f: int, comp: int) -> np.ndarray:
        """Accumulator of frequency f, component comp as an (nx, ny, nz) complex array."""
        out = np.empty(2 * self.N, dtype=np.float64)
        check(self.lib.fit_monitor_read(self.h, int(f), int(comp), dptr(out)), "fit_monitor_read")
        return out.view(np.complex128).reshape((self.nx, self.ny, self.nz), order="F")

    def wake_history(self) -> np.ndarray:
        ixt, iyt, h, nt = self.probe
        nw = (2 * h + 1) ** 2
        out = np.empty(nw * self.nz * nt, dtype=self.np_dtype)
        check(self.lib.fit_wake_history(self.h, vptr(out), out.size), "fit_wake_history")
        return out.reshape(nw, self.nz, nt)

    def wake_history_load(self, ixt: int, iyt: int, h: int, hist: np.ndarray):
        """Load a recorded history hist[(2h+1)^2, nz, nt] (fit_wake_history_load)."""
        nw = (2 * h + 1) ** 2
        hist = np.ascontiguousarray(hist, dtype=self.np_dtype)
        if hist.ndim != 3 or hist.shape[0] != nw or hist.shape[1] != self.nz:
            raise ValueError(f"history shape {hist.shape}: expected ({nw}, {self.nz}, nt)")
        nt = hist.shape[2]
        check(self.lib.fit_wake_history_load(self.h, ixt, iyt, h, nt, vptr(hist)),
              "fit_wake_history_load")
        self.probe = (ixt, iyt, h, int(nt))

    def wake_potential(self, window_ids, carry: bool, cols: np.ndarray, k0: int,
                       dz: np.ndarray, qnorm: float, zrev: bool = False) -> np.ndarray:
        """fit_wake_potential; zrev walks the history rows from the top of the
        window (dz[k] is passed as given: the counter-moving caller negates it)."""
        w = np.ascontiguousarray(window_ids, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        nzw, ns = cols.shape
        dz = np.ascontiguousarray(dz[:nzw], dtype=np.float64)
        out = np.empty(len(w) * ns, dtype=np.float64)
        check(self.lib.fit_wake_potential(self.h, len(w), iptr(w), int(bool(carry)), ns, nzw,
                                          int(k0), int(bool(zrev)), iptr(cols), dptr(dz),
                                          float(qnorm), dptr(out)), "fit_wake_potential")
        return out.reshape(len(w), ns)

    def impedance(self, x: np.ndarray, lam: np.ndarray, N: int, K: int, ds: float,
                  noalloc: bool = False):
        """Z, X*ds, Lam*ds of bins 0 .. K-1 (fit_impedance; noalloc=True uses
        fit_impedance_noalloc, which stages in the library's fixed scratch)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        assert x.size == lam.size
        bufs = [np.empty(K, dtype=np.float64) for _ in range(6)]
        fn, name = ((self.lib.fit_impedance_noalloc, "fit_impedance_noalloc") if noalloc
                    else (self.lib.fit_impedance, "fit_impedance"))
        check(fn(self.h, dptr(x), dptr(lam), x.size, int(N), int(K), float(ds),
                 *[dptr(b) for b in bufs]), name)
        z = bufs[0] + 1j * bufs[1]
        X = bufs[2] + 1j * bufs[3]
        L = bufs[4] + 1j * bufs[5]
        return z, X, L


class Batch:
    """Many contexts stepped by one launch per step (see
    ``fit_batch_*`` in include/wakis_fit.h)."""

    def __init__(self, contexts, conductive=None):
        self.lib = load()
        h = C.c_void_p()
        arr, cond = self._arrays(contexts, conductive)
        check(self.lib.fit_batch_create(arr, len(contexts), iptr(cond), C.byref(h)),
              "fit_batch_create")
        self.h = h
        self.contexts = list(contexts)

    @staticmethod
    def _arrays(contexts, conductive):
        arr = (_P * max(len(contexts), 1))(*[c.h for c in contexts])
        if conductive is None:
            conductive = [0] * len(contexts)
        cond = np.ascontiguousarray(np.asarray(conductive, dtype=np.int32).reshape(-1))
        if cond.size < len(contexts):
            raise ValueError("one conductive flag per member")
        if cond.size == 0:
            cond = np.zeros(1, dtype=np.int32)
        return arr, cond

    def set_members(self, contexts, conductive=None):
        arr, cond = self._arrays(contexts, conductive)
        check(self.lib.fit_batch_set_members(self.h, arr, len(contexts), iptr(cond)),
              "fit_batch_set_members")
        self.contexts = list(contexts)

    def step(self, nsteps: int) -> float:
        """Step every member nsteps; returns the elapsed ms (HIP events)."""
        out = np.zeros(1, dtype=np.float64)
        check(self.lib.fit_batch_step(self.h, int(nsteps), dptr(out)), "fit_batch_step")
        return float(out[0])

    def set_two_step(self, on: bool, cfg: int = -1, kc: int = 0):
        check(self.lib.fit_batch_set_two_step(self.h, int(bool(on)), int(cfg), int(kc)),
              "fit_batch_set_two_step")

    def two_step(self):
        """(pairs by the two-step sweep on the next step call, its tile
        configuration or -1, planes per chunk)."""
        on, cfg, kc = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        check(self.lib.fit_batch_two_step(self.h, C.byref(on), C.byref(cfg), C.byref(kc)),
              "fit_batch_two_step")
        return bool(on.value), int(cfg.value), int(kc.value)

    def tune_two_step(self, reps: int = 3):
        """(tile configuration, planes per chunk, {(cfg, kc): ms per launch})."""
        cfg, kc = C.c_int32(0), C.c_int32(0)
        ms = np.zeros(9, dtype=np.float64)  # FIT_BATCH2_NTUNE
        check(self.lib.fit_batch_tune_two_step(self.h, int(reps), C.byref(cfg), C.byref(kc),
                                               dptr(ms)), "fit_batch_tune_two_step")
        times = {(i // 3, (48, 64, 96)[i % 3]): float(v) for i, v in enumerate(ms) if v > 0}
        return int(cfg.value), int(kc.value), times

    def device_bytes(self) -> int:
        """Device bytes of the batch's own tables."""
        v = C.c_int64(0)
        check(self.lib.fit_batch_device_bytes(self.h, C.byref(v)), "fit_batch_device_bytes")
        return int(v.value)

    def set_tile(self, cfg: int):
        check(self.lib.fit_batch_set_tile(self.h, int(cfg)), "fit_batch_set_tile")

    def get_tile(self):
        v = [C.c_int32() for _ in range(3)]
        check(self.lib.fit_batch_get_tile(self.h, *[C.byref(x) for x in v]), "fit_batch_get_tile")
        return tuple(x.value for x in v)

    def tune(self, reps: int = 5):
        out = np.zeros(NTILE, dtype=np.float64)
        check(self.lib.fit_batch_tune(self.h, int(reps), dptr(out)), "fit_batch_tune")
        return out

    def set_graph(self, on: bool, group: int = 16):
        check(self.lib.fit_batch_set_graph(self.h, int(bool(on)), int(group)), "fit_batch_set_graph")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fit_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
