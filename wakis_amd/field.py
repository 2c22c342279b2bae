"""``Field``: the reference's flat vector-field container, and its device-backed
variant.

Layout (wakis/field.py:19-91): one flat array of 3N values, the x block, then
the y block, then the z block; each block in Fortran order,
n = i + Nx*j + Nx*Ny*k.  Tuple indexing ``F[i, j, k, 'x']`` and the arithmetic
follow wakis/field.py:223-460.

``DeviceField`` keeps the same interface while the values live in HBM inside
a ``wakis_amd._lib.Context``: tuple reads/writes of unit-stride boxes become
box copies, everything else goes through a full-component copy.  Reads on the
device field return copies (as the reference's GPU path does,
wakis/field.py:243-249), never views.
"""

from __future__ import annotations

import copy

import numpy as np

_COMP = {0: 0, 1: 1, 2: 2, "x": 0, "y": 1, "z": 2}


def comp_index(key) -> int:
    try:
        return _COMP[key]
    except (KeyError, TypeError):
        raise IndexError("Component id not valid") from None


class Field:
    """Host field (wakis/field.py:19-570 minus plotting)."""

    def __init__(self, Nx, Ny, Nz, dtype=float, use_ones=False, use_gpu=False):
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.N = Nx * Ny * Nz
        self.dtype = dtype
        self.on_gpu = False
        self.xp = np
        if use_ones:
            self.array = np.ones(self.N * 3, dtype=dtype, order="F")
        else:
            self.array = np.zeros(self.N * 3, dtype=dtype, order="F")

    # component blocks -------------------------------------------------------
    def _block(self, c):
        return self.array[c * self.N:(c + 1) * self.N]

    def _set_block(self, c, value):
        if len(value.shape) > 1:
            self.from_matrix(value, c)
        else:
            self.array[c * self.N:(c + 1) * self.N] = value

    field_x = property(lambda s: s._block(0), lambda s, v: s._set_block(0, v))
    field_y = property(lambda s: s._block(1), lambda s, v: s._set_block(1, v))
    field_z = property(lambda s: s._block(2), lambda s, v: s._set_block(2, v))

    def toarray(self):
        return self.array

    def fromarray(self, array):
        self.array[:] = array

    def to_matrix(self, key):
        c = comp_index(key)
        return np.reshape(self.array[c * self.N:(c + 1) * self.N],
                          (self.Nx, self.Ny, self.Nz), order="F")

    def from_matrix(self, mat, key):
        c = comp_index(key)
        self.array[c * self.N:(c + 1) * self.N] = np.reshape(mat, self.N, order="F")

    def to_gpu(self):
        print("*** use SolverFIT3D fields for device storage")

    def from_gpu(self):
        print("*** GPU is not enabled")

    def __getitem__(self, key):
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            if type(key[3]) is str and key[3].lower() == "abs":
                return self.get_abs()[key[0], key[1], key[2]]
            return self.to_matrix(key[3])[key[0], key[1], key[2]]
        if type(key) is int:
            if key <= self.N:
                return self.array[key]
            raise IndexError("Lexico-graphic index cannot be higher than product of dimensions")
        if type(key) is slice:
            return self.array[key]
        raise ValueError("key must be a 3-tuple or an integer")

    def __setitem__(self, key, value):
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            field = self.to_matrix(key[3])
            field[key[0], key[1], key[2]] = value
            self.from_matrix(field, key[3])
        elif type(key) is int:
            if key <= self.N:
                self.array[key] = value
            else:
                raise IndexError("Lexico-graphic index cannot be higher than product of dimensions")
        elif type(key) is slice:
            self.array[key] = value
        else:
            raise IndexError("key must be a 3-tuple or an integer")

    # arithmetic (wakis/field.py:350-460) ----------------------------------
    def __mul__(self, other, dtype=None):
        dtype = self.dtype if dtype is None else dtype
        out = Field(self.Nx, self.Ny, self.Nz, dtype=dtype)
        a = self.toarray()
        if type(other) is float or type(other) is int:
            out.array = a * other
        elif len(other.shape) > 1:
            for d in "xyz":
                out.from_matrix(self.to_matrix(d) * other, d)
        else:
            out.array = a * other
        return out

    def __div__(self, other, dtype=None):
        dtype = self.dtype if dtype is None else dtype
        out = Field(self.Nx, self.Ny, self.Nz, dtype=dtype)
        a = self.toarray()
        if type(other) is float or type(other) is int:
            out.array = a / other
        elif len(other.shape) > 1:
            for d in "xyz":
                out.from_matrix(self.to_matrix(d) / other, d)
        else:
            out.array = a / other
        return out

    __truediv__ = __div__

    def __add__(self, other, dtype=None):
        dtype = self.dtype if dtype is None else dtype
        out = Field(self.Nx, self.Ny, self.Nz, dtype=dtype)
        if isinstance(other, Field):
            out.field_x = self.field_x + other.field_x
            out.field_y = self.field_y + other.field_y
            out.field_z = self.field_z + other.field_z
        elif len(other.shape) > 1:
            for d in "xyz":
                out.from_matrix(self.to_matrix(d) + other, d)
        else:
            out.array = self.toarray() + other
        return out

    def __repr__(self):
        return ("x:\n" + self.field_x.__repr__() + "\ny:\n" + self.field_y.__repr__()
                + "\nz:\n" + self.field_z.__repr__())

    def __str__(self):
        return ("x:\n" + self.field_x.__str__() + "\ny:\n" + self.field_y.__str__()
                + "\nz:\n" + self.field_z.__str__())

    def copy(self):
        out = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        out.array = np.array(self.toarray(), copy=True)
        return out

    def compute_ijk(self, n):
        if n > self.N:
            raise IndexError("Lexico-graphic index cannot be higher than product of dimensions")
        k = n // (self.Nx * self.Ny)
        i = (n - k * self.Nx * self.Ny) % self.Nx
        j = (n - k * self.Nx * self.Ny) // self.Nx
        return i, j, k

    def get_abs(self, as_matrix=True):
        if as_matrix:
            return np.sqrt(self.to_matrix("x") ** 2 + self.to_matrix("y") ** 2
                           + self.to_matrix("z") ** 2)
        return np.sqrt(self.field_x ** 2 + self.field_y ** 2 + self.field_z ** 2)


def _as_box(key3, dims):
    """(box[6], squeeze_axes) if every index is an int or a unit-step slice."""
    box, squeeze = [], []
    for ax, (idx, n) in enumerate(zip(key3, dims)):
        if isinstance(idx, (int, np.integer)) and not isinstance(idx, bool):
            i = int(idx)
            if i < 0:
                i += n
            if not 0 <= i < n:
                raise IndexError(f"index {idx} out of range for axis {ax} of size {n}")
            box += [i, i + 1]
            squeeze.append(ax)
        elif isinstance(idx, slice):
            start, stop, step = idx.indices(n)
            if step != 1:
                return None, None
            box += [start, max(start, stop)]
        else:
            return None, None
    return box, squeeze


class DeviceField(Field):
    """Field whose storage is a field of a device context (E, H or J)."""

    def __init__(self, ctx, which, owner=None):
        self.ctx = ctx
        self.which = which
        self.owner = owner
        self.Nx, self.Ny, self.Nz = ctx.nx, ctx.ny, ctx.nz
        self.N = ctx.N
        self.dtype = ctx.np_dtype.type
        self.on_gpu = True
        self.xp = np

    # the J field is materialised on read (its mode may be implicit)
    def _before_read(self):
        if self.owner is not None:
            self.owner._before_field_read(self.which)

    def _before_write(self, key=None, value=None):
        if self.owner is not None:
            return self.owner._before_field_write(self.which, key, value)
        return False

    @property
    def array(self):
        self._before_read()
        return self.ctx.read_field(self.which)

    @array.setter
    def array(self, value):
        self._before_write()
        self.ctx.write_field(self.which, np.asarray(value).reshape(-1))

    def toarray(self):
        return self.array

    def fromarray(self, array):
        self.array = array

    def _block(self, c):
        self._before_read()
        full = [0, self.Nx, 0, self.Ny, 0, self.Nz]
        return np.reshape(self.ctx.read_box(self.which, c, full), -1, order="F")

    def _set_block(self, c, value):
        self._before_write()
        full = [0, self.Nx, 0, self.Ny, 0, self.Nz]
        v = np.asarray(value)
        if len(v.shape) <= 1:
            v = np.reshape(v, (self.Nx, self.Ny, self.Nz), order="F")
        self.ctx.write_box(self.which, c, full, v)

    field_x = property(lambda s: s._block(0), lambda s, v: s._set_block(0, v))
    field_y = property(lambda s: s._block(1), lambda s, v: s._set_block(1, v))
    field_z = property(lambda s: s._block(2), lambda s, v: s._set_block(2, v))

    def to_matrix(self, key):
        c = comp_index(key)
        self._before_read()
        return self.ctx.read_box(self.which, c, [0, self.Nx, 0, self.Ny, 0, self.Nz])

    def from_matrix(self, mat, key):
        c = comp_index(key)
        self._before_write()
        self.ctx.write_box(self.which, c, [0, self.Nx, 0, self.Ny, 0, self.Nz],
                           np.reshape(np.asarray(mat), (self.Nx, self.Ny, self.Nz), order="F"))

    def __getitem__(self, key):
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            if type(key[3]) is str and key[3].lower() == "abs":
                return self.get_abs()[key[0], key[1], key[2]]
            c = comp_index(key[3])
            box, squeeze = _as_box(key[:3], (self.Nx, self.Ny, self.Nz))
            self._before_read()
            if box is None:
                return self.to_matrix(c)[key[0], key[1], key[2]]
            if box[0] == box[1] or box[2] == box[3] or box[4] == box[5]:
                shape = [box[1] - box[0], box[3] - box[2], box[5] - box[4]]
                return np.zeros([s for a, s in enumerate(shape) if a not in squeeze],
                                dtype=self.dtype)
            out = self.ctx.read_box(self.which, c, box)
            if squeeze:
                out = np.squeeze(out, axis=tuple(squeeze))
                if out.ndim == 0:
                    return out[()]
            return out
        if type(key) is int or type(key) is slice:
            return self.array[key]
        raise ValueError("key must be a 3-tuple or an integer")

    def __setitem__(self, key, value):
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            c = comp_index(key[3])
            if self._before_write(key, value):
                return  # consumed by the owner (beam line of J)
            box, squeeze = _as_box(key[:3], (self.Nx, self.Ny, self.Nz))
            if box is None:
                m = self.ctx.read_box(self.which, c, [0, self.Nx, 0, self.Ny, 0, self.Nz])
                m[key[0], key[1], key[2]] = value
                self.ctx.write_box(self.which, c, [0, self.Nx, 0, self.Ny, 0, self.Nz], m)
                return
            if box[0] == box[1] or box[2] == box[3] or box[4] == box[5]:
                return
            shape = [box[1] - box[0], box[3] - box[2], box[5] - box[4]]
            v = np.asarray(value, dtype=self.ctx.np_dtype)
            kept = [s for a, s in enumerate(shape) if a not in squeeze]
            v = np.broadcast_to(v, kept).reshape(shape)
            self.ctx.write_box(self.which, c, box, v)
        elif type(key) is int or type(key) is slice:
            a = self.array
            a[key] = value
            self.array = a
        else:
            raise IndexError("key must be a 3-tuple or an integer")

    def copy(self):
        out = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        out.array = self.array
        return out

    def __deepcopy__(self, memo):
        return self.copy()
