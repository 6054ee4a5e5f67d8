"""Device fields on one GPU: the Python mirror of krox/mesh's lattice layer.

Mirrors ``Grid`` (src/lattice/grid.h:13-58), ``Lattice<T>`` and
``LatticeStack<T>`` (src/lattice/lattice.h:23-192).  Storage is owned by
libmesh_hip.so (``mesh_malloc``) in the structure-of-arrays layout described
in include/mesh_hip.h; host copies use the reference's AoS order as numpy
complex arrays of shape (nf, V, 3, 3).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import F32, F64, MeshError, mesh_grid


def _prec(p):
    if p in (F64, "f64", "double", 2, np.float64, np.complex128):
        return F64
    if p in (F32, "f32", "float", 1, "mixed", np.float32, np.complex64):
        return F32
    raise ValueError(f"invalid precision level '{p}'")


class Context:
    """One GPU, one stream, one reduction scratch (mesh_ctx)."""

    def __init__(self, device=0):
        self._L = _lib.lib()
        h = C.c_void_p()
        rc = self._L.mesh_ctx_create(C.byref(h), device)
        if rc != 0:
            raise MeshError(rc, f"cannot create a context on HIP device {device}")
        self.handle = h
        self.device = device

    def check(self, rc):
        if rc != 0:
            msg = self._L.mesh_last_error(self.handle).decode()
            raise MeshError(rc, msg)
        return rc

    def call(self, name, *args):
        return self.check(getattr(self._L, name)(self.handle, *args))

    def synchronize(self):
        self.call("mesh_synchronize")

    def set_option(self, name, value):
        self.call("mesh_set_option", name.encode(), int(value))

    def profile(self, on=True):
        self.call("mesh_profile_enable", int(bool(on)))

    def profile_reset(self):
        self.call("mesh_profile_reset")

    def profile_read(self, kernel):
        """(total milliseconds, launch count) of a profiled kernel."""
        ms, n = C.c_double(), C.c_longlong()
        self.call("mesh_profile_read", kernel.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if self.handle:
            self._L.mesh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class Grid:
    """Lattice geometry, row-major with dimension 0 slowest (grid.h:39-49)."""

    def __init__(self, shape):
        shape = tuple(int(s) for s in shape)
        if not 2 <= len(shape) <= 4:
            raise ValueError("grid must have 2 to 4 dimensions")
        self.shape = shape
        self.c = mesh_grid()
        self.c.nd = len(shape)
        for i, s in enumerate(shape):
            self.c.dims[i] = s
        self.ref = C.byref(self.c)

    def ndim(self):
        return len(self.shape)

    def size(self):
        return int(np.prod(self.shape))

    def flat_index(self, coord):
        r = 0
        for c, s in zip(coord, self.shape):
            if not 0 <= c < s:
                raise IndexError("coordinate out of range")
            r = r * s + c
        return r

    @property
    def stride(self):
        """Sites per direction block of a device field (mesh_site_stride)."""
        return int(_lib.lib().mesh_site_stride(self.ref))

    def __eq__(self, other):
        return isinstance(other, Grid) and self.shape == other.shape

    def __repr__(self):
        return f"Grid({list(self.shape)})"


class Field:
    """nf SU(3) matrices per site on the device (nf = Nd: GaugeField/LatticeStack,
    nf = 1: Lattice<SU3>).  ``owner`` is set for views of one direction."""

    def __init__(self, grid, nf=None, prec=F64, ctx=None, _view=None):
        self.grid = grid
        self.nf = grid.ndim() if nf is None else nf
        self.prec = _prec(prec)
        self.ctx = ctx or default_context()
        self.itemsize = 8 if self.prec == F64 else 4
        self.bytes = self.nf * 9 * grid.stride * 2 * self.itemsize
        if _view is not None:
            self.owner, self.ptr = _view
            return
        self.owner = None
        p = C.c_void_p()
        self.ctx.call("mesh_malloc", C.byref(p), self.bytes)
        self.ptr = p

    def __del__(self):
        try:
            if self.owner is None and self.ptr and self.ctx.handle:
                self.ctx._L.mesh_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass

    # LatticeStack::operator[] (lattice.h:190-191): view of direction mu
    def __getitem__(self, mu):
        if not 0 <= mu < self.nf:
            raise IndexError(mu)
        off = mu * 9 * self.grid.stride * 2 * self.itemsize
        return Field(self.grid, 1, self.prec, self.ctx, _view=(self, C.c_void_p(self.ptr.value + off)))

    def __len__(self):
        return self.nf

    @property
    def cdtype(self):
        return np.complex128 if self.prec == F64 else np.complex64

    def upload(self, a):
        """copy_from_host (lattice.h:102-107); a: complex (nf, V, 3, 3)."""
        a = np.ascontiguousarray(a, dtype=self.cdtype)
        if a.size != self.nf * self.grid.size() * 9:
            raise ValueError("host array does not match the field shape")
        self.ctx.call("mesh_su3_upload", self.grid.ref, self.nf, self.prec, a.ctypes.data_as(C.c_void_p), self.ptr)
        return self

    def download(self):
        """copy_to_host (lattice.h:110-115) -> complex (nf, V, 3, 3)."""
        out = np.empty((self.nf, self.grid.size(), 3, 3), self.cdtype)
        self.ctx.call("mesh_su3_download", self.grid.ref, self.nf, self.prec, self.ptr,
                      out.ctypes.data_as(C.c_void_p))
        return out

    def copy_from(self, other):
        if other.bytes != self.bytes:
            raise ValueError("field size mismatch")
        self.ctx.call("mesh_memcpy", self.ptr, other.ptr, self.bytes)
        return self

    def clone(self):
        return Field(self.grid, self.nf, self.prec, self.ctx).copy_from(self)

    def convert(self, prec):
        out = Field(self.grid, self.nf, prec, self.ctx)
        self.ctx.call("mesh_su3_convert", self.grid.ref, self.nf, out.prec, out.ptr, self.prec, self.ptr)
        return out

    def fill_zeros(self):
        self.ctx.call("mesh_memclear", self.ptr, self.bytes)
        return self

    def fill_unit(self):
        self.ctx.call("mesh_su3_unit", self.grid.ref, self.nf, self.prec, self.ptr)
        return self


def GaugeField(grid, prec=F64, ctx=None):
    return Field(grid, grid.ndim(), prec, ctx)


def Lattice(grid, prec=F64, ctx=None):
    return Field(grid, 1, prec, ctx)


class DeviceArray:
    """Plain device buffer of n doubles (per-site scratch such as k(x))."""

    def __init__(self, n, ctx=None):
        self.ctx = ctx or default_context()
        self.n = n
        p = C.c_void_p()
        self.ctx.call("mesh_malloc", C.byref(p), 8 * n)
        self.ptr = p

    def download(self):
        out = np.empty(self.n)
        self.ctx.call("mesh_memcpy_to_host", out.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n)
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        self.ctx.call("mesh_memcpy_from_host", self.ptr, a.ctypes.data_as(C.c_void_p), 8 * self.n)
        return self

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.ctx._L.mesh_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass
